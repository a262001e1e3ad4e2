// k_sin.hip -- kernel instantiations for the SIN coordinate system.
#include "launch.h"

template int so3g::dispatch<so3g::PROJ_SIN>(const so3g::OpArgs &);
