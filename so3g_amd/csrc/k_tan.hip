// k_tan.hip -- kernel instantiations for the TAN coordinate system.
#include "launch.h"

template int so3g::dispatch<so3g::PROJ_TAN>(const so3g::OpArgs &);
