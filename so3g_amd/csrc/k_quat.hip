// k_quat.hip -- kernel instantiations for the QUAT coordinate system.
#include "launch.h"

template int so3g::dispatch<so3g::PROJ_QUAT>(const so3g::OpArgs &);
