// k_zea.hip -- kernel instantiations for the ZEA coordinate system.
#include "launch.h"

template int so3g::dispatch<so3g::PROJ_ZEA>(const so3g::OpArgs &);
