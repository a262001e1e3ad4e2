// k_cea.hip -- kernel instantiations for the CEA coordinate system.
#include "launch.h"

template int so3g::dispatch<so3g::PROJ_CEA>(const so3g::OpArgs &);
