// k_flat.hip -- kernel instantiations for the FLAT coordinate system.
#include "launch.h"

template int so3g::dispatch<so3g::PROJ_FLAT>(const so3g::OpArgs &);
