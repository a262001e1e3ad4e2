// k_arc.hip -- kernel instantiations for the ARC coordinate system.
#include "launch.h"

template int so3g::dispatch<so3g::PROJ_ARC>(const so3g::OpArgs &);
