// k_car.hip -- kernel instantiations for the CAR coordinate system.
#include "launch.h"

template int so3g::dispatch<so3g::PROJ_CAR>(const so3g::OpArgs &);
