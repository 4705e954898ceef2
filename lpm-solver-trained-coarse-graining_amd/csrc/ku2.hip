// K u launchers and tile kernels: ku2 (a translation unit of its own so the library builds in parallel).
#include "ku_launch.hpp"

LPM_INSTANTIATE_KU2(float, false)
LPM_INSTANTIATE_KU2(float, true)
LPM_INSTANTIATE_KU2(double, false)
LPM_INSTANTIATE_KU2(double, true)
