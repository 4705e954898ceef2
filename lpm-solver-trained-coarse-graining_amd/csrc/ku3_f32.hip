// K u launchers and tile kernels: ku3 f32 (a translation unit of its own so the library builds in parallel).
#include "ku_launch.hpp"

LPM_INSTANTIATE_KU3(float, false)
