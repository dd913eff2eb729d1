// Kernel instantiations for the 3D compressible Euler equations, N = 1..7.
// The kernels themselves live in dg_kernels.h; physics in phys.h.
#include "launch.h"

CANARY_DEFINE_EQUATION(euler3d, canary::Euler3D)
