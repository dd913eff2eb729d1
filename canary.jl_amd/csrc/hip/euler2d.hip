// Kernel instantiations for the 2D compressible Euler equations, N = 1..7.
// The kernels themselves live in dg_kernels.h; physics in phys.h.
#include "launch.h"

CANARY_DEFINE_EQUATION(euler2d, canary::Euler2D)
