// Kernel instantiations for the shallow water equations, N = 1..7.
// The kernels themselves live in dg_kernels.h; physics in phys.h.
#include "launch.h"

CANARY_DEFINE_EQUATION(swe2d, canary::SWE2D)
