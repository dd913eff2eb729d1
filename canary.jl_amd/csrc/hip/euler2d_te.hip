// Kernel instantiations for the 2D Euler equations in total-energy form
// (euler2d_set3c.jl), N = 1..7, in the reference evaluation order.  The
// kernels themselves live in dg_kernels.h; physics in phys.h.
#include "launch.h"

CANARY_DEFINE_EQUATION(euler2d_te, canary::Euler2DTE)
