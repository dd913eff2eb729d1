// The fused LSRK stage for the 2D Euler equations in contracted FP64, N = 1..7:
// built with -ffp-contract=fast, so a*b+c becomes one fused multiply-add
// (CANARY_FP64_FMA).  The exact-order kernels of the same equations are in
// euler2d.hip.
#include "launch.h"

CANARY_DEFINE_FMA_STAGE(euler2d, canary::Euler2DFma)
