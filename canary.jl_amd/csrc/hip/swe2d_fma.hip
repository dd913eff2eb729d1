// The fused LSRK stage for the shallow-water equations in contracted FP64, N = 1..7:
// built with -ffp-contract=fast, so a*b+c becomes one fused multiply-add
// (CANARY_FP64_FMA).  The exact-order kernels of the same equations are in
// swe2d.hip.
#include "launch.h"

CANARY_DEFINE_FMA_STAGE(swe2d, canary::SWE2DFma)
