// The fused LSRK stage for the 2D Euler equations in total-energy form in
// contracted FP64, N = 1..7: built with -ffp-contract=fast (CANARY_FP64_FMA).
// The exact-order kernels of the same equations are in euler2d_te.hip.
#include "launch.h"

CANARY_DEFINE_FMA_STAGE(euler2d_te, canary::Euler2DTEFma)
