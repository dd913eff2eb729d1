// The 3D Euler volume term and fused stages in contracted FP64 with the
// derivative contractions on the matrix cores (v_mfma_f64_16x16x4f64,
// CANARY_FP64_MFMA), N = 1..7.  Everything else is the contracted unit's
// (euler3d_fma.hip).
#include "launch.h"

CANARY_DEFINE_MFMA_UNIT(euler3d, canary::Euler3DFma)
