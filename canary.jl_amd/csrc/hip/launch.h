// Host-side launch plumbing shared by the per-equation translation units:
// argument block, grid shapes and the N -> NQ template dispatch.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <atomic>

#include "dg_kernels.h"

namespace canary {

enum Op { OP_VOLUME = 0, OP_FACE = 1, OP_UPDATE = 2, OP_STAGE = 3, OP_FACE_UPDATE = 4,
          OP_STAGE_PREP = 5 };

struct KArgs {
  void* rhs;
  const void* Q;
  void* Qout;
  const void* vgeo;
  const void* sgeo;
  const double* D;
  int64_t nreal, ntotal;
  const int64_t* vmapM;
  const int64_t* vmapP;
  const int64_t* bndy;
  double rka, rkb, dt;
  int advection;
  PhysParams phys;
  hipStream_t stream;
  const int32_t* pface;  // prepared face codes (OP_STAGE_PREP; vgeo, sgeo prepared too)
};

constexpr int kMaxN = 7;  // N = 1..7 compiled in

// The kernel families of a translation unit: the exact-order unit and the
// contracted unit (built with -ffp-contract=fast) hold every operation; the
// matrix-core unit the contracted unit's volume term and fused stages with
// the derivative contractions on the MFMA units (MfmaSweep).
enum Unit { UNIT_EXACT = 0, UNIT_FMA = 1, UNIT_MFMA = 2 };

inline int launch_status() {
  hipError_t err = hipGetLastError();
  return err == hipSuccess ? 0 : 3 /*CANARY_ERR_LAUNCH*/;
}

// Element order of the in-place kernels (volumerhs, facerhs,
// updatesolution) and of the face + update fusion that follows a volumerhs:
// every launch of one of them walks the other way than the one before, so
// each starts on the elements the previous one touched last, whose lines the
// memory-side cache still holds (three-kernel stage at 3D N=4: 6.6% faster).
// The results of each element do not depend on the order.
inline int inplace_direction() {
  static std::atomic<unsigned> n{0};
  return (int)(n.fetch_add(1, std::memory_order_relaxed) & 1u);
}

// Element order of a ping-pong stage: the two Q buffers swap roles every
// stage, so their address order alternates and so does the direction.
inline int walk_back(const void* qin, const void* qout) {
  return (uintptr_t)qin > (uintptr_t)qout ? 1 : 0;
}

template <class Phys, int NQ, int EPW, bool MFMA = false>
void launch_stage(int64_t nb, const KArgs& a) {
  // measured on 3D N=4 (32^3): one prefetched face item per thread, the
  // nodal pressure cached in LDS, workgroup sized by the volume nodes.
  // 3D N=6 prefetches none: 90 VGPRs instead of 112 make it five waves per
  // SIMD, so three of its 384-lane workgroups fit a CU instead of two
  // (23^3: 298 -> 266 us per stage; every other N is slower without it)
  constexpr bool WIDE = false;
  constexpr int PF = (Phys::DIM == 3 && NQ == 7) ? 0 : 1;
  constexpr bool PM_LDS = true;
  constexpr int NT = FusedCfg<Phys, NQ, EPW, WIDE>::NT;
  const int64_t chunk = (nb + 7) / 8;
  stage_kernel<Phys, NQ, EPW, WIDE, PF, PM_LDS, MFMA><<<dim3((unsigned)(8 * chunk)), dim3(NT), 0, a.stream>>>(
      (double*)a.rhs, (const double*)a.Q, (double*)a.Qout, (const double*)a.vgeo,
      (const double*)a.sgeo, a.D, a.nreal, a.ntotal, a.vmapM, a.vmapP, a.bndy, a.rka, a.rkb,
      a.dt, chunk, walk_back(a.Q, a.Qout), a.phys, a.advection);
}

template <class Phys, int NQ, int EPW, bool MFMA = false>
void launch_stage_prep(int64_t nb, const KArgs& a) {
  constexpr int PF = (Phys::DIM == 3 && NQ == 7) ? 0 : 1;  // as launch_stage
  constexpr int NT = FusedCfg<Phys, NQ, EPW, false>::NT;
  const int64_t chunk = (nb + 7) / 8;
  stage_prep_kernel<Phys, NQ, EPW, PF, MFMA><<<dim3((unsigned)(8 * chunk)), dim3(NT), 0, a.stream>>>(
      (double*)a.rhs, (const double*)a.Q, (double*)a.Qout, (const double*)a.vgeo,
      (const double*)a.sgeo, a.pface, a.D, a.nreal, a.ntotal, a.rka, a.rkb, a.dt, chunk,
      walk_back(a.Q, a.Qout), a.phys, a.advection);
}

// The prepared arrays hold no node heights: equation sets whose pressure
// needs them have no prepared stage.
template <class Phys, int NQ, int EPW, bool MFMA = false>
void launch_stage_prep_if(int64_t nb, const KArgs& a) {
  if constexpr (!NeedsY<Phys>::value) launch_stage_prep<Phys, NQ, EPW, MFMA>(nb, a);
}

template <class Phys, int NQ, int EPW, bool MFMA>
void launch_volume(int64_t nb, const KArgs& a) {
  constexpr int NT = VolumeCfg<Phys, NQ, EPW>::NT;
  const int64_t chunk = (nb + 7) / 8;
  volume_kernel<Phys, NQ, EPW, MFMA><<<dim3((unsigned)(8 * chunk)), dim3(NT), 0, a.stream>>>(
      (double*)a.rhs, (const double*)a.Q, (const double*)a.vgeo, a.D, a.nreal, chunk,
      inplace_direction(), a.phys);
}

template <class Phys, int NQ, int EPW>
void launch_face(int64_t nb, const KArgs& a) {
  constexpr int NT = FusedCfg<Phys, NQ, EPW, false>::NT;
  const int64_t chunk = (nb + 7) / 8;
  face_kernel<Phys, NQ, EPW><<<dim3((unsigned)(8 * chunk)), dim3(NT), 0, a.stream>>>(
      (double*)a.rhs, (const double*)a.Q, (const double*)a.vgeo, (const double*)a.sgeo, a.nreal,
      a.ntotal, a.vmapM, a.vmapP, a.bndy, chunk, inplace_direction(), a.phys);
}

// Nonzero: the grid would not fit a launch.
template <class Phys, int NQ>
int launch_update(const KArgs& a) {
  const int64_t total = a.nreal * Shape<Phys, NQ>::Np;
  const int64_t chunk = ((total + 255) / 256 + 7) / 8;
  if (8 * chunk > 0x7fffffffLL) return 1;
  update_kernel<Phys, NQ><<<dim3((unsigned)(8 * chunk)), dim3(256), 0, a.stream>>>(
      (double*)a.rhs, (double*)a.Qout, (const double*)a.vgeo, a.nreal, a.rka, a.rkb, a.dt,
      a.advection, chunk, inplace_direction());
  return 0;
}

template <class Phys, int NQ, int EPW>
void launch_face_update(int64_t nb, const KArgs& a) {
  constexpr int NT = FusedCfg<Phys, NQ, EPW, false>::NT;
  const int64_t chunk = (nb + 7) / 8;
  face_update_kernel<Phys, NQ, EPW><<<dim3((unsigned)(8 * chunk)), dim3(NT), 0, a.stream>>>(
      (double*)a.rhs, (const double*)a.Q, (double*)a.Qout, (const double*)a.vgeo,
      (const double*)a.sgeo, a.nreal, a.ntotal, a.vmapM, a.vmapP, a.bndy, a.rka, a.rkb, a.dt,
      chunk, inplace_direction(), a.phys, a.advection);
}

// The operations a unit does not hold report CANARY_ERR_UNSUPPORTED.
template <class Phys, int NQ, int UNIT>
int launch(int op, const KArgs& a) {
  using S = Shape<Phys, NQ>;
  constexpr int EPW = default_epw<S::Np>();
  constexpr bool MFMA = UNIT == UNIT_MFMA;
  if (a.nreal <= 0) return 0;
  const int64_t nb = (a.nreal + EPW - 1) / EPW;
  if constexpr (UNIT != UNIT_EXACT) {
    if (op == OP_STAGE)
      launch_stage<Phys, NQ, EPW, MFMA>(nb, a);
    else if (op == OP_STAGE_PREP && !NeedsY<Phys>::value)
      launch_stage_prep_if<Phys, NQ, EPW, MFMA>(nb, a);
    else if (op == OP_VOLUME)
      launch_volume<Phys, NQ, EPW, MFMA>(nb, a);
    else if (op == OP_FACE_UPDATE || op == OP_FACE || op == OP_UPDATE) {
      // no derivative contraction in these: the contracted unit holds them,
      // the matrix-core unit does not (each kernel lives in one unit)
      if constexpr (MFMA) {
        return 2 /*CANARY_ERR_UNSUPPORTED*/;
      } else {
        if (op == OP_FACE_UPDATE) launch_face_update<Phys, NQ, EPW>(nb, a);
        if (op == OP_FACE) launch_face<Phys, NQ, EPW>(nb, a);
        if (op == OP_UPDATE && launch_update<Phys, NQ>(a)) return 1 /*CANARY_ERR_ARG*/;
      }
    } else {
      return 2 /*CANARY_ERR_UNSUPPORTED*/;
    }
  } else {
    switch (op) {
      case OP_VOLUME:
        launch_volume<Phys, NQ, EPW, false>(nb, a);
        break;
      case OP_FACE:
        launch_face<Phys, NQ, EPW>(nb, a);
        break;
      case OP_UPDATE:
        if (launch_update<Phys, NQ>(a)) return 1 /*CANARY_ERR_ARG*/;
        break;
      case OP_STAGE:
        launch_stage<Phys, NQ, EPW>(nb, a);
        break;
      case OP_STAGE_PREP:
        if (NeedsY<Phys>::value) return 2 /*CANARY_ERR_UNSUPPORTED*/;
        launch_stage_prep_if<Phys, NQ, EPW>(nb, a);
        break;
      case OP_FACE_UPDATE:
        launch_face_update<Phys, NQ, EPW>(nb, a);
        break;
      default:
        return 1 /*CANARY_ERR_ARG*/;
    }
  }
  return launch_status();
}

template <class Phys, int UNIT = UNIT_EXACT>
int dispatch(int op, int N, const KArgs& a) {
  switch (N) {
    case 1: return launch<Phys, 2, UNIT>(op, a);
    case 2: return launch<Phys, 3, UNIT>(op, a);
    case 3: return launch<Phys, 4, UNIT>(op, a);
    case 4: return launch<Phys, 5, UNIT>(op, a);
    case 5: return launch<Phys, 6, UNIT>(op, a);
    case 6: return launch<Phys, 7, UNIT>(op, a);
    case 7: return launch<Phys, 8, UNIT>(op, a);
  }
  return 2 /*CANARY_ERR_UNSUPPORTED*/;
}

// Read (and optionally clear) this translation unit's device error record.
inline int read_status(int64_t* info, int clear) {
  unsigned long long h[4] = {0, 0, 0, 0};
  if (hipMemcpyFromSymbol(h, HIP_SYMBOL(g_status), sizeof(h), 0, hipMemcpyDeviceToHost) != hipSuccess)
    return 3;
  if (clear && h[0] != 0) {
    unsigned long long z[4] = {0, 0, 0, 0};
    if (hipMemcpyToSymbol(HIP_SYMBOL(g_status), z, sizeof(z), 0, hipMemcpyHostToDevice) != hipSuccess)
      return 3;
  }
  if (info)
    for (int i = 0; i < 4; ++i) info[i] = (int64_t)h[i];
  return (int)h[0];
}

}  // namespace canary

// The contracted-FP64 unit of an equation set (every operation), built with
// -ffp-contract=fast from a physics policy whose pressure avoids pow.
#define CANARY_DEFINE_FMA_STAGE(NAME, PHYS)                                       \
  extern "C" int canary_k_##NAME##_fma(int op, int N, const canary::KArgs* a) {  \
    return canary::dispatch<PHYS, canary::UNIT_FMA>(op, N, *a);                   \
  }                                                                               \
  extern "C" int canary_k_##NAME##_fma_status(int64_t* info, int clear) {        \
    return canary::read_status(info, clear);                                      \
  }

// The matrix-core unit: the contracted unit's volume term and fused stages
// with the derivative contractions on the MFMA units.
#define CANARY_DEFINE_MFMA_UNIT(NAME, PHYS)                                       \
  extern "C" int canary_k_##NAME##_mfma(int op, int N, const canary::KArgs* a) { \
    return canary::dispatch<PHYS, canary::UNIT_MFMA>(op, N, *a);                  \
  }                                                                               \
  extern "C" int canary_k_##NAME##_mfma_status(int64_t* info, int clear) {       \
    return canary::read_status(info, clear);                                      \
  }

#define CANARY_DEFINE_EQUATION(NAME, PHYS)                                        \
  extern "C" int canary_k_##NAME(int op, int N, const canary::KArgs* a) {        \
    return canary::dispatch<PHYS>(op, N, *a);                                     \
  }                                                                               \
  extern "C" int canary_k_##NAME##_status(int64_t* info, int clear) {            \
    return canary::read_status(info, clear);                                      \
  }
