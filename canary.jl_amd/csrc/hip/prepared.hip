// Prepared geometry of the fused stage: compact copies of the reference
// arrays, built once per mesh on the device, that the stage reads in place
// of vgeo, sgeo, vmapM, vmapP and elemtobndy (the layout is described at
// `Prepared` in dg_kernels.h).  What the compact arrays leave out is either
// unused by the stage (x, y, z, vMJI) or a function of what they keep, which
// this kernel checks on every entry before the stage may rely on it:
//   * vmapM[:, f, e] is Np (e - 1) + fmask[:, f] (src/mesh.jl:994), so the
//     stage computes it;
//   * vmapP[:, f, e] is Np (eP - 1) + fmask[:, fP] for one neighbour element
//     eP and face fP (src/mesh.jl:995, orientation 1 being the only one
//     `mappings` supports), so one int32 per face holds it;
//   * MJI is 1 ./ MJ (IEEE division; euler3d_set2c.jl:101-139), so the stage
//     divides instead of reading it;
//   * elemtobndy is 0 (interior or periodic) or 1 (wall), the only codes the
//     face kernels accept (euler3d_set2c.jl:279-283).
// A failed check is recorded in this unit's status word (canary_check_errors)
// with the face and element, and the entry's outputs must not be used.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../../include/canary_dg.h"
#include "dg_kernels.h"

namespace canary {
namespace prep {

static __device__ unsigned long long g_prep_status[4];

__device__ __forceinline__ void prep_error(unsigned long long code, long long face, long long elem,
                                           long long value) {
  if (atomicCAS(&g_prep_status[0], 0ull, code) == 0ull) {
    g_prep_status[1] = (unsigned long long)face;
    g_prep_status[2] = (unsigned long long)elem;
    g_prep_status[3] = (unsigned long long)value;
  }
}

template <int DIM, int NQ>
__global__ __launch_bounds__(256) void prepare_geometry_kernel(
    int64_t nelem, int64_t ntotal, const double* __restrict__ vgeo,
    const double* __restrict__ sgeo, const int64_t* __restrict__ vmapM,
    const int64_t* __restrict__ vmapP, const int64_t* __restrict__ bndy,
    double* __restrict__ pvgeo, double* __restrict__ psgeo, int32_t* __restrict__ pface) {
  constexpr int Np = ipow(NQ, DIM), Nfp = ipow(NQ, DIM - 1), NF = 2 * DIM, NFN = NF * Nfp;
  constexpr int NVG = DIM == 3 ? 14 : 8, NSG = DIM == 3 ? 5 : 4;  // reference field counts
  constexpr int PV = DIM * DIM + 1, PS = DIM + 1;                 // prepared field counts
  const int64_t tid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  // volume nodes: metric terms and MJ; MJI must be 1 / MJ
  for (int64_t g = tid; g < nelem * Np; g += stride) {
    const int64_t e = g / Np;
    const int n = (int)(g - e * Np);
    const double* v = vgeo + (size_t)NVG * Np * e;
    double* o = pvgeo + (size_t)PV * Np * e;
#pragma unroll
    for (int c = 0; c < PV; ++c) o[n + Np * c] = v[n + Np * c];
    const double MJ = v[n + Np * (DIM * DIM)], MJI = v[n + Np * (DIM * DIM + 1)];
    if (!(MJI == 1.0 / MJ)) prep_error(1 /*CANARY_ERR_ARG*/, 0, e + 1, n + 1);
  }
  // face nodes: normal and sMJ; vmapM must be the canonical face mask
  for (int64_t g = tid; g < nelem * NFN; g += stride) {
    const int64_t e = g / NFN;
    const int fi = (int)(g - e * NFN);
    const int f = fi / Nfp, fn = fi - f * Nfp;
    const double* s = sgeo + (size_t)NSG * g;
    double* o = psgeo + (size_t)PS * g;
#pragma unroll
    for (int c = 0; c < PS; ++c) o[c] = s[c];
    const int64_t idM = vmapM[g];
    if (idM != (int64_t)Np * e + canon_vid<DIM, NQ>(f, fn) + 1)
      prep_error(5 /*CANARY_ERR_VMAP*/, f + 1, e + 1, idM);
  }
  // faces: wall (-1) or the neighbour element and face of the whole vmapP column
  for (int64_t g = tid; g < nelem * NF; g += stride) {
    const int64_t e = g / NF;
    const int f = (int)(g - e * NF);
    const int64_t bc = bndy[g];
    int32_t code = -1;
    if (bc == 0) {
      const int64_t* vp = vmapP + (size_t)Nfp * g;
      const int64_t id0 = vp[0];
      if (id0 < 1 || id0 > ntotal * Np) {
        prep_error(5 /*CANARY_ERR_VMAP*/, f + 1, e + 1, id0);
      } else {
        const int64_t eP = (id0 - 1) / Np;
        int fP = -1;
        for (int c = 0; c < NF && fP < 0; ++c) {
          bool all = true;
          for (int k = 0; k < Nfp && all; ++k)
            all = vp[k] == (int64_t)Np * eP + canon_vid<DIM, NQ>(c, k) + 1;
          if (all) fP = c;
        }
        if (fP < 0)
          prep_error(9 /*CANARY_ERR_ORIENTATION*/, f + 1, e + 1, id0);
        else if (eP > (int64_t)(0x7fffffff / NF) - 1)
          prep_error(11 /*CANARY_ERR_CAPACITY*/, f + 1, e + 1, eP + 1);
        else
          code = (int32_t)(NF * eP + fP);
      }
    } else if (bc != 1) {
      prep_error(4 /*CANARY_ERR_BC*/, f + 1, e + 1, bc);
    }
    pface[g] = code;
  }
}

template <int DIM, int NQ>
int prepare(int64_t nelem, int64_t ntotal, const void* vgeo, const void* sgeo,
            const int64_t* vmapM, const int64_t* vmapP, const int64_t* bndy, void* pvgeo,
            void* psgeo, int32_t* pface, hipStream_t st) {
  const int64_t work = nelem * ipow(NQ, DIM);
  int64_t blocks = (work + 255) / 256;
  if (blocks > 65536) blocks = 65536;
  prepare_geometry_kernel<DIM, NQ><<<dim3((unsigned)blocks), dim3(256), 0, st>>>(
      nelem, ntotal, (const double*)vgeo, (const double*)sgeo, vmapM, vmapP, bndy,
      (double*)pvgeo, (double*)psgeo, pface);
  return hipGetLastError() == hipSuccess ? 0 : CANARY_ERR_LAUNCH;
}

}  // namespace prep
}  // namespace canary

extern "C" int canary_prepare_geometry(int dim, int N, int64_t nelem, int64_t nelem_total,
                                       const void* vgeo, const void* sgeo, const int64_t* vmapM,
                                       const int64_t* vmapP, const int64_t* elemtobndy,
                                       void* pvgeo, void* psgeo, int32_t* pface, void* stream) {
  if ((dim != 2 && dim != 3) || N < 1 || N > 7) return CANARY_ERR_UNSUPPORTED;
  if (nelem < 0 || nelem_total < nelem) return CANARY_ERR_ARG;
  if (nelem == 0) return 0;
  if (!vgeo || !sgeo || !vmapM || !vmapP || !elemtobndy || !pvgeo || !psgeo || !pface)
    return CANARY_ERR_ARG;
  hipStream_t st = (hipStream_t)stream;
#define CANARY_PREP_N(D, NN)                                                                     \
  case NN:                                                                                       \
    return canary::prep::prepare<D, NN + 1>(nelem, nelem_total, vgeo, sgeo, vmapM, vmapP,        \
                                            elemtobndy, pvgeo, psgeo, pface, st);
  if (dim == 3) {
    switch (N) {
      CANARY_PREP_N(3, 1) CANARY_PREP_N(3, 2) CANARY_PREP_N(3, 3) CANARY_PREP_N(3, 4)
      CANARY_PREP_N(3, 5) CANARY_PREP_N(3, 6) CANARY_PREP_N(3, 7)
    }
  } else {
    switch (N) {
      CANARY_PREP_N(2, 1) CANARY_PREP_N(2, 2) CANARY_PREP_N(2, 3) CANARY_PREP_N(2, 4)
      CANARY_PREP_N(2, 5) CANARY_PREP_N(2, 6) CANARY_PREP_N(2, 7)
    }
  }
#undef CANARY_PREP_N
  return CANARY_ERR_UNSUPPORTED;
}

extern "C" int canary_k_prepared_status(int64_t* info, int clear) {
  unsigned long long h[4] = {0, 0, 0, 0};
  if (hipMemcpyFromSymbol(h, HIP_SYMBOL(canary::prep::g_prep_status), sizeof(h), 0,
                          hipMemcpyDeviceToHost) != hipSuccess)
    return CANARY_ERR_LAUNCH;
  if (clear && h[0] != 0) {
    unsigned long long z[4] = {0, 0, 0, 0};
    if (hipMemcpyToSymbol(HIP_SYMBOL(canary::prep::g_prep_status), z, sizeof(z), 0,
                          hipMemcpyHostToDevice) != hipSuccess)
      return CANARY_ERR_LAUNCH;
  }
  if (info)
    for (int i = 0; i < 4; ++i) info[i] = (int64_t)h[i];
  return (int)h[0];
}
