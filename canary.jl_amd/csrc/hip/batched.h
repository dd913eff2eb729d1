// Batched form, shared by the exact-order unit (batched.hip) and the
// contracted-FP64 unit (batched_fma.hip): many independent solves (mixed N,
// element counts, meshes, dt and constants) advanced through a device work
// table of element blocks.  Each table entry names a solve and the first
// element of a block; the kernel for one (N, elements-per-workgroup) class
// runs the fused stage body of dg_kernels.h on every entry of its slice of
// the table.  One kernel per class keeps each variant's register allocation
// its own; the host launches the classes back to back, or on forked streams
// inside one graph.  batched_all_kernel instead runs the entries of every
// order in one launch (each order at its fullest 256-lane variant), sized by
// its largest order.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../../include/canary_dg.h"
#include "dg_kernels.h"
#include "pow_registry.h"

namespace canary {
namespace batch {

constexpr int MAXT = 256;  // threads per workgroup ceiling for several elements

// Candidate `WHICH` of elements per workgroup for (DIM, NQ): 0 = the most
// elements that fit MAXT threads (one element, with as many threads as it
// has nodes, when it alone exceeds MAXT: 3D N = 6, 7), 1 = about half of
// that, 2 = the single-solve choice (first lane fill >= 90%).  0 = no such
// candidate.
template <int DIM, int NQ, int WHICH>
constexpr int epw_candidate() {
  constexpr int Np = ipow(NQ, DIM);
  constexpr int emax = MAXT / Np;
  if (WHICH == 0) return emax > 0 ? emax : 1;  // more than MAXT nodes: one element alone
  if (WHICH == 1) return emax >= 2 ? (emax + 1) / 2 : 0;
  if (WHICH == 2) {
    constexpr int d = default_epw<Np>();
    return d * Np <= MAXT ? d : 0;
  }
  return 0;
}

__device__ __forceinline__ PhysParams params_of(const canary_phys& c) {
  PhysParams p;
  p.gamma = c.gamma;
  p.p0 = c.p0;
  p.R = c.R_gas;
  p.cp = c.c_p;
  p.cv = c.c_v;
  p.g = c.gravity;
  p.kappa = c.c_p / c.c_v;
  p.swe_g = c.swe_gravity;
  p.dnl = c.delta_nl;
  p.ktab = nullptr;
  p.ser = kPowSeries;
  return p;
}

template <class Phys, int NQ, int EPW>
struct BCfg {
  using C = FusedCfg<Phys, NQ, EPW, false>;
  static constexpr int NT = C::NT;
};

// One stage of the block at e0 of solve sv, over the reference arrays or,
// PREP, over the prepared ones (pvgeo / psgeo in vgeo / sgeo, the face codes
// in pface; the Prepared layout of dg_kernels.h).
template <class Phys, int NQ, int EPW, int NT, int PF, bool PREP>
__device__ __forceinline__ void solve_stage(char* lds, const canary_batch_solve& sv, int64_t e0,
                                            int parity, double rka, double rkb,
                                            const PhysParams& p) {
  using Real = typename Phys::Real;
  if constexpr (PREP)
    stage_body<Phys, NQ, EPW, NT, PF, true, true, true, true>(
        lds, (Real*)sv.rhs, (const Real*)sv.Q[parity], (Real*)sv.Q[1 - parity],
        (const Real*)sv.vgeo, (const Real*)sv.sgeo, sv.D, e0, sv.nreal, sv.nelem_total, nullptr,
        reinterpret_cast<const int64_t*>(sv.pface), nullptr, rka, rkb, sv.dt, p, sv.advection);
  else
    stage_body<Phys, NQ, EPW, NT, PF, true>(
        lds, (Real*)sv.rhs, (const Real*)sv.Q[parity], (Real*)sv.Q[1 - parity],
        (const Real*)sv.vgeo, (const Real*)sv.sgeo, sv.D, e0, sv.nreal, sv.nelem_total,
        sv.vmapM, sv.vmapP, sv.elemtobndy, rka, rkb, sv.dt, p, sv.advection);
}

// One workgroup per work-table entry; workgroup b takes entry
// (b & 7) * chunk + (b >> 3) (grid = 8 * chunk), so XCD b & 7 gets a
// contiguous run of the table.  (A grid-stride loop over entries keeps the
// per-thread face and LDS addressing live across iterations and costs
// about 100 VGPRs in 3D; one entry per workgroup does not.)
template <class Phys, int NQ, int EPW, bool PREP>
__global__ __launch_bounds__((BCfg<Phys, NQ, EPW>::NT)) void batched_kernel(
    const canary_batch_solve* __restrict__ solves, const canary_batch_item* __restrict__ items,
    int64_t nitems, int64_t chunk, int parity, double rka, double rkb,
    const crpow::KappaTable* __restrict__ ktabs, int nktab) {
  using C = FusedCfg<Phys, NQ, EPW, false>;
  __shared__ __attribute__((aligned(16))) char lds[C::LDS];
  const int64_t w = stage_block(blockIdx.x, chunk, parity);  // direction alternates per stage
  if (w >= nitems) return;
  const canary_batch_item it = items[w];
  const canary_batch_solve& sv = solves[it.solve];
  if (it.N != NQ - 1 || it.epw != EPW || sv.N != NQ - 1 || (sv.pface != nullptr) != PREP) {
    if (threadIdx.x == 0) record_error(1 /*CANARY_ERR_ARG*/, it.solve, it.e0 + 1, it.N);
    return;
  }
  if (it.e0 < 0 || it.e0 >= sv.nreal) {
    if (threadIdx.x == 0) record_error(1 /*CANARY_ERR_ARG*/, it.solve, it.e0 + 1, sv.nreal);
    return;
  }
  PhysParams p = params_of(sv.phys);
  if (Phys::HAS_P)  // the registered pow table of this solve's kappa, if any
    for (int i = 0; i < nktab; ++i)
      if (ktabs[i].kappa == p.kappa) {
        p.ktab = ktabs + i;
        break;
      }
  // 3D N=6 prefetches no face item (fewer VGPRs: a third 384-lane
  // workgroup per CU), as the single-solve stage does
  solve_stage<Phys, NQ, EPW, C::NT, (Phys::DIM == 3 && NQ == 7 && EPW == 1) ? 0 : 1, PREP>(
      lds, sv, it.e0, parity, rka, rkb, p);
}

// ---------------------------------------------- every class in one launch
// All entries of a batch whose order fits 256 lanes with candidate 0 (the
// most elements that fit 256 lanes: every 2D order, 3D N <= 5) run in ONE
// launch per stage: 256-lane workgroups, one work-table entry each, a
// switch on the entry's N into that order's stage body.
template <class Phys, int NQ>
constexpr bool fits_all() {
  return epw_candidate<Phys::DIM, NQ, 0>() * ipow(NQ, Phys::DIM) <= MAXT;
}

template <class Phys, int NQ>
constexpr int all_lds_nq() {
  if constexpr (fits_all<Phys, NQ>())
    return FusedCfg<Phys, NQ, epw_candidate<Phys::DIM, NQ, 0>(), false>::LDS;
  else
    return 0;
}

template <class Phys>
constexpr int all_lds() {
  constexpr int a[7] = {all_lds_nq<Phys, 2>(), all_lds_nq<Phys, 3>(), all_lds_nq<Phys, 4>(),
                        all_lds_nq<Phys, 5>(), all_lds_nq<Phys, 6>(), all_lds_nq<Phys, 7>(),
                        all_lds_nq<Phys, 8>()};
  int m = 16;
  for (int i = 0; i < 7; ++i) m = a[i] > m ? a[i] : m;
  return m;
}

template <class Phys, int NQ, bool PREP>
__device__ __forceinline__ void all_item(char* lds, const canary_batch_solve& sv,
                                         const canary_batch_item& it, int parity, double rka,
                                         double rkb, const PhysParams& p) {
  if constexpr (fits_all<Phys, NQ>()) {
    constexpr int E = epw_candidate<Phys::DIM, NQ, 0>();
    solve_stage<Phys, NQ, E, MAXT, 1, PREP>(lds, sv, it.e0, parity, rka, rkb, p);
  }
}

template <class Phys, bool PREP>
__global__ __launch_bounds__(MAXT) void batched_all_kernel(
    const canary_batch_solve* __restrict__ solves, const canary_batch_item* __restrict__ items,
    int64_t nitems, int64_t chunk, int parity, double rka, double rkb,
    const crpow::KappaTable* __restrict__ ktabs, int nktab) {
  __shared__ __attribute__((aligned(16))) char lds[all_lds<Phys>()];
  const int64_t w = stage_block(blockIdx.x, chunk, parity);
  if (w >= nitems) return;
  const canary_batch_item it = items[w];
  const canary_batch_solve& sv = solves[it.solve];
  int want = 0;
  switch (it.N) {
    case 1: want = fits_all<Phys, 2>() ? epw_candidate<Phys::DIM, 2, 0>() : 0; break;
    case 2: want = fits_all<Phys, 3>() ? epw_candidate<Phys::DIM, 3, 0>() : 0; break;
    case 3: want = fits_all<Phys, 4>() ? epw_candidate<Phys::DIM, 4, 0>() : 0; break;
    case 4: want = fits_all<Phys, 5>() ? epw_candidate<Phys::DIM, 5, 0>() : 0; break;
    case 5: want = fits_all<Phys, 6>() ? epw_candidate<Phys::DIM, 6, 0>() : 0; break;
    case 6: want = fits_all<Phys, 7>() ? epw_candidate<Phys::DIM, 7, 0>() : 0; break;
    case 7: want = fits_all<Phys, 8>() ? epw_candidate<Phys::DIM, 8, 0>() : 0; break;
  }
  if (want == 0 || it.epw != want || sv.N != it.N || (sv.pface != nullptr) != PREP) {
    if (threadIdx.x == 0) record_error(1 /*CANARY_ERR_ARG*/, it.solve, it.e0 + 1, it.N);
    return;
  }
  if (it.e0 < 0 || it.e0 >= sv.nreal) {
    if (threadIdx.x == 0) record_error(1 /*CANARY_ERR_ARG*/, it.solve, it.e0 + 1, sv.nreal);
    return;
  }
  PhysParams p = params_of(sv.phys);
  if (Phys::HAS_P)
    for (int i = 0; i < nktab; ++i)
      if (ktabs[i].kappa == p.kappa) {
        p.ktab = ktabs + i;
        break;
      }
  switch (it.N) {
    case 1: all_item<Phys, 2, PREP>(lds, sv, it, parity, rka, rkb, p); break;
    case 2: all_item<Phys, 3, PREP>(lds, sv, it, parity, rka, rkb, p); break;
    case 3: all_item<Phys, 4, PREP>(lds, sv, it, parity, rka, rkb, p); break;
    case 4: all_item<Phys, 5, PREP>(lds, sv, it, parity, rka, rkb, p); break;
    case 5: all_item<Phys, 6, PREP>(lds, sv, it, parity, rka, rkb, p); break;
    case 6: all_item<Phys, 7, PREP>(lds, sv, it, parity, rka, rkb, p); break;
    case 7: all_item<Phys, 8, PREP>(lds, sv, it, parity, rka, rkb, p); break;
  }
}

template <class Phys, bool PREP>
int run_all(const canary_batch_solve* solves, const canary_batch_item* items, int64_t nitems,
            int parity, double rka, double rkb, hipStream_t st) {
  const int64_t chunk = (nitems + 7) / 8;
  if (chunk > ((int64_t)1 << 27)) return CANARY_ERR_ARG;
  int nktab = 0;
  const crpow::KappaTable* ktabs = Phys::HAS_P ? kappa_tables(&nktab) : nullptr;
  batched_all_kernel<Phys, PREP><<<dim3((unsigned)(8 * chunk)), dim3(MAXT), 0, st>>>(
      solves, items, nitems, chunk, parity, rka, rkb, ktabs, nktab);
  return hipGetLastError() == hipSuccess ? 0 : CANARY_ERR_LAUNCH;
}

template <class Phys, int NQ, int WHICH, bool PREP>
int run_which(int* epw, int* threads, const canary_batch_solve* solves,
              const canary_batch_item* items, int64_t nitems, int parity, double rka, double rkb,
              hipStream_t st) {
  constexpr int E = epw_candidate<Phys::DIM, NQ, WHICH>();
  if constexpr (E < 1) {
    return CANARY_ERR_UNSUPPORTED;
  } else {
    constexpr int NT = BCfg<Phys, NQ, E>::NT;
    if (epw) {  // query
      *epw = E;
      *threads = NT;
      return 0;
    }
    const int64_t chunk = (nitems + 7) / 8;
    if (chunk > ((int64_t)1 << 27)) return CANARY_ERR_ARG;
    int nktab = 0;
    const crpow::KappaTable* ktabs = Phys::HAS_P ? kappa_tables(&nktab) : nullptr;
    batched_kernel<Phys, NQ, E, PREP><<<dim3((unsigned)(8 * chunk)), dim3(NT), 0, st>>>(
        solves, items, nitems, chunk, parity, rka, rkb, ktabs, nktab);
    return hipGetLastError() == hipSuccess ? 0 : CANARY_ERR_LAUNCH;
  }
}

// which >= 0: query candidate `which`; which < 0: launch the candidate whose
// elements per workgroup equal `want`.
template <class Phys, int NQ>
int run_nq(int which, int want, int* epw, int* threads, const canary_batch_solve* solves,
           const canary_batch_item* items, int64_t nitems, int parity, double rka, double rkb,
           bool prep, hipStream_t st) {
  if (which >= 0) {
    if (which == 0) return run_which<Phys, NQ, 0, false>(epw, threads, 0, 0, 0, 0, 0, 0, 0);
    if (which == 1) return run_which<Phys, NQ, 1, false>(epw, threads, 0, 0, 0, 0, 0, 0, 0);
    if (which == 2) return run_which<Phys, NQ, 2, false>(epw, threads, 0, 0, 0, 0, 0, 0, 0);
    return CANARY_ERR_UNSUPPORTED;
  }
  constexpr int E0 = epw_candidate<Phys::DIM, NQ, 0>();
  constexpr int E1 = epw_candidate<Phys::DIM, NQ, 1>();
  constexpr int E2 = epw_candidate<Phys::DIM, NQ, 2>();
#define CANARY_RUN_W(W)                                                                         \
  return prep ? run_which<Phys, NQ, W, true>(nullptr, nullptr, solves, items, nitems, parity,   \
                                             rka, rkb, st)                                      \
              : run_which<Phys, NQ, W, false>(nullptr, nullptr, solves, items, nitems, parity,  \
                                              rka, rkb, st);
  if (want == E0 && E0 > 0) CANARY_RUN_W(0)
  if (want == E1 && E1 > 0) CANARY_RUN_W(1)
  if (want == E2 && E2 > 0) CANARY_RUN_W(2)
#undef CANARY_RUN_W
  return CANARY_ERR_UNSUPPORTED;
}

template <class Phys>
int run(int N, int which, int want, int* epw, int* threads, const canary_batch_solve* solves,
        const canary_batch_item* items, int64_t nitems, int parity, double rka, double rkb,
        bool prep, hipStream_t st) {
#define CANARY_BATCH_N(NN)                                                                    \
  case NN:                                                                                    \
    return run_nq<Phys, NN + 1>(which, want, epw, threads, solves, items, nitems, parity, rka, \
                                rkb, prep, st);
  switch (N) {
    CANARY_BATCH_N(1)
    CANARY_BATCH_N(2)
    CANARY_BATCH_N(3)
    CANARY_BATCH_N(4)
    CANARY_BATCH_N(5)
    CANARY_BATCH_N(6)
    CANARY_BATCH_N(7)
  }
#undef CANARY_BATCH_N
  return CANARY_ERR_UNSUPPORTED;
}

}  // namespace batch
}  // namespace canary
