// Batched form, exact-order unit: the FP64 kernels in the reference
// evaluation order and the FP32 shallow-water kernels; the C ABI of the
// batched form (CANARY_FP64_FMA is routed to batched_fma.hip).
#include "batched.h"

extern "C" int canary_k_batch_fma(int eqn, int dim, int N, int which, int want, int* epw,
                                  int* threads, const canary_batch_solve* solves,
                                  const canary_batch_item* items, int64_t nitems, int parity,
                                  double rka, double rkb, int prep, void* stream);
extern "C" int canary_k_batch_fma_all(int eqn, int dim, const canary_batch_solve* solves,
                                      const canary_batch_item* items, int64_t nitems, int parity,
                                      double rka, double rkb, int prep, void* stream);

namespace canary {
namespace batch {

int route(int eqn, int dim, int prec, int N, int which, int want, int* epw, int* threads,
          const canary_batch_solve* solves, const canary_batch_item* items, int64_t nitems,
          int parity, double rka, double rkb, bool prep, hipStream_t st) {
  if (prec == CANARY_FP64) {
    if (eqn == CANARY_EULER && dim == 3)
      return run<Euler3D>(N, which, want, epw, threads, solves, items, nitems, parity, rka, rkb,
                          prep, st);
    if (eqn == CANARY_EULER && dim == 2)
      return run<Euler2D>(N, which, want, epw, threads, solves, items, nitems, parity, rka, rkb,
                          prep, st);
    if (eqn == CANARY_SWE && dim == 2)
      return run<SWE2D>(N, which, want, epw, threads, solves, items, nitems, parity, rka, rkb,
                        prep, st);
  } else if (prec == CANARY_FP32) {
    if (eqn == CANARY_SWE && dim == 2)
      return run<SWE2DF32>(N, which, want, epw, threads, solves, items, nitems, parity, rka, rkb,
                           prep, st);
  } else if (prec == CANARY_FP64_FMA) {
    return canary_k_batch_fma(eqn, dim, N, which, want, epw, threads, solves, items, nitems,
                              parity, rka, rkb, prep ? 1 : 0, st);
  }
  return CANARY_ERR_UNSUPPORTED;
}

template <class Phys>
int run_all_layout(const canary_batch_solve* solves, const canary_batch_item* items,
                   int64_t nitems, int parity, double rka, double rkb, bool prep, hipStream_t st) {
  return prep ? run_all<Phys, true>(solves, items, nitems, parity, rka, rkb, st)
              : run_all<Phys, false>(solves, items, nitems, parity, rka, rkb, st);
}

int route_all(int eqn, int dim, int prec, const canary_batch_solve* solves,
              const canary_batch_item* items, int64_t nitems, int parity, double rka, double rkb,
              bool prep, hipStream_t st) {
  if (prec == CANARY_FP64) {
    if (eqn == CANARY_EULER && dim == 3)
      return run_all_layout<Euler3D>(solves, items, nitems, parity, rka, rkb, prep, st);
    if (eqn == CANARY_EULER && dim == 2)
      return run_all_layout<Euler2D>(solves, items, nitems, parity, rka, rkb, prep, st);
    if (eqn == CANARY_SWE && dim == 2)
      return run_all_layout<SWE2D>(solves, items, nitems, parity, rka, rkb, prep, st);
  } else if (prec == CANARY_FP32) {
    if (eqn == CANARY_SWE && dim == 2)
      return run_all_layout<SWE2DF32>(solves, items, nitems, parity, rka, rkb, prep, st);
  } else if (prec == CANARY_FP64_FMA) {
    return canary_k_batch_fma_all(eqn, dim, solves, items, nitems, parity, rka, rkb,
                                  prep ? 1 : 0, st);
  }
  return CANARY_ERR_UNSUPPORTED;
}

}  // namespace batch
}  // namespace canary

extern "C" int canary_batch_all_variant(int eqn, int dim, int prec, int N, int* epw,
                                        int* threads) {
  if (!epw || !threads) return CANARY_ERR_ARG;
  int e = 0, t = 0;
  const int rc = canary::batch::route(eqn, dim, prec, N, 0, 0, &e, &t, nullptr, nullptr, 0, 0,
                                      0.0, 0.0, false, nullptr);
  if (rc) return rc;
  if (t > canary::batch::MAXT) return CANARY_ERR_UNSUPPORTED;
  *epw = e;
  *threads = canary::batch::MAXT;
  return 0;
}

static int stage_all(int eqn, int dim, int prec, const canary_batch_solve* solves,
                     const canary_batch_item* items, int64_t nitems, int parity, double rka,
                     double rkb, bool prep, void* stream) {
  if (nitems < 0 || (parity != 0 && parity != 1)) return CANARY_ERR_ARG;
  if (nitems == 0) return 0;
  if (!solves || !items) return CANARY_ERR_ARG;
  return canary::batch::route_all(eqn, dim, prec, solves, items, nitems, parity, rka, rkb, prep,
                                  (hipStream_t)stream);
}

extern "C" int canary_batch_stage_all(int eqn, int dim, int prec,
                                      const canary_batch_solve* solves,
                                      const canary_batch_item* items, int64_t nitems, int parity,
                                      double rka, double rkb, void* stream) {
  return stage_all(eqn, dim, prec, solves, items, nitems, parity, rka, rkb, false, stream);
}

extern "C" int canary_batch_stage_all_prepared(int eqn, int dim, int prec,
                                               const canary_batch_solve* solves,
                                               const canary_batch_item* items, int64_t nitems,
                                               int parity, double rka, double rkb,
                                               void* stream) {
  return stage_all(eqn, dim, prec, solves, items, nitems, parity, rka, rkb, true, stream);
}

extern "C" int canary_batch_variant(int eqn, int dim, int prec, int N, int which, int* epw,
                                    int* threads) {
  if (!epw || !threads || which < 0) return CANARY_ERR_ARG;
  return canary::batch::route(eqn, dim, prec, N, which, 0, epw, threads, nullptr, nullptr, 0, 0,
                              0.0, 0.0, false, nullptr);
}

static int stage_class(int eqn, int dim, int prec, int N, int epw,
                       const canary_batch_solve* solves, const canary_batch_item* items,
                       int64_t nitems, int parity, double rka, double rkb, bool prep,
                       void* stream) {
  if (nitems < 0 || (parity != 0 && parity != 1) || epw < 1) return CANARY_ERR_ARG;
  if (nitems == 0) return 0;
  if (!solves || !items) return CANARY_ERR_ARG;
  return canary::batch::route(eqn, dim, prec, N, -1, epw, nullptr, nullptr, solves, items, nitems,
                              parity, rka, rkb, prep, (hipStream_t)stream);
}

extern "C" int canary_batch_stage(int eqn, int dim, int prec, int N, int epw,
                                  const canary_batch_solve* solves,
                                  const canary_batch_item* items, int64_t nitems, int parity,
                                  double rka, double rkb, void* stream) {
  return stage_class(eqn, dim, prec, N, epw, solves, items, nitems, parity, rka, rkb, false,
                     stream);
}

extern "C" int canary_batch_stage_prepared(int eqn, int dim, int prec, int N, int epw,
                                           const canary_batch_solve* solves,
                                           const canary_batch_item* items, int64_t nitems,
                                           int parity, double rka, double rkb, void* stream) {
  return stage_class(eqn, dim, prec, N, epw, solves, items, nitems, parity, rka, rkb, true,
                     stream);
}

extern "C" int canary_k_batch_status(int64_t* info, int clear) {
  unsigned long long h[4] = {0, 0, 0, 0};
  if (hipMemcpyFromSymbol(h, HIP_SYMBOL(canary::g_status), sizeof(h), 0, hipMemcpyDeviceToHost) !=
      hipSuccess)
    return CANARY_ERR_LAUNCH;
  if (clear && h[0] != 0) {
    unsigned long long z[4] = {0, 0, 0, 0};
    if (hipMemcpyToSymbol(HIP_SYMBOL(canary::g_status), z, sizeof(z), 0, hipMemcpyHostToDevice) !=
        hipSuccess)
      return CANARY_ERR_LAUNCH;
  }
  if (info)
    for (int i = 0; i < 4; ++i) info[i] = (int64_t)h[i];
  return (int)h[0];
}
