// Batched form, contracted-FP64 unit (CANARY_FP64_FMA): the same kernels as
// batched.hip with the contracted physics policies, built with
// -ffp-contract=fast.
#include "batched.h"

extern "C" int canary_k_batch_fma(int eqn, int dim, int N, int which, int want, int* epw,
                                  int* threads, const canary_batch_solve* solves,
                                  const canary_batch_item* items, int64_t nitems, int parity,
                                  double rka, double rkb, int prep, void* stream) {
  using namespace canary;
  using namespace canary::batch;
  hipStream_t st = (hipStream_t)stream;
  const bool pr = prep != 0;
  if (eqn == CANARY_EULER && dim == 3)
    return run<Euler3DFma>(N, which, want, epw, threads, solves, items, nitems, parity, rka, rkb,
                           pr, st);
  if (eqn == CANARY_EULER && dim == 2)
    return run<Euler2DFma>(N, which, want, epw, threads, solves, items, nitems, parity, rka, rkb,
                           pr, st);
  if (eqn == CANARY_SWE && dim == 2)
    return run<SWE2DFma>(N, which, want, epw, threads, solves, items, nitems, parity, rka, rkb,
                         pr, st);
  return CANARY_ERR_UNSUPPORTED;
}

extern "C" int canary_k_batch_fma_all(int eqn, int dim, const canary_batch_solve* solves,
                                      const canary_batch_item* items, int64_t nitems, int parity,
                                      double rka, double rkb, int prep, void* stream) {
  using namespace canary;
  using namespace canary::batch;
  hipStream_t st = (hipStream_t)stream;
#define CANARY_ALL(P)                                                                   \
  return prep ? run_all<P, true>(solves, items, nitems, parity, rka, rkb, st)           \
              : run_all<P, false>(solves, items, nitems, parity, rka, rkb, st);
  if (eqn == CANARY_EULER && dim == 3) CANARY_ALL(Euler3DFma)
  if (eqn == CANARY_EULER && dim == 2) CANARY_ALL(Euler2DFma)
  if (eqn == CANARY_SWE && dim == 2) CANARY_ALL(SWE2DFma)
#undef CANARY_ALL
  return CANARY_ERR_UNSUPPORTED;
}

extern "C" int canary_k_batch_fma_status(int64_t* info, int clear) {
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
