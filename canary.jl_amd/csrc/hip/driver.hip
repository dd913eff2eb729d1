// Native LSRK54 time loop: lowstorageRK of the reference drivers
// (euler3d_set2c.jl:754-891, euler2d_set2c.jl:639-775,
// shallow_water.jl:615-747) over the device entry points of this library,
// with the reference's stage loop (:819-855): for s = 1..5
//   volumerhs!; facerhs!; updatesolution!(RKA[s % 5 + 1], RKB[s], dt)
// or its fused forms.  Optionally the steps are captured once into a
// hipGraph on the caller's stream and replayed.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../../include/canary_dg.h"

namespace {

// LSRK54 (Carpenter & Kennedy 1994) as DFloat(num) / DFloat(den), the
// reference's tables (euler3d_set2c.jl:772-788).
void lsrk54_tables(double* a, double* b, double* c) {
  a[0] = 0.0;
  a[1] = (double)-567301805773 / (double)1357537059087;
  a[2] = (double)-2404267990393 / (double)2016746695238;
  a[3] = (double)-3550918686646 / (double)2091501179385;
  a[4] = (double)-1275806237668 / (double)842570457699;
  b[0] = (double)1432997174477 / (double)9575080441755;
  b[1] = (double)5161836677717 / (double)13612068292357;
  b[2] = (double)1720146321549 / (double)2090206949498;
  b[3] = (double)3134564353537 / (double)4481467310338;
  b[4] = (double)2277821191437 / (double)14882151754819;
  c[0] = 0.0;
  c[1] = (double)1432997174477 / (double)9575080441755;
  c[2] = (double)2526269341429 / (double)6820363962896;
  c[3] = (double)2006345519317 / (double)3224310063776;
  c[4] = (double)2802321613138 / (double)2924317926251;
}

constexpr int PATH_PREPARED = 3;  // canary_stage_prepared over the prepared arrays

struct Run {
  int eqn, dim, N, prec, path, advection;
  const int32_t* pface;
  double dt;
  void* Q[2];
  void* rhs;
  const void* vgeo;
  const void* sgeo;
  const double* D;
  int64_t nreal, ntotal;
  const int64_t* vmapM;
  const int64_t* vmapP;
  const int64_t* bndy;
  const canary_phys* phys;
  void* stream;
  double A[5], B[5], C[5];
};

// One LSRK step from buffer `cur`; returns the status and flips `cur` for
// the ping-pong paths.
int step(Run& r, int& cur) {
  for (int s = 0; s < 5; ++s) {
    const double rka = r.A[(s + 1) % 5], rkb = r.B[s];
    int rc = 0;
    if (r.path == CANARY_PATH_KERNELS) {
      // matrix-core form: its volume term, then facerhs and the update
      // contracted (they have no derivative contraction)
      void* q = r.Q[0];
      const int pf = r.prec == CANARY_FP64_MFMA ? CANARY_FP64_FMA : r.prec;
      rc = canary_volumerhs(r.eqn, r.dim, r.N, r.prec, r.rhs, q, r.vgeo, r.D, r.nreal, r.phys,
                            r.stream);
      if (!rc)
        rc = canary_facerhs_geo(r.eqn, r.dim, r.N, pf, r.rhs, q, r.sgeo, r.vgeo, r.nreal,
                                r.ntotal, r.vmapM, r.vmapP, r.bndy, r.phys, r.stream);
      if (!rc)
        rc = canary_updatesolution(r.eqn, r.dim, r.N, pf, r.rhs, q, r.vgeo, r.nreal, rka, rkb,
                                   r.dt, r.advection, r.stream);
    } else if (r.path == PATH_PREPARED) {
      rc = canary_stage_prepared(r.eqn, r.dim, r.N, r.prec, r.rhs, r.Q[cur], r.Q[1 - cur], r.vgeo,
                                 r.sgeo, r.pface, r.D, r.nreal, r.ntotal, rka, rkb, r.dt,
                                 r.advection, r.phys, r.stream);
      cur = 1 - cur;
    } else if (r.path == CANARY_PATH_FUSED) {
      rc = canary_stage_fused(r.eqn, r.dim, r.N, r.prec, r.rhs, r.Q[cur], r.Q[1 - cur], r.vgeo,
                              r.sgeo, r.D, r.nreal, r.ntotal, r.vmapM, r.vmapP, r.bndy, rka, rkb,
                              r.dt, r.advection, r.phys, r.stream);
      cur = 1 - cur;
    } else {
      // contracted: the volume term in the exact order; matrix-core: its
      // matrix-core form, and face + update (no contraction) contracted
      rc = canary_volumerhs(r.eqn, r.dim, r.N, r.prec == CANARY_FP64_FMA ? CANARY_FP64 : r.prec,
                            r.rhs, r.Q[cur], r.vgeo, r.D, r.nreal, r.phys, r.stream);
      if (!rc)
        rc = canary_facerhs_update(r.eqn, r.dim, r.N,
                                   r.prec == CANARY_FP64_MFMA ? CANARY_FP64_FMA : r.prec, r.rhs,
                                   r.Q[cur], r.Q[1 - cur],
                                   r.vgeo, r.sgeo, r.nreal, r.ntotal, r.vmapM, r.vmapP, r.bndy,
                                   rka, rkb, r.dt, r.advection, r.phys, r.stream);
      cur = 1 - cur;
    }
    if (rc) return rc;
  }
  return 0;
}

int run_loop(Run& r, int64_t nsteps, int graph, int* result_in_alt);

// The Euler pressure's pow table, made before any graph capture begins (a
// capture in progress cannot make one, and its kernels would then take the
// general path).
void register_kappa(int eqn, const canary_phys* phys, void* stream) {
  if (eqn != CANARY_EULER) return;
  canary_phys d;
  if (!phys) {
    canary_phys_default(&d);
    phys = &d;
  }
  canary_pow_register(phys->c_p / phys->c_v, stream);
}

}  // namespace

extern "C" int canary_batch_run(int kind, int eqn, int dim, int prec, const void* solves,
                                const canary_batch_class* classes, int nclasses, int64_t nsteps,
                                int parity, void* stream) {
  if (kind < CANARY_BATCH_DROPIN || kind > CANARY_BATCH_PREPARED_ALL) return CANARY_ERR_ARG;
  const bool all = kind == CANARY_BATCH_DROPIN_ALL || kind == CANARY_BATCH_PREPARED_ALL;
  if (nclasses < 0 || nsteps < 0 || (parity != 0 && parity != 1)) return CANARY_ERR_ARG;
  if (nsteps == 0 || nclasses == 0) return 0;
  if (!solves || !classes) return CANARY_ERR_ARG;
  // every class checked before the first launch, so a bad one launches nothing
  for (int c = 0; c < nclasses; ++c) {
    const canary_batch_class& cl = classes[c];
    if (cl.nitems < 0) return CANARY_ERR_ARG;
    if (cl.nitems == 0) continue;
    if (!cl.items) return CANARY_ERR_ARG;
    if (all) {  // entries of any order, checked by the kernel
      int e = 0, t = 0;
      const int rc = canary_batch_all_variant(eqn, dim, prec, 1, &e, &t);
      if (rc) return rc;
      continue;
    }
    bool found = false;
    for (int which = 0; which < 3 && !found; ++which) {
      int e = 0, t = 0;
      const int rc = kind == CANARY_BATCH_P32
                         ? canary_p32_batch_variant(dim, cl.N, which, &e, &t)
                         : canary_batch_variant(eqn, dim, prec, cl.N, which, &e, &t);
      found = rc == 0 && e == cl.epw;
    }
    if (!found) return CANARY_ERR_UNSUPPORTED;
  }
  double A[5], B[5], C[5];
  lsrk54_tables(A, B, C);
  int par = parity;
  for (int64_t k = 0; k < nsteps; ++k)
    for (int s = 0; s < 5; ++s) {
      for (int c = 0; c < nclasses; ++c) {
        const canary_batch_class& cl = classes[c];
        if (cl.nitems == 0) continue;
        int rc;
        if (kind == CANARY_BATCH_DROPIN)
          rc = canary_batch_stage(eqn, dim, prec, cl.N, cl.epw, (const canary_batch_solve*)solves,
                                  cl.items, cl.nitems, par, A[(s + 1) % 5], B[s], stream);
        else if (kind == CANARY_BATCH_DROPIN_ALL)
          rc = canary_batch_stage_all(eqn, dim, prec, (const canary_batch_solve*)solves, cl.items,
                                      cl.nitems, par, A[(s + 1) % 5], B[s], stream);
        else if (kind == CANARY_BATCH_PREPARED)
          rc = canary_batch_stage_prepared(eqn, dim, prec, cl.N, cl.epw,
                                           (const canary_batch_solve*)solves, cl.items, cl.nitems,
                                           par, A[(s + 1) % 5], B[s], stream);
        else if (kind == CANARY_BATCH_PREPARED_ALL)
          rc = canary_batch_stage_all_prepared(eqn, dim, prec, (const canary_batch_solve*)solves,
                                               cl.items, cl.nitems, par, A[(s + 1) % 5], B[s],
                                               stream);
        else
          rc = canary_p32_batch_stage(dim, cl.N, cl.epw, (const canary_p32_solve*)solves,
                                      cl.items, cl.nitems, par, A[(s + 1) % 5], B[s], stream);
        if (rc) return rc;
      }
      par = 1 - par;
    }
  return 0;
}

extern "C" void canary_lsrk54_coefs(double* rka, double* rkb, double* rkc) {
  double a[5], b[5], c[5];
  lsrk54_tables(a, b, c);
  for (int i = 0; i < 5; ++i) {
    if (rka) rka[i] = a[i];
    if (rkb) rkb[i] = b[i];
    if (rkc) rkc[i] = c[i];
  }
}

extern "C" int canary_lsrk54(int eqn, int dim, int N, int prec, int path, int64_t nsteps,
                             double dt, void* Q, void* Qalt, void* rhs, const void* vgeo,
                             const void* sgeo, const double* D, int64_t nreal,
                             int64_t nelem_total, const int64_t* vmapM, const int64_t* vmapP,
                             const int64_t* elemtobndy, const canary_phys* phys, int advection,
                             int graph, int* result_in_alt, void* stream) {
  if (!result_in_alt || nsteps < 0) return CANARY_ERR_ARG;
  if (path != CANARY_PATH_KERNELS && path != CANARY_PATH_FUSED && path != CANARY_PATH_FACE_UPDATE)
    return CANARY_ERR_ARG;
  if (prec != CANARY_FP64 && prec != CANARY_FP64_FMA && prec != CANARY_FP64_MFMA)
    return CANARY_ERR_UNSUPPORTED;
  if (advection && eqn != CANARY_SWE) return CANARY_ERR_ARG;
  if (path != CANARY_PATH_KERNELS && (!Qalt || Qalt == Q)) return CANARY_ERR_ARG;
  // the loop exchanges no ghost elements: they would be frozen in Q and, on
  // the ping-pong paths, never written in Qalt
  if (nelem_total != nreal) return CANARY_ERR_ARG;
  register_kappa(eqn, phys, stream);
  *result_in_alt = 0;
  if (nsteps == 0 || nreal == 0) return 0;
  Run r;
  r.pface = nullptr;
  r.eqn = eqn;
  r.dim = dim;
  r.N = N;
  r.prec = prec;
  r.path = path;
  r.advection = advection;
  r.dt = dt;
  r.Q[0] = Q;
  r.Q[1] = Qalt;
  r.rhs = rhs;
  r.vgeo = vgeo;
  r.sgeo = sgeo;
  r.D = D;
  r.nreal = nreal;
  r.ntotal = nelem_total;
  r.vmapM = vmapM;
  r.vmapP = vmapP;
  r.bndy = elemtobndy;
  r.phys = phys;
  r.stream = stream;
  return run_loop(r, nsteps, graph, result_in_alt);
}

extern "C" int canary_lsrk54_prepared(int eqn, int dim, int N, int prec, int64_t nsteps,
                                      double dt, void* Q, void* Qalt, void* rhs,
                                      const void* pvgeo, const void* psgeo,
                                      const int32_t* pface, const double* D, int64_t nreal,
                                      const canary_phys* phys, int advection, int graph,
                                      int* result_in_alt, void* stream) {
  if (!result_in_alt || nsteps < 0) return CANARY_ERR_ARG;
  if (advection && eqn != CANARY_SWE) return CANARY_ERR_ARG;
  if (!Qalt || Qalt == Q) return CANARY_ERR_ARG;
  register_kappa(eqn, phys, stream);
  *result_in_alt = 0;
  if (nsteps == 0 || nreal == 0) return 0;
  Run r;
  r.eqn = eqn;
  r.dim = dim;
  r.N = N;
  r.prec = prec;
  r.path = PATH_PREPARED;
  r.advection = advection;
  r.pface = pface;
  r.dt = dt;
  r.Q[0] = Q;
  r.Q[1] = Qalt;
  r.rhs = rhs;
  r.vgeo = pvgeo;
  r.sgeo = psgeo;
  r.D = D;
  r.nreal = nreal;
  r.ntotal = nreal;
  r.vmapM = nullptr;
  r.vmapP = nullptr;
  r.bndy = nullptr;
  r.phys = phys;
  r.stream = stream;
  return run_loop(r, nsteps, graph, result_in_alt);
}

namespace {

int run_loop(Run& r, int64_t nsteps, int graph, int* result_in_alt) {
  void* stream = r.stream;
  lsrk54_tables(r.A, r.B, r.C);
  int cur = 0;
  // graph: two steps (an even number of stages, so the ping-pong buffers are
  // back in place) captured once and replayed; the legacy null stream
  // cannot be captured, so it always runs eagerly
  const int64_t per = 2;
  int64_t done = 0;
  if (graph && stream && nsteps >= per) {
    hipStream_t st = (hipStream_t)stream;
    hipGraph_t g = nullptr;
    hipGraphExec_t ge = nullptr;
    if (hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal) != hipSuccess)
      return CANARY_ERR_LAUNCH;
    int c = 0, rc = 0;
    for (int64_t k = 0; k < per && !rc; ++k) rc = step(r, c);
    const hipError_t ec = hipStreamEndCapture(st, &g);
    if (rc) {
      if (g) hipGraphDestroy(g);
      return rc;
    }
    if (ec != hipSuccess || !g) return CANARY_ERR_LAUNCH;
    if (hipGraphInstantiate(&ge, g, nullptr, nullptr, 0) != hipSuccess) {
      hipGraphDestroy(g);
      return CANARY_ERR_LAUNCH;
    }
    int rc2 = 0;
    for (; done + per <= nsteps; done += per)
      if (hipGraphLaunch(ge, st) != hipSuccess) {
        rc2 = CANARY_ERR_LAUNCH;
        break;
      }
    // the graph is released only after its launches have run
    if (hipStreamSynchronize(st) != hipSuccess && !rc2) rc2 = CANARY_ERR_LAUNCH;
    hipGraphExecDestroy(ge);
    hipGraphDestroy(g);
    if (rc2) return rc2;
  }
  for (; done < nsteps; ++done) {
    const int rc = step(r, cur);
    if (rc) return rc;
  }
  *result_in_alt = cur;
  return 0;
}

}  // namespace
