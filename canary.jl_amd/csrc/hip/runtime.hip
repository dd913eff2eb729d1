// Public C ABI of the device path: argument checks, (equation, dim)
// dispatch to the kernel translation units, device error reporting.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <mutex>

#include "../../../include/canary_dg.h"
#include "launch.h"
#include "pow_registry.h"

extern "C" {
int canary_k_euler3d(int op, int N, const canary::KArgs* a);
int canary_k_euler2d(int op, int N, const canary::KArgs* a);
int canary_k_swe2d(int op, int N, const canary::KArgs* a);
int canary_k_euler2d_te(int op, int N, const canary::KArgs* a);
int canary_k_euler2d_te_fma(int op, int N, const canary::KArgs* a);
int canary_k_euler2d_te_status(int64_t* info, int clear);
int canary_k_euler2d_te_fma_status(int64_t* info, int clear);
int canary_k_euler3d_fma(int op, int N, const canary::KArgs* a);
int canary_k_euler3d_mfma(int op, int N, const canary::KArgs* a);
int canary_k_euler3d_mfma_status(int64_t* info, int clear);
int canary_k_euler2d_fma(int op, int N, const canary::KArgs* a);
int canary_k_swe2d_fma(int op, int N, const canary::KArgs* a);
int canary_k_euler3d_status(int64_t* info, int clear);
int canary_k_euler2d_status(int64_t* info, int clear);
int canary_k_swe2d_status(int64_t* info, int clear);
int canary_k_euler3d_fma_status(int64_t* info, int clear);
int canary_k_euler2d_fma_status(int64_t* info, int clear);
int canary_k_swe2d_fma_status(int64_t* info, int clear);
int canary_k_perturb32_status(int64_t* info, int clear);
int canary_k_batch_status(int64_t* info, int clear);
int canary_k_batch_fma_status(int64_t* info, int clear);
int canary_k_geometry_status(int64_t* info, int clear);
int canary_k_aux_status(int64_t* info, int clear);
int canary_k_prepared_status(int64_t* info, int clear);
}

namespace canary {

// The pow tables of every registered kappa, per device (see pow_registry.h).
static __device__ crpow::KappaTable g_ktab[kMaxKappa];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_ktab_mu;
crpow::KappaTable* g_ktab_base[kMaxDevices];
double g_ktab_kappa[kMaxDevices][kMaxKappa];
int g_ktab_count[kMaxDevices];

bool same_bits(double a, double b) {
  uint64_t x, y;
  memcpy(&x, &a, 8);
  memcpy(&y, &b, 8);
  return x == y;
}

// device index and the array's address on it (looked up once per device)
int ktab_device(int* dev) {
  if (hipGetDevice(dev) != hipSuccess || *dev < 0 || *dev >= kMaxDevices) return -1;
  if (!g_ktab_base[*dev]) {
    void* p = nullptr;
    if (hipGetSymbolAddress(&p, HIP_SYMBOL(g_ktab)) != hipSuccess || !p) return -1;
    g_ktab_base[*dev] = static_cast<crpow::KappaTable*>(p);
  }
  return 0;
}
}  // namespace

const crpow::KappaTable* kappa_table(double kappa, hipStream_t stream) {
  std::lock_guard<std::mutex> lock(g_ktab_mu);
  int dev = -1;
  if (ktab_device(&dev)) return nullptr;
  for (int i = 0; i < g_ktab_count[dev]; ++i)
    if (same_bits(g_ktab_kappa[dev][i], kappa)) return g_ktab_base[dev] + i;
  if (g_ktab_count[dev] >= kMaxKappa) return nullptr;
  hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
  if (stream && (hipStreamIsCapturing(stream, &cs) != hipSuccess ||
                 cs != hipStreamCaptureStatusNone))
    return nullptr;
  static crpow::KappaTable host;  // built under the lock
  if (!crpow::kappa_table_build(kappa, &host)) return nullptr;
  const int slot = g_ktab_count[dev];
  if (hipMemcpy(g_ktab_base[dev] + slot, &host, sizeof(host), hipMemcpyHostToDevice) != hipSuccess)
    return nullptr;
  g_ktab_kappa[dev][slot] = kappa;
  g_ktab_count[dev] = slot + 1;
  return g_ktab_base[dev] + slot;
}

const crpow::KappaTable* kappa_tables(int* n) {
  std::lock_guard<std::mutex> lock(g_ktab_mu);
  int dev = -1;
  *n = 0;
  if (ktab_device(&dev)) return nullptr;
  *n = g_ktab_count[dev];
  return g_ktab_base[dev];
}

}  // namespace canary

extern "C" int canary_pow_register(double kappa, void* stream) {
  return canary::kappa_table(kappa, (hipStream_t)stream) ? 0 : CANARY_ERR_UNSUPPORTED;
}

namespace {

using KFn = int (*)(int, int, const canary::KArgs*);

// CANARY_FP64 selects the exact-order units, CANARY_FP64_FMA the contracted
// ones (every operation), CANARY_FP64_MFMA the matrix-core unit (3D Euler:
// volume term and fused stages).
KFn select(int eqn, int dim, int prec) {
  if (prec == CANARY_FP64_MFMA)
    return eqn == CANARY_EULER && dim == 3 ? canary_k_euler3d_mfma : nullptr;
  const bool fma = prec == CANARY_FP64_FMA;
  if (prec != CANARY_FP64 && !fma) return nullptr;
  if (eqn == CANARY_EULER && dim == 3) return fma ? canary_k_euler3d_fma : canary_k_euler3d;
  if (eqn == CANARY_EULER && dim == 2) return fma ? canary_k_euler2d_fma : canary_k_euler2d;
  if (eqn == CANARY_SWE && dim == 2) return fma ? canary_k_swe2d_fma : canary_k_swe2d;
  if (eqn == CANARY_EULER_TE && dim == 2)
    return fma ? canary_k_euler2d_te_fma : canary_k_euler2d_te;
  return nullptr;
}

canary::PhysParams to_params(const canary_phys* ph, void* stream = nullptr, int eqn = -1) {
  canary_phys d;
  if (!ph) {
    canary_phys_default(&d);
    ph = &d;
  }
  canary::PhysParams p;
  p.gamma = ph->gamma;
  p.p0 = ph->p0;
  p.R = ph->R_gas;
  p.cp = ph->c_p;
  p.cv = ph->c_v;
  p.g = ph->gravity;
  p.kappa = ph->c_p / ph->c_v;
  p.swe_g = ph->swe_gravity;
  p.dnl = ph->delta_nl;
  p.ktab = eqn == CANARY_EULER ? canary::kappa_table(p.kappa, (hipStream_t)stream) : nullptr;
  p.ser = canary::kPowSeries;
  return p;
}

// contracted: the entry point accepts the contracted forms (CANARY_FP64_FMA,
// CANARY_FP64_MFMA); an operation the selected unit does not hold reports
// CANARY_ERR_UNSUPPORTED from the unit.
int check_common(int eqn, int dim, int N, int prec, int64_t nreal, KFn* fn,
                 bool contracted = false) {
  if ((prec == CANARY_FP64_FMA || prec == CANARY_FP64_MFMA) && !contracted)
    return CANARY_ERR_UNSUPPORTED;
  *fn = select(eqn, dim, prec);
  if (!*fn) return CANARY_ERR_UNSUPPORTED;
  if (N < 1 || N > canary::kMaxN) return CANARY_ERR_UNSUPPORTED;
  if (nreal < 0) return CANARY_ERR_ARG;
  return 0;
}

}  // namespace

extern "C" int canary_volumerhs(int eqn, int dim, int N, int prec, void* rhs, const void* Q,
                                const void* vgeo, const double* D, int64_t nreal,
                                const canary_phys* phys, void* stream) {
  KFn fn;
  int rc = check_common(eqn, dim, N, prec, nreal, &fn, true);
  if (rc) return rc;
  if (nreal == 0) return 0;
  if (!rhs || !Q || !vgeo || !D) return CANARY_ERR_ARG;
  canary::KArgs a = {};
  a.rhs = rhs;
  a.Q = Q;
  a.vgeo = vgeo;
  a.D = D;
  a.nreal = nreal;
  a.ntotal = nreal;
  a.phys = to_params(phys, stream, eqn);
  a.stream = (hipStream_t)stream;
  return fn(canary::OP_VOLUME, N, &a);
}

extern "C" int canary_facerhs_geo(int eqn, int dim, int N, int prec, void* rhs, const void* Q,
                                  const void* sgeo, const void* vgeo, int64_t nreal,
                                  int64_t nelem_total, const int64_t* vmapM,
                                  const int64_t* vmapP, const int64_t* elemtobndy,
                                  const canary_phys* phys, void* stream) {
  KFn fn;
  int rc = check_common(eqn, dim, N, prec, nreal, &fn, true);
  if (rc) return rc;
  if (nreal == 0) return 0;
  if (nelem_total < nreal) return CANARY_ERR_ARG;
  if (!rhs || !Q || !sgeo || !vmapP || !elemtobndy) return CANARY_ERR_ARG;
  if (eqn == CANARY_EULER_TE && !vgeo) return CANARY_ERR_ARG;  // the pressure reads y
  canary::KArgs a = {};
  a.rhs = rhs;
  a.Q = Q;
  a.vgeo = vgeo;
  a.sgeo = sgeo;
  a.nreal = nreal;
  a.ntotal = nelem_total;
  a.vmapM = vmapM;
  a.vmapP = vmapP;
  a.bndy = elemtobndy;
  a.phys = to_params(phys, stream, eqn);
  a.stream = (hipStream_t)stream;
  return fn(canary::OP_FACE, N, &a);
}

extern "C" int canary_facerhs(int eqn, int dim, int N, int prec, void* rhs, const void* Q,
                              const void* sgeo, int64_t nreal, int64_t nelem_total,
                              const int64_t* vmapM, const int64_t* vmapP,
                              const int64_t* elemtobndy, const canary_phys* phys, void* stream) {
  return canary_facerhs_geo(eqn, dim, N, prec, rhs, Q, sgeo, nullptr, nreal, nelem_total, vmapM,
                            vmapP, elemtobndy, phys, stream);
}

extern "C" int canary_updatesolution(int eqn, int dim, int N, int prec, void* rhs, void* Q,
                                     const void* vgeo, int64_t nreal, double rka, double rkb,
                                     double dt, int advection, void* stream) {
  KFn fn;
  int rc = check_common(eqn, dim, N, prec, nreal, &fn, true);
  if (rc) return rc;
  if (nreal == 0) return 0;
  if (!rhs || !Q || !vgeo) return CANARY_ERR_ARG;
  if (advection && eqn != CANARY_SWE) return CANARY_ERR_ARG;
  canary::KArgs a = {};
  a.rhs = rhs;
  a.Q = Q;
  a.Qout = Q;
  a.vgeo = vgeo;
  a.nreal = nreal;
  a.ntotal = nreal;
  a.rka = rka;
  a.rkb = rkb;
  a.dt = dt;
  a.advection = advection;
  a.stream = (hipStream_t)stream;
  return fn(canary::OP_UPDATE, N, &a);
}

extern "C" int canary_stage_fused(int eqn, int dim, int N, int prec, void* rhs, const void* Qin,
                                  void* Qout, const void* vgeo, const void* sgeo,
                                  const double* D, int64_t nreal, int64_t nelem_total,
                                  const int64_t* vmapM, const int64_t* vmapP,
                                  const int64_t* elemtobndy, double rka, double rkb, double dt,
                                  int advection, const canary_phys* phys, void* stream) {
  KFn fn;
  int rc = check_common(eqn, dim, N, prec, nreal, &fn, true);
  if (rc) return rc;
  if (advection && eqn != CANARY_SWE) return CANARY_ERR_ARG;
  if (nreal == 0) return 0;
  if (nelem_total < nreal) return CANARY_ERR_ARG;
  if (!rhs || !Qin || !Qout || !vgeo || !sgeo || !D || !vmapP || !elemtobndy)
    return CANARY_ERR_ARG;
  if (Qin == (const void*)Qout) return CANARY_ERR_ARG;
  canary::KArgs a = {};
  a.rhs = rhs;
  a.Q = Qin;
  a.Qout = Qout;
  a.vgeo = vgeo;
  a.sgeo = sgeo;
  a.D = D;
  a.nreal = nreal;
  a.ntotal = nelem_total;
  a.vmapM = vmapM;
  a.vmapP = vmapP;
  a.bndy = elemtobndy;
  a.rka = rka;
  a.rkb = rkb;
  a.dt = dt;
  a.advection = advection;
  a.phys = to_params(phys, stream, eqn);
  a.stream = (hipStream_t)stream;
  return fn(canary::OP_STAGE, N, &a);
}

extern "C" int canary_facerhs_update(int eqn, int dim, int N, int prec, void* rhs,
                                     const void* Qin, void* Qout, const void* vgeo,
                                     const void* sgeo, int64_t nreal, int64_t nelem_total,
                                     const int64_t* vmapM, const int64_t* vmapP,
                                     const int64_t* elemtobndy, double rka, double rkb, double dt,
                                     int advection, const canary_phys* phys, void* stream) {
  KFn fn;
  int rc = check_common(eqn, dim, N, prec, nreal, &fn, true);
  if (rc) return rc;
  if (advection && eqn != CANARY_SWE) return CANARY_ERR_ARG;
  if (nreal == 0) return 0;
  if (nelem_total < nreal) return CANARY_ERR_ARG;
  if (!rhs || !Qin || !Qout || !vgeo || !sgeo || !vmapP || !elemtobndy)
    return CANARY_ERR_ARG;
  if (Qin == (const void*)Qout) return CANARY_ERR_ARG;
  canary::KArgs a = {};
  a.rhs = rhs;
  a.Q = Qin;
  a.Qout = Qout;
  a.vgeo = vgeo;
  a.sgeo = sgeo;
  a.nreal = nreal;
  a.ntotal = nelem_total;
  a.vmapM = vmapM;
  a.vmapP = vmapP;
  a.bndy = elemtobndy;
  a.rka = rka;
  a.rkb = rkb;
  a.dt = dt;
  a.advection = advection;
  a.phys = to_params(phys, stream, eqn);
  a.stream = (hipStream_t)stream;
  return fn(canary::OP_FACE_UPDATE, N, &a);
}

extern "C" int canary_stage_prepared(int eqn, int dim, int N, int prec, void* rhs,
                                     const void* Qin, void* Qout, const void* pvgeo,
                                     const void* psgeo, const int32_t* pface, const double* D,
                                     int64_t nreal, int64_t nelem_total, double rka, double rkb,
                                     double dt, int advection, const canary_phys* phys,
                                     void* stream) {
  KFn fn;
  int rc = check_common(eqn, dim, N, prec, nreal, &fn, true);
  if (rc) return rc;
  if (advection && eqn != CANARY_SWE) return CANARY_ERR_ARG;
  if (nreal == 0) return 0;
  if (nelem_total < nreal) return CANARY_ERR_ARG;
  if (!rhs || !Qin || !Qout || !pvgeo || !psgeo || !pface || !D) return CANARY_ERR_ARG;
  if (Qin == (const void*)Qout) return CANARY_ERR_ARG;
  canary::KArgs a = {};
  a.rhs = rhs;
  a.Q = Qin;
  a.Qout = Qout;
  a.vgeo = pvgeo;
  a.sgeo = psgeo;
  a.pface = pface;
  a.D = D;
  a.nreal = nreal;
  a.ntotal = nelem_total;
  a.rka = rka;
  a.rkb = rkb;
  a.dt = dt;
  a.advection = advection;
  a.phys = to_params(phys, stream, eqn);
  a.stream = (hipStream_t)stream;
  return fn(canary::OP_STAGE_PREP, N, &a);
}

extern "C" int canary_check_errors(void* stream, int64_t* info) {
  if (hipStreamSynchronize((hipStream_t)stream) != hipSuccess) return CANARY_ERR_LAUNCH;
  int64_t tmp[4];
  int first = 0;
  int (*readers[15])(int64_t*, int) = {
      canary_k_euler3d_status,        canary_k_euler2d_status,
      canary_k_swe2d_status,          canary_k_euler2d_te_status,
      canary_k_euler3d_fma_status,    canary_k_euler2d_fma_status,
      canary_k_swe2d_fma_status,      canary_k_euler2d_te_fma_status,
      canary_k_perturb32_status,      canary_k_batch_status,
      canary_k_batch_fma_status,      canary_k_geometry_status,
      canary_k_aux_status,            canary_k_prepared_status,
      canary_k_euler3d_mfma_status};
  for (auto rd : readers) {
    int rc = rd(tmp, 1);
    if (rc && !first) {
      first = rc;
      if (info)
        for (int i = 0; i < 4; ++i) info[i] = tmp[i];
    }
  }
  if (!first && info)
    for (int i = 0; i < 4; ++i) info[i] = 0;
  return first;
}

extern "C" int canary_device_count(void) {
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess) return 0;
  return n;
}

// Streams owned by the batched / served runtimes: created here so that every
// runtime stream is distinct (a stream borrowed from a shared pool can alias
// another user's stream, which then joins a graph capture with unrelated
// work pending).  Default priority, not synchronising with the null stream.
extern "C" int canary_stream_create(void** out) {
  if (!out) return CANARY_ERR_ARG;
  hipStream_t s = nullptr;
  if (hipStreamCreateWithFlags(&s, hipStreamNonBlocking) != hipSuccess) return CANARY_ERR_LAUNCH;
  *out = (void*)s;
  return 0;
}

extern "C" int canary_stream_destroy(void* s) {
  if (!s) return CANARY_ERR_ARG;
  hipStream_t st = (hipStream_t)s;
  if (hipStreamSynchronize(st) != hipSuccess) return CANARY_ERR_LAUNCH;
  return hipStreamDestroy(st) == hipSuccess ? 0 : CANARY_ERR_LAUNCH;
}
