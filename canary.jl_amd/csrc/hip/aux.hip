// Kernels beside the stage: ghost-element exchange buffers and the two
// reductions the drivers print.
//
//   fillsendQ     sendQ[:, :, e] = Q[:, :, sendelems[e]]   (euler3d_set2c.jl:645-659)
//   transferrecvQ Q[:, :, nrealelem + e] = recvQ[:, :, e]  (euler3d_set2c.jl:663-676)
//   pack_nodes / unpack_nodes: the same for a list of single nodes (the face
//                 traces a neighbour reads through vmapP), buf[k, s] = Q[node_k, s]
//   l2energysquared  sum_e sum_s sum_i MJ q^2               (euler3d_set2c.jl:738-750)
//   cfl              min over nodes of the local step       (euler3d_set2c.jl:67-97)
//   courant          Courant-number step of the Euler sets (no reference counterpart)
//
// The copies move whole elements (Np * nstate contiguous words) with one
// lane per word.  The reductions are deterministic: a fixed grid of
// RED_BLOCKS workgroups each sums a fixed contiguous range in LDS tree
// order, then one workgroup combines the RED_BLOCKS partials; no atomics.
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>

#include <mutex>

#include "../../../include/canary_dg.h"
#include "crpow.h"

namespace canary {
namespace aux {

static __device__ unsigned long long g_aux_status[4];

__device__ __forceinline__ void aux_error(unsigned long long code, long long a, long long b,
                                          long long value) {
  if (atomicCAS(&g_aux_status[0], 0ull, code) == 0ull) {
    g_aux_status[1] = (unsigned long long)a;
    g_aux_status[2] = (unsigned long long)b;
    g_aux_status[3] = (unsigned long long)value;
  }
}

constexpr int RED_BLOCKS = CANARY_REDUCE_WORK;
constexpr int RED_THREADS = 256;

template <class W>
__global__ __launch_bounds__(256) void fill_send_kernel(W* __restrict__ sendQ,
                                                        const W* __restrict__ Q,
                                                        const int64_t* __restrict__ sendelems,
                                                        int64_t nsend, int64_t words,
                                                        int64_t nelem_total) {
  const int64_t total = nsend * words;
  for (int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; g < total;
       g += (int64_t)gridDim.x * blockDim.x) {
    const int64_t e = g / words, w = g - e * words;
    const int64_t re = sendelems[e];
    if (re < 1 || re > nelem_total) {
      if (w == 0) aux_error(CANARY_ERR_ARG, e + 1, -1, re);
      continue;
    }
    sendQ[g] = Q[(size_t)(re - 1) * words + w];
  }
}

template <class W>
__global__ __launch_bounds__(256) void recv_kernel(W* __restrict__ Q, const W* __restrict__ recvQ,
                                                   int64_t total, int64_t offset) {
  for (int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; g < total;
       g += (int64_t)gridDim.x * blockDim.x)
    Q[offset + g] = recvQ[g];
}

// node ids are 1-based over the (Np, nstate, nelem_total) array, as vmapP's;
// one lane per (node, state) word, consecutive lanes on consecutive nodes
template <class W, bool PACK>
__global__ __launch_bounds__(256) void nodes_kernel(W* __restrict__ Q, W* __restrict__ buf,
                                                    const int64_t* __restrict__ nodes,
                                                    int64_t nnode, int64_t Np, int64_t nstate,
                                                    int64_t nelem_total) {
  const int64_t total = nnode * nstate;
  for (int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; g < total;
       g += (int64_t)gridDim.x * blockDim.x) {
    const int64_t s = g / nnode, k = g - s * nnode;
    const int64_t id = nodes[k];
    if (id < 1 || id > nelem_total * Np) {
      if (s == 0) aux_error(CANARY_ERR_VMAP, k + 1, -1, id);
      continue;
    }
    const int64_t e = (id - 1) / Np, n = (id - 1) - e * Np;
    const size_t q = (size_t)(e * nstate + s) * Np + n;
    if (PACK)
      buf[g] = Q[q];
    else
      Q[q] = buf[g];
  }
}

// LDS tree sum / min of the block's values; result in lane 0
template <bool MIN>
__device__ __forceinline__ double block_reduce(double v, double* sh) {
  const int t = threadIdx.x;
  sh[t] = v;
  __syncthreads();
  for (int s = RED_THREADS / 2; s > 0; s >>= 1) {
    if (t < s) sh[t] = MIN ? fmin(sh[t], sh[t + s]) : sh[t] + sh[t + s];
    __syncthreads();
  }
  return sh[0];
}

__global__ __launch_bounds__(RED_THREADS) void energy_kernel(
    const double* __restrict__ Q, const double* __restrict__ vgeo, int64_t Np, int64_t nstate,
    int64_t nvgeo, int64_t iMJ, int64_t nelem, double* __restrict__ work) {
  __shared__ double sh[RED_THREADS];
  const int64_t total = Np * nstate * nelem;
  const int64_t chunk = (total + RED_BLOCKS - 1) / RED_BLOCKS;
  const int64_t lo = blockIdx.x * chunk;
  const int64_t hi = lo + chunk < total ? lo + chunk : total;
  double acc = 0.0;
  for (int64_t g = lo + threadIdx.x; g < hi; g += RED_THREADS) {
    const int64_t e = g / (Np * nstate), r = g - e * Np * nstate;
    const int64_t i = r % Np;
    const double q = Q[g];
    acc += vgeo[i + Np * (iMJ + nvgeo * e)] * (q * q);
  }
  const double s = block_reduce<false>(acc, sh);
  if (threadIdx.x == 0) work[blockIdx.x] = s;
}

// local time step of node n of element e (host/problems.cpp canary_cfl)
__device__ __forceinline__ double local_dt(int eqn, int dim, int64_t Np, const double* Q,
                                           const double* vgeo, int64_t e, int64_t n,
                                           const canary_phys& ph) {
  if (eqn == CANARY_EULER && dim == 3) {
    const double* q = Q + (size_t)Np * 5 * e;
    const double* vg = vgeo + (size_t)Np * 14 * e;
    const double rho = q[n + Np * 3], U = q[n], V = q[n + Np], W = q[n + 2 * Np];
    const double E = q[n + Np * 4];
    const double P = ph.p0 * canary::crpow::pow_cr(ph.R_gas * E / ph.p0, ph.c_p / ph.c_v);
    const double c = rho * sqrt(ph.gamma * P / rho);
    const double a = fabs(U * vg[n] + V * vg[n + 3 * Np] + W * vg[n + 6 * Np]) + c;
    const double b = fabs(U * vg[n + Np] + V * vg[n + 4 * Np] + W * vg[n + 7 * Np]) + c;
    const double d = fabs(U * vg[n + 2 * Np] + V * vg[n + 5 * Np] + W * vg[n + 8 * Np]) + c;
    double m = a > b ? a : b;
    m = m > d ? m : d;
    return 2 * rho / m;
  }
  if (eqn == CANARY_EULER && dim == 2) {
    const double* q = Q + (size_t)Np * 4 * e;
    const double* vg = vgeo + (size_t)Np * 8 * e;
    const double rho = q[n + Np * 2], U = q[n], V = q[n + Np], E = q[n + Np * 3];
    const double P = ph.p0 * canary::crpow::pow_cr(ph.R_gas * E / ph.p0, ph.c_p / ph.c_v);
    const double c = rho * sqrt(ph.gamma * P / rho);
    const double a = fabs(U * vg[n] + V * vg[n + 2 * Np]) + c;
    const double b = fabs(U * vg[n + Np] + V * vg[n + 3 * Np]) + c;
    return 2 * rho / (a > b ? a : b);
  }
  const double* q = Q + (size_t)Np * 4 * e;
  const double* vg = vgeo + (size_t)Np * 8 * e;
  const double h = q[n + 2 * Np], b = q[n + 3 * Np], U = q[n], V = q[n + Np];
  const double H = h + b;
  const double s = H * sqrt(ph.swe_gravity * H) * ph.delta_nl;
  const double a = fabs(U * vg[n] + V * vg[n + 2 * Np]) + s;
  const double c = fabs(U * vg[n + Np] + V * vg[n + 3 * Np]) + s;
  return 2 * H / (a > c ? a : c);
}

__global__ __launch_bounds__(RED_THREADS) void cfl_kernel(int eqn, int dim, int64_t Np,
                                                          const double* __restrict__ vgeo,
                                                          const double* __restrict__ Q,
                                                          int64_t nelem, canary_phys ph,
                                                          double* __restrict__ work) {
  __shared__ double sh[RED_THREADS];
  const int64_t total = Np * nelem;
  const int64_t chunk = (total + RED_BLOCKS - 1) / RED_BLOCKS;
  const int64_t lo = blockIdx.x * chunk;
  const int64_t hi = lo + chunk < total ? lo + chunk : total;
  double acc = 1.7976931348623157e308;
  for (int64_t g = lo + threadIdx.x; g < hi; g += RED_THREADS) {
    const int64_t e = g / Np, n = g - e * Np;
    const double v = local_dt(eqn, dim, Np, Q, vgeo, e, n, ph);
    acc = v < acc ? v : acc;
  }
  const double s = block_reduce<true>(acc, sh);
  if (threadIdx.x == 0) work[blockIdx.x] = s;
}

// Courant-number step of the Euler sets (problems.euler_courant): min over
// nodes of courant * dx / (|u| + c) / N, dx = sqrt(sum_d (1 / (2 xi^d_d))^2)
__global__ __launch_bounds__(RED_THREADS) void courant_kernel(int dim, int N, int64_t Np,
                                                              const double* __restrict__ vgeo,
                                                              const double* __restrict__ Q,
                                                              int64_t nelem, canary_phys ph,
                                                              double courant,
                                                              double* __restrict__ work) {
  __shared__ double sh[RED_THREADS];
  const int ns = dim + 2, nv = dim == 3 ? 14 : 8;
  const int64_t total = Np * nelem;
  const int64_t chunk = (total + RED_BLOCKS - 1) / RED_BLOCKS;
  const int64_t lo = blockIdx.x * chunk;
  const int64_t hi = lo + chunk < total ? lo + chunk : total;
  double acc = 1.7976931348623157e308;
  for (int64_t g = lo + threadIdx.x; g < hi; g += RED_THREADS) {
    const int64_t e = g / Np, n = g - e * Np;
    const double* q = Q + (size_t)Np * ns * e;
    const double* vg = vgeo + (size_t)Np * nv * e;
    const double rho = q[n + Np * dim], rt = q[n + Np * (dim + 1)];
    double vel2 = 0.0, dx2 = 0.0;
    for (int c = 0; c < dim; ++c) {
      const double u = q[n + Np * c] / rho;
      vel2 = vel2 + u * u;
      const double h = 1.0 / (2 * vg[n + Np * (c * (dim + 1))]);  // xi_x, eta_y(, zeta_z)
      dx2 = dx2 + h * h;
    }
    const double P = ph.p0 * canary::crpow::pow_cr(ph.R_gas * rt / ph.p0, ph.c_p / ph.c_v);
    const double cs = sqrt(ph.gamma * P / rho);
    const double v = courant * sqrt(dx2) / (sqrt(vel2) + cs) / N;
    acc = v < acc ? v : acc;
  }
  const double s = block_reduce<true>(acc, sh);
  if (threadIdx.x == 0) work[blockIdx.x] = s;
}

template <bool MIN>
__global__ __launch_bounds__(RED_THREADS) void finish_kernel(const double* __restrict__ work,
                                                             double* __restrict__ out) {
  __shared__ double sh[RED_THREADS];
  double acc = MIN ? 1.7976931348623157e308 : 0.0;
  for (int b = threadIdx.x; b < RED_BLOCKS; b += RED_THREADS)
    acc = MIN ? fmin(acc, work[b]) : acc + work[b];
  const double s = block_reduce<MIN>(acc, sh);
  if (threadIdx.x == 0) out[0] = s;
}

int status() { return hipGetLastError() == hipSuccess ? 0 : CANARY_ERR_LAUNCH; }

// Partials of canary_cfl_device_noalloc: the library's own copy of the
// `work` buffer, one per device (the loader places it with the code object;
// it stays until the process ends).  Host code must not pass this name to a
// kernel -- on the host it is only a stand-in for the symbol; the device
// address comes from hipGetSymbolAddress, kept per device below.
static __device__ __align__(256) double scratch[CANARY_REDUCE_WORK];
static_assert(RED_BLOCKS <= CANARY_REDUCE_WORK, "one partial per reduction workgroup");

constexpr int MAX_DEVICES = 64;
std::mutex g_scratch_mu;
double* g_scratch_ptr[MAX_DEVICES];

// The current device's address of `scratch`, looked up on the first call
// made with that device current and kept.
int scratch_ptr(double** out) {
  int dev = -1;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0) return CANARY_ERR_LAUNCH;
  if (dev >= MAX_DEVICES) return CANARY_ERR_CAPACITY;
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void* p = nullptr;
    if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p) return CANARY_ERR_LAUNCH;
    g_scratch_ptr[dev] = static_cast<double*>(p);
  }
  *out = g_scratch_ptr[dev];
  return 0;
}

int64_t grid_for(int64_t work) {
  int64_t b = (work + 255) / 256;
  return b < 1 ? 1 : (b > 65536 ? 65536 : b);
}

}  // namespace aux
}  // namespace canary

using namespace canary::aux;

extern "C" int canary_fillsendQ(int dim, int N, int nstate, void* sendQ, const void* Q,
                                const int64_t* sendelems, int64_t nsend, int64_t nelem_total,
                                int word_bytes, void* stream) {
  if ((dim != 2 && dim != 3) || N < 1 || nstate < 1 || nsend < 0 || nelem_total < 0)
    return CANARY_ERR_ARG;
  if (nsend == 0) return 0;
  if (!sendQ || !Q || !sendelems) return CANARY_ERR_ARG;
  int64_t Np = 1;
  for (int d = 0; d < dim; ++d) Np *= N + 1;
  const int64_t words = Np * nstate;
  hipStream_t st = (hipStream_t)stream;
  const unsigned g = (unsigned)grid_for(nsend * words);
  if (word_bytes == 8)
    fill_send_kernel<double><<<g, 256, 0, st>>>((double*)sendQ, (const double*)Q, sendelems,
                                                nsend, words, nelem_total);
  else if (word_bytes == 4)
    fill_send_kernel<float><<<g, 256, 0, st>>>((float*)sendQ, (const float*)Q, sendelems, nsend,
                                               words, nelem_total);
  else
    return CANARY_ERR_ARG;
  return status();
}

extern "C" int canary_transferrecvQ(int dim, int N, int nstate, void* Q, const void* recvQ,
                                    int64_t nrecv, int64_t nrealelem, int word_bytes,
                                    void* stream) {
  if ((dim != 2 && dim != 3) || N < 1 || nstate < 1 || nrecv < 0 || nrealelem < 0)
    return CANARY_ERR_ARG;
  if (nrecv == 0) return 0;
  if (!Q || !recvQ) return CANARY_ERR_ARG;
  int64_t Np = 1;
  for (int d = 0; d < dim; ++d) Np *= N + 1;
  const int64_t words = Np * nstate;
  hipStream_t st = (hipStream_t)stream;
  const unsigned g = (unsigned)grid_for(nrecv * words);
  if (word_bytes == 8)
    recv_kernel<double><<<g, 256, 0, st>>>((double*)Q, (const double*)recvQ, nrecv * words,
                                           nrealelem * words);
  else if (word_bytes == 4)
    recv_kernel<float><<<g, 256, 0, st>>>((float*)Q, (const float*)recvQ, nrecv * words,
                                          nrealelem * words);
  else
    return CANARY_ERR_ARG;
  return status();
}

namespace {
template <bool PACK>
int nodes_entry(int dim, int N, int nstate, void* Q, void* buf, const int64_t* nodes,
                int64_t nnode, int64_t nelem_total, int word_bytes, void* stream) {
  if ((dim != 2 && dim != 3) || N < 1 || nstate < 1 || nnode < 0 || nelem_total < 0)
    return CANARY_ERR_ARG;
  if (nnode == 0) return 0;
  if (!Q || !buf || !nodes) return CANARY_ERR_ARG;
  int64_t Np = 1;
  for (int d = 0; d < dim; ++d) Np *= N + 1;
  hipStream_t st = (hipStream_t)stream;
  const unsigned g = (unsigned)grid_for(nnode * nstate);
  if (word_bytes == 8)
    nodes_kernel<double, PACK><<<g, 256, 0, st>>>((double*)Q, (double*)buf, nodes, nnode, Np,
                                                  nstate, nelem_total);
  else if (word_bytes == 4)
    nodes_kernel<float, PACK><<<g, 256, 0, st>>>((float*)Q, (float*)buf, nodes, nnode, Np,
                                                 nstate, nelem_total);
  else
    return CANARY_ERR_ARG;
  return status();
}
}  // namespace

extern "C" int canary_pack_nodes(int dim, int N, int nstate, void* buf, const void* Q,
                                 const int64_t* nodes, int64_t nnode, int64_t nelem_total,
                                 int word_bytes, void* stream) {
  return nodes_entry<true>(dim, N, nstate, const_cast<void*>(Q), buf, nodes, nnode, nelem_total,
                           word_bytes, stream);
}

extern "C" int canary_unpack_nodes(int dim, int N, int nstate, void* Q, const void* buf,
                                   const int64_t* nodes, int64_t nnode, int64_t nelem_total,
                                   int word_bytes, void* stream) {
  return nodes_entry<false>(dim, N, nstate, Q, const_cast<void*>(buf), nodes, nnode, nelem_total,
                            word_bytes, stream);
}

extern "C" int canary_l2energysquared_device(int Np, int nstate, int nvgeo, int iMJ,
                                             const double* Q, const double* vgeo, int64_t nelem,
                                             double* work, double* out, void* stream) {
  if (Np < 1 || nstate < 1 || nvgeo < 1 || iMJ < 0 || iMJ >= nvgeo || nelem < 0)
    return CANARY_ERR_ARG;
  if (!work || !out || (nelem > 0 && (!Q || !vgeo))) return CANARY_ERR_ARG;
  hipStream_t st = (hipStream_t)stream;
  energy_kernel<<<RED_BLOCKS, RED_THREADS, 0, st>>>(Q, vgeo, Np, nstate, nvgeo, iMJ, nelem, work);
  finish_kernel<false><<<1, RED_THREADS, 0, st>>>(work, out);
  return status();
}

extern "C" int canary_cfl_device(int eqn, int dim, int N, const double* vgeo, const double* Q,
                                 int64_t nelem, const canary_phys* phys, double* work,
                                 double* out, void* stream) {
  const bool ok = (eqn == CANARY_EULER && (dim == 2 || dim == 3)) || (eqn == CANARY_SWE && dim == 2);
  if (!ok || N < 1 || nelem < 0 || !work || !out || (nelem > 0 && (!Q || !vgeo)))
    return CANARY_ERR_ARG;
  canary_phys ph;
  if (phys)
    ph = *phys;
  else
    canary_phys_default(&ph);
  int64_t Np = 1;
  for (int d = 0; d < dim; ++d) Np *= N + 1;
  hipStream_t st = (hipStream_t)stream;
  cfl_kernel<<<RED_BLOCKS, RED_THREADS, 0, st>>>(eqn, dim, Np, vgeo, Q, nelem, ph, work);
  finish_kernel<true><<<1, RED_THREADS, 0, st>>>(work, out);
  return status();
}

// canary_cfl_device with the partials in the per-device static array: no
// allocation and no workspace argument.  The reduction writes all
// RED_BLOCKS partials on every call before the finishing kernel reads them,
// so nothing carries over between calls and nothing needs clearing; the
// partial count does not depend on nelem, so no input size exceeds the
// array.  Calls that overlap on one device share the array and race.
extern "C" int canary_cfl_device_noalloc(int eqn, int dim, int N, const double* vgeo,
                                         const double* Q, int64_t nelem,
                                         const canary_phys* phys, double* out, void* stream) {
  const bool ok = (eqn == CANARY_EULER && (dim == 2 || dim == 3)) || (eqn == CANARY_SWE && dim == 2);
  if (!ok || N < 1 || nelem < 0 || !out || (nelem > 0 && (!Q || !vgeo))) return CANARY_ERR_ARG;
  double* work = nullptr;
  const int rc = scratch_ptr(&work);
  if (rc != 0) return rc;
  return canary_cfl_device(eqn, dim, N, vgeo, Q, nelem, phys, work, out, stream);
}

extern "C" int canary_courant_device(int dim, int N, const double* vgeo, const double* Q,
                                     int64_t nelem, const canary_phys* phys, double courant,
                                     double* work, double* out, void* stream) {
  if ((dim != 2 && dim != 3) || N < 1 || nelem < 0 || !work || !out ||
      (nelem > 0 && (!Q || !vgeo)))
    return CANARY_ERR_ARG;
  canary_phys ph;
  if (phys)
    ph = *phys;
  else
    canary_phys_default(&ph);
  int64_t Np = 1;
  for (int d = 0; d < dim; ++d) Np *= N + 1;
  hipStream_t st = (hipStream_t)stream;
  courant_kernel<<<RED_BLOCKS, RED_THREADS, 0, st>>>(dim, N, Np, vgeo, Q, nelem, ph, courant,
                                                     work);
  finish_kernel<true><<<1, RED_THREADS, 0, st>>>(work, out);
  return status();
}

extern "C" int canary_k_aux_status(int64_t* info, int clear) {
  unsigned long long h[4] = {0, 0, 0, 0};
  if (hipMemcpyFromSymbol(h, HIP_SYMBOL(canary::aux::g_aux_status), sizeof(h), 0,
                          hipMemcpyDeviceToHost) != hipSuccess)
    return CANARY_ERR_LAUNCH;
  if (clear && h[0] != 0) {
    unsigned long long z[4] = {0, 0, 0, 0};
    if (hipMemcpyToSymbol(HIP_SYMBOL(canary::aux::g_aux_status), z, sizeof(z), 0,
                          hipMemcpyHostToDevice) != hipSuccess)
      return CANARY_ERR_LAUNCH;
  }
  if (info)
    for (int i = 0; i < 4; ++i) info[i] = (int64_t)h[i];
  return (int)h[0];
}
