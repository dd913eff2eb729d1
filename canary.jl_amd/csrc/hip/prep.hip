// Device-side problem preparation: the initial states of host/problems.cpp
// (rising thermal bubble, euler3d_set2c.jl:1189-1218 / euler2d_set2c.jl:
// 1056-1083 plus the seeded jitter; the exact isentropic vortex) evaluated
// on the GPU from the node coordinates in vgeo, one lane per node, with the
// host's formulas term for term.  With the device geometry of geometry.hip
// a request is prepared without building vgeo on the host.  cos / pow / exp
// / sqrt are the device library's, so values can differ from the host ones
// in the last bit.
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>

#include "../../../include/canary_dg.h"
#include "crpow.h"

namespace canary {
namespace prep {

__device__ __forceinline__ uint64_t splitmix64(uint64_t x) {
  x += 0x9E3779B97F4A7C15ull;
  x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
  x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
  return x ^ (x >> 31);
}

__device__ __forceinline__ uint64_t dbits(double v) {
  if (v == 0.0) v = 0.0;  // -0 and +0 hash alike
  return (uint64_t)__double_as_longlong(v);
}

// canary_hash_jitter of the host: uniform in [-1, 1) from (seed, x, y, z)
__device__ __forceinline__ double hash_jitter(uint64_t seed, double x, double y, double z) {
  uint64_t h = splitmix64(seed);
  h = splitmix64(h ^ dbits(x));
  h = splitmix64(h ^ dbits(y));
  h = splitmix64(h ^ dbits(z));
  const double u = (double)(h >> 11) * (1.0 / 9007199254740992.0);
  return 2.0 * u - 1.0;
}

__global__ __launch_bounds__(256) void bubble_kernel(int dim, int64_t Np, int64_t nelem,
                                                     const double* __restrict__ vgeo,
                                                     canary_bubble b, canary_phys ph,
                                                     double* __restrict__ Q) {
  const int ns = dim == 3 ? 5 : 4, nv = dim == 3 ? 14 : 8, ix = dim == 3 ? 11 : 6;
  const double p0 = ph.p0, R_gas = ph.R_gas, c_p = ph.c_p, c_v = ph.c_v, g = ph.gravity;
  for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < Np * nelem;
       t += (int64_t)gridDim.x * blockDim.x) {
    const int64_t e = t / Np, n = t - e * Np;
    const double* vg = vgeo + (size_t)Np * nv * e;
    const double x = vg[n + Np * ix], y = vg[n + Np * (ix + 1)];
    const double z = dim == 3 ? vg[n + Np * (ix + 2)] : 0.0;
    const double vert = dim == 3 ? z : y;
    double r;
    if (dim == 3 && b.spherical)
      r = sqrt((x - b.xc) * (x - b.xc) + (y - b.yc) * (y - b.yc) + (z - b.zc) * (z - b.zc));
    else
      r = sqrt((x - b.xc) * (x - b.xc) + (vert - b.zc) * (vert - b.zc));
    double dth = 0.0;
    if (r <= b.rc) dth = 0.5 * b.dtheta * (1.0 + cos(M_PI * r / b.rc));
    double th = b.theta0 + dth;
    if (b.jitter != 0.0) th = th + b.jitter * hash_jitter(b.seed, x, y, z);
    const double pik = 1.0 - g / (c_p * th) * vert;
    const double c = c_v / R_gas;
    const double rho = p0 / (R_gas * th) * canary::crpow::pow_cr(pik, c);
    double* q = Q + (size_t)Np * ns * e;
    const int RHO = dim, E = dim + 1;
    q[n + Np * RHO] = rho;
    for (int s = 0; s < dim; ++s) q[n + Np * s] = rho * 0.0;
    q[n + Np * E] = rho * th;
  }
}

__global__ __launch_bounds__(256) void vortex_kernel(int64_t Np, int64_t nelem,
                                                     const double* __restrict__ vgeo,
                                                     canary_vortex v, double t,
                                                     double* __restrict__ Q) {
  const int ns = 4, nv = 8;
  const double gam = v.gamma;
  const double xc = v.x0 + v.u0 * t, yc = v.y0 + v.v0 * t;
  for (int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; g < Np * nelem;
       g += (int64_t)gridDim.x * blockDim.x) {
    const int64_t e = g / Np, n = g - e * Np;
    const double* vg = vgeo + (size_t)Np * nv * e;
    const double x = vg[n + Np * 6], y = vg[n + Np * 7];
    double dx = x - xc, dy = y - yc;
    if (v.Lx > 0) dx -= v.Lx * rint(dx / v.Lx);
    if (v.Ly > 0) dy -= v.Ly * rint(dy / v.Ly);
    const double r2 = dx * dx + dy * dy;
    const double f = exp(0.5 * (1.0 - r2));
    const double du = -v.beta / (2 * M_PI) * dy * f;
    const double dv = v.beta / (2 * M_PI) * dx * f;
    const double T = v.p_inf / v.rho_inf - (gam - 1) * v.beta * v.beta / (8 * gam * M_PI * M_PI) * f * f;
    const double Kent = v.p_inf / canary::crpow::pow_cr(v.rho_inf, gam);
    const double rho = canary::crpow::pow_cr(T / Kent, 1.0 / (gam - 1));
    const double p = Kent * canary::crpow::pow_cr(rho, gam);
    const double u = v.u0 + du, w = v.v0 + dv;
    const double rth = canary::crpow::pow_cr(p, 1.0 / gam);
    double* q = Q + (size_t)Np * ns * e;
    q[n + Np * 0] = rho * u;
    q[n + Np * 1] = rho * w;
    q[n + Np * 2] = rho;
    q[n + Np * 3] = rth;
  }
}

int64_t blocks_for(int64_t work) {
  int64_t b = (work + 255) / 256;
  return b < 1 ? 1 : (b > 65536 ? 65536 : b);
}

}  // namespace prep
}  // namespace canary

extern "C" int canary_ic_bubble_device(int dim, int N, int64_t nelem, const double* vgeo,
                                       const canary_bubble* b, const canary_phys* ph, double* Q,
                                       void* stream) {
  if ((dim != 2 && dim != 3) || N < 1 || nelem < 0 || !b || !ph) return CANARY_ERR_ARG;
  if (nelem == 0) return 0;
  if (!vgeo || !Q) return CANARY_ERR_ARG;
  int64_t Np = 1;
  for (int d = 0; d < dim; ++d) Np *= N + 1;
  canary::prep::bubble_kernel<<<(unsigned)canary::prep::blocks_for(Np * nelem), 256, 0,
                                (hipStream_t)stream>>>(dim, Np, nelem, vgeo, *b, *ph, Q);
  return hipGetLastError() == hipSuccess ? 0 : CANARY_ERR_LAUNCH;
}

extern "C" int canary_ic_vortex2d_device(int N, int64_t nelem, const double* vgeo,
                                         const canary_vortex* v, double t, double* Q,
                                         void* stream) {
  if (N < 1 || nelem < 0 || !v) return CANARY_ERR_ARG;
  if (nelem == 0) return 0;
  if (!vgeo || !Q) return CANARY_ERR_ARG;
  const int64_t Np = (int64_t)(N + 1) * (N + 1);
  canary::prep::vortex_kernel<<<(unsigned)canary::prep::blocks_for(Np * nelem), 256, 0,
                                (hipStream_t)stream>>>(Np, nelem, vgeo, *v, t, Q);
  return hipGetLastError() == hipSuccess ? 0 : CANARY_ERR_LAUNCH;
}
