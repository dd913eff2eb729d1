// Device geometry: creategrid + computemetric + the vgeo / sgeo assembly of
// computegeometry (euler3d_set2c.jl:101-139, euler2d_set2c.jl:97-134;
// src/metric.jl:11-107 creategrid!, :120-400 computemetric!) on the GPU,
// one workgroup per element, one thread per volume node.  Every sum runs
// in the order of the host restatement (host/metric.cpp) and the file is
// compiled without FMA contraction, so the arrays are the host's bit for
// bit (division and sqrt are IEEE correctly rounded on both sides).
//
//   vgeo (Np, nvgeo, nelem): metric terms, MJ, MJI, x, y(, z)
//   sgeo (nsgeo, Nfp, nface, nelem): n_x .. n_d, sMJ, vMJI
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../../include/canary_dg.h"
#include "dg_kernels.h"

namespace canary {
namespace geo {

static __device__ unsigned long long g_geo_status[4];

__device__ __forceinline__ void geo_error(unsigned long long code, long long face, long long elem,
                                          long long value) {
  if (atomicCAS(&g_geo_status[0], 0ull, code) == 0ull) {
    g_geo_status[1] = (unsigned long long)face;
    g_geo_status[2] = (unsigned long long)elem;
    g_geo_status[3] = (unsigned long long)value;
  }
}

constexpr int MAXQ = 8;

struct Nodes {
  double r[MAXQ], w[MAXQ], D[MAXQ * MAXQ];  // LGL points, weights, D[a + Nq*b]
};

__device__ __forceinline__ double hypot2(double x, double y) {
  double ax = fabs(x), ay = fabs(y);
  if (ax < ay) {
    const double t = ax;
    ax = ay;
    ay = t;
  }
  const double r = (ax == 0.0) ? ay : ay / ax;
  if (isnan(r)) {
    if (isinf(ax) || isinf(ay)) return INFINITY;
    return r;
  }
  return ax * sqrt(1.0 + r * r);
}

__device__ __forceinline__ double hypot3(double x, double y, double z) {
  const double ax = fabs(x), ay = fabs(y), az = fabs(z);
  const double m = fmax(ax, fmax(ay, az));
  if (isinf(m)) return INFINITY;
  if (m == 0.0 || isnan(m)) return m;
  const double a = ax / m, b = ay / m, c = az / m;
  return m * sqrt(a * a + b * b + c * c);
}

template <int DIM, int NQ>
struct GCfg {
  static constexpr int Np = ipow(NQ, DIM), Nfp = ipow(NQ, DIM - 1), NF = 2 * DIM;
  static constexpr int NT = round64(Np);
  static constexpr int NV = DIM == 3 ? 14 : 8, NSG = DIM == 3 ? 5 : 4;
  static constexpr int NTMP = DIM == 3 ? 9 : 1;  // cross products (3D)
};

template <int DIM, int NQ>
__global__ __launch_bounds__((GCfg<DIM, NQ>::NT)) void geometry_kernel(
    int64_t nelem, const double* __restrict__ e2c, Nodes nd, double* __restrict__ vgeo,
    double* __restrict__ sgeo) {
  using G = GCfg<DIM, NQ>;
  constexpr int Np = G::Np, Nfp = G::Nfp, NF = G::NF, NT = G::NT, NV = G::NV, NSG = G::NSG;
  constexpr int NVERT = 1 << DIM;
  __shared__ double sX[DIM][Np];
  __shared__ double sT[G::NTMP][Np];
  __shared__ double sJ[Np];
  __shared__ double sM[DIM * DIM][Np];
  const int t = threadIdx.x;
  const int i = t % NQ, j = (t / NQ) % NQ, k = DIM == 3 ? t / (NQ * NQ) : 0;
  auto Dm = [&](int a, int b) { return nd.D[a + NQ * b]; };
  auto L = [&](int a, int b, int c) { return a + NQ * (b + NQ * c); };
  const double* r = nd.r;
  for (int64_t e = blockIdx.x; e < nelem; e += gridDim.x) {
    const double* c = e2c + (size_t)DIM * NVERT * e;
    // creategrid
    if (t < Np) {
#pragma unroll
      for (int n = 0; n < DIM; ++n) {
        auto C = [&](int v) { return c[n + DIM * v]; };
        double x;
        if (DIM == 2) {
          x = ((1 - r[i]) * (1 - r[j]) * C(0) + (1 + r[i]) * (1 - r[j]) * C(1) +
               (1 - r[i]) * (1 + r[j]) * C(2) + (1 + r[i]) * (1 + r[j]) * C(3)) / 4;
        } else {
          x = ((1 - r[i]) * (1 - r[j]) * (1 - r[k]) * C(0) +
               (1 + r[i]) * (1 - r[j]) * (1 - r[k]) * C(1) +
               (1 - r[i]) * (1 + r[j]) * (1 - r[k]) * C(2) +
               (1 + r[i]) * (1 + r[j]) * (1 - r[k]) * C(3) +
               (1 - r[i]) * (1 - r[j]) * (1 + r[k]) * C(4) +
               (1 + r[i]) * (1 - r[j]) * (1 + r[k]) * C(5) +
               (1 - r[i]) * (1 + r[j]) * (1 + r[k]) * C(6) +
               (1 + r[i]) * (1 + r[j]) * (1 + r[k]) * C(7)) / 8;
        }
        sX[n][t] = x;
      }
    }
    __syncthreads();
    double JI2 = 0.0;
    if (t < Np) {
      if constexpr (DIM == 2) {
        double xr = 0, xs = 0, yr = 0, ys = 0;
        for (int n = 0; n < NQ; ++n) {
          xr += Dm(i, n) * sX[0][L(n, j, 0)];
          xs += Dm(j, n) * sX[0][L(i, n, 0)];
          yr += Dm(i, n) * sX[1][L(n, j, 0)];
          ys += Dm(j, n) * sX[1][L(i, n, 0)];
        }
        const double J = xr * ys - yr * xs;
        sJ[t] = J;
        sM[0][t] = ys / J;   // xi_x
        sM[1][t] = -yr / J;  // eta_x
        sM[2][t] = -xs / J;  // xi_y
        sM[3][t] = xr / J;   // eta_y
      } else {
        double xr = 0, xs = 0, xt = 0, yr = 0, ys = 0, yt = 0, zr = 0, zs = 0, zt = 0;
        for (int n = 0; n < NQ; ++n) {
          xr += Dm(i, n) * sX[0][L(n, j, k)];
          xs += Dm(j, n) * sX[0][L(i, n, k)];
          xt += Dm(k, n) * sX[0][L(i, j, n)];
          yr += Dm(i, n) * sX[1][L(n, j, k)];
          ys += Dm(j, n) * sX[1][L(i, n, k)];
          yt += Dm(k, n) * sX[1][L(i, j, n)];
          zr += Dm(i, n) * sX[DIM - 1][L(n, j, k)];
          zs += Dm(j, n) * sX[DIM - 1][L(i, n, k)];
          zt += Dm(k, n) * sX[DIM - 1][L(i, j, n)];
        }
        const double J =
            (xr * (ys * zt - zs * yt) + yr * (zs * xt - xs * zt) + zr * (xs * yt - ys * xt));
        sJ[t] = J;
        JI2 = 1 / (2 * J);
        const double x = sX[0][t], y = sX[1][t], z = sX[DIM - 1][t];
        sT[0][t] = y * zr - z * yr;  // yzr
        sT[1][t] = y * zs - z * ys;  // yzs
        sT[2][t] = y * zt - z * yt;  // yzt
        sT[3][t] = z * xr - x * zr;  // zxr
        sT[4][t] = z * xs - x * zs;  // zxs
        sT[5][t] = z * xt - x * zt;  // zxt
        sT[6][t] = x * yr - y * xr;  // xyr
        sT[7][t] = x * ys - y * xs;  // xys
        sT[8][t] = x * yt - y * xt;  // xyt
      }
    }
    __syncthreads();
    if constexpr (DIM == 3) {
      if (t < Np) {
        const double* yzr = sT[0];
        const double* yzs = sT[1];
        const double* yzt = sT[2];
        const double* zxr = sT[3];
        const double* zxs = sT[4];
        const double* zxt = sT[5];
        const double* xyr = sT[6];
        const double* xys = sT[7];
        const double* xyt = sT[8];
        double a1 = 0, a2 = 0, a3 = 0, a4 = 0, a5 = 0, a6 = 0, a7 = 0, a8 = 0, a9 = 0;
        for (int n = 0; n < NQ; ++n) {
          a1 += Dm(j, n) * yzt[L(i, n, k)];
          a1 -= Dm(k, n) * yzs[L(i, j, n)];
          a2 += Dm(k, n) * yzr[L(i, j, n)];
          a2 -= Dm(i, n) * yzt[L(n, j, k)];
          a3 += Dm(i, n) * yzs[L(n, j, k)];
          a3 -= Dm(j, n) * yzr[L(i, n, k)];
          a4 += Dm(j, n) * zxt[L(i, n, k)];
          a4 -= Dm(k, n) * zxs[L(i, j, n)];
          a5 += Dm(k, n) * zxr[L(i, j, n)];
          a5 -= Dm(i, n) * zxt[L(n, j, k)];
          a6 += Dm(i, n) * zxs[L(n, j, k)];
          a6 -= Dm(j, n) * zxr[L(i, n, k)];
          a7 += Dm(j, n) * xyt[L(i, n, k)];
          a7 -= Dm(k, n) * xys[L(i, j, n)];
          a8 += Dm(k, n) * xyr[L(i, j, n)];
          a8 -= Dm(i, n) * xyt[L(n, j, k)];
          a9 += Dm(i, n) * xys[L(n, j, k)];
          a9 -= Dm(j, n) * xyr[L(i, n, k)];
        }
        sM[0][t] = a1 * JI2;
        sM[1][t] = a2 * JI2;
        sM[2][t] = a3 * JI2;
        sM[3][t] = a4 * JI2;
        sM[4][t] = a5 * JI2;
        sM[5][t] = a6 * JI2;
        sM[6][t] = a7 * JI2;
        sM[7][t] = a8 * JI2;
        sM[8][t] = a9 * JI2;
      }
    }
    __syncthreads();
    // vgeo: metric, MJ = M J with M = (w_k w_j) w_i, MJI, coordinates
    if (t < Np) {
      double* v = vgeo + (size_t)NV * Np * e + t;
#pragma unroll
      for (int f = 0; f < DIM * DIM; ++f) v[Np * f] = sM[f][t];
      const double Mw = DIM == 3 ? (nd.w[k] * nd.w[j]) * nd.w[i] : nd.w[j] * nd.w[i];
      const double mj = Mw * sJ[t];
      v[Np * (DIM * DIM)] = mj;
      v[Np * (DIM * DIM + 1)] = 1 / mj;
#pragma unroll
      for (int cc = 0; cc < DIM; ++cc) v[Np * (DIM * DIM + 2 + cc)] = sX[cc][t];
    }
    // sgeo: unit normals, sMJ = sM sJ (vMJI by sgeo_vmji_kernel)
    for (int it = t; it < NF * Nfp; it += NT) {
      const int f = it / Nfp, n = it - f * Nfp;
      const int a = n % NQ, b = n / NQ;  // face node (a, b); 2D: b = 0
      int vid;
      if (DIM == 2) {
        vid = (f >> 1) == 0 ? L((f & 1) * (NQ - 1), a, 0) : L(a, (f & 1) * (NQ - 1), 0);
      } else {
        vid = (f >> 1) == 0 ? L((f & 1) * (NQ - 1), a, b)
              : (f >> 1) == 1 ? L(a, (f & 1) * (NQ - 1), b)
                              : L(a, b, (f & 1) * (NQ - 1));
      }
      const int dir = f >> 1;  // xi, eta or zeta
      const double J = sJ[vid];
      const double mJ = (f & 1) ? J : -J;
      double nrm[3] = {0.0, 0.0, 0.0};
#pragma unroll
      for (int cc = 0; cc < DIM; ++cc) nrm[cc] = mJ * sM[dir + DIM * cc][vid];
      double sj;
      if (DIM == 2) {
        sj = hypot2(nrm[0], nrm[1]);
      } else {
        sj = hypot3(nrm[0], nrm[1], nrm[2]);
      }
      double* s = sgeo + (size_t)NSG * ((size_t)n + Nfp * (f + (size_t)NF * e));
#pragma unroll
      for (int cc = 0; cc < DIM; ++cc) s[cc] = nrm[cc] / sj;
      const double sMw = DIM == 3 ? nd.w[b] * nd.w[a] : nd.w[a];
      s[DIM] = sMw * sj;
    }
    __syncthreads();  // LDS is reused by the next element
  }
}

// sgeo[d+1] = vgeo MJI at vmapM (a gather over the whole array)
template <int DIM>
__global__ __launch_bounds__(256) void vmji_kernel(int64_t nfn_total, int64_t Np, int64_t nelem,
                                                   const int64_t* __restrict__ vmapM,
                                                   const double* __restrict__ vgeo,
                                                   double* __restrict__ sgeo) {
  constexpr int NV = DIM == 3 ? 14 : 8, NSG = DIM == 3 ? 5 : 4;
  for (int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; q < nfn_total;
       q += (int64_t)gridDim.x * blockDim.x) {
    const int64_t id = vmapM[q] - 1;
    double v = 0.0;
    if (id < 0 || id >= Np * nelem) {
      geo_error(5 /*CANARY_ERR_VMAP*/, q, -1, id + 1);
    } else {
      const int64_t ve = id / Np, vn = id - ve * Np;
      v = vgeo[vn + Np * ((DIM * DIM + 1) + (int64_t)NV * ve)];
    }
    sgeo[(size_t)NSG * q + DIM + 1] = v;
  }
}

template <int DIM, int NQ>
int run(int64_t nelem, const double* e2c, const int64_t* vmapM, double* vgeo, double* sgeo,
        hipStream_t st) {
  using G = GCfg<DIM, NQ>;
  Nodes nd{};
  int rc = canary_lglpoints(NQ - 1, nd.r, nd.w);
  if (rc) return rc;
  double Dc[NQ * NQ];
  rc = canary_spectralderivative(NQ, nd.r, Dc);
  if (rc) return rc;
  for (int a = 0; a < NQ * NQ; ++a) nd.D[a] = Dc[a];
  int64_t grid = nelem < 65536 * 4 ? nelem : 65536 * 4;
  geometry_kernel<DIM, NQ><<<dim3((unsigned)grid), dim3(G::NT), 0, st>>>(nelem, e2c, nd, vgeo,
                                                                        sgeo);
  if (hipGetLastError() != hipSuccess) return CANARY_ERR_LAUNCH;
  const int64_t nfn = (int64_t)G::Nfp * G::NF * nelem;
  int64_t blocks = (nfn + 255) / 256;
  if (blocks > 65536) blocks = 65536;
  vmji_kernel<DIM><<<dim3((unsigned)blocks), dim3(256), 0, st>>>(nfn, G::Np, nelem, vmapM, vgeo,
                                                                 sgeo);
  return hipGetLastError() == hipSuccess ? 0 : CANARY_ERR_LAUNCH;
}

}  // namespace geo
}  // namespace canary

extern "C" int canary_computegeometry_device(int d, int N, int64_t nelem,
                                             const double* elemtocoord, const int64_t* vmapM,
                                             double* vgeo, double* sgeo, void* stream) {
  if ((d != 2 && d != 3) || nelem < 0) return CANARY_ERR_ARG;
  if (nelem == 0) return 0;
  if (!elemtocoord || !vmapM || !vgeo || !sgeo) return CANARY_ERR_ARG;
  hipStream_t st = (hipStream_t)stream;
  using namespace canary::geo;
#define CANARY_GEO(NN)                                                            \
  case NN:                                                                        \
    return d == 3 ? run<3, NN + 1>(nelem, elemtocoord, vmapM, vgeo, sgeo, st)     \
                  : run<2, NN + 1>(nelem, elemtocoord, vmapM, vgeo, sgeo, st);
  switch (N) {
    CANARY_GEO(1)
    CANARY_GEO(2)
    CANARY_GEO(3)
    CANARY_GEO(4)
    CANARY_GEO(5)
    CANARY_GEO(6)
    CANARY_GEO(7)
  }
#undef CANARY_GEO
  return CANARY_ERR_UNSUPPORTED;
}

extern "C" int canary_k_geometry_status(int64_t* info, int clear) {
  unsigned long long h[4] = {0, 0, 0, 0};
  if (hipMemcpyFromSymbol(h, HIP_SYMBOL(canary::geo::g_geo_status), sizeof(h), 0,
                          hipMemcpyDeviceToHost) != hipSuccess)
    return CANARY_ERR_LAUNCH;
  if (clear && h[0] != 0) {
    unsigned long long z[4] = {0, 0, 0, 0};
    if (hipMemcpyToSymbol(HIP_SYMBOL(canary::geo::g_geo_status), z, sizeof(z), 0,
                          hipMemcpyHostToDevice) != hipSuccess)
      return CANARY_ERR_LAUNCH;
  }
  if (info)
    for (int i = 0; i < 4; ++i) info[i] = (int64_t)h[i];
  return (int)h[0];
}
