// Element grids, metric terms and the vgeo/sgeo geometry arrays.
//
// Reference behaviour followed (evaluation order kept: left-to-right sums
// and products, no contraction -- this file is built with -ffp-contract=off):
//   src/metric.jl:11-107    creategrid! 1D/2D/3D (linear/bi/trilinear blend)
//   src/metric.jl:120-141   computemetric! 1D
//   src/metric.jl:155-208   computemetric! 2D
//   src/metric.jl:230-400   computemetric! 3D, curl-invariant form
//                           (Kopriva 2006, eq. 37), normals n = +-J grad(xi)
//                           normalised by sJ = |.|
//   examples/3d_kernels/euler3d_set2c.jl:101-139  computegeometry (3D)
//   examples/2d_kernels/euler2d_set2c.jl:97-134   computegeometry (2D)
//     M = kron(1, w, w[, w]) so M[i,j,k] = (w_k*w_j)*w_i; MJ = M.*J;
//     MJI = 1 ./ MJ; vMJI = MJI[vmapM]; sMJ = sM .* sJ.
#include <cmath>
#include <cstdint>
#include <vector>

#include "../../../include/canary_dg.h"

namespace {

double hypot2(double x, double y) {
  double ax = std::fabs(x), ay = std::fabs(y);
  if (ax < ay) std::swap(ax, ay);
  double r = (ax == 0.0) ? ay : ay / ax;
  if (std::isnan(r)) {
    if (std::isinf(ax) || std::isinf(ay)) return INFINITY;
    return r;
  }
  return ax * std::sqrt(1.0 + r * r);
}

// Scaled three-argument hypot: m * sqrt((x/m)^2 + (y/m)^2 + (z/m)^2).
// Its last-ulp agreement with Julia 1.0's hypot(x, y, z) is not pinned.
double hypot3(double x, double y, double z) {
  double ax = std::fabs(x), ay = std::fabs(y), az = std::fabs(z);
  double m = std::fmax(ax, std::fmax(ay, az));
  if (std::isinf(m)) return INFINITY;
  if (m == 0.0 || std::isnan(m)) return m;
  double a = ax / m, b = ay / m, c = az / m;
  return m * std::sqrt(a * a + b * b + c * c);
}

}  // namespace

extern "C" int canary_creategrid(int d, int Nq, int64_t nelem, const double* e2c,
                                 const double* r, double* X) {
  if (d < 1 || d > 3 || Nq < 1 || nelem < 0) return CANARY_ERR_ARG;
  const int nvert = 1 << d;
  size_t Np = 1;
  for (int j = 0; j < d; ++j) Np *= Nq;
  const size_t nn = Np * nelem;
  for (int n = 0; n < d; ++n) {
    double* f = X + n * nn;
#pragma omp parallel for schedule(static)
    for (int64_t e = 0; e < nelem; ++e) {
      const double* c = e2c + (size_t)d * nvert * e;
      auto C = [&](int v) { return c[n + d * v]; };
      if (d == 1) {
        for (int i = 0; i < Nq; ++i)
          f[i + Nq * e] = ((1 - r[i]) * C(0) + (1 + r[i]) * C(1)) / 2;
      } else if (d == 2) {
        for (int j = 0; j < Nq; ++j)
          for (int i = 0; i < Nq; ++i)
            f[i + Nq * (j + (size_t)Nq * e)] =
                ((1 - r[i]) * (1 - r[j]) * C(0) + (1 + r[i]) * (1 - r[j]) * C(1) +
                 (1 - r[i]) * (1 + r[j]) * C(2) + (1 + r[i]) * (1 + r[j]) * C(3)) / 4;
      } else {
        for (int k = 0; k < Nq; ++k)
          for (int j = 0; j < Nq; ++j)
            for (int i = 0; i < Nq; ++i)
              f[i + Nq * (j + Nq * (k + (size_t)Nq * e))] =
                  ((1 - r[i]) * (1 - r[j]) * (1 - r[k]) * C(0) +
                   (1 + r[i]) * (1 - r[j]) * (1 - r[k]) * C(1) +
                   (1 - r[i]) * (1 + r[j]) * (1 - r[k]) * C(2) +
                   (1 + r[i]) * (1 + r[j]) * (1 - r[k]) * C(3) +
                   (1 - r[i]) * (1 - r[j]) * (1 + r[k]) * C(4) +
                   (1 + r[i]) * (1 - r[j]) * (1 + r[k]) * C(5) +
                   (1 - r[i]) * (1 + r[j]) * (1 + r[k]) * C(6) +
                   (1 + r[i]) * (1 + r[j]) * (1 + r[k]) * C(7)) / 8;
      }
    }
  }
  return 0;
}

extern "C" int canary_computemetric(int d, int Nq, int64_t nelem, const double* X,
                                    const double* D, double* vol, double* surf) {
  if (d < 1 || d > 3 || Nq < 1 || nelem < 0) return CANARY_ERR_ARG;
  auto Dm = [&](int a, int b) { return D[a + Nq * b]; };
  if (d == 1) {
    const size_t nn = (size_t)Nq * nelem;
    const double* x = X;
    double *J = vol, *xix = vol + nn;
    double *sJ = surf, *nx = surf + 2 * nelem;
    for (int64_t e = 0; e < nelem; ++e) {
      for (int i = 0; i < Nq; ++i) {
        double s = 0.0;  // (D * x[:, e])[i], a matrix-vector product
        for (int n = 0; n < Nq; ++n) s += Dm(i, n) * x[n + Nq * e];
        J[i + Nq * e] = s;
      }
      for (int i = 0; i < Nq; ++i) xix[i + Nq * e] = 1 / J[i + Nq * e];
      auto sgn = [](double v) { return v > 0 ? 1.0 : (v < 0 ? -1.0 : 0.0); };
      nx[0 + 2 * e] = -sgn(J[0 + Nq * e]);
      nx[1 + 2 * e] = sgn(J[Nq - 1 + Nq * e]);
      sJ[0 + 2 * e] = 1;
      sJ[1 + 2 * e] = 1;
    }
    return 0;
  }
  if (d == 2) {
    const size_t nn = (size_t)Nq * Nq * nelem;
    const size_t ns = (size_t)Nq * 4 * nelem;
    const double *x = X, *y = X + nn;
    double *J = vol, *xix = vol + nn, *etax = vol + 2 * nn, *xiy = vol + 3 * nn, *etay = vol + 4 * nn;
    double *sJ = surf, *nx = surf + ns, *ny = surf + 2 * ns;
#pragma omp parallel for schedule(static)
    for (int64_t e = 0; e < nelem; ++e) {
      auto V = [&](int i, int j) { return (size_t)i + Nq * (j + (size_t)Nq * e); };
      for (int j = 0; j < Nq; ++j)
        for (int i = 0; i < Nq; ++i) {
          double xr = 0, xs = 0, yr = 0, ys = 0;
          for (int n = 0; n < Nq; ++n) {
            xr += Dm(i, n) * x[V(n, j)];
            xs += Dm(j, n) * x[V(i, n)];
            yr += Dm(i, n) * y[V(n, j)];
            ys += Dm(j, n) * y[V(i, n)];
          }
          size_t p = V(i, j);
          J[p] = xr * ys - yr * xs;
          xix[p] = ys / J[p];
          etax[p] = -yr / J[p];
          xiy[p] = -xs / J[p];
          etay[p] = xr / J[p];
        }
      for (int i = 0; i < Nq; ++i) {
        auto S = [&](int f) { return (size_t)i + Nq * (f + 4 * (size_t)e); };
        nx[S(0)] = -J[V(0, i)] * xix[V(0, i)];
        ny[S(0)] = -J[V(0, i)] * xiy[V(0, i)];
        nx[S(1)] = J[V(Nq - 1, i)] * xix[V(Nq - 1, i)];
        ny[S(1)] = J[V(Nq - 1, i)] * xiy[V(Nq - 1, i)];
        nx[S(2)] = -J[V(i, 0)] * etax[V(i, 0)];
        ny[S(2)] = -J[V(i, 0)] * etay[V(i, 0)];
        nx[S(3)] = J[V(i, Nq - 1)] * etax[V(i, Nq - 1)];
        ny[S(3)] = J[V(i, Nq - 1)] * etay[V(i, Nq - 1)];
        for (int f = 0; f < 4; ++f) {
          sJ[S(f)] = hypot2(nx[S(f)], ny[S(f)]);
          nx[S(f)] /= sJ[S(f)];
          ny[S(f)] /= sJ[S(f)];
        }
      }
    }
    return 0;
  }
  // d == 3
  const size_t Np = (size_t)Nq * Nq * Nq, nn = Np * nelem;
  const size_t Nfp = (size_t)Nq * Nq, ns = Nfp * 6 * nelem;
  const double *x = X, *y = X + nn, *z = X + 2 * nn;
  double* J = vol;
  double *xix = vol + nn, *etax = vol + 2 * nn, *zetax = vol + 3 * nn;
  double *xiy = vol + 4 * nn, *etay = vol + 5 * nn, *zetay = vol + 6 * nn;
  double *xiz = vol + 7 * nn, *etaz = vol + 8 * nn, *zetaz = vol + 9 * nn;
  double *sJ = surf, *nx = surf + ns, *ny = surf + 2 * ns, *nz = surf + 3 * ns;
  auto L = [&](int i, int j, int k) { return (size_t)i + Nq * (j + (size_t)Nq * k); };
#pragma omp parallel
  {
  std::vector<double> JI2(Np), yzr(Np), yzs(Np), yzt(Np), zxr(Np), zxs(Np), zxt(Np), xyr(Np),
      xys(Np), xyt(Np);
#pragma omp for schedule(static)
  for (int64_t e = 0; e < nelem; ++e) {
    const size_t eo = Np * e;
    for (int k = 0; k < Nq; ++k)
      for (int j = 0; j < Nq; ++j)
        for (int i = 0; i < Nq; ++i) {
          double xr = 0, xs = 0, xt = 0, yr = 0, ys = 0, yt = 0, zr = 0, zs = 0, zt = 0;
          for (int n = 0; n < Nq; ++n) {
            xr += Dm(i, n) * x[eo + L(n, j, k)];
            xs += Dm(j, n) * x[eo + L(i, n, k)];
            xt += Dm(k, n) * x[eo + L(i, j, n)];
            yr += Dm(i, n) * y[eo + L(n, j, k)];
            ys += Dm(j, n) * y[eo + L(i, n, k)];
            yt += Dm(k, n) * y[eo + L(i, j, n)];
            zr += Dm(i, n) * z[eo + L(n, j, k)];
            zs += Dm(j, n) * z[eo + L(i, n, k)];
            zt += Dm(k, n) * z[eo + L(i, j, n)];
          }
          size_t p = L(i, j, k), g = eo + p;
          J[g] = (xr * (ys * zt - zs * yt) + yr * (zs * xt - xs * zt) + zr * (xs * yt - ys * xt));
          JI2[p] = 1 / (2 * J[g]);
          yzr[p] = y[g] * zr - z[g] * yr;
          yzs[p] = y[g] * zs - z[g] * ys;
          yzt[p] = y[g] * zt - z[g] * yt;
          zxr[p] = z[g] * xr - x[g] * zr;
          zxs[p] = z[g] * xs - x[g] * zs;
          zxt[p] = z[g] * xt - x[g] * zt;
          xyr[p] = x[g] * yr - y[g] * xr;
          xys[p] = x[g] * ys - y[g] * xs;
          xyt[p] = x[g] * yt - y[g] * xt;
        }
    for (int k = 0; k < Nq; ++k)
      for (int j = 0; j < Nq; ++j)
        for (int i = 0; i < Nq; ++i) {
          size_t g = eo + L(i, j, k);
          double a1 = 0, a2 = 0, a3 = 0, a4 = 0, a5 = 0, a6 = 0, a7 = 0, a8 = 0, a9 = 0;
          for (int n = 0; n < Nq; ++n) {
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
          double s = JI2[L(i, j, k)];
          xix[g] = a1 * s;
          etax[g] = a2 * s;
          zetax[g] = a3 * s;
          xiy[g] = a4 * s;
          etay[g] = a5 * s;
          zetay[g] = a6 * s;
          xiz[g] = a7 * s;
          etaz[g] = a8 * s;
          zetaz[g] = a9 * s;
        }
    for (int j = 0; j < Nq; ++j)
      for (int i = 0; i < Nq; ++i) {
        size_t ije = i + (size_t)Nq * j + e * 6 * Nfp;
        size_t A1 = eo + L(0, i, j), A2 = eo + L(Nq - 1, i, j);
        size_t B1 = eo + L(i, 0, j), B2 = eo + L(i, Nq - 1, j);
        size_t C1 = eo + L(i, j, 0), C2 = eo + L(i, j, Nq - 1);
        nx[ije + 0 * Nfp] = -J[A1] * xix[A1];
        nx[ije + 1 * Nfp] = J[A2] * xix[A2];
        nx[ije + 2 * Nfp] = -J[B1] * etax[B1];
        nx[ije + 3 * Nfp] = J[B2] * etax[B2];
        nx[ije + 4 * Nfp] = -J[C1] * zetax[C1];
        nx[ije + 5 * Nfp] = J[C2] * zetax[C2];
        ny[ije + 0 * Nfp] = -J[A1] * xiy[A1];
        ny[ije + 1 * Nfp] = J[A2] * xiy[A2];
        ny[ije + 2 * Nfp] = -J[B1] * etay[B1];
        ny[ije + 3 * Nfp] = J[B2] * etay[B2];
        ny[ije + 4 * Nfp] = -J[C1] * zetay[C1];
        ny[ije + 5 * Nfp] = J[C2] * zetay[C2];
        nz[ije + 0 * Nfp] = -J[A1] * xiz[A1];
        nz[ije + 1 * Nfp] = J[A2] * xiz[A2];
        nz[ije + 2 * Nfp] = -J[B1] * etaz[B1];
        nz[ije + 3 * Nfp] = J[B2] * etaz[B2];
        nz[ije + 4 * Nfp] = -J[C1] * zetaz[C1];
        nz[ije + 5 * Nfp] = J[C2] * zetaz[C2];
        for (int f = 0; f < 6; ++f) {
          size_t q = ije + f * Nfp;
          sJ[q] = hypot3(nx[q], ny[q], nz[q]);
          nx[q] /= sJ[q];
          ny[q] /= sJ[q];
          nz[q] /= sJ[q];
        }
      }
  }
  }  // omp parallel
  return 0;
}

extern "C" int canary_computegeometry(int d, int N, int64_t nelem, const double* e2c,
                                      const int64_t* vmapM, double* vgeo, double* sgeo) {
  if ((d != 2 && d != 3) || N < 1 || nelem < 0) return CANARY_ERR_ARG;
  const int Nq = N + 1, nface = 2 * d;
  size_t Np = 1;
  for (int j = 0; j < d; ++j) Np *= Nq;
  const size_t Nfp = Np / Nq, nn = Np * nelem, ns = Nfp * nface * nelem;
  std::vector<double> r(Nq), w(Nq), D((size_t)Nq * Nq);
  int rc = canary_lglpoints(N, r.data(), w.data());
  if (rc) return rc;
  canary_spectralderivative(Nq, r.data(), D.data());
  std::vector<double> X(d * nn), vol((1 + d * d) * nn), surf((1 + d) * ns);
  canary_creategrid(d, Nq, nelem, e2c, r.data(), X.data());
  canary_computemetric(d, Nq, nelem, X.data(), D.data(), vol.data(), surf.data());
  const int nvgeo = d == 3 ? 14 : 8, nsgeo = d == 3 ? 5 : 4;
  const int iMJ = d * d, iMJI = d * d + 1, ix = d * d + 2;
  std::vector<double> M(Np), sM(Nfp);
  for (size_t p = 0; p < Np; ++p) {
    int i = p % Nq, j = (p / Nq) % Nq, k = (int)(p / ((size_t)Nq * Nq));
    M[p] = d == 3 ? (w[k] * w[j]) * w[i] : w[j] * w[i];
  }
  for (size_t p = 0; p < Nfp; ++p) {
    int i = p % Nq, j = (int)(p / Nq);
    sM[p] = d == 3 ? w[j] * w[i] : w[i];
  }
#pragma omp parallel for schedule(static)
  for (int64_t e = 0; e < nelem; ++e)
    for (size_t p = 0; p < Np; ++p) {
      size_t g = p + Np * e;
      double* v = vgeo + p + Np * nvgeo * e;
      for (int f = 0; f < d * d; ++f) v[Np * f] = vol[(1 + f) * nn + g];
      double mj = M[p] * vol[g];
      v[Np * iMJ] = mj;
      v[Np * iMJI] = 1 / mj;
      for (int c = 0; c < d; ++c) v[Np * (ix + c)] = X[c * nn + g];
    }
#pragma omp parallel for schedule(static)
  for (int64_t e = 0; e < nelem; ++e)
    for (int f = 0; f < nface; ++f)
      for (size_t n = 0; n < Nfp; ++n) {
        size_t q = n + Nfp * (f + (size_t)nface * e);
        double* s = sgeo + (size_t)nsgeo * q;
        for (int c = 0; c < d; ++c) s[c] = surf[(1 + c) * ns + q];
        s[d] = sM[n] * surf[q];
        int64_t id = vmapM[q] - 1;
        size_t ve = (size_t)(id / (int64_t)Np), vn = (size_t)(id % (int64_t)Np);
        s[d + 1] = vgeo[vn + Np * (iMJI + (size_t)nvgeo * ve)];
      }
  return 0;
}
