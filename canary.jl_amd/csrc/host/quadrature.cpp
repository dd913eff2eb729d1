// 1-D operators of the nodal DG discretisation: Gauss-Legendre(-Lobatto)
// points and weights, barycentric weights, the spectral differentiation
// matrix and the barycentric interpolation matrix.
//
// Reference behaviour followed (steps and the endpoint fix-up, not code):
//   src/GaussQuadrature/quadrature.jl:126-131  legendre / legendre_coefs
//   src/GaussQuadrature/quadrature.jl:175-214  custom_gauss_rule (Lobatto
//       modification of the Jacobi matrix, weights w_i = (b_1 z_i)^2, sort,
//       and only the LEFT end point forced exactly to lo -- the `elseif`
//       at :208-212 means `both` never forces +1)
//   src/GaussQuadrature/quadrature.jl:216-231  solve (last component of
//       (J - shift I)^{-1} e_n by forward elimination)
//   src/GaussQuadrature/quadrature.jl:235-326  implicit QL with Wilkinson
//       shift, tracking only the first eigenvector components
//   src/operators.jl:30-33, 56-69, 84-104, 118-137
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <numeric>
#include <vector>

#include "../../../include/canary_dg.h"

namespace {

// Two-argument hypot in the scaled form ax*sqrt(1+(ay/ax)^2) (Julia 1.0's
// generic method); ulp-level agreement with Julia is not pinned by any test.
double hypot2(double x, double y) {
  double ax = std::fabs(x), ay = std::fabs(y);
  if (ax < ay) std::swap(ax, ay);
  double r = (ax == 0.0) ? ay / 1.0 : ay / ax;
  double rr = ax * std::sqrt(1.0 + r * r);
  if (std::isnan(r)) {
    if (std::isinf(ax) || std::isinf(ay)) return INFINITY;
    return r;
  }
  return rr;
}

double solve_last(int n, double shift, const std::vector<double>& a,
                  const std::vector<double>& b) {
  double t = a[0] - shift;
  for (int i = 1; i < n - 1; ++i) t = a[i] - shift - (b[i] * b[i]) / t;
  return 1.0 / t;
}

// Eigenvalues of the symmetric tridiagonal (d, e) and first components z of
// the normalised eigenvectors.  e[1..n-1] hold the off-diagonals, e has n+1
// entries.  Returns 0, or CANARY_ERR_NOCONVERGENCE after maxits sweeps.
int tridiag_ql(std::vector<double>& d, std::vector<double>& e,
               std::vector<double>& z, int maxits) {
  const int n = (int)z.size();
  const double eps = std::ldexp(1.0, -52);
  z[0] = 1.0;
  for (int i = 1; i < n; ++i) z[i] = 0.0;
  e[n] = 0.0;
  if (n == 1) return 0;
  for (int l = 0; l < n; ++l) {
    for (int it = 1; it <= maxits; ++it) {
      int m = n - 1;
      for (int i = l; i < n - 1; ++i) {
        if (std::fabs(e[i + 1]) <= eps * (std::fabs(d[i]) + std::fabs(d[i + 1]))) {
          m = i;
          break;
        }
      }
      double p = d[l];
      if (m == l) break;
      if (it == maxits) return CANARY_ERR_NOCONVERGENCE;
      double g = (d[l + 1] - p) / (2 * e[l + 1]);
      double r = hypot2(g, 1.0);
      g = d[m] - p + e[l + 1] / (g + std::copysign(r, g));
      double s = 1.0, c = 1.0;
      p = 0.0;
      for (int i = m - 1; i >= l; --i) {
        double f = s * e[i + 1];
        double b = c * e[i + 1];
        if (std::fabs(f) < std::fabs(g)) {
          s = f / g;
          r = hypot2(s, 1.0);
          e[i + 2] = g * r;
          c = 1.0 / r;
          s *= c;
        } else {
          c = g / f;
          r = hypot2(c, 1.0);
          e[i + 2] = f * r;
          s = 1.0 / r;
          c *= s;
        }
        g = d[i + 1] - p;
        r = (d[i] - g) * s + 2 * c * b;
        p = s * r;
        d[i + 1] = g + p;
        g = c * r - b;
        f = z[i + 1];
        z[i + 1] = s * z[i] + c * f;
        z[i] = c * z[i] - s * f;
      }
      d[l] -= p;
      e[l + 1] = g;
      e[m + 1] = 0.0;
    }
  }
  return 0;
}

}  // namespace

extern "C" int canary_gauss_rule(int n, int endpt, int maxits, double* x, double* w) {
  if (n < 1 || !x || !w) return CANARY_ERR_ARG;
  const double lo = -1.0, hi = 1.0;
  std::vector<double> a(n, 0.0), b(n + 1, 0.0), z(n, 0.0);
  b[0] = std::sqrt(2.0);
  for (int k = 2; k <= n + 1; ++k)
    b[k - 1] = (double)(k - 1) / std::sqrt((double)((2 * k - 1) * (2 * k - 3)));
  if (endpt == CANARY_ENDPT_LEFT) {
    if (n == 1) a[0] = lo;
    else a[n - 1] = solve_last(n, lo, a, b) * (b[n - 1] * b[n - 1]) + lo;
  } else if (endpt == CANARY_ENDPT_RIGHT) {
    if (n == 1) a[0] = hi;
    else a[n - 1] = solve_last(n, hi, a, b) * (b[n - 1] * b[n - 1]) + hi;
  } else if (endpt == CANARY_ENDPT_BOTH) {
    if (n == 1) return CANARY_ERR_LOBATTO_N1;
    double g = solve_last(n, lo, a, b);
    double t1 = (hi - lo) / (g - solve_last(n, hi, a, b));
    b[n - 1] = std::sqrt(t1);
    a[n - 1] = lo + g * t1;
  }
  int rc = tridiag_ql(a, b, z, maxits > 0 ? maxits : 30);
  if (rc) return rc;
  std::vector<double> wt(n);
  for (int i = 0; i < n; ++i) {
    double q = b[0] * z[i];
    wt[i] = q * q;
  }
  std::vector<int> idx(n);
  std::iota(idx.begin(), idx.end(), 0);
  std::stable_sort(idx.begin(), idx.end(), [&](int i, int j) { return a[i] < a[j]; });
  if (endpt == CANARY_ENDPT_LEFT || endpt == CANARY_ENDPT_BOTH) a[idx[0]] = lo;
  else if (endpt == CANARY_ENDPT_RIGHT) a[idx[n - 1]] = hi;
  for (int i = 0; i < n; ++i) {
    x[i] = a[idx[i]];
    w[i] = wt[idx[i]];
  }
  return 0;
}

extern "C" int canary_lglpoints(int N, double* r, double* w) {
  if (N < 1) return CANARY_ERR_ARG;
  return canary_gauss_rule(N + 1, CANARY_ENDPT_BOTH, 30, r, w);
}

extern "C" int canary_lgpoints(int N, double* r, double* w) {
  if (N < 1) return CANARY_ERR_ARG;
  return canary_gauss_rule(N + 1, CANARY_ENDPT_NEITHER, 30, r, w);
}

extern "C" int canary_baryweights(int n, const double* r, double* wb) {
  if (n < 1) return CANARY_ERR_ARG;
  for (int j = 0; j < n; ++j) {
    double p = 1.0;
    for (int i = 0; i < n; ++i)
      if (i != j) p = p * (r[j] - r[i]);
    wb[j] = 1.0 / p;
  }
  return 0;
}

// D is n x n, column-major: D[j + n*k] = l_k'(r_j).
extern "C" int canary_spectralderivative(int n, const double* r, double* D) {
  if (n < 1) return CANARY_ERR_ARG;
  std::vector<double> wb(n);
  canary_baryweights(n, r, wb.data());
  for (int k = 0; k < n; ++k) {
    for (int j = 0; j < n; ++j) {
      double v = 0.0;
      if (k == j) {
        for (int l = 0; l < n; ++l)
          if (l != k) v = v + 1.0 / (r[k] - r[l]);
      } else {
        v = (wb[k] / wb[j]) / (r[j] - r[k]);
      }
      D[j + (size_t)n * k] = v;
    }
  }
  return 0;
}

// I is ndst x nsrc, column-major.
extern "C" int canary_interpolationmatrix(int nsrc, const double* rsrc, int ndst,
                                          const double* rdst, double* I) {
  if (nsrc < 1 || ndst < 1) return CANARY_ERR_ARG;
  std::vector<double> wb(nsrc);
  canary_baryweights(nsrc, rsrc, wb.data());
  for (int k = 0; k < ndst; ++k) {
    for (int j = 0; j < nsrc; ++j) {
      double v = wb[j] / (rdst[k] - rsrc[j]);
      I[k + (size_t)ndst * j] = v;
      if (!std::isfinite(v)) {
        for (int jj = 0; jj < nsrc; ++jj) I[k + (size_t)ndst * jj] = 0.0;
        I[k + (size_t)ndst * j] = 1.0;
        break;
      }
    }
    double d = 0.0;
    for (int j = 0; j < nsrc; ++j) d += I[k + (size_t)ndst * j];
    for (int j = 0; j < nsrc; ++j) I[k + (size_t)ndst * j] = I[k + (size_t)ndst * j] / d;
  }
  return 0;
}
