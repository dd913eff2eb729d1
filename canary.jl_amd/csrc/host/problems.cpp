// Initial conditions, time-step estimate and energy diagnostic on the host.
//
// Reference behaviour followed:
//   examples/3d_kernels/euler3d_set2c.jl:1189-1218  rising thermal bubble
//     (theta = theta_ref + 0.5*theta_c*(1+cos(pi r/rc)) inside r <= rc,
//      Exner pi = 1 - g z/(c_p theta), rho = p0/(R theta) pi^(c_v/R)),
//     then convert_set2nc_to_set2c :895-915 (U = rho u, E = rho theta)
//   examples/2d_kernels/euler2d_set2c.jl:1056-1083  same in 2D (y up)
//   euler3d_set2c.jl:67-97 / euler2d_set2c.jl:66-93  cfl (per-node
//     2 rho / max_d(|U . grad xi_d| + rho c))
//   euler3d_set2c.jl:738-750  L2energysquared
//   examples/old_fxg_files/shallow_water.jl:43-67  SWE cfl
// Additions with no reference counterpart (documented where used):
//   * a spherical bubble option and a seeded theta jitter hashed from the
//     node coordinates (so shared face nodes get the same value and the
//     field does not depend on element order);
//   * the 2D isentropic vortex (exact solution), written in set2c
//     variables with R = 1, p0 = 1, c_p/c_v = gamma, for which rho*theta
//     equals rho;
//   * Gaussian-bump shallow-water initial data over still depth b.
#include <cmath>
#include <cstdint>
#include <cstring>

#include "../../../include/canary_dg.h"
#include "../hip/crpow.h"

namespace {

uint64_t splitmix64(uint64_t x) {
  x += 0x9E3779B97F4A7C15ull;
  x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
  x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
  return x ^ (x >> 31);
}

uint64_t dbits(double v) {
  if (v == 0.0) v = 0.0;  // -0 and +0 hash alike
  uint64_t b;
  std::memcpy(&b, &v, 8);
  return b;
}

}  // namespace

extern "C" void canary_phys_default(canary_phys* p) {
  p->gamma = 14.0 / 10.0;
  p->p0 = 100000.0;
  p->R_gas = 28717.0 / 100.0;
  p->c_p = 100467.0 / 100.0;
  p->c_v = 7175.0 / 10.0;
  p->gravity = 10.0;
  p->swe_gravity = 10.0;
  p->delta_nl = 1.0;
}

// Uniform value in [-1, 1) from (seed, coordinates).
extern "C" double canary_hash_jitter(uint64_t seed, double x, double y, double z) {
  uint64_t h = splitmix64(seed);
  h = splitmix64(h ^ dbits(x));
  h = splitmix64(h ^ dbits(y));
  h = splitmix64(h ^ dbits(z));
  double u = (double)(h >> 11) * (1.0 / 9007199254740992.0);  // [0, 1)
  return 2.0 * u - 1.0;
}

extern "C" int canary_ic_bubble(int dim, int N, int64_t nelem, const double* vgeo,
                                const canary_bubble* b, const canary_phys* ph, double* Q) {
  if ((dim != 2 && dim != 3) || N < 1 || nelem < 0 || !b || !ph) return CANARY_ERR_ARG;
  const int Nq = N + 1;
  const int Np = dim == 3 ? Nq * Nq * Nq : Nq * Nq;
  const int ns = dim == 3 ? 5 : 4, nv = dim == 3 ? 14 : 8, ix = dim == 3 ? 11 : 6;
  const double p0 = ph->p0, R_gas = ph->R_gas, c_p = ph->c_p, c_v = ph->c_v, g = ph->gravity;
#pragma omp parallel for schedule(static)
  for (int64_t e = 0; e < nelem; ++e)
    for (int n = 0; n < Np; ++n) {
      const double* vg = vgeo + (size_t)Np * nv * e;
      double x = vg[n + Np * ix], y = vg[n + Np * (ix + 1)];
      double z = dim == 3 ? vg[n + Np * (ix + 2)] : 0.0;
      double vert = dim == 3 ? z : y;  // vertical coordinate
      double r;
      if (dim == 3 && b->spherical)
        r = std::sqrt((x - b->xc) * (x - b->xc) + (y - b->yc) * (y - b->yc) +
                      (z - b->zc) * (z - b->zc));
      else
        r = std::sqrt((x - b->xc) * (x - b->xc) + (vert - b->zc) * (vert - b->zc));
      double dth = 0.0;
      if (r <= b->rc) dth = 0.5 * b->dtheta * (1.0 + std::cos(M_PI * r / b->rc));
      double th = b->theta0 + dth;
      if (b->jitter != 0.0) th = th + b->jitter * canary_hash_jitter(b->seed, x, y, z);
      double pik = 1.0 - g / (c_p * th) * vert;
      double c = c_v / R_gas;
      double rho = p0 / (R_gas * th) * canary::crpow::pow_cr(pik, c);
      double* q = Q + (size_t)Np * ns * e;
      // set2nc (rho, u=0, v=0, w=0, theta) converted to set2c
      if (dim == 3) {
        q[n + Np * 3] = rho;
        q[n + Np * 0] = rho * 0.0;
        q[n + Np * 1] = rho * 0.0;
        q[n + Np * 2] = rho * 0.0;
        q[n + Np * 4] = rho * th;
      } else {
        q[n + Np * 2] = rho;
        q[n + Np * 0] = rho * 0.0;
        q[n + Np * 1] = rho * 0.0;
        q[n + Np * 3] = rho * th;
      }
    }
  return 0;
}

// Exact isentropic vortex at time t (periodic box of size Lx x Ly, nearest
// image of the moving centre), in set2c variables (rho u, rho v, rho, rho theta)
// with p0 = R = 1, so rho*theta = p^(1/gamma) = rho.
extern "C" int canary_ic_vortex2d(int N, int64_t nelem, const double* vgeo,
                                  const canary_vortex* v, double t, double* Q) {
  if (N < 1 || nelem < 0 || !v) return CANARY_ERR_ARG;
  const int Nq = N + 1, Np = Nq * Nq, ns = 4, nv = 8;
  const double gam = v->gamma;
  double xc = v->x0 + v->u0 * t, yc = v->y0 + v->v0 * t;
#pragma omp parallel for schedule(static)
  for (int64_t e = 0; e < nelem; ++e)
    for (int n = 0; n < Np; ++n) {
      const double* vg = vgeo + (size_t)Np * nv * e;
      double x = vg[n + Np * 6], y = vg[n + Np * 7];
      double dx = x - xc, dy = y - yc;
      if (v->Lx > 0) dx -= v->Lx * std::nearbyint(dx / v->Lx);
      if (v->Ly > 0) dy -= v->Ly * std::nearbyint(dy / v->Ly);
      double r2 = dx * dx + dy * dy;
      double f = std::exp(0.5 * (1.0 - r2));
      double du = -v->beta / (2 * M_PI) * dy * f;
      double dv = v->beta / (2 * M_PI) * dx * f;
      double T = v->p_inf / v->rho_inf -
                 (gam - 1) * v->beta * v->beta / (8 * gam * M_PI * M_PI) * f * f;
      // isentropic: p / rho^gamma = p_inf / rho_inf^gamma
      double Kent = v->p_inf / canary::crpow::pow_cr(v->rho_inf, gam);
      double rho = canary::crpow::pow_cr(T / Kent, 1.0 / (gam - 1));
      double p = Kent * canary::crpow::pow_cr(rho, gam);
      double u = v->u0 + du, w = v->v0 + dv;
      double rth = canary::crpow::pow_cr(p, 1.0 / gam);  // p0 = R = 1
      double* q = Q + (size_t)Np * ns * e;
      q[n + Np * 0] = rho * u;
      q[n + Np * 1] = rho * w;
      q[n + Np * 2] = rho;
      q[n + Np * 3] = rth;
    }
  return 0;
}

// Shallow water: b = depth, h = sum of Gaussian bumps A*exp(-(r/w)^2) on a
// periodic box (nearest image), at rest.
extern "C" int canary_ic_swe_bumps(int N, int64_t nelem, const double* vgeo, double depth,
                                   int nbump, const double* bumps, double Lx, double Ly,
                                   double* Q) {
  if (N < 1 || nelem < 0 || nbump < 0) return CANARY_ERR_ARG;
  const int Nq = N + 1, Np = Nq * Nq, ns = 4, nv = 8;
#pragma omp parallel for schedule(static)
  for (int64_t e = 0; e < nelem; ++e)
    for (int n = 0; n < Np; ++n) {
      const double* vg = vgeo + (size_t)Np * nv * e;
      double x = vg[n + Np * 6], y = vg[n + Np * 7];
      double h = 0.0;
      for (int k = 0; k < nbump; ++k) {
        const double* bp = bumps + 4 * k;  // (xc, yc, amplitude, width)
        double dx = x - bp[0], dy = y - bp[1];
        if (Lx > 0) dx -= Lx * std::nearbyint(dx / Lx);
        if (Ly > 0) dy -= Ly * std::nearbyint(dy / Ly);
        double rr = (dx * dx + dy * dy) / (bp[3] * bp[3]);
        h += bp[2] * std::exp(-rr);
      }
      double* q = Q + (size_t)Np * ns * e;
      q[n + Np * 0] = 0.0;
      q[n + Np * 1] = 0.0;
      q[n + Np * 2] = h;
      q[n + Np * 3] = depth;
    }
  return 0;
}

// Reference time-step estimate (before the reference's `/ N^sqrt(2)`).
extern "C" double canary_cfl(int eqn, int dim, int N, const double* vgeo, const double* Q,
                             int64_t nelem, const canary_phys* ph) {
  const int Nq = N + 1;
  const int Np = dim == 3 ? Nq * Nq * Nq : Nq * Nq;
  double dt = 1.7976931348623157e308;
  for (int64_t e = 0; e < nelem; ++e)
    for (int n = 0; n < Np; ++n) {
      double loc;
      if (eqn == CANARY_EULER && dim == 3) {
        const double* q = Q + (size_t)Np * 5 * e;
        const double* vg = vgeo + (size_t)Np * 14 * e;
        double rho = q[n + Np * 3], U = q[n], V = q[n + Np], W = q[n + 2 * Np], E = q[n + Np * 4];
        double P = ph->p0 * canary::crpow::pow_cr(ph->R_gas * E / ph->p0, ph->c_p / ph->c_v);
        double c = rho * std::sqrt(ph->gamma * P / rho);
        double a = std::fabs(U * vg[n] + V * vg[n + 3 * Np] + W * vg[n + 6 * Np]) + c;
        double b = std::fabs(U * vg[n + Np] + V * vg[n + 4 * Np] + W * vg[n + 7 * Np]) + c;
        double d = std::fabs(U * vg[n + 2 * Np] + V * vg[n + 5 * Np] + W * vg[n + 8 * Np]) + c;
        double m = a > b ? a : b;
        m = m > d ? m : d;
        loc = 2 * rho / m;
      } else if (eqn == CANARY_EULER && dim == 2) {
        const double* q = Q + (size_t)Np * 4 * e;
        const double* vg = vgeo + (size_t)Np * 8 * e;
        double rho = q[n + Np * 2], U = q[n], V = q[n + Np], E = q[n + Np * 3];
        double P = ph->p0 * canary::crpow::pow_cr(ph->R_gas * E / ph->p0, ph->c_p / ph->c_v);
        double c = rho * std::sqrt(ph->gamma * P / rho);
        double a = std::fabs(U * vg[n] + V * vg[n + 2 * Np]) + c;
        double b = std::fabs(U * vg[n + Np] + V * vg[n + 3 * Np]) + c;
        loc = 2 * rho / (a > b ? a : b);
      } else if (eqn == CANARY_SWE && dim == 2) {
        const double* q = Q + (size_t)Np * 4 * e;
        const double* vg = vgeo + (size_t)Np * 8 * e;
        double h = q[n + 2 * Np], b = q[n + 3 * Np], U = q[n], V = q[n + Np];
        double H = h + b;
        double g = ph->swe_gravity, dnl = ph->delta_nl;
        double s = H * std::sqrt(g * H) * dnl;
        double a = std::fabs(U * vg[n] + V * vg[n + 2 * Np]) + s;
        double c = std::fabs(U * vg[n + Np] + V * vg[n + 3 * Np]) + s;
        loc = 2 * H / (a > c ? a : c);
      } else {
        return -1.0;
      }
      if (loc < dt) dt = loc;
    }
  return dt;
}

extern "C" double canary_l2energysquared(int Np, int nstate, int nvgeo, int iMJ,
                                         const double* Q, const double* vgeo, int64_t nelem) {
  double energy = 0.0;
  for (int64_t e = 0; e < nelem; ++e)
    for (int s = 0; s < nstate; ++s)
      for (int i = 0; i < Np; ++i) {
        double q = Q[i + (size_t)Np * (s + (size_t)nstate * e)];
        energy += vgeo[i + (size_t)Np * (iMJ + (size_t)nvgeo * e)] * (q * q);
      }
  return energy;
}

// x^y correctly rounded: the pow of the exact-order kernels, on the host.
extern "C" double canary_pow_cr(double x, double y) { return canary::crpow::pow_cr(x, y); }

// The same through the fixed-exponent table of y (the kernels' path for a
// gas's kappa), the table built on the host for this call.
extern "C" double canary_pow_cr_table(double x, double y) {
  static canary::crpow::KappaTable kt;
  static double built = -1.0;
  static bool ok = false;
#pragma omp critical(canary_pow_table)
  {
    if (built != y) {
      ok = canary::crpow::kappa_table_build(y, &kt);
      built = y;
    }
  }
  return canary::crpow::pow_cr_tab(x, y, ok ? &kt : nullptr);
}
