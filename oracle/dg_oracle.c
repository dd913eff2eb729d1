/* dg_oracle.c -- CPU oracle for the nodal-DG hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load this library, and only as the
 * checker / the timed CPU baseline.  The product path never calls it.
 *
 * A plain-C restatement of the reference's Julia CPU kernels, following
 * their loop order and expression association operation for operation
 * (Julia emits no FMA contraction; build with -ffp-contract=off):
 *   euler3d_*  examples/3d_kernels/euler3d_set2c.jl
 *              volumerhs! :144-231, facerhs! :234-348, updatesolution! :352-358,
 *              cfl :67-97, L2energysquared :738-750, RK tables :772-788,
 *              stage loop :819-855
 *   euler2d_*  examples/2d_kernels/euler2d_set2c.jl
 *              volumerhs! :140-205, facerhs! :208-300, updatesolution! :304-310
 *   swe_*      examples/old_fxg_files/shallow_water.jl
 *              volumerhs! :121-173, fluxrhs! :178-263 (including its bP
 *              read from element eM, :202), updatesolution! :267-298
 *              (all four states, optional velocity freeze), cfl :43-67
 *   euler2dte_* examples/2d_kernels/euler2d_set3c.jl (total energy E)
 *              volumerhs! :214-279, fluxrhs! :282-378 (the pressure reads
 *              vgeo's y at the M and P nodes), courantnumber :121-168; the
 *              update is the device one, knl_updatesolution! :886-901 (no
 *              viscous term: the CPU update adds visc * rhs_gradQ, which at
 *              the example's visc = 0 can only turn a -0.0 rhs entry into
 *              +0.0, and the viscous LDG operators are not on this path)
 * Layouts are the reference's column-major ones (see include/canary_dg.h).
 * The per-face-node heap slice sgeo[:, n, f, e] of the Julia face kernel
 * (:251) is not reproduced; it does not change any value.
 * `nthreads` > 1 splits the element loops with OpenMP (each element only
 * writes its own rhs/Q entries, so the result is identical).
 *
 * How each function is pinned (the reference's own tests hold no vectors
 * for these kernels; Julia cannot run here):
 *   volumerhs / facerhs (all three equation sets): equal to an independent
 *     dense strong-form assembly in NumPy on tiny meshes, free-stream
 *     preservation, discrete conservation, buoyancy-only residual at rest,
 *     wall mirror symmetry, bad-bc error (tests/test_oracle.py);
 *   updatesolution: the formula on random data, the SWE velocity freeze;
 *   RK tables and stage loop: the stability polynomial of LSRK54 on
 *     y' = z y (1, 1, 1/2, 1/6, 1/24, 1/200) and h-convergence of the
 *     isentropic vortex at order >= N + 1/2 (tests/test_convergence.py);
 *   cfl, L2energysquared: closed forms on uniform states.
 * pow: the reference calls Julia's libm pow, whose last bit no reference
 * test fixes.  The pressure pow here is correctly rounded by default
 * (oracle_pow below), so the device's correctly rounded pow gives the same bits;
 * mode 0 uses the C library's pow instead (timing of the CPU baseline).
 * oracle_pow is pinned against mpmath (tests/test_crpow.py).
 */
#include <math.h>
#include <quadmath.h>
#include <stdint.h>
#include <stdlib.h>
#ifdef _OPENMP
#include <omp.h>
#endif

typedef struct {
  double gamma, p0, R_gas, c_p, c_v, gravity, swe_gravity, delta_nl;
} ophys;

#define RET_BADBC 4

/* x^y for the pressure.  Mode 1 (default): correctly rounded -- libquadmath's
 * powq (error about one unit of its 113-bit significand) rounded once to
 * double, which is the correctly rounded double unless x^y lies within
 * about 2^-110 (relative) of a midpoint between two doubles.  Mode 0: the C
 * library's pow (glibc, not correctly rounded; about 40x faster). */
static int g_pow_mode = 1;
void oracle_set_pow_mode(int mode) { g_pow_mode = mode; }
int oracle_get_pow_mode(void) { return g_pow_mode; }
double oracle_pow(double x, double y) {
  return g_pow_mode ? (double)powq((__float128)x, (__float128)y) : pow(x, y);
}

static int ipow(int a, int b) {
  int r = 1;
  while (b-- > 0) r *= a;
  return r;
}

static void set_threads(int nthreads) {
#ifdef _OPENMP
  omp_set_num_threads(nthreads > 0 ? nthreads : 1);
#else
  (void)nthreads;
#endif
}

/* ------------------------------------------------------------------ 3D Euler */
enum { U3 = 0, V3, W3, R3, E3 };
enum { XIX = 0, ETAX, ZETAX, XIY, ETAY, ZETAY, XIZ, ETAZ, ZETAZ, MJ3, MJI3 };

int oracle_euler3d_volumerhs(int N, double* rhs, const double* Q, const double* vgeo,
                             const double* D, int64_t nelem, const ophys* ph, int nthreads) {
  const int Nq = N + 1, Np = Nq * Nq * Nq, ns = 5, nv = 14;
  const double gamma = ph->gamma, p0 = ph->p0, R_gas = ph->R_gas, c_p = ph->c_p,
               c_v = ph->c_v, gravity = ph->gravity;
  (void)gamma;
  set_threads(nthreads);
#pragma omp parallel
  {
    double* sF = (double*)malloc(sizeof(double) * Np * ns);
    double* sG = (double*)malloc(sizeof(double) * Np * ns);
    double* sH = (double*)malloc(sizeof(double) * Np * ns);
#pragma omp for schedule(static)
    for (int64_t e = 0; e < nelem; ++e) {
      const double* q = Q + (size_t)Np * ns * e;
      const double* vg = vgeo + (size_t)Np * nv * e;
      double* r = rhs + (size_t)Np * ns * e;
      for (int k = 0; k < Nq; ++k)
        for (int j = 0; j < Nq; ++j)
          for (int i = 0; i < Nq; ++i) {
            int n = i + Nq * (j + Nq * k);
            double MJ = vg[n + Np * MJ3];
            double xix = vg[n + Np * XIX], xiy = vg[n + Np * XIY], xiz = vg[n + Np * XIZ];
            double etax = vg[n + Np * ETAX], etay = vg[n + Np * ETAY], etaz = vg[n + Np * ETAZ];
            double zetax = vg[n + Np * ZETAX], zetay = vg[n + Np * ZETAY], zetaz = vg[n + Np * ZETAZ];
            double U = q[n + Np * U3], V = q[n + Np * V3], W = q[n + Np * W3];
            double rho = q[n + Np * R3], E = q[n + Np * E3];
            double P = p0 * oracle_pow(R_gas * E / p0, c_p / c_v);
            double rinv = 1 / rho;
            double frx = U, fUx = rinv * U * U + P, fVx = rinv * V * U, fWx = rinv * W * U,
                   fEx = E * rinv * U;
            double fry = V, fUy = rinv * U * V, fVy = rinv * V * V + P, fWy = rinv * W * V,
                   fEy = E * rinv * V;
            double frz = W, fUz = rinv * U * W, fVz = rinv * V * W, fWz = rinv * W * W + P,
                   fEz = E * rinv * W;
            sF[n + Np * R3] = MJ * (xix * frx + xiy * fry + xiz * frz);
            sF[n + Np * U3] = MJ * (xix * fUx + xiy * fUy + xiz * fUz);
            sF[n + Np * V3] = MJ * (xix * fVx + xiy * fVy + xiz * fVz);
            sF[n + Np * W3] = MJ * (xix * fWx + xiy * fWy + xiz * fWz);
            sF[n + Np * E3] = MJ * (xix * fEx + xiy * fEy + xiz * fEz);
            sG[n + Np * R3] = MJ * (etax * frx + etay * fry + etaz * frz);
            sG[n + Np * U3] = MJ * (etax * fUx + etay * fUy + etaz * fUz);
            sG[n + Np * V3] = MJ * (etax * fVx + etay * fVy + etaz * fVz);
            sG[n + Np * W3] = MJ * (etax * fWx + etay * fWy + etaz * fWz);
            sG[n + Np * E3] = MJ * (etax * fEx + etay * fEy + etaz * fEz);
            sH[n + Np * R3] = MJ * (zetax * frx + zetay * fry + zetaz * frz);
            sH[n + Np * U3] = MJ * (zetax * fUx + zetay * fUy + zetaz * fUz);
            sH[n + Np * V3] = MJ * (zetax * fVx + zetay * fVy + zetaz * fVz);
            sH[n + Np * W3] = MJ * (zetax * fWx + zetay * fWy + zetaz * fWz);
            sH[n + Np * E3] = MJ * (zetax * fEx + zetay * fEy + zetaz * fEz);
            r[n + Np * W3] -= MJ * rho * gravity;
          }
      for (int s = 0; s < ns; ++s)
        for (int k = 0; k < Nq; ++k)
          for (int j = 0; j < Nq; ++j)
            for (int i = 0; i < Nq; ++i)
              for (int m = 0; m < Nq; ++m)
                r[i + Nq * (j + Nq * k) + Np * s] += D[m + Nq * i] * sF[m + Nq * (j + Nq * k) + Np * s];
      for (int s = 0; s < ns; ++s)
        for (int k = 0; k < Nq; ++k)
          for (int j = 0; j < Nq; ++j)
            for (int i = 0; i < Nq; ++i)
              for (int m = 0; m < Nq; ++m)
                r[i + Nq * (j + Nq * k) + Np * s] += D[m + Nq * j] * sG[i + Nq * (m + Nq * k) + Np * s];
      for (int s = 0; s < ns; ++s)
        for (int k = 0; k < Nq; ++k)
          for (int j = 0; j < Nq; ++j)
            for (int i = 0; i < Nq; ++i)
              for (int m = 0; m < Nq; ++m)
                r[i + Nq * (j + Nq * k) + Np * s] += D[m + Nq * k] * sH[i + Nq * (j + Nq * m) + Np * s];
    }
    free(sF);
    free(sG);
    free(sH);
  }
  return 0;
}

int oracle_euler3d_facerhs(int N, double* rhs, const double* Q, const double* sgeo,
                           int64_t nelem, const int64_t* vmapM, const int64_t* vmapP,
                           const int64_t* elemtobndy, const ophys* ph, int nthreads) {
  const int Nq = N + 1, Np = Nq * Nq * Nq, Nfp = Nq * Nq, nface = 6, ns = 5;
  const double gamma = ph->gamma, p0 = ph->p0, R_gas = ph->R_gas, c_p = ph->c_p, c_v = ph->c_v;
  int bad = 0;
  set_threads(nthreads);
#pragma omp parallel for schedule(static) reduction(| : bad)
  for (int64_t e = 0; e < nelem; ++e) {
    for (int f = 0; f < nface; ++f) {
      int64_t bc = elemtobndy[f + nface * e];
      if (bc != 0 && bc != 1) {
        bad = 1;
        continue;
      }
      for (int n = 0; n < Nfp; ++n) {
        size_t fo = n + (size_t)Nfp * (f + (size_t)nface * e);
        const double* sg = sgeo + 5 * fo;
        double nxM = sg[0], nyM = sg[1], nzM = sg[2], sMJ = sg[3];
        int64_t idM = vmapM[fo], idP = vmapP[fo];
        int64_t eM = e, eP = (idP - 1) / Np;
        int64_t vidM = (idM - 1) % Np, vidP = (idP - 1) % Np;
        const double* qM = Q + (size_t)Np * ns * eM;
        const double* qP = Q + (size_t)Np * ns * eP;
        double rhoM = qM[vidM + Np * R3], UM = qM[vidM + Np * U3], VM = qM[vidM + Np * V3],
               WM = qM[vidM + Np * W3], EM = qM[vidM + Np * E3];
        double PM = p0 * oracle_pow(R_gas * EM / p0, c_p / c_v);
        double rhoP, UP, VP, WP, EP, PP;
        if (bc == 0) {
          rhoP = qP[vidP + Np * R3];
          UP = qP[vidP + Np * U3];
          VP = qP[vidP + Np * V3];
          WP = qP[vidP + Np * W3];
          EP = qP[vidP + Np * E3];
          PP = p0 * oracle_pow(R_gas * EP / p0, c_p / c_v);
        } else {
          double UnM = nxM * UM + nyM * VM + nzM * WM;
          UP = UM - 2 * UnM * nxM;
          VP = VM - 2 * UnM * nyM;
          WP = WM - 2 * UnM * nzM;
          rhoP = rhoM;
          EP = EM;
          PP = PM;
        }
        double rMi = 1 / rhoM;
        double frMx = UM, fUMx = rMi * UM * UM + PM, fVMx = rMi * VM * UM, fWMx = rMi * WM * UM,
               fEMx = rMi * UM * EM;
        double frMy = VM, fUMy = rMi * UM * VM, fVMy = rMi * VM * VM + PM, fWMy = rMi * WM * VM,
               fEMy = rMi * VM * EM;
        double frMz = WM, fUMz = rMi * UM * WM, fVMz = rMi * VM * WM, fWMz = rMi * WM * WM + PM,
               fEMz = rMi * WM * EM;
        double rPi = 1 / rhoP;
        double frPx = UP, fUPx = rPi * UP * UP + PP, fVPx = rPi * VP * UP, fWPx = rPi * WP * UP,
               fEPx = rPi * UP * EP;
        double frPy = VP, fUPy = rPi * UP * VP, fVPy = rPi * VP * VP + PP, fWPy = rPi * WP * VP,
               fEPy = rPi * VP * EP;
        double frPz = WP, fUPz = rPi * UP * WP, fVPz = rPi * VP * WP, fWPz = rPi * WP * WP + PP,
               fEPz = rPi * WP * EP;
        double lM = rMi * fabs(nxM * UM + nyM * VM + nzM * WM) + sqrt(rMi * gamma * PM);
        double lP = rPi * fabs(nxM * UP + nyM * VP + nzM * WP) + sqrt(rPi * gamma * PP);
        double l = lM > lP ? lM : lP;
        /* Julia max(a, b): NaN-propagating; both finite here. */
        if (isnan(lM) || isnan(lP)) l = NAN;
        double frS = (nxM * (frMx + frPx) + nyM * (frMy + frPy) + nzM * (frMz + frPz) - l * (rhoP - rhoM)) / 2;
        double fUS = (nxM * (fUMx + fUPx) + nyM * (fUMy + fUPy) + nzM * (fUMz + fUPz) - l * (UP - UM)) / 2;
        double fVS = (nxM * (fVMx + fVPx) + nyM * (fVMy + fVPy) + nzM * (fVMz + fVPz) - l * (VP - VM)) / 2;
        double fWS = (nxM * (fWMx + fWPx) + nyM * (fWMy + fWPy) + nzM * (fWMz + fWPz) - l * (WP - WM)) / 2;
        double fES = (nxM * (fEMx + fEPx) + nyM * (fEMy + fEPy) + nzM * (fEMz + fEPz) - l * (EP - EM)) / 2;
        double* r = rhs + (size_t)Np * ns * eM;
        r[vidM + Np * R3] -= sMJ * frS;
        r[vidM + Np * U3] -= sMJ * fUS;
        r[vidM + Np * V3] -= sMJ * fVS;
        r[vidM + Np * W3] -= sMJ * fWS;
        r[vidM + Np * E3] -= sMJ * fES;
      }
    }
  }
  return bad ? RET_BADBC : 0;
}

/* updatesolution! for any equation set: all `ns` states, MJI at vgeo field
 * `iMJI` of `nvgeo`. */
int oracle_updatesolution(int Np, int ns, int nvgeo, int iMJI, double* rhs, double* Q,
                          const double* vgeo, int64_t nelem, double rka, double rkb,
                          double dt, int nthreads) {
  set_threads(nthreads);
#pragma omp parallel for schedule(static)
  for (int64_t e = 0; e < nelem; ++e)
    for (int s = 0; s < ns; ++s)
      for (int i = 0; i < Np; ++i) {
        size_t o = i + (size_t)Np * (s + (size_t)ns * e);
        Q[o] += rkb * dt * rhs[o] * vgeo[i + (size_t)Np * (iMJI + (size_t)nvgeo * e)];
        rhs[o] *= rka;
      }
  return 0;
}

/* ------------------------------------------------------------------ 2D Euler */
enum { U2 = 0, V2, R2, E2 };
enum { XIX2 = 0, ETAX2, XIY2, ETAY2, MJ2, MJI2 };

int oracle_euler2d_volumerhs(int N, double* rhs, const double* Q, const double* vgeo,
                             const double* D, int64_t nelem, const ophys* ph, int nthreads) {
  const int Nq = N + 1, Np = Nq * Nq, ns = 4, nv = 8;
  const double p0 = ph->p0, R_gas = ph->R_gas, c_p = ph->c_p, c_v = ph->c_v,
               gravity = ph->gravity;
  set_threads(nthreads);
#pragma omp parallel
  {
    double* sF = (double*)malloc(sizeof(double) * Np * ns);
    double* sG = (double*)malloc(sizeof(double) * Np * ns);
#pragma omp for schedule(static)
    for (int64_t e = 0; e < nelem; ++e) {
      const double* q = Q + (size_t)Np * ns * e;
      const double* vg = vgeo + (size_t)Np * nv * e;
      double* r = rhs + (size_t)Np * ns * e;
      for (int j = 0; j < Nq; ++j)
        for (int i = 0; i < Nq; ++i) {
          int n = i + Nq * j;
          double MJ = vg[n + Np * MJ2];
          double xix = vg[n + Np * XIX2], xiy = vg[n + Np * XIY2];
          double etax = vg[n + Np * ETAX2], etay = vg[n + Np * ETAY2];
          double U = q[n + Np * U2], V = q[n + Np * V2], rho = q[n + Np * R2], E = q[n + Np * E2];
          double P = p0 * oracle_pow(R_gas * E / p0, c_p / c_v);
          double rinv = 1 / rho;
          double frx = U, fUx = rinv * U * U + P, fVx = rinv * V * U, fEx = rinv * U * E;
          double fry = V, fUy = rinv * U * V, fVy = rinv * V * V + P, fEy = rinv * V * E;
          sF[n + Np * R2] = MJ * (xix * frx + xiy * fry);
          sF[n + Np * U2] = MJ * (xix * fUx + xiy * fUy);
          sF[n + Np * V2] = MJ * (xix * fVx + xiy * fVy);
          sF[n + Np * E2] = MJ * (xix * fEx + xiy * fEy);
          sG[n + Np * R2] = MJ * (etax * frx + etay * fry);
          sG[n + Np * U2] = MJ * (etax * fUx + etay * fUy);
          sG[n + Np * V2] = MJ * (etax * fVx + etay * fVy);
          sG[n + Np * E2] = MJ * (etax * fEx + etay * fEy);
          r[n + Np * V2] -= MJ * rho * gravity;
        }
      for (int s = 0; s < ns; ++s)
        for (int j = 0; j < Nq; ++j)
          for (int i = 0; i < Nq; ++i)
            for (int k = 0; k < Nq; ++k)
              r[i + Nq * j + Np * s] += D[k + Nq * i] * sF[k + Nq * j + Np * s];
      for (int s = 0; s < ns; ++s)
        for (int j = 0; j < Nq; ++j)
          for (int i = 0; i < Nq; ++i)
            for (int k = 0; k < Nq; ++k)
              r[i + Nq * j + Np * s] += D[k + Nq * j] * sG[i + Nq * k + Np * s];
    }
    free(sF);
    free(sG);
  }
  return 0;
}

int oracle_euler2d_facerhs(int N, double* rhs, const double* Q, const double* sgeo,
                           int64_t nelem, const int64_t* vmapM, const int64_t* vmapP,
                           const int64_t* elemtobndy, const ophys* ph, int nthreads) {
  const int Nq = N + 1, Np = Nq * Nq, Nfp = Nq, nface = 4, ns = 4;
  const double gamma = ph->gamma, p0 = ph->p0, R_gas = ph->R_gas, c_p = ph->c_p, c_v = ph->c_v;
  int bad = 0;
  set_threads(nthreads);
#pragma omp parallel for schedule(static) reduction(| : bad)
  for (int64_t e = 0; e < nelem; ++e) {
    for (int f = 0; f < nface; ++f) {
      int64_t bc = elemtobndy[f + nface * e];
      if (bc != 0 && bc != 1) {
        bad = 1;
        continue;
      }
      for (int n = 0; n < Nfp; ++n) {
        size_t fo = n + (size_t)Nfp * (f + (size_t)nface * e);
        const double* sg = sgeo + 4 * fo;
        double nxM = sg[0], nyM = sg[1], sMJ = sg[2];
        int64_t idM = vmapM[fo], idP = vmapP[fo];
        int64_t eM = e, eP = (idP - 1) / Np;
        int64_t vidM = (idM - 1) % Np, vidP = (idP - 1) % Np;
        const double* qM = Q + (size_t)Np * ns * eM;
        const double* qP = Q + (size_t)Np * ns * eP;
        double rhoM = qM[vidM + Np * R2], UM = qM[vidM + Np * U2], VM = qM[vidM + Np * V2],
               EM = qM[vidM + Np * E2];
        double PM = p0 * oracle_pow(R_gas * EM / p0, c_p / c_v);
        double rhoP, UP, VP, EP, PP;
        if (bc == 0) {
          rhoP = qP[vidP + Np * R2];
          UP = qP[vidP + Np * U2];
          VP = qP[vidP + Np * V2];
          EP = qP[vidP + Np * E2];
          PP = p0 * oracle_pow(R_gas * EP / p0, c_p / c_v);
        } else {
          double UnM = nxM * UM + nyM * VM;
          UP = UM - 2 * UnM * nxM;
          VP = VM - 2 * UnM * nyM;
          rhoP = rhoM;
          EP = EM;
          PP = PM;
        }
        double rMi = 1 / rhoM;
        double frMx = UM, fUMx = rMi * UM * UM + PM, fVMx = rMi * VM * UM, fEMx = rMi * UM * EM;
        double frMy = VM, fUMy = rMi * UM * VM, fVMy = rMi * VM * VM + PM, fEMy = rMi * VM * EM;
        double rPi = 1 / rhoP;
        double frPx = UP, fUPx = rPi * UP * UP + PP, fVPx = rPi * VP * UP, fEPx = rPi * UP * EP;
        double frPy = VP, fUPy = rPi * UP * VP, fVPy = rPi * VP * VP + PP, fEPy = rPi * VP * EP;
        double lM = rMi * fabs(nxM * UM + nyM * VM) + sqrt(rMi * gamma * PM);
        double lP = rPi * fabs(nxM * UP + nyM * VP) + sqrt(rPi * gamma * PP);
        double l = lM > lP ? lM : lP;
        if (isnan(lM) || isnan(lP)) l = NAN;
        double frS = (nxM * (frMx + frPx) + nyM * (frMy + frPy) + -(l * (rhoP - rhoM))) / 2;
        double fUS = (nxM * (fUMx + fUPx) + nyM * (fUMy + fUPy) + -(l * (UP - UM))) / 2;
        double fVS = (nxM * (fVMx + fVPx) + nyM * (fVMy + fVPy) + -(l * (VP - VM))) / 2;
        double fES = (nxM * (fEMx + fEPx) + nyM * (fEMy + fEPy) + -(l * (EP - EM))) / 2;
        double* r = rhs + (size_t)Np * ns * eM;
        r[vidM + Np * R2] -= sMJ * frS;
        r[vidM + Np * U2] -= sMJ * fUS;
        r[vidM + Np * V2] -= sMJ * fVS;
        r[vidM + Np * E2] -= sMJ * fES;
      }
    }
  }
  return bad ? RET_BADBC : 0;
}

/* ------------------------------------------- 2D Euler, total energy (set3c) */
enum { YV2 = 7 };

/* P = (R/cv) (E - (U^2 + V^2) / (2 rho) - rho g y)  (euler2d_set3c.jl:226) */
static double te_pressure(double R_gas, double c_v, double gravity, double U, double V, double rho,
                          double E, double y) {
  return (R_gas / c_v) * (E - (U * U + V * V) / (2 * rho) - rho * gravity * y);
}

int oracle_euler2dte_volumerhs(int N, double* rhs, const double* Q, const double* vgeo,
                               const double* D, int64_t nelem, const ophys* ph, int nthreads) {
  const int Nq = N + 1, Np = Nq * Nq, ns = 4, nv = 8;
  const double R_gas = ph->R_gas, c_v = ph->c_v, gravity = ph->gravity;
  set_threads(nthreads);
#pragma omp parallel
  {
    double* sF = (double*)malloc(sizeof(double) * Np * ns);
    double* sG = (double*)malloc(sizeof(double) * Np * ns);
#pragma omp for schedule(static)
    for (int64_t e = 0; e < nelem; ++e) {
      const double* q = Q + (size_t)Np * ns * e;
      const double* vg = vgeo + (size_t)Np * nv * e;
      double* r = rhs + (size_t)Np * ns * e;
      for (int j = 0; j < Nq; ++j)
        for (int i = 0; i < Nq; ++i) {
          int n = i + Nq * j;
          double MJ = vg[n + Np * MJ2];
          double xix = vg[n + Np * XIX2], xiy = vg[n + Np * XIY2];
          double etax = vg[n + Np * ETAX2], etay = vg[n + Np * ETAY2];
          double y = vg[n + Np * YV2];
          double U = q[n + Np * U2], V = q[n + Np * V2], rho = q[n + Np * R2], E = q[n + Np * E2];
          double P = te_pressure(R_gas, c_v, gravity, U, V, rho, E, y);
          double rinv = 1 / rho;
          double frx = U, fUx = rinv * U * U + P, fVx = rinv * V * U, fEx = rinv * U * (E + P);
          double fry = V, fUy = rinv * U * V, fVy = rinv * V * V + P, fEy = rinv * V * (E + P);
          sF[n + Np * R2] = MJ * (xix * frx + xiy * fry);
          sF[n + Np * U2] = MJ * (xix * fUx + xiy * fUy);
          sF[n + Np * V2] = MJ * (xix * fVx + xiy * fVy);
          sF[n + Np * E2] = MJ * (xix * fEx + xiy * fEy);
          sG[n + Np * R2] = MJ * (etax * frx + etay * fry);
          sG[n + Np * U2] = MJ * (etax * fUx + etay * fUy);
          sG[n + Np * V2] = MJ * (etax * fVx + etay * fVy);
          sG[n + Np * E2] = MJ * (etax * fEx + etay * fEy);
          r[n + Np * V2] -= MJ * rho * gravity;
        }
      for (int s = 0; s < ns; ++s)
        for (int j = 0; j < Nq; ++j)
          for (int i = 0; i < Nq; ++i)
            for (int k = 0; k < Nq; ++k)
              r[i + Nq * j + Np * s] += D[k + Nq * i] * sF[k + Nq * j + Np * s];
      for (int s = 0; s < ns; ++s)
        for (int j = 0; j < Nq; ++j)
          for (int i = 0; i < Nq; ++i)
            for (int k = 0; k < Nq; ++k)
              r[i + Nq * j + Np * s] += D[k + Nq * j] * sG[i + Nq * k + Np * s];
    }
    free(sF);
    free(sG);
  }
  return 0;
}

int oracle_euler2dte_fluxrhs(int N, double* rhs, const double* Q, const double* sgeo,
                             const double* vgeo, int64_t nelem, const int64_t* vmapM,
                             const int64_t* vmapP, const int64_t* elemtobndy, const ophys* ph,
                             int nthreads) {
  const int Nq = N + 1, Np = Nq * Nq, Nfp = Nq, nface = 4, ns = 4, nv = 8;
  const double gamma = ph->gamma, R_gas = ph->R_gas, c_v = ph->c_v, gravity = ph->gravity;
  int bad = 0;
  set_threads(nthreads);
#pragma omp parallel for schedule(static) reduction(| : bad)
  for (int64_t e = 0; e < nelem; ++e) {
    for (int f = 0; f < nface; ++f) {
      int64_t bc = elemtobndy[f + nface * e];
      if (bc != 0 && bc != 1) {
        bad = 1;
        continue;
      }
      for (int n = 0; n < Nfp; ++n) {
        size_t fo = n + (size_t)Nfp * (f + (size_t)nface * e);
        const double* sg = sgeo + 4 * fo;
        double nxM = sg[0], nyM = sg[1], sMJ = sg[2];
        int64_t idM = vmapM[fo], idP = vmapP[fo];
        int64_t eM = e, eP = (idP - 1) / Np;
        int64_t vidM = (idM - 1) % Np, vidP = (idP - 1) % Np;
        const double* qM = Q + (size_t)Np * ns * eM;
        const double* qP = Q + (size_t)Np * ns * eP;
        double rhoM = qM[vidM + Np * R2], UM = qM[vidM + Np * U2], VM = qM[vidM + Np * V2],
               EM = qM[vidM + Np * E2];
        double yM = vgeo[vidM + (size_t)Np * (YV2 + (size_t)nv * eM)];
        double PM = te_pressure(R_gas, c_v, gravity, UM, VM, rhoM, EM, yM);
        double rhoP, UP, VP, EP, PP;
        if (bc == 0) {
          rhoP = qP[vidP + Np * R2];
          UP = qP[vidP + Np * U2];
          VP = qP[vidP + Np * V2];
          EP = qP[vidP + Np * E2];
          double yP = vgeo[vidP + (size_t)Np * (YV2 + (size_t)nv * eP)];
          PP = te_pressure(R_gas, c_v, gravity, UP, VP, rhoP, EP, yP);
        } else {
          double UnM = nxM * UM + nyM * VM;
          UP = UM - 2 * UnM * nxM;
          VP = VM - 2 * UnM * nyM;
          rhoP = rhoM;
          EP = EM;
          PP = PM;
        }
        double rMi = 1 / rhoM;
        double frMx = UM, fUMx = rMi * UM * UM + PM, fVMx = rMi * VM * UM,
               fEMx = rMi * UM * (EM + PM);
        double frMy = VM, fUMy = rMi * UM * VM, fVMy = rMi * VM * VM + PM,
               fEMy = rMi * VM * (EM + PM);
        double rPi = 1 / rhoP;
        double frPx = UP, fUPx = rPi * UP * UP + PP, fVPx = rPi * VP * UP,
               fEPx = rPi * UP * (EP + PP);
        double frPy = VP, fUPy = rPi * UP * VP, fVPy = rPi * VP * VP + PP,
               fEPy = rPi * VP * (EP + PP);
        double lM = rMi * fabs(nxM * UM + nyM * VM) + sqrt(rMi * gamma * PM);
        double lP = rPi * fabs(nxM * UP + nyM * VP) + sqrt(rPi * gamma * PP);
        double l = lM > lP ? lM : lP;
        if (isnan(lM) || isnan(lP)) l = NAN;
        double frS = (nxM * (frMx + frPx) + nyM * (frMy + frPy) + -(l * (rhoP - rhoM))) / 2;
        double fUS = (nxM * (fUMx + fUPx) + nyM * (fUMy + fUPy) + -(l * (UP - UM))) / 2;
        double fVS = (nxM * (fVMx + fVPx) + nyM * (fVMy + fVPy) + -(l * (VP - VM))) / 2;
        double fES = (nxM * (fEMx + fEPx) + nyM * (fEMy + fEPy) + -(l * (EP - EM))) / 2;
        double* r = rhs + (size_t)Np * ns * eM;
        r[vidM + Np * R2] -= sMJ * frS;
        r[vidM + Np * U2] -= sMJ * fUS;
        r[vidM + Np * V2] -= sMJ * fVS;
        r[vidM + Np * E2] -= sMJ * fES;
      }
    }
  }
  return bad ? RET_BADBC : 0;
}

/* courantnumber's time step (euler2d_set3c.jl:121-148): the minimum over
 * nodes of dx / (|u| + sqrt(gamma P / rho)) / N, dx = sqrt((1/(2 xi_x))^2 +
 * (1/(2 eta_y))^2). */
double oracle_euler2dte_dt(int N, const double* vgeo, const double* Q, int64_t nelem,
                           const ophys* ph) {
  const int Np = (N + 1) * (N + 1), ns = 4, nv = 8;
  const double gamma = ph->gamma, R_gas = ph->R_gas, c_v = ph->c_v, gravity = ph->gravity;
  double dt = 1.7976931348623157e308;
  for (int64_t e = 0; e < nelem; ++e)
    for (int n = 0; n < Np; ++n) {
      const double* q = Q + (size_t)Np * ns * e;
      const double* vg = vgeo + (size_t)Np * nv * e;
      double rho = q[n + Np * R2], U = q[n + Np * U2], V = q[n + Np * V2], E = q[n + Np * E2];
      double xix = vg[n + Np * XIX2], etay = vg[n + Np * ETAY2], y = vg[n + Np * YV2];
      double P = te_pressure(R_gas, c_v, gravity, U, V, rho, E, y);
      double u = U / rho, v = V / rho;
      double a = 1.0 / (2 * xix), b = 1.0 / (2 * etay);
      double dx = sqrt(a * a + b * b);
      double vel = sqrt(u * u + v * v);
      double ws = vel + sqrt(gamma * P / rho);
      double loc = 1.0 * dx / ws / N;
      if (loc < dt) dt = loc;
    }
  return dt;
}

/* ------------------------------------------------------------- shallow water */
enum { SU = 0, SV, SH, SB };

int oracle_swe_volumerhs(int N, double* rhs, const double* Q, const double* vgeo,
                         const double* D, int64_t nelem, const ophys* ph, int nthreads) {
  const int Nq = N + 1, Np = Nq * Nq, ns = 4, nv = 8;
  const double gravity = ph->swe_gravity, dnl = ph->delta_nl;
  set_threads(nthreads);
#pragma omp parallel
  {
    double* sF = (double*)malloc(sizeof(double) * Np * 3);
    double* sG = (double*)malloc(sizeof(double) * Np * 3);
#pragma omp for schedule(static)
    for (int64_t e = 0; e < nelem; ++e) {
      const double* q = Q + (size_t)Np * ns * e;
      const double* vg = vgeo + (size_t)Np * nv * e;
      double* r = rhs + (size_t)Np * ns * e;
      for (int j = 0; j < Nq; ++j)
        for (int i = 0; i < Nq; ++i) {
          int n = i + Nq * j;
          double MJ = vg[n + Np * MJ2];
          double xix = vg[n + Np * XIX2], xiy = vg[n + Np * XIY2];
          double etax = vg[n + Np * ETAX2], etay = vg[n + Np * ETAY2];
          double U = q[n + Np * SU], V = q[n + Np * SV], h = q[n + Np * SH], b = q[n + Np * SB];
          double H = h + b;
          double u = U / H, v = V / H;
          double fhx = U, fhy = V;
          double fUx = (H * u * u + 0.5 * gravity * (h * h)) * dnl + gravity * h * b;
          double fUy = (H * u * v) * dnl;
          double fVx = (H * v * u) * dnl;
          double fVy = (H * v * v + 0.5 * gravity * (h * h)) * dnl + gravity * h * b;
          sF[n + Np * SH] = MJ * (xix * fhx + xiy * fhy);
          sF[n + Np * SU] = MJ * (xix * fUx + xiy * fUy);
          sF[n + Np * SV] = MJ * (xix * fVx + xiy * fVy);
          sG[n + Np * SH] = MJ * (etax * fhx + etay * fhy);
          sG[n + Np * SU] = MJ * (etax * fUx + etay * fUy);
          sG[n + Np * SV] = MJ * (etax * fVx + etay * fVy);
        }
      for (int s = 0; s < 3; ++s)
        for (int j = 0; j < Nq; ++j)
          for (int i = 0; i < Nq; ++i)
            for (int m = 0; m < Nq; ++m)
              r[i + Nq * j + Np * s] += D[m + Nq * i] * sF[m + Nq * j + Np * s];
      for (int s = 0; s < 3; ++s)
        for (int j = 0; j < Nq; ++j)
          for (int i = 0; i < Nq; ++i)
            for (int m = 0; m < Nq; ++m)
              r[i + Nq * j + Np * s] += D[m + Nq * j] * sG[i + Nq * m + Np * s];
    }
    free(sF);
    free(sG);
  }
  return 0;
}

int oracle_swe_fluxrhs(int N, double* rhs, const double* Q, const double* sgeo, int64_t nelem,
                       const int64_t* vmapM, const int64_t* vmapP, const int64_t* elemtobndy,
                       const ophys* ph, int nthreads) {
  const int Nq = N + 1, Np = Nq * Nq, Nfp = Nq, nface = 4, ns = 4;
  const double gravity = ph->swe_gravity, dnl = ph->delta_nl;
  int bad = 0;
  set_threads(nthreads);
#pragma omp parallel for schedule(static) reduction(| : bad)
  for (int64_t e = 0; e < nelem; ++e) {
    for (int f = 0; f < nface; ++f) {
      int64_t bc = elemtobndy[f + nface * e];
      if (bc != 0 && bc != 1) {
        bad = 1;
        continue;
      }
      for (int n = 0; n < Nfp; ++n) {
        size_t fo = n + (size_t)Nfp * (f + (size_t)nface * e);
        const double* sg = sgeo + 4 * fo;
        double nxM = sg[0], nyM = sg[1], sMJ = sg[2];
        int64_t idM = vmapM[fo], idP = vmapP[fo];
        int64_t eM = e, eP = (idP - 1) / Np;
        int64_t vidM = (idM - 1) % Np, vidP = (idP - 1) % Np;
        const double* qM = Q + (size_t)Np * ns * eM;
        const double* qP = Q + (size_t)Np * ns * eP;
        double hM = qM[vidM + Np * SH], bM = qM[vidM + Np * SB], UM = qM[vidM + Np * SU],
               VM = qM[vidM + Np * SV];
        double hP, bP, UP, VP;
        if (bc == 0) {
          hP = qP[vidP + Np * SH];
          bP = qM[vidP + Np * SB]; /* the reference reads b from element eM */
          UP = qP[vidP + Np * SU];
          VP = qP[vidP + Np * SV];
        } else {
          double UnM = nxM * UM + nyM * VM;
          UP = UM - 2 * UnM * nxM;
          VP = VM - 2 * UnM * nyM;
          hP = hM;
          bP = bM;
        }
        double HM = hM + bM, uM = UM / HM, vM = VM / HM;
        double HP = hP + bP, uP = UP / HP, vP = VP / HP;
        double fhMx = UM, fhMy = VM;
        double fUMx = (HM * uM * uM + 0.5 * gravity * (hM * hM)) * dnl + gravity * hM * bM;
        double fUMy = HM * uM * vM * dnl;
        double fVMx = HM * vM * uM * dnl;
        double fVMy = (HM * vM * vM + 0.5 * gravity * (hM * hM)) * dnl + gravity * hM * bM;
        double fhPx = UP, fhPy = VP;
        double fUPx = (HP * uP * uP + 0.5 * gravity * (hP * hP)) * dnl + gravity * hP * bP;
        double fUPy = HP * uP * vP * dnl;
        double fVPx = HP * vP * uP * dnl;
        double fVPy = (HP * vP * vP + 0.5 * gravity * (hP * hP)) * dnl + gravity * hP * bP;
        double lM = (fabs(nxM * uM + nyM * vM) + sqrt(gravity * HM)) * dnl + (sqrt(gravity * bM)) * (1.0 - dnl);
        double lP = (fabs(nxM * uP + nyM * vP) + sqrt(gravity * HP)) * dnl + (sqrt(gravity * bP)) * (1.0 - dnl);
        double l = lM > lP ? lM : lP;
        if (isnan(lM) || isnan(lP)) l = NAN;
        double fhS = (nxM * (fhMx + fhPx) + nyM * (fhMy + fhPy) + -(l * (hP - hM))) / 2;
        double fUS = (nxM * (fUMx + fUPx) + nyM * (fUMy + fUPy) + -(l * (UP - UM))) / 2;
        double fVS = (nxM * (fVMx + fVPx) + nyM * (fVMy + fVPy) + -(l * (VP - VM))) / 2;
        double* r = rhs + (size_t)Np * ns * eM;
        r[vidM + Np * SH] -= sMJ * fhS;
        r[vidM + Np * SU] -= sMJ * fUS;
        r[vidM + Np * SV] -= sMJ * fVS;
      }
    }
  }
  return bad ? RET_BADBC : 0;
}

int oracle_swe_updatesolution(int N, double* rhs, double* Q, const double* vgeo, int64_t nelem,
                              double rka, double rkb, double dt, int advection, int nthreads) {
  const int Np = (N + 1) * (N + 1), ns = 4, nv = 8;
  set_threads(nthreads);
#pragma omp parallel for schedule(static)
  for (int64_t e = 0; e < nelem; ++e) {
    double* q = Q + (size_t)Np * ns * e;
    double* r = rhs + (size_t)Np * ns * e;
    const double* vg = vgeo + (size_t)Np * nv * e;
    double u[512], v[512];
    if (advection)
      for (int i = 0; i < Np; ++i) {
        u[i] = q[i + Np * SU] / (q[i + Np * SH] + q[i + Np * SB]);
        v[i] = q[i + Np * SV] / (q[i + Np * SH] + q[i + Np * SB]);
      }
    for (int s = 0; s < ns; ++s)
      for (int i = 0; i < Np; ++i) {
        q[i + Np * s] += rkb * dt * r[i + Np * s] * vg[i + Np * MJI2];
        r[i + Np * s] *= rka;
      }
    if (advection)
      for (int i = 0; i < Np; ++i) {
        q[i + Np * SU] = (q[i + Np * SH] + q[i + Np * SB]) * u[i];
        q[i + Np * SV] = (q[i + Np * SH] + q[i + Np * SB]) * v[i];
      }
  }
  return 0;
}

/* -------------------------------------------------------------- time loop */
/* LSRK54 coefficients as DFloat(num)/DFloat(den): euler3d_set2c.jl:772-788 */
void oracle_lsrk54_coefs(double* rka, double* rkb, double* rkc) {
  rka[0] = 0.0;
  rka[1] = (double)-567301805773 / (double)1357537059087;
  rka[2] = (double)-2404267990393 / (double)2016746695238;
  rka[3] = (double)-3550918686646 / (double)2091501179385;
  rka[4] = (double)-1275806237668 / (double)842570457699;
  rkb[0] = (double)1432997174477 / (double)9575080441755;
  rkb[1] = (double)5161836677717 / (double)13612068292357;
  rkb[2] = (double)1720146321549 / (double)2090206949498;
  rkb[3] = (double)3134564353537 / (double)4481467310338;
  rkb[4] = (double)2277821191437 / (double)14882151754819;
  rkc[0] = 0.0;
  rkc[1] = (double)1432997174477 / (double)9575080441755;
  rkc[2] = (double)2526269341429 / (double)6820363962896;
  rkc[3] = (double)2006345519317 / (double)3224310063776;
  rkc[4] = (double)2802321613138 / (double)2924317926251;
}

/* eqn: 0 = Euler (dim 2 or 3), 1 = SWE (dim 2), 2 = Euler total energy (dim
 * 2).  nsteps full LSRK54 steps of
 * the reference stage loop (volume; face; update with RKA[s%5+1], RKB[s]). */
int oracle_lsrk(int eqn, int dim, int N, double* Q, double* rhs, const double* vgeo,
                const double* sgeo, const double* D, int64_t nelem, const int64_t* vmapM,
                const int64_t* vmapP, const int64_t* elemtobndy, double dt, int nsteps,
                const ophys* ph, int advection, int nthreads) {
  double A[5], B[5], C[5];
  oracle_lsrk54_coefs(A, B, C);
  const int Np = ipow(N + 1, dim);
  for (int step = 0; step < nsteps; ++step)
    for (int s = 0; s < 5; ++s) {
      int rc = 0;
      if (eqn == 0 && dim == 3) {
        oracle_euler3d_volumerhs(N, rhs, Q, vgeo, D, nelem, ph, nthreads);
        rc = oracle_euler3d_facerhs(N, rhs, Q, sgeo, nelem, vmapM, vmapP, elemtobndy, ph, nthreads);
        oracle_updatesolution(Np, 5, 14, MJI3, rhs, Q, vgeo, nelem, A[(s + 1) % 5], B[s], dt, nthreads);
      } else if (eqn == 0 && dim == 2) {
        oracle_euler2d_volumerhs(N, rhs, Q, vgeo, D, nelem, ph, nthreads);
        rc = oracle_euler2d_facerhs(N, rhs, Q, sgeo, nelem, vmapM, vmapP, elemtobndy, ph, nthreads);
        oracle_updatesolution(Np, 4, 8, MJI2, rhs, Q, vgeo, nelem, A[(s + 1) % 5], B[s], dt, nthreads);
      } else if (eqn == 2 && dim == 2) {
        oracle_euler2dte_volumerhs(N, rhs, Q, vgeo, D, nelem, ph, nthreads);
        rc = oracle_euler2dte_fluxrhs(N, rhs, Q, sgeo, vgeo, nelem, vmapM, vmapP, elemtobndy, ph,
                                      nthreads);
        oracle_updatesolution(Np, 4, 8, MJI2, rhs, Q, vgeo, nelem, A[(s + 1) % 5], B[s], dt, nthreads);
      } else if (eqn == 1 && dim == 2) {
        oracle_swe_volumerhs(N, rhs, Q, vgeo, D, nelem, ph, nthreads);
        rc = oracle_swe_fluxrhs(N, rhs, Q, sgeo, nelem, vmapM, vmapP, elemtobndy, ph, nthreads);
        oracle_swe_updatesolution(N, rhs, Q, vgeo, nelem, A[(s + 1) % 5], B[s], dt, advection, nthreads);
      } else {
        return 1;
      }
      if (rc) return rc;
    }
  return 0;
}

/* L2energysquared: sum over elems, states, nodes of MJ * Q^2
 * (euler3d_set2c.jl:738-750); MJ at vgeo field iMJ of nvgeo. */
double oracle_l2energysquared(int Np, int ns, int nvgeo, int iMJ, const double* Q,
                              const double* vgeo, int64_t nelem) {
  double energy = 0.0;
  for (int64_t e = 0; e < nelem; ++e)
    for (int q = 0; q < ns; ++q)
      for (int i = 0; i < Np; ++i) {
        double v = Q[i + (size_t)Np * (q + (size_t)ns * e)];
        energy += vgeo[i + (size_t)Np * (iMJ + (size_t)nvgeo * e)] * (v * v);
      }
  return energy;
}

/* cfl for set2c Euler (euler3d_set2c.jl:67-97, euler2d_set2c.jl:66-93) */
double oracle_euler_cfl(int dim, int N, const double* vgeo, const double* Q, int64_t nelem,
                        const ophys* ph) {
  const int Np = ipow(N + 1, dim), ns = dim == 3 ? 5 : 4, nv = dim == 3 ? 14 : 8;
  const double gamma = ph->gamma, p0 = ph->p0, R_gas = ph->R_gas, c_p = ph->c_p, c_v = ph->c_v;
  double dt = 1.7976931348623157e308;
  for (int64_t e = 0; e < nelem; ++e)
    for (int n = 0; n < Np; ++n) {
      const double* q = Q + (size_t)Np * ns * e;
      const double* vg = vgeo + (size_t)Np * nv * e;
      double loc;
      if (dim == 3) {
        double rho = q[n + Np * R3], U = q[n + Np * U3], V = q[n + Np * V3], W = q[n + Np * W3];
        double E = q[n + Np * E3];
        double P = p0 * oracle_pow(R_gas * E / p0, c_p / c_v);
        double xix = vg[n + Np * XIX], xiy = vg[n + Np * XIY], xiz = vg[n + Np * XIZ];
        double etax = vg[n + Np * ETAX], etay = vg[n + Np * ETAY], etaz = vg[n + Np * ETAZ];
        double zetax = vg[n + Np * ZETAX], zetay = vg[n + Np * ZETAY], zetaz = vg[n + Np * ZETAZ];
        double c = rho * sqrt(gamma * P / rho);
        double a = fabs(U * xix + V * xiy + W * xiz) + c;
        double b = fabs(U * etax + V * etay + W * etaz) + c;
        double d = fabs(U * zetax + V * zetay + W * zetaz) + c;
        double m = a > b ? a : b;
        m = m > d ? m : d;
        loc = 2 * rho / m;
      } else {
        double rho = q[n + Np * R2], U = q[n + Np * U2], V = q[n + Np * V2], E = q[n + Np * E2];
        double P = p0 * oracle_pow(R_gas * E / p0, c_p / c_v);
        double xix = vg[n + Np * XIX2], xiy = vg[n + Np * XIY2];
        double etax = vg[n + Np * ETAX2], etay = vg[n + Np * ETAY2];
        double c = rho * sqrt(gamma * P / rho);
        double a = fabs(U * xix + V * xiy) + c;
        double b = fabs(U * etax + V * etay) + c;
        loc = 2 * rho / (a > b ? a : b);
      }
      if (loc < dt) dt = loc;
    }
  return dt;
}
