// Physics policies of the DG kernels: fluxes, sources and the Rusanov face
// flux of the set2c compressible Euler equations (2D/3D) and of the
// shallow-water equations, in FP64 with the reference's evaluation order.
// The exact-order units (euler3d.hip, euler2d.hip, swe2d.hip, batched.hip)
// are compiled with -ffp-contract=off, so a*b+c stays two correctly rounded
// operations; the *_fma.hip units use the contracted policies below.
//
// Reference formulas (same association, left to right):
//   3D Euler volume  examples/3d_kernels/euler3d_set2c.jl:175-215
//   3D Euler face    examples/3d_kernels/euler3d_set2c.jl:262-339
//   2D Euler volume  examples/2d_kernels/euler2d_set2c.jl:163-184
//   2D Euler face    examples/2d_kernels/euler2d_set2c.jl:238-292
//   SWE volume       examples/old_fxg_files/shallow_water.jl:140-160
//   SWE face         examples/old_fxg_files/shallow_water.jl:196-255
#pragma once
#include <hip/hip_runtime.h>
#include <math.h>

#include "crpow.h"

namespace canary {

// Coefficients of the contracted pressure's series (FastOps::powk): the
// 2 atanh series of the logarithm, 1/19, 1/17, ..., 1/3, and the Taylor
// series of the exponential, 1/14!, 1/13!, ..., 1/6!.
constexpr int kLogTerms = 9, kExpTerms = 9;
struct PowSeries {
  double lg[kLogTerms];
  double ex[kExpTerms];
};
constexpr PowSeries kPowSeries = {
    {1.0 / 19, 1.0 / 17, 1.0 / 15, 1.0 / 13, 1.0 / 11, 1.0 / 9, 1.0 / 7, 1.0 / 5, 1.0 / 3},
    {1.0 / 87178291200.0, 1.0 / 6227020800.0, 1.0 / 479001600.0, 1.0 / 39916800.0,
     1.0 / 3628800.0, 1.0 / 362880.0, 1.0 / 40320.0, 1.0 / 5040.0, 1.0 / 720.0}};

struct PhysParams {
  double gamma, p0, R, cp, cv, g, kappa;  // kappa = cp / cv
  double swe_g, dnl;
  const crpow::KappaTable* ktab;  // pow table of kappa (device memory), or null
  // kPowSeries, carried as a kernel argument by the host-launched kernels:
  // the coefficients then sit in scalar registers and each series step is
  // one multiply-add (as literals, each double costs two vector moves)
  PowSeries ser;
};

// Julia's max(a, b) for Float64: NaN if either is NaN, else the larger.
__device__ __forceinline__ double jmax(double a, double b) {
  if (isnan(a) || isnan(b)) return a + b;
  return b > a ? b : a;
}

// ------------------------------------------------------- arithmetic policies
// FastOps: the contracted kernels (CANARY_FP64_FMA, -ffp-contract=fast).
//   recip: the hardware reciprocal estimate refined by two Newton steps
//     r <- r + r (1 - x r), each step squaring the relative error, so the
//     estimate's error falls below the last bit (within an ulp of 1/x).
//   powk: x^k = exp(k log x) with both functions as short series where the
//     arguments of an atmosphere put them (x = R rho theta / p0 near 1):
//     log x = 2 atanh(s) = 2 (s + s^3/3 + s^5/5 + ...), s = (x - 1)/(x + 1),
//     for 1/sqrt2 < x < sqrt2 (|s| < 0.1716: ten terms leave < 2^-55);
//     e^y as its Taylor series to y^14/14! for |y| <= 1/2 (the next term is
//     < 2^-55 of e^y).  Outside those ranges, exp(k log x) by the device
//     library, out of line (pow_by_exp_log).
//     Evaluated by Horner's rule with fused multiply-adds; a few ulp in all.
// exp(k log x) by the device library, out of line: the contracted pow's
// path outside the series ranges (rare for an atmosphere), kept out of the
// stage kernels' instruction stream.
__device__ __attribute__((noinline)) inline double pow_by_exp_log(double x, double k) {
  return exp(k * log(x));
}

struct FastOps {
  __device__ static double recip(double x) {
    double r = __builtin_amdgcn_rcp(x);
    double e = fma(-x, r, 1.0);
    r = fma(r, e, r);
    e = fma(-x, r, 1.0);
    return fma(r, e, r);
  }
  __device__ static double powk(double x, const PhysParams& p) {
    const double k = p.kappa;
    if (!(x > 0.70710678118654752 && x < 1.4142135623730950)) return pow_by_exp_log(x, k);
    const double s = (x - 1) * recip(x + 1);
    const double z = s * s;
    const PowSeries& c = p.ser;
    double a = c.lg[0];
#pragma unroll
    for (int i = 1; i < kLogTerms; ++i) a = fma(a, z, c.lg[i]);
    a = fma(a, z, 1.0);
    const double lx = (2 * s) * a;
    const double y = k * lx;
    if (!(fabs(y) <= 0.5)) return pow_by_exp_log(x, k);
    double b = c.ex[0];
#pragma unroll
    for (int i = 1; i < kExpTerms; ++i) b = fma(b, y, c.ex[i]);
    b = fma(b, y, 1.0 / 120.0);
    b = fma(b, y, 1.0 / 24.0);
    b = fma(b, y, 1.0 / 6.0);
    b = fma(b, y, 0.5);
    b = fma(b, y, 1.0);
    return fma(b, y, 1.0);
  }
};

// ExactOps: the exact-order kernels (compiled with -ffp-contract=off): IEEE
// division, and every expression of the reference in its association with
// no multiply-add fused.  pow is a library routine the reference does not
// fix (Julia's libm pow, glibc's and OCML's differ in their last bits); here
// it is the correctly rounded x^y of crpow.h, so the pressure is a function
// of the state alone and the oracle, whose pow is correctly rounded too,
// gets the same bits (with kappa's table, when the host has made one, the
// same bits come from the cheaper fast phase).
struct ExactOps {
  __device__ static double recip(double x) { return 1 / x; }
  __device__ static double powk(double x, const PhysParams& p) {
    return crpow::pow_cr_tab(x, p.kappa, p.ktab);
  }
};

// ------------------------------------------------------------ 3D Euler set2c
template <class Ops>
struct Euler3DT {
  using Real = double;
  static constexpr int DIM = 3, NS = 5, NSR = 5, NVGEO = 14, NSGEO = 5;
  static constexpr int iMJ = 9, iMJI = 10;
  static constexpr bool HAS_P = true, B_FROM_M = false, HAS_ADVECTION = false;
  enum { U = 0, V = 1, W = 2, RHO = 3, E = 4 };

  __device__ static double pressure(const double* q, const PhysParams& p) {
    return p.p0 * Ops::powk(p.R * q[E] / p.p0, p);
  }
  // met[d + DIM*c] = d(xi_d)/d(x_c) in the vgeo field order.
  __device__ static void volume_flux(const double* q, double P, const double* met, double MJ,
                                     const PhysParams&, double F[DIM][NSR]) {
    const double Uq = q[U], Vq = q[V], Wq = q[W], rho = q[RHO], Eq = q[E];
    const double ri = Ops::recip(rho);
    const double frx = Uq, fUx = ri * Uq * Uq + P, fVx = ri * Vq * Uq, fWx = ri * Wq * Uq, fEx = Eq * ri * Uq;
    const double fry = Vq, fUy = ri * Uq * Vq, fVy = ri * Vq * Vq + P, fWy = ri * Wq * Vq, fEy = Eq * ri * Vq;
    const double frz = Wq, fUz = ri * Uq * Wq, fVz = ri * Vq * Wq, fWz = ri * Wq * Wq + P, fEz = Eq * ri * Wq;
#pragma unroll
    for (int d = 0; d < DIM; ++d) {
      const double mx = met[d], my = met[d + 3], mz = met[d + 6];
      F[d][RHO] = MJ * (mx * frx + my * fry + mz * frz);
      F[d][U] = MJ * (mx * fUx + my * fUy + mz * fUz);
      F[d][V] = MJ * (mx * fVx + my * fVy + mz * fVz);
      F[d][W] = MJ * (mx * fWx + my * fWy + mz * fWz);
      F[d][E] = MJ * (mx * fEx + my * fEy + mz * fEz);
    }
  }
  __device__ static void source(const double* q, double MJ, const PhysParams& p, double* r) {
    r[W] -= MJ * q[RHO] * p.g;
  }
  // Wall state: momentum reflected, rho, rho*theta and P copied.
  __device__ static void reflect(const double* qM, double PM, const double* n, double* qP, double& PP) {
    const double UnM = n[0] * qM[U] + n[1] * qM[V] + n[2] * qM[W];
    qP[U] = qM[U] - 2 * UnM * n[0];
    qP[V] = qM[V] - 2 * UnM * n[1];
    qP[W] = qM[W] - 2 * UnM * n[2];
    qP[RHO] = qM[RHO];
    qP[E] = qM[E];
    PP = PM;
  }
  // out[s] = sMJ * (Rusanov flux)_s, the amount the reference subtracts.
  __device__ static void face_flux(const double* qM, double PM, const double* qP, double PP,
                                   const double* n, double sMJ, const PhysParams& p, double* out) {
    const double nx = n[0], ny = n[1], nz = n[2];
    const double UM = qM[U], VM = qM[V], WM = qM[W], rM = qM[RHO], EM = qM[E];
    const double UP = qP[U], VP = qP[V], WP = qP[W], rP = qP[RHO], EP = qP[E];
    const double rMi = Ops::recip(rM);
    const double frMx = UM, fUMx = rMi * UM * UM + PM, fVMx = rMi * VM * UM, fWMx = rMi * WM * UM, fEMx = rMi * UM * EM;
    const double frMy = VM, fUMy = rMi * UM * VM, fVMy = rMi * VM * VM + PM, fWMy = rMi * WM * VM, fEMy = rMi * VM * EM;
    const double frMz = WM, fUMz = rMi * UM * WM, fVMz = rMi * VM * WM, fWMz = rMi * WM * WM + PM, fEMz = rMi * WM * EM;
    const double rPi = Ops::recip(rP);
    const double frPx = UP, fUPx = rPi * UP * UP + PP, fVPx = rPi * VP * UP, fWPx = rPi * WP * UP, fEPx = rPi * UP * EP;
    const double frPy = VP, fUPy = rPi * UP * VP, fVPy = rPi * VP * VP + PP, fWPy = rPi * WP * VP, fEPy = rPi * VP * EP;
    const double frPz = WP, fUPz = rPi * UP * WP, fVPz = rPi * VP * WP, fWPz = rPi * WP * WP + PP, fEPz = rPi * WP * EP;
    const double lM = rMi * fabs(nx * UM + ny * VM + nz * WM) + sqrt(rMi * p.gamma * PM);
    const double lP = rPi * fabs(nx * UP + ny * VP + nz * WP) + sqrt(rPi * p.gamma * PP);
    const double l = jmax(lM, lP);
    out[RHO] = sMJ * ((nx * (frMx + frPx) + ny * (frMy + frPy) + nz * (frMz + frPz) - l * (rP - rM)) / 2);
    out[U] = sMJ * ((nx * (fUMx + fUPx) + ny * (fUMy + fUPy) + nz * (fUMz + fUPz) - l * (UP - UM)) / 2);
    out[V] = sMJ * ((nx * (fVMx + fVPx) + ny * (fVMy + fVPy) + nz * (fVMz + fVPz) - l * (VP - VM)) / 2);
    out[W] = sMJ * ((nx * (fWMx + fWPx) + ny * (fWMy + fWPy) + nz * (fWMz + fWPz) - l * (WP - WM)) / 2);
    out[E] = sMJ * ((nx * (fEMx + fEPx) + ny * (fEMy + fEPy) + nz * (fEMz + fEPz) - l * (EP - EM)) / 2);
  }
};

struct Euler3D : Euler3DT<ExactOps> {};
struct Euler3DFma : Euler3DT<FastOps> {};

// ------------------------------------------------------------ 2D Euler set2c
template <class Ops>
struct Euler2DT {
  using Real = double;
  static constexpr int DIM = 2, NS = 4, NSR = 4, NVGEO = 8, NSGEO = 4;
  static constexpr int iMJ = 4, iMJI = 5;
  static constexpr bool HAS_P = true, B_FROM_M = false, HAS_ADVECTION = false;
  enum { U = 0, V = 1, RHO = 2, E = 3 };

  __device__ static double pressure(const double* q, const PhysParams& p) {
    return p.p0 * Ops::powk(p.R * q[E] / p.p0, p);
  }
  __device__ static void volume_flux(const double* q, double P, const double* met, double MJ,
                                     const PhysParams&, double F[DIM][NSR]) {
    const double Uq = q[U], Vq = q[V], rho = q[RHO], Eq = q[E];
    const double ri = Ops::recip(rho);
    const double frx = Uq, fUx = ri * Uq * Uq + P, fVx = ri * Vq * Uq, fEx = ri * Uq * Eq;
    const double fry = Vq, fUy = ri * Uq * Vq, fVy = ri * Vq * Vq + P, fEy = ri * Vq * Eq;
#pragma unroll
    for (int d = 0; d < DIM; ++d) {
      const double mx = met[d], my = met[d + 2];
      F[d][RHO] = MJ * (mx * frx + my * fry);
      F[d][U] = MJ * (mx * fUx + my * fUy);
      F[d][V] = MJ * (mx * fVx + my * fVy);
      F[d][E] = MJ * (mx * fEx + my * fEy);
    }
  }
  __device__ static void source(const double* q, double MJ, const PhysParams& p, double* r) {
    r[V] -= MJ * q[RHO] * p.g;
  }
  __device__ static void reflect(const double* qM, double PM, const double* n, double* qP, double& PP) {
    const double UnM = n[0] * qM[U] + n[1] * qM[V];
    qP[U] = qM[U] - 2 * UnM * n[0];
    qP[V] = qM[V] - 2 * UnM * n[1];
    qP[RHO] = qM[RHO];
    qP[E] = qM[E];
    PP = PM;
  }
  __device__ static void face_flux(const double* qM, double PM, const double* qP, double PP,
                                   const double* n, double sMJ, const PhysParams& p, double* out) {
    const double nx = n[0], ny = n[1];
    const double UM = qM[U], VM = qM[V], rM = qM[RHO], EM = qM[E];
    const double UP = qP[U], VP = qP[V], rP = qP[RHO], EP = qP[E];
    const double rMi = Ops::recip(rM);
    const double frMx = UM, fUMx = rMi * UM * UM + PM, fVMx = rMi * VM * UM, fEMx = rMi * UM * EM;
    const double frMy = VM, fUMy = rMi * UM * VM, fVMy = rMi * VM * VM + PM, fEMy = rMi * VM * EM;
    const double rPi = Ops::recip(rP);
    const double frPx = UP, fUPx = rPi * UP * UP + PP, fVPx = rPi * VP * UP, fEPx = rPi * UP * EP;
    const double frPy = VP, fUPy = rPi * UP * VP, fVPy = rPi * VP * VP + PP, fEPy = rPi * VP * EP;
    const double lM = rMi * fabs(nx * UM + ny * VM) + sqrt(rMi * p.gamma * PM);
    const double lP = rPi * fabs(nx * UP + ny * VP) + sqrt(rPi * p.gamma * PP);
    const double l = jmax(lM, lP);
    out[RHO] = sMJ * ((nx * (frMx + frPx) + ny * (frMy + frPy) + -(l * (rP - rM))) / 2);
    out[U] = sMJ * ((nx * (fUMx + fUPx) + ny * (fUMy + fUPy) + -(l * (UP - UM))) / 2);
    out[V] = sMJ * ((nx * (fVMx + fVPx) + ny * (fVMy + fVPy) + -(l * (VP - VM))) / 2);
    out[E] = sMJ * ((nx * (fEMx + fEPx) + ny * (fEMy + fEPy) + -(l * (EP - EM))) / 2);
  }
};

struct Euler2D : Euler2DT<ExactOps> {};
struct Euler2DFma : Euler2DT<FastOps> {};

// ------------------------------------------- 2D Euler, total energy (set3c)
// Q = (U, V, rho, E) with E the total energy (internal + kinetic +
// potential); the pressure needs the node's height y (vgeo field 7):
// P = (R/cv) (E - (U^2 + V^2) / (2 rho) - rho g y) (euler2d_set3c.jl:226,
// :308, :320), the energy flux is U (E + P) / rho (:233-238).  The face
// kernel reads y at the P node through vgeo (:318-320).
template <class Ops>
struct Euler2DTET {
  using Real = double;
  static constexpr int DIM = 2, NS = 4, NSR = 4, NVGEO = 8, NSGEO = 4;
  static constexpr int iMJ = 4, iMJI = 5, iY = 7;
  static constexpr bool HAS_P = true, B_FROM_M = false, HAS_ADVECTION = false, NEEDS_Y = true;
  enum { U = 0, V = 1, RHO = 2, E = 3 };

  __device__ static double pressure_y(const double* q, double y, const PhysParams& p) {
    const double Uq = q[U], Vq = q[V], rho = q[RHO];
    return (p.R / p.cv) * (q[E] - (Uq * Uq + Vq * Vq) / (2 * rho) - rho * p.g * y);
  }
  __device__ static void volume_flux(const double* q, double P, const double* met, double MJ,
                                     const PhysParams&, double F[DIM][NSR]) {
    const double Uq = q[U], Vq = q[V], rho = q[RHO], Eq = q[E];
    const double ri = Ops::recip(rho);
    const double EP = Eq + P;
    const double frx = Uq, fUx = ri * Uq * Uq + P, fVx = ri * Vq * Uq, fEx = ri * Uq * EP;
    const double fry = Vq, fUy = ri * Uq * Vq, fVy = ri * Vq * Vq + P, fEy = ri * Vq * EP;
#pragma unroll
    for (int d = 0; d < DIM; ++d) {
      const double mx = met[d], my = met[d + 2];
      F[d][RHO] = MJ * (mx * frx + my * fry);
      F[d][U] = MJ * (mx * fUx + my * fUy);
      F[d][V] = MJ * (mx * fVx + my * fVy);
      F[d][E] = MJ * (mx * fEx + my * fEy);
    }
  }
  __device__ static void source(const double* q, double MJ, const PhysParams& p, double* r) {
    r[V] -= MJ * q[RHO] * p.g;
  }
  __device__ static void reflect(const double* qM, double PM, const double* n, double* qP, double& PP) {
    const double UnM = n[0] * qM[U] + n[1] * qM[V];
    qP[U] = qM[U] - 2 * UnM * n[0];
    qP[V] = qM[V] - 2 * UnM * n[1];
    qP[RHO] = qM[RHO];
    qP[E] = qM[E];
    PP = PM;
  }
  __device__ static void face_flux(const double* qM, double PM, const double* qP, double PP,
                                   const double* n, double sMJ, const PhysParams& p, double* out) {
    const double nx = n[0], ny = n[1];
    const double UM = qM[U], VM = qM[V], rM = qM[RHO], EM = qM[E];
    const double UP = qP[U], VP = qP[V], rP = qP[RHO], EP = qP[E];
    const double rMi = Ops::recip(rM);
    const double EPM = EM + PM, EPP = EP + PP;
    const double frMx = UM, fUMx = rMi * UM * UM + PM, fVMx = rMi * VM * UM, fEMx = rMi * UM * EPM;
    const double frMy = VM, fUMy = rMi * UM * VM, fVMy = rMi * VM * VM + PM, fEMy = rMi * VM * EPM;
    const double rPi = Ops::recip(rP);
    const double frPx = UP, fUPx = rPi * UP * UP + PP, fVPx = rPi * VP * UP, fEPx = rPi * UP * EPP;
    const double frPy = VP, fUPy = rPi * UP * VP, fVPy = rPi * VP * VP + PP, fEPy = rPi * VP * EPP;
    const double lM = rMi * fabs(nx * UM + ny * VM) + sqrt(rMi * p.gamma * PM);
    const double lP = rPi * fabs(nx * UP + ny * VP) + sqrt(rPi * p.gamma * PP);
    const double l = jmax(lM, lP);
    out[RHO] = sMJ * ((nx * (frMx + frPx) + ny * (frMy + frPy) + -(l * (rP - rM))) / 2);
    out[U] = sMJ * ((nx * (fUMx + fUPx) + ny * (fUMy + fUPy) + -(l * (UP - UM))) / 2);
    out[V] = sMJ * ((nx * (fVMx + fVPx) + ny * (fVMy + fVPy) + -(l * (VP - VM))) / 2);
    out[E] = sMJ * ((nx * (fEMx + fEPx) + ny * (fEMy + fEPy) + -(l * (EP - EM))) / 2);
  }
};

struct Euler2DTE : Euler2DTET<ExactOps> {};
struct Euler2DTEFma : Euler2DTET<FastOps> {};

// Does the pressure of this equation set need the node's height y?
template <class Phys, class = void>
struct NeedsY {
  static constexpr bool value = false;
};
template <class Phys>
struct NeedsY<Phys, decltype((void)Phys::NEEDS_Y)> {
  static constexpr bool value = Phys::NEEDS_Y;
};

// The pressure at a node of height y (ignored where the pressure does not
// depend on it).
template <class Phys>
__device__ __forceinline__ typename Phys::Real pressure_at(const typename Phys::Real* q,
                                                           typename Phys::Real y,
                                                           const PhysParams& p) {
  if constexpr (NeedsY<Phys>::value)
    return Phys::pressure_y(q, y, p);
  else
    return Phys::pressure(q, p);
}

// ------------------------------------------------------------ shallow water
// Q = (H u, H v, h, b); b (bathymetry / still depth) has no tendency.
struct SWE2D {
  using Real = double;
  static constexpr int DIM = 2, NS = 4, NSR = 3, NVGEO = 8, NSGEO = 4;
  static constexpr int iMJ = 4, iMJI = 5;
  // B_FROM_M: the reference face kernel reads the P-side b from element eM
  // (shallow_water.jl:202); kept for parity, harmless for a continuous b.
  static constexpr bool HAS_P = false, B_FROM_M = true, HAS_ADVECTION = true;
  enum { U = 0, V = 1, H = 2, B = 3 };

  __device__ static double pressure(const double*, const PhysParams&) { return 0.0; }
  __device__ static void volume_flux(const double* q, double, const double* met, double MJ,
                                     const PhysParams& p, double F[DIM][NSR]) {
    const double g = p.swe_g, dnl = p.dnl;
    const double Uq = q[U], Vq = q[V], h = q[H], b = q[B];
    const double Ht = h + b;
    const double u = Uq / Ht, v = Vq / Ht;
    const double fhx = Uq, fhy = Vq;
    const double fUx = (Ht * u * u + 0.5 * g * (h * h)) * dnl + g * h * b;
    const double fUy = (Ht * u * v) * dnl;
    const double fVx = (Ht * v * u) * dnl;
    const double fVy = (Ht * v * v + 0.5 * g * (h * h)) * dnl + g * h * b;
#pragma unroll
    for (int d = 0; d < DIM; ++d) {
      const double mx = met[d], my = met[d + 2];
      F[d][H] = MJ * (mx * fhx + my * fhy);
      F[d][U] = MJ * (mx * fUx + my * fUy);
      F[d][V] = MJ * (mx * fVx + my * fVy);
    }
  }
  __device__ static void source(const double*, double, const PhysParams&, double*) {}
  __device__ static void reflect(const double* qM, double, const double* n, double* qP, double& PP) {
    const double UnM = n[0] * qM[U] + n[1] * qM[V];
    qP[U] = qM[U] - 2 * UnM * n[0];
    qP[V] = qM[V] - 2 * UnM * n[1];
    qP[H] = qM[H];
    qP[B] = qM[B];
    PP = 0.0;
  }
  __device__ static void face_flux(const double* qM, double, const double* qP, double,
                                   const double* n, double sMJ, const PhysParams& p, double* out) {
    const double g = p.swe_g, dnl = p.dnl;
    const double nx = n[0], ny = n[1];
    const double hM = qM[H], bM = qM[B], UM = qM[U], VM = qM[V];
    const double hP = qP[H], bP = qP[B], UP = qP[U], VP = qP[V];
    const double HM = hM + bM, uM = UM / HM, vM = VM / HM;
    const double HP = hP + bP, uP = UP / HP, vP = VP / HP;
    const double fhMx = UM, fhMy = VM;
    const double fUMx = (HM * uM * uM + 0.5 * g * (hM * hM)) * dnl + g * hM * bM;
    const double fUMy = HM * uM * vM * dnl;
    const double fVMx = HM * vM * uM * dnl;
    const double fVMy = (HM * vM * vM + 0.5 * g * (hM * hM)) * dnl + g * hM * bM;
    const double fhPx = UP, fhPy = VP;
    const double fUPx = (HP * uP * uP + 0.5 * g * (hP * hP)) * dnl + g * hP * bP;
    const double fUPy = HP * uP * vP * dnl;
    const double fVPx = HP * vP * uP * dnl;
    const double fVPy = (HP * vP * vP + 0.5 * g * (hP * hP)) * dnl + g * hP * bP;
    const double lM = (fabs(nx * uM + ny * vM) + sqrt(g * HM)) * dnl + (sqrt(g * bM)) * (1.0 - dnl);
    const double lP = (fabs(nx * uP + ny * vP) + sqrt(g * HP)) * dnl + (sqrt(g * bP)) * (1.0 - dnl);
    const double l = jmax(lM, lP);
    out[H] = sMJ * ((nx * (fhMx + fhPx) + ny * (fhMy + fhPy) + -(l * (hP - hM))) / 2);
    out[U] = sMJ * ((nx * (fUMx + fUPx) + ny * (fUMy + fUPy) + -(l * (UP - UM))) / 2);
    out[V] = sMJ * ((nx * (fVMx + fVPx) + ny * (fVMy + fVPy) + -(l * (VP - VM))) / 2);
  }
};

// The same physics for the contracted unit: a type of its own, so that the
// kernels instantiated there get their own names.  Kernel names are shared by
// every unit that instantiates a template with the same arguments, and the
// host launch stub is one symbol; with one type for both units, launches
// from the contracted unit ran the exact-order unit's code.
struct SWE2DFma : SWE2D {};

// --------------------------------------------- shallow water, FP32 variant
// The same fluxes as SWE2D evaluated in FP32 (batched FP32 ensembles).
// b = still depth is read only and h is already the deviation from it, so
// FP32 keeps h's relative precision.
struct SWE2DF32 {
  using Real = float;
  static constexpr int DIM = 2, NS = 4, NSR = 3, NVGEO = 8, NSGEO = 4;
  static constexpr int iMJ = 4, iMJI = 5;
  static constexpr bool HAS_P = false, B_FROM_M = true, HAS_ADVECTION = true;
  enum { U = 0, V = 1, H = 2, B = 3 };

  __device__ static float pressure(const float*, const PhysParams&) { return 0.f; }
  __device__ static void volume_flux(const float* q, float, const float* met, float MJ,
                                     const PhysParams& p, float F[DIM][NSR]) {
    const float g = (float)p.swe_g, dnl = (float)p.dnl;
    const float Uq = q[U], Vq = q[V], h = q[H], b = q[B];
    const float Ht = h + b;
    const float u = Uq / Ht, v = Vq / Ht;
    const float fUx = (Ht * u * u + 0.5f * g * (h * h)) * dnl + g * h * b;
    const float fUy = (Ht * u * v) * dnl;
    const float fVx = (Ht * v * u) * dnl;
    const float fVy = (Ht * v * v + 0.5f * g * (h * h)) * dnl + g * h * b;
#pragma unroll
    for (int d = 0; d < DIM; ++d) {
      const float mx = met[d], my = met[d + 2];
      F[d][H] = MJ * (mx * Uq + my * Vq);
      F[d][U] = MJ * (mx * fUx + my * fUy);
      F[d][V] = MJ * (mx * fVx + my * fVy);
    }
  }
  __device__ static void source(const float*, float, const PhysParams&, float*) {}
  __device__ static void reflect(const float* qM, float, const float* n, float* qP, float& PP) {
    const float UnM = n[0] * qM[U] + n[1] * qM[V];
    qP[U] = qM[U] - 2.f * UnM * n[0];
    qP[V] = qM[V] - 2.f * UnM * n[1];
    qP[H] = qM[H];
    qP[B] = qM[B];
    PP = 0.f;
  }
  __device__ static void face_flux(const float* qM, float, const float* qP, float,
                                   const float* n, float sMJ, const PhysParams& p, float* out) {
    const float g = (float)p.swe_g, dnl = (float)p.dnl;
    const float nx = n[0], ny = n[1];
    const float hM = qM[H], bM = qM[B], UM = qM[U], VM = qM[V];
    const float hP = qP[H], bP = qP[B], UP = qP[U], VP = qP[V];
    const float HM = hM + bM, uM = UM / HM, vM = VM / HM;
    const float HP = hP + bP, uP = UP / HP, vP = VP / HP;
    const float fUMx = (HM * uM * uM + 0.5f * g * (hM * hM)) * dnl + g * hM * bM;
    const float fUMy = HM * uM * vM * dnl;
    const float fVMx = HM * vM * uM * dnl;
    const float fVMy = (HM * vM * vM + 0.5f * g * (hM * hM)) * dnl + g * hM * bM;
    const float fUPx = (HP * uP * uP + 0.5f * g * (hP * hP)) * dnl + g * hP * bP;
    const float fUPy = HP * uP * vP * dnl;
    const float fVPx = HP * vP * uP * dnl;
    const float fVPy = (HP * vP * vP + 0.5f * g * (hP * hP)) * dnl + g * hP * bP;
    const float lM = (fabsf(nx * uM + ny * vM) + sqrtf(g * HM)) * dnl + sqrtf(g * bM) * (1.f - dnl);
    const float lP = (fabsf(nx * uP + ny * vP) + sqrtf(g * HP)) * dnl + sqrtf(g * bP) * (1.f - dnl);
    const float l = fmaxf(lM, lP);
    out[H] = sMJ * ((nx * (UM + UP) + ny * (VM + VP) - l * (hP - hM)) * 0.5f);
    out[U] = sMJ * ((nx * (fUMx + fUPx) + ny * (fUMy + fUPy) - l * (UP - UM)) * 0.5f);
    out[V] = sMJ * ((nx * (fVMx + fVPx) + ny * (fVMy + fVPy) - l * (VP - VM)) * 0.5f);
  }
};

}  // namespace canary
