// Correctly rounded x^y in double precision for the exact-order kernels.
//
// The pressure P = p0 (R rho theta / p0)^(cp/cv) of the Euler sets needs a
// pow; the reference calls Julia's libm pow (euler3d_set2c.jl:175, :262,
// :270), whose last bit no test fixes.  Correct rounding makes the result a
// function of (x, y) alone, so any two correctly rounded implementations --
// this one on the device, the oracle's quad-precision one on the host --
// give the same bits.
//
// Method: x = 2^E m, m*r - 1 = t exact (one FMA, table r ~ 1/m, |t| <= 2^-7),
// log x = E ln2 - log r + log1p(t) (Tang's table-driven logarithm, ACM TOMS
// 16(4), 1990); z = y log x; e^z = 2^(k/128) e^s with |s| <= ln2/256 (Tang's
// table-driven exponential, ACM TOMS 15(2), 1989); sums and products carried
// as double-double values with the error-free transformations of Knuth
// (TwoSum), Dekker (FastTwoSum) and the FMA product.  Ziv's strategy ("Fast
// evaluation of elementary mathematical functions with correctly rounded
// last bit", ACM TOMS 17(3), 1991) decides the rounding: a fast phase with
// relative error below (1 + |y|) 2^-71 returns when both ends of its error
// interval round to the same double (for y = cp/cv all but about 1 call in
// 2^16); otherwise an accurate phase with error below
// (1 + |y| + |y log x|) 2^-100 repeats the test, and returns its own
// rounding if even that fails (about 1 call in 2^45).
//
// Claim: correctly rounded for positive finite x (subnormals included) and
// finite y whenever |y log x| < 700 (the result is then a normal double),
// except exactly representable results and rounding midpoints (the
// accurate phase cannot separate those; they need y with a small power of
// two denominator, which the ratio cp/cv rounded to a double is not).
// Outside that domain the device library pow is returned.
//
// Every operation is written out: no multiply-add may be contracted here,
// and the explicit fma() calls are the only fused ones.
#pragma once
#include <math.h>
#include <stdint.h>
#include <string.h>

#include "crpow_tables.h"

#ifndef CANARY_HD
#if defined(__HIPCC__)
#define CANARY_HD __host__ __device__
#else
#define CANARY_HD
#endif
#endif

namespace canary {
namespace crpow {

CANARY_HD inline uint64_t as_bits(double x) {
  uint64_t u;
  memcpy(&u, &x, sizeof u);
  return u;
}
CANARY_HD inline double from_bits(uint64_t u) {
  double x;
  memcpy(&x, &u, sizeof x);
  return x;
}

// s + e = a + b exactly (Knuth)
CANARY_HD inline void two_sum(double a, double b, double& s, double& e) {
#pragma STDC FP_CONTRACT OFF
  s = a + b;
  const double bb = s - a;
  e = (a - (s - bb)) + (b - bb);
}
// s + e = a + b exactly when |a| >= |b| or a == 0 (Dekker)
CANARY_HD inline void fast_two_sum(double a, double b, double& s, double& e) {
#pragma STDC FP_CONTRACT OFF
  s = a + b;
  e = b - (s - a);
}

struct DD {
  double hi, lo;
};

// a * b for a double-double a and a double b (relative error ~2^-104)
CANARY_HD inline DD dd_mul_d(DD a, double b) {
#pragma STDC FP_CONTRACT OFF
  const double p = a.hi * b;
  double e = fma(a.hi, b, -p);
  e = e + a.lo * b;
  DD r;
  fast_two_sum(p, e, r.hi, r.lo);
  return r;
}
// a * b for double-doubles
CANARY_HD inline DD dd_mul(DD a, DD b) {
#pragma STDC FP_CONTRACT OFF
  const double p = a.hi * b.hi;
  double e = fma(a.hi, b.hi, -p);
  e = e + (a.hi * b.lo + a.lo * b.hi);
  DD r;
  fast_two_sum(p, e, r.hi, r.lo);
  return r;
}
// a + b for double-doubles (the accurate sum of two double-doubles)
CANARY_HD inline DD dd_add(DD a, DD b) {
#pragma STDC FP_CONTRACT OFF
  double s, e, t, f;
  two_sum(a.hi, b.hi, s, e);
  two_sum(a.lo, b.lo, t, f);
  e = e + t;
  fast_two_sum(s, e, s, e);
  e = e + f;
  DD r;
  fast_two_sum(s, e, r.hi, r.lo);
  return r;
}

// Ziv's rounding test: the double nearest to every value within u of
// hi + lo (|lo| <= ulp(hi)/2), or 0 if they round differently.
CANARY_HD inline bool round_test(double hi, double lo, double u, double& out) {
#pragma STDC FP_CONTRACT OFF
  const double a = hi + (lo + u), b = hi + (lo - u);
  out = a;
  return a == b;
}

// Accurate phase: x = 2^E (m as reduced), t = m r - 1 exact, log table entry L.
// Returns 2^-scale x^y as hi + lo; `err` bounds its relative error:
// (1 + |y| + |y log x|) 2^-100.  (The double-double logarithm is good to
// about 2^-104 of |log x| plus 2^-106, which y multiplies into the
// exponent; everything after adds about 2^-104.)
CANARY_HD inline DD pow_accurate(double t, int E, const LogEntry& L, double y, int& scale,
                                 double& err) {
#pragma STDC FP_CONTRACT OFF
  // log1p(t) = t + t^2 A(t), A = -1/2 + t/3 - t^2/4 + ... - t^12/14 (|t| <= 2^-7:
  // the first dropped term is below 2^-108)
  double a = -1.0 / 14;
  a = fma(a, t, 1.0 / 13);
  a = fma(a, t, -1.0 / 12);
  a = fma(a, t, 1.0 / 11);
  a = fma(a, t, -1.0 / 10);
  a = fma(a, t, 1.0 / 9);
  a = fma(a, t, -1.0 / 8);
  DD A = {a, 0.0};
  A = dd_add(DD{kInv7Hi, kInv7Lo}, dd_mul_d(A, t));
  A = dd_add(DD{-kInv6Hi, -kInv6Lo}, dd_mul_d(A, t));
  A = dd_add(DD{kInv5Hi, kInv5Lo}, dd_mul_d(A, t));
  A = dd_add(DD{-0.25, 0.0}, dd_mul_d(A, t));
  A = dd_add(DD{kInv3Hi, kInv3Lo}, dd_mul_d(A, t));
  A = dd_add(DD{-0.5, 0.0}, dd_mul_d(A, t));
  const DD l1 = dd_add(DD{t, 0.0}, dd_mul_d(dd_mul_d(A, t), t));
  // E ln2 - log r + log1p(t)
  const double Ed = (double)E;
  double ph, pe;
  ph = Ed * kLn2Mid;
  pe = fma(Ed, kLn2Mid, -ph);
  DD eln2;
  fast_two_sum(Ed * kLn2Hi, ph, eln2.hi, eln2.lo);
  eln2.lo = eln2.lo + (pe + Ed * kLn2Lo);
  const DD lg = dd_add(dd_add(eln2, DD{L.lhi, L.llo}), l1);
  // z = y log x
  DD z;
  z.hi = y * lg.hi;
  z.lo = fma(y, lg.hi, -z.hi) + y * lg.lo;
  fast_two_sum(z.hi, z.lo, z.hi, z.lo);
  // z = k ln2/128 + s
  const double kd = rint(z.hi * kInvC);
  const double s0 = fma(-kd, kCHi, z.hi);  // exact
  double mh = -kd * kCMid;
  double me = fma(-kd, kCMid, -mh);
  DD s = dd_add(DD{s0, 0.0}, DD{mh, me});
  s = dd_add(s, DD{z.lo, -kd * kCLo});
  // e^s.hi - 1 = sh (1 + sh (1/2 + sh (1/6 + ...))) to sh^10/10! (|sh| <= 2^-8.5:
  // the first dropped term is below 2^-114); e^s.lo = 1 + s.lo to 2^-120
  const double sh = s.hi;
  double b = 1.0 / 3628800.0;
  b = fma(b, sh, 1.0 / 362880.0);
  b = fma(b, sh, 1.0 / 40320.0);
  b = fma(b, sh, 1.0 / 5040.0);
  b = fma(b, sh, 1.0 / 720.0);
  DD B = {b, 0.0};
  B = dd_add(DD{kFact5Hi, kFact5Lo}, dd_mul_d(B, sh));
  B = dd_add(DD{kFact4Hi, kFact4Lo}, dd_mul_d(B, sh));
  B = dd_add(DD{kFact3Hi, kFact3Lo}, dd_mul_d(B, sh));
  B = dd_add(DD{0.5, 0.0}, dd_mul_d(B, sh));
  B = dd_add(DD{1.0, 0.0}, dd_mul_d(B, sh));
  const DD em1 = dd_mul_d(B, sh);
  // (1 + em1)(1 + s.lo)
  DD ex = dd_add(DD{1.0, 0.0}, em1);
  ex = dd_add(ex, DD{s.lo, s.lo * em1.hi});
  const int k = (int)kd;
  scale = k >> 7;
  err = ((1.0 + fabs(y)) + fabs(z.hi)) * 0x1p-100;
  const ExpEntry& T = kExp[k & 127];
  return dd_mul(DD{T.hi, T.lo}, ex);
}

// 2^k for a normal exponent k
CANARY_HD inline double exp2i(int k) { return from_bits((uint64_t)(k + 1023) << 52); }

// Reduction of a positive finite x: x = 2^E m, t = m r - 1 exact for the
// table entry `cell` (the cell of the fraction's top 7 bits; cells from 53
// on use m/2 and E + 1, so m lies in [1/sqrt2, sqrt2)).
struct Reduced {
  double t;
  int E, cell;
};
CANARY_HD inline Reduced reduce(double x) {
#pragma STDC FP_CONTRACT OFF
  uint64_t ix = as_bits(x);
  int E = 0;
  if (ix < 0x0010000000000000ull) {  // subnormal: scale into the normal range
    ix = as_bits(x * 0x1p54);
    E = -54;
  }
  const int i = (int)((ix >> 45) & 127);
  const bool upper = i >= 53;
  E += (int)(ix >> 52) - 1023 + (upper ? 1 : 0);
  const double m =
      from_bits((ix & 0x000fffffffffffffull) | (upper ? 0x3fe0000000000000ull : 0x3ff0000000000000ull));
  return Reduced{fma(m, kLog[i].r, -1.0), E, i};  // exact
}

// Fast phase: 2^-scale x^y as hi + lo; `err` bounds its relative error:
// (1 + |y|) 2^-71.  (The logarithm's absolute error stays below 2^-73, and
// y multiplies it into the exponent; the exponential adds below 2^-73.)
// Returns false when |y log x| >= 700 (outside the claim).
CANARY_HD inline bool pow_fast(const Reduced& rd, double y, DD& R, int& scale, double& err) {
#pragma STDC FP_CONTRACT OFF
  const double t = rd.t;
  const LogEntry& L = kLog[rd.cell];
  // log1p(t) = t - t^2/2 + t^3 c(t), c to t^10/10 (the first dropped term,
  // t^11/11, is below 2^-80)
  double c = 0.1;
  c = fma(c, t, -1.0 / 9);
  c = fma(c, t, 0.125);
  c = fma(c, t, -1.0 / 7);
  c = fma(c, t, 1.0 / 6);
  c = fma(c, t, -0.2);
  c = fma(c, t, 0.25);
  c = fma(c, t, -1.0 / 3);
  c = -c;  // 1/3 - t/4 + t^2/5 - ...
  const double t2 = t * t;
  const double t2e = fma(t, t, -t2);
  double lh, ll;
  fast_two_sum(t, -0.5 * t2, lh, ll);
  ll = (ll - 0.5 * t2e) + t2 * (t * c);
  // log x = E ln2 - log r + log1p(t)
  const double Ed = (double)rd.E;
  double ah, al, bh, bl;
  two_sum(Ed * kLn2Hi, L.lhi, ah, al);
  two_sum(ah, lh, bh, bl);
  const double lo = ((al + bl) + (L.llo + Ed * kLn2Mid)) + ll;
  double gh, gl;
  fast_two_sum(bh, lo, gh, gl);
  // z = y log x
  const double zh = y * gh;
  const double zl = fma(y, gh, -zh) + y * gl;
  if (!(fabs(zh) < 700.0)) return false;
  // z = k ln2/128 + s, |s| <= ln2/256 (1 + 2^-40)
  const double kd = rint(zh * kInvC);
  const double s0 = fma(-kd, kCHi, zh);   // exact
  const double s1 = fma(-kd, kCMid, zl);  // below 2^-25
  double sh, sl;
  two_sum(s0, s1, sh, sl);
  // e^s - 1 = sh + sh^2/2 + sh^3 p(sh) + sl (1 + sh), p to sh^7/7! (the
  // first dropped term is below 2^-83)
  double p = 1.0 / 5040;
  p = fma(p, sh, 1.0 / 720);
  p = fma(p, sh, 1.0 / 120);
  p = fma(p, sh, 1.0 / 24);
  p = fma(p, sh, 1.0 / 6);
  const double q = sh * sh;
  const double qe = fma(sh, sh, -q);
  double e1, e2;
  fast_two_sum(sh, 0.5 * q, e1, e2);
  const double el = ((e2 + sl) + (0.5 * qe + sh * sl)) + q * (sh * p);
  double f1, f2;
  fast_two_sum(1.0, e1, f1, f2);
  f2 = f2 + el;
  const int k = (int)kd;
  const ExpEntry& T = kExp[k & 127];
  const double ph = T.hi * f1;
  const double pl = fma(T.hi, f1, -ph) + (T.hi * f2 + T.lo * f1);
  fast_two_sum(ph, pl, R.hi, R.lo);
  scale = k >> 7;
  err = (1.0 + fabs(y)) * 0x1p-71;
  return true;
}

CANARY_HD inline double pow_cr(double x, double y) {
#pragma STDC FP_CONTRACT OFF
  if (!(x > 0.0) || !(x <= 1.7976931348623157e308) || !(fabs(y) <= 1.7976931348623157e308))
    return pow(x, y);
  const Reduced rd = reduce(x);
  DD R;
  int k = 0;
  double err;
  if (!pow_fast(rd, y, R, k, err)) return pow(x, y);
  double out;
  if (round_test(R.hi, R.lo, R.hi * err, out)) return out * exp2i(k);
  // accurate phase
  R = pow_accurate(rd.t, rd.E, kLog[rd.cell], y, k, err);
  round_test(R.hi, R.lo, R.hi * err, out);
  return out * exp2i(k);
}

// ------------------------------------------------ one gas's exponent
// For y = kappa of one gas (cp/cv in [1, 2]) and x in [1/sqrt2, sqrt2) --
// every pressure of an atmosphere -- the reduction above has E = 0 and
//   x^kappa = r^-kappa (1 + t)^kappa = r^-kappa sum_k binom(kappa, k) t^k,
// so with r^-kappa tabulated per cell (double-double, from the accurate
// phase) the fast phase needs no logarithm, exponential or second table:
// the binomial series to t^9 (the first dropped term, binom(kappa, 10) t^10,
// is below 2^-75 for kappa in [1, 2], |t| <= 2^-7) plus one double-double
// product.  Its relative error stays below 2^-73; the rounding test uses
// 2^-70.  The table is built once per kappa (kappa_table_build).
struct KappaTable {
  double kappa;
  double b2lo;  // binom(kappa, 2) = b[0] + b2lo exactly
  double b[8];  // binom(kappa, k) for k = 2..9 (b[0]: high part of k = 2)
  DD T[128];    // r^-kappa of each cell's r
};

// Fills `kt` for `kappa`; false (and kt unusable) outside kappa in [1, 2].
CANARY_HD inline bool kappa_table_build(double kappa, KappaTable* kt) {
#pragma STDC FP_CONTRACT OFF
  if (!(kappa >= 1.0 && kappa <= 2.0)) return false;
  kt->kappa = kappa;
  const double km1 = kappa - 1.0;  // exact for kappa in [1, 2]
  const double p = kappa * km1, pe = fma(kappa, km1, -p);
  kt->b[0] = 0.5 * p;  // binom(kappa, 2) = kappa (kappa - 1) / 2, exactly hi + lo
  kt->b2lo = 0.5 * pe;
  double bk = kt->b[0] + kt->b2lo;
  for (int k = 3; k <= 9; ++k) {
    bk = bk * (kappa - (double)(k - 1)) / (double)k;
    kt->b[k - 2] = bk;
  }
  for (int i = 0; i < 128; ++i) {
    const double r = kLog[i].r;
    if (r == 1.0) {
      kt->T[i] = DD{1.0, 0.0};
      continue;
    }
    const Reduced rd = reduce(r);
    int sc = 0;
    double err;
    DD v = pow_accurate(rd.t, rd.E, kLog[rd.cell], -kappa, sc, err);
    const double f = exp2i(sc);
    kt->T[i] = DD{v.hi * f, v.lo * f};
  }
  return true;
}

// Fast phase for y = kt.kappa and rd.E == 0: x^kappa as hi + lo.
CANARY_HD inline DD pow_kappa_fast(const Reduced& rd, const KappaTable& kt) {
#pragma STDC FP_CONTRACT OFF
  const double t = rd.t, kappa = kt.kappa;
  // S = binom(kappa, 3) + t binom(kappa, 4) + ... + t^6 binom(kappa, 9)
  double S = kt.b[7];
  S = fma(S, t, kt.b[6]);
  S = fma(S, t, kt.b[5]);
  S = fma(S, t, kt.b[4]);
  S = fma(S, t, kt.b[3]);
  S = fma(S, t, kt.b[2]);
  S = fma(S, t, kt.b[1]);
  // H = binom(kappa, 2) + t S as Hh + Hl
  double Hh, Hl;
  two_sum(kt.b[0], t * S, Hh, Hl);
  Hl = Hl + kt.b2lo;
  // (1 + t)^kappa - 1 = kappa t + t^2 H: both products exactly, summed
  const double p1 = kappa * t, e1 = fma(kappa, t, -p1);
  const double q = t * t, qe = fma(t, t, -q);
  const double u = q * Hh, ue = fma(q, Hh, -u);
  double s, se;
  fast_two_sum(p1, u, s, se);  // |kappa t| >= t^2 >= |u|
  const double lo = ((se + e1) + ue) + (q * Hl + qe * Hh);
  double B1, B2;
  fast_two_sum(1.0, s, B1, B2);
  B2 = B2 + lo;
  const DD T = kt.T[rd.cell];
  const double ph = T.hi * B1;
  const double pl = fma(T.hi, B1, -ph) + (T.hi * B2 + T.lo * B1);
  DD R;
  fast_two_sum(ph, pl, R.hi, R.lo);
  return R;
}

// Everything but the table's fast phase, out of line: the general fast
// phase, the accurate phase and the out-of-domain fallback hold most of the
// code, and inlined at every pressure they made the exact-order stage
// kernel several times its contracted twin's size (instruction-cache
// misses); with a table they run for about one call in 2^16.
__attribute__((noinline)) CANARY_HD inline double pow_cr_slow(double x, double y) {
  return pow_cr(x, y);
}

// pow_cr with the table of one exponent: the same correctly rounded result,
// by the cheaper fast phase when y == kt->kappa and x lies in
// [1/sqrt2, sqrt2) (kt may be null).
CANARY_HD inline double pow_cr_tab(double x, double y, const KappaTable* kt) {
#pragma STDC FP_CONTRACT OFF
  if (kt && y == kt->kappa && x >= 0x1.6a09e667f3bcdp-1 && x < 0x1.6a09e667f3bcdp+0) {
    // 1/sqrt2 <= x < sqrt2 (both bounds rounded inward of the cells' edges)
    // is positive, normal and reduces with E = 0
    const Reduced rd = reduce(x);
    if (rd.E == 0) {
      const DD R = pow_kappa_fast(rd, *kt);
      double out;
      if (round_test(R.hi, R.lo, R.hi * 0x1p-70, out)) return out;
    }
  }
  return pow_cr_slow(x, y);
}

}  // namespace crpow
}  // namespace canary
