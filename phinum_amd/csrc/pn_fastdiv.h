// pn_fastdiv.h -- correctly rounded FP64 division by a divisor fixed for a
// whole run, in four full-rate FMA-pipe operations instead of the hardware
// division sequence (two v_div_scale, a quarter-rate v_rcp, five FMAs,
// v_div_fmas and v_div_fixup).
//
// Method, for a/b with zh = RN(1/b) and zl = RN(1/b - zh) precomputed:
//   1. q = RN(a*zh + RN(a*zl))      Brisebarre, Muller, Raina, "Accelerating
//      correctly rounded floating-point division when the divisor is known in
//      advance", IEEE Trans. Computers 53(8), 2004.  a*(zh+zl) is within about
//      2^-105 relative of a/b, so q is within ulp/2 + 2^-100 ulp of a/b: q is a
//      faithful quotient.
//   2. e = a - b*q                  exact in one FMA: for a faithful q the
//      remainder is a multiple of ulp(b)*ulp(q) below 2^53 of them.
//   3. RN(q + e*zh) = RN(a/b)       Markstein's correction (P. Markstein,
//      "IA-64 and Elementary Functions", 2000, Thm. 8.10): q + e*zh equals
//      a/b + (a/b - q)(b*zh - 1) with |b*zh - 1| < 2^-53, and a/b is never
//      closer than ulp/(2B) (B the divisor's integer significand < 2^53) to a
//      rounding midpoint, so the perturbation never moves it across one.
// The argument needs every intermediate to be normal and the remainder to be
// representable: |a*zl| >= 2^-1022 holds for |a/b| >= 2^-914, ulp(b)*ulp(q) >=
// 2^-1074 for |a| >= 2^-968, and nothing overflows for |a/b| <= 2^1020.  The
// fast form is therefore taken only for divisors b in [2^-100, 2^100] and
// dividends |a| in [2^-600, 2^600] or a == +0 (steps 1-3 map +0 to +0 exactly
// when b > 0); anything else (-0, Inf, NaN, negative divisors) goes through
// the IEEE division, so the result is RN(a/b) for every input, bit for bit.
#pragma once

#include <cmath>

#ifdef __HIPCC__
#define PN_HD __host__ __device__ __forceinline__
#else
#define PN_HD inline
#endif

struct pn_recip {
  double b, zh, zl;  // divisor, RN(1/b), RN(1/b - zh)
};

// host side: the two-term reciprocal; fast == 0 when b is outside the range
// the proof covers or not positive (the caller then keeps the IEEE division)
inline pn_recip pn_make_recip(double b, int *fast) {
  pn_recip r;
  r.b = b;
  r.zh = 1.0 / b;
  // 1 - b*zh is exact in one fma (remainder of a faithful reciprocal)
  const double rem = std::fma(-b, r.zh, 1.0);
  r.zl = rem / b;
  *fast = (b >= 0x1p-100 && b <= 0x1p+100) ? 1 : 0;
  return r;
}

// RN(a/b) for |a| in [2^-600, 2^600] or a == +0, given pn_make_recip(b) with fast == 1
PN_HD double pn_div_fast(double a, const pn_recip &r) {
  const double q = __builtin_fma(a, r.zh, a * r.zl);
  const double e = __builtin_fma(-r.b, q, a);
  return __builtin_fma(e, r.zh, q);
}

// true when `a` lies where pn_div_fast is exact
PN_HD bool pn_div_fast_ok(double a) {
  const double m = __builtin_fabs(a);
  return (m >= 0x1p-600 && m <= 0x1p+600) || (a == 0.0 && !__builtin_signbit(a));
}
