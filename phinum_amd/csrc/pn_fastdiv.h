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

// Markstein's correction (steps 2-3) of a faithful quotient q of a/b
PN_HD double pn_div_correct(double a, const pn_recip &r, double q) {
  const double e = __builtin_fma(-r.b, q, a);
  return __builtin_fma(e, r.zh, q);
}

// RN(a/b) for |a| in [2^-600, 2^600] or a == +0, given pn_make_recip(b) with fast == 1
PN_HD double pn_div_fast(double a, const pn_recip &r) {
  return pn_div_correct(a, r, __builtin_fma(a, r.zh, a * r.zl));
}

// RN(a/b) in two operations, Brisebarre-Muller-Raina's step 1 alone, for
// divisors that pass pn_bmr_exact (same range as pn_div_fast).  s = a*zh +
// RN(a*zl) is within |a/b| 2^-105 (1 + 2^-54) of a/b, so RN(s) can differ from
// RN(a/b) only where a/b lies that close to a rounding midpoint m.  With
// a = A 2^ea, b = B 2^eb (A, B 53-bit integers) and m = (2K+1) 2^(E-53),
// |a/b - m| = 2^(E-53) |N1| / B for the integer N1 = A 2^t - B (2K+1), t = 53
// or 54; the bound then needs |N1| <= 4.  N1 is a multiple of 2^s (2^s the
// largest power of two dividing B), so s >= 3 (6 = 3 2^1 and 120 = 15 2^3 give
// s = 51 and 49) rules every a out; otherwise the few A with |N1| <= 4 are
// solved for and tested.
PN_HD double pn_div_bmr(double a, const pn_recip &r) { return __builtin_fma(a, r.zh, a * r.zl); }

// host side: 1 when pn_div_bmr is correctly rounded for every dividend in the
// fast range (requires pn_make_recip's fast == 1)
inline int pn_bmr_exact(const pn_recip &r) {
  int e;
  const double fr = std::frexp(r.b, &e);
  unsigned long long B = (unsigned long long)std::ldexp(fr, 53);
  int s = 0;
  while (!(B & 1ull)) {
    B >>= 1;
    s++;
  }
  if (s >= 3) return 1;
  const unsigned long long Bo = B;  // odd part
  for (int t = 53; t <= 54; t++) {
    for (int c = -4; c <= 4; c++) {
      if (c == 0 || c % (1 << s) != 0) continue;
      const long long cc = c / (1 << s);
      const int tt = t - s;  // A 2^tt - Bo M = cc with M odd
      // inverse of 2^tt modulo Bo by exponentiation of the inverse of 2
      const unsigned __int128 inv2 = (Bo + 1) / 2;
      unsigned __int128 inv = 1;
      for (int k = 0; k < tt; k++) inv = inv * inv2 % Bo;
      const unsigned __int128 cm = (unsigned __int128)((cc % (long long)Bo + (long long)Bo) % (long long)Bo);
      const unsigned long long A0 = (unsigned long long)(cm * inv % Bo);
      const unsigned long long lo = 1ull << 52, hi = 1ull << 53;
      unsigned long long A = A0;
      if (A < lo) A += (lo - A + Bo - 1) / Bo * Bo;
      for (; A < hi; A += Bo) {
        const __int128 num = ((__int128)A << tt) - cc;
        if (num % (__int128)Bo != 0 || ((num / (__int128)Bo) & 1) == 0) continue;
        const double a = std::ldexp((double)A, -52);  // in [1, 2): a normal dividend
        if (pn_div_bmr(a, r) != a / r.b) return 0;
        if (pn_div_bmr(-a, r) != -a / r.b) return 0;
        if (Bo == 1) break;
      }
    }
  }
  return 1;
}

// true when `a` lies where pn_div_fast is exact
PN_HD bool pn_div_fast_ok(double a) {
  const double m = __builtin_fabs(a);
  return (m >= 0x1p-600 && m <= 0x1p+600) || (a == 0.0 && !__builtin_signbit(a));
}
