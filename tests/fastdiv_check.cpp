// Checks pn_div_fast (phinum_amd/csrc/pn_fastdiv.h) against IEEE division bit
// for bit: log-uniform dividends over the whole fast range, dividends whose
// quotient lies next to a rounding midpoint (the hard cases for correct
// rounding), the range ends, zero, and the divisors of the leapfrog force
// (6, 120, kappa) plus random ones.
// pn_div_bmr (two operations) is checked the same way on every divisor that
// pn_bmr_exact accepts, which must include 6 and 120.
// usage: fastdiv_check <samples per divisor> <seed>; prints
// "mismatches N checked M bmr_divisors D bmr_rejected R"
#include <cfloat>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>

#include "../phinum_amd/csrc/pn_fastdiv.h"

static uint64_t g_s;
static uint64_t rnd() {  // splitmix64
  uint64_t z = (g_s += 0x9E3779B97F4A7C15ull);
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}
static double rand_in(double lo_exp, double hi_exp) {  // log-uniform
  const double u = (rnd() >> 11) * 0x1p-53;
  const double m = 1.0 + (rnd() >> 12) * 0x1p-52;
  return std::ldexp(m, (int)std::floor(lo_exp + u * (hi_exp - lo_exp)));
}
// inverse of x modulo odd m (extended Euclid)
static uint64_t inv_mod(uint64_t x, uint64_t m) {
  __int128 t = 0, nt = 1, r = m, nr = x % m;
  while (nr) {
    const __int128 q = r / nr, a = t - q * nt, b = r - q * nr;
    t = nt; nt = a; r = nr; nr = b;
  }
  return (uint64_t)(t < 0 ? t + m : t);
}
static uint64_t bits(double x) {
  uint64_t u;
  std::memcpy(&u, &x, 8);
  return u;
}

int main(int argc, char **argv) {
  const long n = argc > 1 ? atol(argv[1]) : 1000000;
  g_s = argc > 2 ? strtoull(argv[2], nullptr, 10) : 1;
  double divisors[64];
  // fast range ends are refused outside [2^-100, 2^100]
  {
    int f1 = 1, f2 = 1;
    pn_make_recip(0x1.fffffffffffffp-101, &f1);
    pn_make_recip(0x1.0000000000001p+100, &f2);
    if (f1 || f2 || pn_div_fast_ok(0x1.0000000000001p+600) || pn_div_fast_ok(0x1.fffffffffffffp-601)) {
      printf("range check failed\n");
      return 1;
    }
  }
  int nd = 0;
  divisors[nd++] = 6.0;
  divisors[nd++] = 120.0;
  divisors[nd++] = 1.224744871391589;   // kappa for lambda=1, g=0.2 (6 lambda units)
  divisors[nd++] = std::sqrt(36.0 / (120.0 * 0.37));
  divisors[nd++] = 0.1369;              // a non-power-of-two a^2
  divisors[nd++] = 3.0;
  divisors[nd++] = 0x1.fffffffffffffp+0;  // significand all ones
  divisors[nd++] = 0x1.0000000000001p+0;
  divisors[nd++] = 7.0;
  // two divisors for which the two-operation form fails on some dividend:
  // pn_bmr_exact must find that and reject them
  const double bmr_fail[2] = {0x1.fff9af89f7dd3p+5, 0x1.defa6767e7f8bp-2};
  divisors[nd++] = bmr_fail[0];
  divisors[nd++] = bmr_fail[1];
  while (nd < 40) divisors[nd++] = rand_in(-99, 99);
  divisors[nd++] = 0x1p-100;
  divisors[nd++] = 0x1.8p+99;
  long bad = 0, checked = 0, bmr_div = 0, bmr_rej = 0;
  {  // refused: negative divisor, -0 dividend
    int fast = 1;
    pn_make_recip(-7.0, &fast);
    if (fast || pn_div_fast_ok(-0.0)) bad++;
  }
  for (int d = 0; d < nd; d++) {
    int fast = 0;
    const pn_recip r = pn_make_recip(divisors[d], &fast);
    if (!fast) continue;
    const double b = divisors[d];
    const int bmr = pn_bmr_exact(r);
    bmr_div += bmr;
    bmr_rej += !bmr;
    if ((b == 6.0 || b == 120.0) && !bmr) {
      printf("pn_bmr_exact rejects %g\n", b);
      bad++;
    }
    if ((b == bmr_fail[0] || b == bmr_fail[1]) && bmr) {
      printf("pn_bmr_exact accepts %a\n", b);
      bad++;
    }
    auto check = [&](double a) {
      if (!pn_div_fast_ok(a)) return;
      const double want = a / b;
      const double got = pn_div_fast(a, r);
      checked++;
      if (bits(want) != bits(got)) {
        if (bad < 10) printf("mismatch a=%a b=%a want=%a got=%a\n", a, b, want, got);
        bad++;
      }
      if (bmr) {
        const double g2 = pn_div_bmr(a, r);
        checked++;
        if (bits(want) != bits(g2)) {
          if (bad < 10) printf("bmr mismatch a=%a b=%a want=%a got=%a\n", a, b, want, g2);
          bad++;
        }
      }
    };
    check(0.0);
    check(0x1p-600);
    check(0x1p+600);
    check(std::nextafter(0x1p+600, 0.0));
    check(0x1.fffffffffffffp-600);
    check(-0x1p-600);
    // dividends whose quotient is as close to a midpoint as the divisor
    // allows: with b = B_odd*2^s*2^eb, A*2^t - B_odd*M = c for an odd M and a
    // small c makes A/b lie c*2^(e-53)/B_odd from the midpoint M*2^(e-53-...)
    {
      int eb;
      const double fr = std::frexp(b, &eb);  // b = fr*2^eb, fr in [0.5, 1)
      uint64_t B = (uint64_t)std::ldexp(fr, 53);
      int sh = 0;
      while (!(B & 1)) { B >>= 1; sh++; }
      if (B > 1) {
        for (int t = 53 - sh; t <= 54 - sh; t++) {
          const uint64_t inv2t = inv_mod((uint64_t)(((unsigned __int128)1 << t) % B), B);
          for (long c = -41; c <= 41; c += 2) {
            const uint64_t cm = (uint64_t)((c % (long)B + (long)B) % (long)B);
            const uint64_t A0 = (uint64_t)(((unsigned __int128)cm * inv2t) % B);
            const uint64_t lo = 1ull << 52, hi = 1ull << 53;
            const uint64_t j0 = A0 >= lo ? 0 : (lo - A0 + B - 1) / B;
            const uint64_t span = (hi - 1 - A0) / B;
            for (int k = 0; k < 64 && j0 + k <= span; k++) {
              const uint64_t j = (span - j0 < 64) ? j0 + k : j0 + rnd() % (span - j0 + 1);
              const uint64_t A = A0 + j * B;
              const __int128 num = ((__int128)A << t) - c;
              if (num % B != 0 || ((num / B) & 1) == 0) continue;
              const int e = (int)(rnd() % 1190) - 595;
              check(std::ldexp((double)A, e - 52));
            }
          }
        }
      }
    }
    for (long i = 0; i < n; i++) {
      // uniform over exponents, both signs
      double a = rand_in(-600, 600);
      check((rnd() & 1) ? a : -a);
      // near a rounding midpoint: m = q + ulp(q)/2 exactly in long double,
      // a = b*m rounded, so a/b sits within a few ulp(a)/|b| of m
      const double q = rand_in(-480, 480);
      const long double m = (long double)q + (long double)(std::nextafter(q, INFINITY) - q) / 2;
      const double a2 = (double)((long double)b * m);
      check(a2);
      check(std::nextafter(a2, INFINITY));
      check(std::nextafter(a2, -INFINITY));
      // small integers and exact quotients
      const double k = (double)(rnd() % 4096);
      check(k);
      check(k * b);
      check(-k * b);
    }
  }
  printf("mismatches %ld checked %ld bmr_divisors %ld bmr_rejected %ld\n", bad, checked, bmr_div, bmr_rej);
  return bad != 0;
}
