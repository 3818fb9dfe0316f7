// pn_kernels.h -- device pieces shared by the single-lattice and batched
// paths, so that both produce bit-identical site updates and sums.
#pragma once
#include <hip/hip_runtime.h>

#include <type_traits>

#include "pn_fastdiv.h"

// Blocks are dispatched round-robin over the 8 XCDs of an MI355X.  The remap
// gives XCD k a contiguous run of logical tiles, so tiles that share y and z
// halo rows are processed on the same XCD and find each other's planes in
// that XCD's L2.
__device__ __forceinline__ int xcd_tile(int b, int nblocks) {
  const int q = nblocks >> 3, r = nblocks & 7;
  const int k = b & 7, j = b >> 3;
  return k * q + (k < r ? k : r) + j;
}

struct CoefF64 {
  double a2, inv_a2, c2, lmode, kappa;
  pn_recip r6, rk, r120;  // two-term reciprocals of 6, kappa and 120
  int fdiv;               // kappa lies where pn_div_fast is exact
};

// host side: the reciprocals for the force divisions
inline void coef64_recips(CoefF64 &k) {
  int f6 = 0, fk = 0, f120 = 0;
  k.r6 = pn_make_recip(6., &f6);
  k.rk = pn_make_recip(k.kappa, &fk);
  k.r120 = pn_make_recip(120., &f120);
  k.fdiv = f6 && fk && f120;
}
struct CoefF32 {
  float inv_a2, c2, c4, c6;
};

// |f| in [2^-100, 2^100] or f == 0 puts q = f*f in [2^-200, 2^200] or +0,
// which keeps q/6, q/kappa (kappa in [2^-100, 2^100]) and t*t/120 inside the
// range where pn_div_fast equals the IEEE quotient.
__device__ __forceinline__ bool fast_div_site(double f) {
  const double a = __builtin_fabs(f);
  return (a >= 0x1p-100 && a <= 0x1p+100) || a == 0.0;
}

enum { DIV_IEEE = 0, DIV_FAST = 1, DIV_PER_SITE = 2 };

// Theory-0 update of one site, reference order (dynamics.cpp:80,84,183).
// DIV_FAST needs k.fdiv and fast_div_site(f), checked by the caller (the
// step kernels check a whole wave at once); DIV_PER_SITE checks here.
template <bool POW2, int DIV>
__device__ __forceinline__ void site_update_f64(const CoefF64 &k, double xp, double xm, double yp,
                                                double ym, double zp, double zm, double f,
                                                double &p, double dtk) {
  double s = xp + xm;
  s = s + yp;
  s = s + ym;
  s = s + zp;
  s = s + zm;
  double lap = POW2 ? (s - 6. * f) * k.inv_a2 : (s - 6. * f) / k.a2;
  // the reference's 0.0 + f*f: f*f is never -0, so adding +0 changes no bit
  const double q = f * f;
  double B, C;
  const bool fast = DIV == DIV_FAST || (DIV == DIV_PER_SITE && k.fdiv && fast_div_site(f));
  if (fast) {
    B = k.lmode * pn_div_fast(q, k.r6);
    const double t = pn_div_fast(q, k.rk);
    C = pn_div_fast(t * t, k.r120);
  } else {
    B = k.lmode * (q / 6.);
    const double t = q / k.kappa;
    C = (t * t) / 120.;
  }
  p = p + (lap - ((k.c2 + B) + C) * f) * dtk;
}

// Runs body(std::integral_constant<int, DIV_FAST>) when every active lane of
// the wave passes `lane_ok` and the coefficients allow it, else the IEEE form:
// one wave-uniform branch instead of a divergent one per site.
template <typename F>
__device__ __forceinline__ void with_div_mode(const CoefF64 &k, bool lane_ok, F &&body) {
  if (k.fdiv && __all(lane_ok)) body(std::integral_constant<int, DIV_FAST>());
  else body(std::integral_constant<int, DIV_IEEE>());
}

// FP32 form: same stencil, potential with host-rounded coefficients.
__device__ __forceinline__ void site_update_f32(const CoefF32 &k, float xp, float xm, float yp,
                                                float ym, float zp, float zm, float f, float &p,
                                                float dtk) {
  float s = xp + xm;
  s = s + yp;
  s = s + ym;
  s = s + zp;
  s = s + zm;
  float lap = (s - 6.0f * f) * k.inv_a2;
  float q = f * f;
  float force = ((k.c2 + k.c4 * q) + k.c6 * (q * q)) * f;
  p = p + (lap - force) * dtk;
}

template <typename T, int W>
struct VecT {
  typedef T type __attribute__((ext_vector_type(W)));
};

// Exact two-sum accumulation: (hi, lo) pairs updated without loss, merged in
// a fixed tree, give sums that are deterministic and within a few ulp.
__device__ __forceinline__ void two_acc(double &hi, double &lo, double x) {
  const double s = hi + x;
  const double bp = s - hi;
  const double e = (hi - (s - bp)) + (x - bp);
  hi = s;
  lo = lo + e;
}

__device__ __forceinline__ void two_merge(double &h1, double &l1, double h2, double l2) {
  two_acc(h1, l1, h2);
  l1 = l1 + l2;
}

struct RedCoef {
  double a2;    // sqr(a_t)
  double cm;    // (Mass_Mode*0.5)*BareM_sqr_long        (fieldanalysis.cpp:457)
  double lmode; // Lambda_Mode                           (:571)
  double d6;    // 720.0*Nc*Nc*Kappa*Kappa               (:571)
};

