// pn_kernels.h -- device pieces shared by the single-lattice and batched
// paths, so that both produce bit-identical site updates and sums.
#pragma once
#include <hip/hip_runtime.h>

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
};
struct CoefF32 {
  float inv_a2, c2, c4, c6;
};

// Theory-0 update of one site, reference order (dynamics.cpp:80,84,183).
template <bool POW2>
__device__ __forceinline__ void site_update_f64(const CoefF64 &k, double xp, double xm, double yp,
                                                double ym, double zp, double zm, double f,
                                                double &p, double dtk) {
  double s = xp + xm;
  s = s + yp;
  s = s + ym;
  s = s + zp;
  s = s + zm;
  double lap = POW2 ? (s - 6. * f) * k.inv_a2 : (s - 6. * f) / k.a2;
  double q = 0.0 + f * f;
  double B = k.lmode * (q / 6.);
  double t = q / k.kappa;
  double C = (t * t) / 120.;
  p = p + (lap - ((k.c2 + B) + C) * f) * dtk;
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

