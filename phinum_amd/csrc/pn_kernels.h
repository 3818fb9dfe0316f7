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
  int rk_bmr;             // q/kappa needs no correction step (pn_bmr_exact)
};

// host side: the reciprocals for the force divisions
inline void coef64_recips(CoefF64 &k) {
  int f6 = 0, fk = 0, f120 = 0;
  k.r6 = pn_make_recip(6., &f6);
  k.rk = pn_make_recip(k.kappa, &fk);
  k.r120 = pn_make_recip(120., &f120);
  // 6 and 120 always pass pn_bmr_exact (their significands carry 2^1 and 2^3)
  k.fdiv = f6 && fk && f120 && pn_bmr_exact(k.r6) && pn_bmr_exact(k.r120);
  k.rk_bmr = k.fdiv && pn_bmr_exact(k.rk);
}
struct CoefF32 {
  float inv_a2, c2, c4, c6;
};

// |f| in [2^-100, 2^100) puts q = f*f in [2^-200, 2^200), which keeps q/6,
// q/kappa (kappa in [2^-100, 2^100]) and t*t/120 inside the range where
// pn_div_fast equals the IEEE quotient.  The test reads the exponent bits:
// zero, subnormals, Inf and NaN fall outside and take the IEEE divisions.
__device__ __forceinline__ bool fast_div_site(double f) {
  const unsigned hi = (unsigned)((unsigned long long)__double_as_longlong(f) >> 32) & 0x7fffffffu;
  return hi - 0x39B00000u < 0x0C800000u;
}

// DIV_FAST: the fast divisions, with the correction step for q/kappa chosen
// at run time from k.rk_bmr; DIV_FAST_2OP / DIV_FAST_KC: the same with that
// choice made at compile time (q/kappa in two operations / corrected);
// DIV_PER_SITE: DIV_FAST or IEEE, decided per site.
enum { DIV_IEEE = 0, DIV_FAST = 1, DIV_PER_SITE = 2, DIV_FAST_2OP = 3, DIV_FAST_KC = 4 };

// Theory-0 update of one site, reference order (dynamics.cpp:80,84,183).
// DIV_FAST and DIV_FAST_KC need k.fdiv and fast_div_site(f), checked by the
// caller (the step kernels check a whole wave at once); DIV_PER_SITE checks
// here.
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
  const bool fast = DIV != DIV_IEEE && (DIV != DIV_PER_SITE || (k.fdiv && fast_div_site(f)));
  if (fast) {
    B = k.lmode * pn_div_bmr(q, k.r6);
    double t = pn_div_bmr(q, k.rk);
    if (DIV == DIV_FAST_KC || ((DIV == DIV_FAST || DIV == DIV_PER_SITE) && !k.rk_bmr))
      t = pn_div_correct(q, k.rk, t);
    C = pn_div_bmr(t * t, k.r120);
  } else {
    B = k.lmode * (q / 6.);
    const double t = q / k.kappa;
    C = (t * t) / 120.;
  }
  p = p + (lap - ((k.c2 + B) + C) * f) * dtk;
}

// Runs body(std::integral_constant<int, FAST>) (FAST one of the DIV_FAST
// modes) when every active lane of the wave passes `lane_ok` and the
// coefficients allow it, else the IEEE form: one wave-uniform branch instead
// of a divergent one per site.
template <int FAST = DIV_FAST, typename F>
__device__ __forceinline__ void with_div_mode(const CoefF64 &k, bool lane_ok, F &&body) {
  if (k.fdiv && __all(lane_ok)) body(std::integral_constant<int, FAST>());
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
  double inv_a2;  // 1/a2, used only when a2 is a power of two (the product is then exact)
  int pow2;
};


// Two leapfrog steps on one tile of full x rows (temporal blocking), shared by
// the single-lattice and batched two-step kernels.  Thread (tx, ty) owns W
// sites of row y0-1+ty; the tile has RT rows, of which the inner RT-2 are
// output rows.  Stage A advances step n on all RT rows, stage B step n+1 on
// the inner rows, reading the intermediate Phi of its plane from LDS (two
// planes of RT*N values) and the planes above and below from registers.  The
// z loop runs stage A one plane ahead of stage B, plus one extra plane at each
// end of [z0, z1).  Halo rows and planes are recomputed by both neighbouring
// tiles with the same arithmetic, so the result is bit-identical to two single
// steps.  Pi is read from `pi` and written to `pio`: a neighbouring tile's
// stage A still needs the old Pi of the rows this tile writes.  Threads with
// active == false only join the barriers (one per plane), which every thread
// of the block must reach.  PF issues the next plane's HBM reads before the
// barrier (two more vector registers of W values per thread).  LP keeps the
// tile's Phi planes in LDS too (two more planes of RT*N values after the
// intermediate ones): stage A then reads its x and y neighbours there instead
// of from L2, except the one outside row of each edge row.
template <typename T, typename C, bool POW2, int W, bool PF = true, bool LP = true, int FAST = DIV_FAST>
__device__ __forceinline__ void step2_tile(const T *__restrict__ phi, const T *__restrict__ pi,
                                           T *__restrict__ pio, T *__restrict__ phin, T *lds_g, int N,
                                           int RT, int tx, int ty, bool active, int y0, int z0,
                                           int z1, const C &k, T dk0, T dd0, T dk1, T dd1) {
  typedef typename VecT<T, W>::type V;
  // LDS through address-space-3 pointers: ds_read/ds_write without a flat path
  typedef __attribute__((address_space(3))) T LT;
  typedef __attribute__((address_space(3))) V LV;
  LT *lds = (LT *)lds_g;
  const int TY = RT - 2;
  const int x = W * tx;
  int y = y0 - 1 + ty;
  y = y < 0 ? y + N : (y >= N ? y - N : y);
  const bool inner = active && ty >= 1 && ty <= TY && (y0 + ty - 1) < N;
  const int xp = (x + W == N) ? 0 : x + W, xm = (x == 0) ? N - 1 : x - 1;
  const int yp = (y + 1 == N) ? 0 : y + 1, ym = (y == 0) ? N - 1 : y - 1;
  const long long P = (long long)N * N;
  const int row = y * N, rowp = yp * N, rowm = ym * N;
  auto wz = [N](int z) { return z < 0 ? z + N : (z >= N ? z - N : z); };
  // an_n, p_n: Phi of plane za+1 and Pi of plane za for the coming
  // iteration, read from HBM before the previous iteration's barrier so that
  // their latency overlaps the wait (the in-plane neighbours that follow are
  // mostly L2 hits: the neighbouring rows' own reads)
  // plane offsets of za, za+1, za+2, advanced by one plane per iteration
  const long long NP = (long long)N * P;
  long long om = (long long)wz(z0 - 2) * P, o0 = (long long)wz(z0 - 1) * P,
            o1 = (long long)wz(z0) * P, o2 = (long long)wz(z0 + 1) * P;
  V an_m = {}, an_c = {}, an_n = {}, p_n = {};
  if (active) {
    an_m = *(const V *)&phi[om + row + x];
    an_c = *(const V *)&phi[o0 + row + x];
    if (PF) {
      an_n = *(const V *)&phi[o1 + row + x];
      p_n = *(const V *)&pi[o0 + row + x];
    }
  }
  // Phi planes for LP: buffer (z - z0) & 1 holds plane z; the first plane is
  // stored before the loop, each later one an iteration ahead of its use
  LT *lphi = lds + (size_t)2 * RT * N;
  if (LP) {
    if (active) *(LV *)&lphi[(size_t)RT * N + ty * N + x] = an_c;
    __syncthreads();
  }
  V b_m = an_m, b_c = an_c, pih = an_c;
  for (int za = z0 - 1; za <= z1; za++) {  // stage A on plane za, stage B on za-1
    V an_p = an_n, p = p_n, f1 = {};
    const long long base = o0;
    if (active) {
      if (!PF) {
        an_p = *(const V *)&phi[o1 + row + x];
        p = *(const V *)&pi[base + row + x];
      }
      T vxp, vxm;
      V vyp, vym;
      if (LP) {
        *(LV *)&lphi[(size_t)((za + 1 - z0) & 1) * RT * N + ty * N + x] = an_p;
        const LT *lp = lphi + (size_t)((za - z0) & 1) * RT * N;
        vxp = lp[ty * N + xp];
        vxm = lp[ty * N + xm];
        vyp = ty + 1 < RT ? *(const LV *)&lp[(ty + 1) * N + x] : *(const V *)&phi[base + rowp + x];
        vym = ty > 0 ? *(const LV *)&lp[(ty - 1) * N + x] : *(const V *)&phi[base + rowm + x];
      } else {
        vxp = phi[base + row + xp];
        vxm = phi[base + row + xm];
        vyp = *(const V *)&phi[base + rowp + x];
        vym = *(const V *)&phi[base + rowm + x];
      }
      auto stage_a = [&](auto div) {
#pragma unroll
        for (int j = 0; j < W; j++) {
          const T right = (j == W - 1) ? vxp : an_c[(j + 1) % W];
          const T left = (j == 0) ? vxm : an_c[(j + W - 1) % W];
          T pj = p[j];
          if constexpr (sizeof(T) == 8)
            site_update_f64<POW2, decltype(div)::value>(k, right, left, vyp[j], vym[j], an_p[j],
                                                        an_m[j], an_c[j], pj, dk0);
          else
            site_update_f32(k, right, left, vyp[j], vym[j], an_p[j], an_m[j], an_c[j], pj, dk0);
          p[j] = pj;
          f1[j] = an_c[j] + pj * dd0;
        }
      };
      if constexpr (sizeof(T) == 8) {
        bool ok = true;
#pragma unroll
        for (int j = 0; j < W; j++) ok = ok && fast_div_site(an_c[j]);
        with_div_mode<FAST>(k, ok, stage_a);
      } else {
        stage_a(std::integral_constant<int, DIV_IEEE>());
      }
      LT *buf = lds + (size_t)((za - z0 + 2) & 1) * RT * N;
      *(LV *)&buf[ty * N + x] = f1;
    }
    if (za >= z0 + 1 && inner) {
      const int zb = za - 1;
      const LT *bb = lds + (size_t)((zb - z0 + 2) & 1) * RT * N;
      const T bxp = bb[ty * N + xp], bxm = bb[ty * N + xm];
      const V byp = *(const LV *)&bb[(ty + 1) * N + x], bym = *(const LV *)&bb[(ty - 1) * N + x];
      V q = pih, f2;
      auto stage_b = [&](auto div) {
#pragma unroll
        for (int j = 0; j < W; j++) {
          const T right = (j == W - 1) ? bxp : b_c[(j + 1) % W];
          const T left = (j == 0) ? bxm : b_c[(j + W - 1) % W];
          T qj = q[j];
          if constexpr (sizeof(T) == 8)
            site_update_f64<POW2, decltype(div)::value>(k, right, left, byp[j], bym[j], f1[j],
                                                        b_m[j], b_c[j], qj, dk1);
          else
            site_update_f32(k, right, left, byp[j], bym[j], f1[j], b_m[j], b_c[j], qj, dk1);
          q[j] = qj;
          f2[j] = b_c[j] + qj * dd1;
        }
      };
      if constexpr (sizeof(T) == 8) {
        bool ok = true;
#pragma unroll
        for (int j = 0; j < W; j++) ok = ok && fast_div_site(b_c[j]);
        with_div_mode<FAST>(k, ok, stage_b);
      } else {
        stage_b(std::integral_constant<int, DIV_IEEE>());
      }
      const long long bbase = om;  // plane zb = za - 1
      __builtin_nontemporal_store(q, (V *)&pio[bbase + row + x]);
      __builtin_nontemporal_store(f2, (V *)&phin[bbase + row + x]);
    }
    if (PF && active && za < z1) {
      an_n = *(const V *)&phi[o2 + row + x];
      p_n = *(const V *)&pi[o1 + row + x];
    }
    om = o0;
    o0 = o1;
    o1 = o2;
    o2 = o2 + P == NP ? 0 : o2 + P;
    // The only barrier of the iteration.  LDS holds two planes of stage A's
    // results: stage A of plane za writes buffer (za & 1) while stage B reads
    // buffer ((za-1) & 1), written one iteration earlier; with LP two planes
    // of Phi likewise: plane za+1 is written while stage A reads plane za.
    // This barrier orders those writes before the next iteration reads them,
    // and this iteration's reads before the next one overwrites the buffers
    // they used.
    __syncthreads();
    b_m = b_c;
    b_c = f1;
    pih = p;
    an_m = an_c;
    an_c = an_p;
  }
}

// Three leapfrog steps on one tile of full x rows: step2_tile with a third
// stage.  Thread (tx, ty) owns W sites of row y0-2+ty; stage A advances step
// n on all RT rows, stage B step n+1 on rows 1..RT-2, stage C step n+2 on the
// output rows 2..RT-3.  The z loop runs A on plane za, B on za-1 and C on
// za-2, plus two extra planes at each end of [z0, z1) for A and one for B.
// LDS holds two planes each of A's and B's results and of Phi (six planes of
// RT*N values); stage B reads A's plane written an iteration earlier and C
// reads B's likewise, so one barrier per plane still orders every LDS write
// before its reads and every read before the buffer is written again.  Halo
// rows and planes are recomputed by neighbouring tiles with the same
// arithmetic: the result is bit-identical to three single steps.
template <typename T, typename C, bool POW2, int W, int FAST = DIV_FAST>
__device__ __forceinline__ void step3_tile(const T *__restrict__ phi, const T *__restrict__ pi,
                                           T *__restrict__ pio, T *__restrict__ phin, T *lds_g, int N,
                                           int RT, int tx, int ty, int y0, int z0, int z1, const C &k,
                                           T dk0, T dd0, T dk1, T dd1, T dk2, T dd2) {
  typedef typename VecT<T, W>::type V;
  typedef __attribute__((address_space(3))) T LT;
  typedef __attribute__((address_space(3))) V LV;
  LT *lds = (LT *)lds_g;
  const size_t PL = (size_t)RT * N;  // one LDS plane
  LT *bufA = lds, *bufB = lds + 2 * PL, *lphi = lds + 4 * PL;
  const int x = W * tx;
  int y = y0 - 2 + ty;
  y = y < 0 ? y + N : (y >= N ? y - N : y);
  const bool brow = ty >= 1 && ty <= RT - 2;
  const bool crow = ty >= 2 && ty <= RT - 3 && (y0 + ty - 2) < N;
  const int xp = (x + W == N) ? 0 : x + W, xm = (x == 0) ? N - 1 : x - 1;
  const int yp = (y + 1 == N) ? 0 : y + 1, ym = (y == 0) ? N - 1 : y - 1;
  const long long P = (long long)N * N;
  const int row = y * N, rowp = yp * N, rowm = ym * N;
  auto wz = [N](int z) { return z < 0 ? z + N : (z >= N ? z - N : z); };
  auto slot = [z0](int z) { return (size_t)((z - z0 + 4) & 1); };
  const long long NP = (long long)N * P;
  // plane offsets of za-2 .. za+2, advanced by one plane per iteration
  long long om2 = (long long)wz(z0 - 4) * P, om1 = (long long)wz(z0 - 3) * P,
            o0 = (long long)wz(z0 - 2) * P, o1 = (long long)wz(z0 - 1) * P,
            o2 = (long long)wz(z0) * P;
  V an_m = *(const V *)&phi[om1 + row + x];
  V an_c = *(const V *)&phi[o0 + row + x];
  V an_n = *(const V *)&phi[o1 + row + x];
  V p_n = *(const V *)&pi[o0 + row + x];
  *(LV *)&lphi[slot(z0 - 2) * PL + ty * N + x] = an_c;
  __syncthreads();
  V b_m = an_m, b_c = an_c, pih = an_c, c_m = an_c, c_c = an_c, qh = an_c;
  for (int za = z0 - 2; za <= z1 + 1; za++) {
    V an_p = an_n, p = p_n, f1 = {}, f2 = {}, q = {};
    // stage A, plane za, all rows
    {
      *(LV *)&lphi[slot(za + 1) * PL + ty * N + x] = an_p;
      const LT *lp = lphi + slot(za) * PL;
      const T vxp = lp[ty * N + xp], vxm = lp[ty * N + xm];
      const V vyp = ty + 1 < RT ? *(const LV *)&lp[(ty + 1) * N + x] : *(const V *)&phi[o0 + rowp + x];
      const V vym = ty > 0 ? *(const LV *)&lp[(ty - 1) * N + x] : *(const V *)&phi[o0 + rowm + x];
      auto stage_a = [&](auto div) {
#pragma unroll
        for (int j = 0; j < W; j++) {
          const T right = (j == W - 1) ? vxp : an_c[(j + 1) % W];
          const T left = (j == 0) ? vxm : an_c[(j + W - 1) % W];
          T pj = p[j];
          if constexpr (sizeof(T) == 8)
            site_update_f64<POW2, decltype(div)::value>(k, right, left, vyp[j], vym[j], an_p[j], an_m[j],
                                                        an_c[j], pj, dk0);
          else
            site_update_f32(k, right, left, vyp[j], vym[j], an_p[j], an_m[j], an_c[j], pj, dk0);
          p[j] = pj;
          f1[j] = an_c[j] + pj * dd0;
        }
      };
      if constexpr (sizeof(T) == 8) {
        bool ok = true;
#pragma unroll
        for (int j = 0; j < W; j++) ok = ok && fast_div_site(an_c[j]);
        with_div_mode<FAST>(k, ok, stage_a);
      } else {
        stage_a(std::integral_constant<int, DIV_IEEE>());
      }
      *(LV *)&bufA[slot(za) * PL + ty * N + x] = f1;
    }
    // stage B, plane za-1, rows 1..RT-2
    if (za >= z0 && brow) {
      const LT *ba = bufA + slot(za - 1) * PL;
      const T bxp = ba[ty * N + xp], bxm = ba[ty * N + xm];
      const V byp = *(const LV *)&ba[(ty + 1) * N + x], bym = *(const LV *)&ba[(ty - 1) * N + x];
      q = pih;
      auto stage_b = [&](auto div) {
#pragma unroll
        for (int j = 0; j < W; j++) {
          const T right = (j == W - 1) ? bxp : b_c[(j + 1) % W];
          const T left = (j == 0) ? bxm : b_c[(j + W - 1) % W];
          T qj = q[j];
          if constexpr (sizeof(T) == 8)
            site_update_f64<POW2, decltype(div)::value>(k, right, left, byp[j], bym[j], f1[j], b_m[j],
                                                        b_c[j], qj, dk1);
          else
            site_update_f32(k, right, left, byp[j], bym[j], f1[j], b_m[j], b_c[j], qj, dk1);
          q[j] = qj;
          f2[j] = b_c[j] + qj * dd1;
        }
      };
      if constexpr (sizeof(T) == 8) {
        bool ok = true;
#pragma unroll
        for (int j = 0; j < W; j++) ok = ok && fast_div_site(b_c[j]);
        with_div_mode<FAST>(k, ok, stage_b);
      } else {
        stage_b(std::integral_constant<int, DIV_IEEE>());
      }
      *(LV *)&bufB[slot(za - 1) * PL + ty * N + x] = f2;
    }
    // stage C, plane za-2, output rows
    if (za >= z0 + 2 && crow) {
      const LT *bb = bufB + slot(za - 2) * PL;
      const T cxp = bb[ty * N + xp], cxm = bb[ty * N + xm];
      const V cyp = *(const LV *)&bb[(ty + 1) * N + x], cym = *(const LV *)&bb[(ty - 1) * N + x];
      V r = qh, f3;
      auto stage_c = [&](auto div) {
#pragma unroll
        for (int j = 0; j < W; j++) {
          const T right = (j == W - 1) ? cxp : c_c[(j + 1) % W];
          const T left = (j == 0) ? cxm : c_c[(j + W - 1) % W];
          T rj = r[j];
          if constexpr (sizeof(T) == 8)
            site_update_f64<POW2, decltype(div)::value>(k, right, left, cyp[j], cym[j], f2[j], c_m[j],
                                                        c_c[j], rj, dk2);
          else
            site_update_f32(k, right, left, cyp[j], cym[j], f2[j], c_m[j], c_c[j], rj, dk2);
          r[j] = rj;
          f3[j] = c_c[j] + rj * dd2;
        }
      };
      if constexpr (sizeof(T) == 8) {
        bool ok = true;
#pragma unroll
        for (int j = 0; j < W; j++) ok = ok && fast_div_site(c_c[j]);
        with_div_mode<FAST>(k, ok, stage_c);
      } else {
        stage_c(std::integral_constant<int, DIV_IEEE>());
      }
      __builtin_nontemporal_store(r, (V *)&pio[om2 + row + x]);
      __builtin_nontemporal_store(f3, (V *)&phin[om2 + row + x]);
    }
    if (za <= z1) {  // Phi of plane za+2 and Pi of plane za+1 for the next iteration
      an_n = *(const V *)&phi[o2 + row + x];
      p_n = *(const V *)&pi[o1 + row + x];
    }
    om2 = om1;
    om1 = o0;
    o0 = o1;
    o1 = o2;
    o2 = o2 + P == NP ? 0 : o2 + P;
    __syncthreads();
    an_m = an_c;
    an_c = an_p;
    b_m = b_c;
    b_c = f1;
    pih = p;
    c_m = c_c;
    c_c = f2;
    qh = q;
  }
}

// The same three-step tile with the plane loop unrolled by three.  Every value
// that passes from plane to plane lives for three iterations: a Phi plane
// (read at za+1, za, za-1 by stage A), the Pi of a plane between its three
// kicks (stage A at za, B at za+1, C at za+2), stage A's and stage B's
// results (read by the next stage over three planes).  With an LDS ring of
// three planes per buffer, the slot and the register of plane z are both
// (z - z0 + 3) % 3, a constant in each of the three copies of the loop body,
// so nothing is copied between iterations and no LDS slot is computed.  The
// arithmetic and its order are those of step3_tile.  LDS: 9 planes of RT*N.
// Threads with active == false only join the barriers (one per plane).
template <typename T, typename C, bool POW2, int W, int FAST = DIV_FAST>
__device__ __forceinline__ void step3_tile_r3(const T *__restrict__ phi, const T *__restrict__ pi,
                                              T *__restrict__ pio, T *__restrict__ phin, T *lds_g,
                                              int N, int RT, int tx, int ty, bool active, int y0, int z0,
                                              int z1, const C &k, T dk0, T dd0, T dk1, T dd1, T dk2,
                                              T dd2) {
  typedef typename VecT<T, W>::type V;
  typedef __attribute__((address_space(3))) T LT;
  typedef __attribute__((address_space(3))) V LV;
  LT *lds = (LT *)lds_g;
  const int PL = RT * N;  // one LDS plane
  LT *lphi = lds, *bufA = lds + 3 * PL, *bufB = lds + 6 * PL;
  const int x = W * tx;
  int y = y0 - 2 + ty;
  y = y < 0 ? y + N : (y >= N ? y - N : y);
  const bool brow = active && ty >= 1 && ty <= RT - 2;
  const bool crow = active && ty >= 2 && ty <= RT - 3 && (y0 + ty - 2) < N;
  const int xp = (x + W == N) ? 0 : x + W, xm = (x == 0) ? N - 1 : x - 1;
  const int yp = (y + 1 == N) ? 0 : y + 1, ym = (y == 0) ? N - 1 : y - 1;
  const long long P = (long long)N * N;
  const int row = y * N, rowp = yp * N, rowm = ym * N;
  // this thread's element and its x/y neighbours within an LDS plane
  const int me = ty * N + x, mxp = ty * N + xp, mxm = ty * N + xm;
  const int myp = (ty + 1) * N + x, mym = (ty - 1) * N + x;
  auto wz = [N](int z) { return z < 0 ? z + N : (z >= N ? z - N : z); };
  const long long NP = (long long)N * P;
  // plane offsets of za-2 .. za+2, advanced by one plane per iteration
  long long om2 = (long long)wz(z0 - 4) * P, om1 = (long long)wz(z0 - 3) * P,
            o0 = (long long)wz(z0 - 2) * P, o1 = (long long)wz(z0 - 1) * P,
            o2 = (long long)wz(z0) * P;
  // by plane z, index (z - z0 + 3) % 3: Phi, Pi, stage A's and stage B's results
  V F[3], Q[3], RA[3], RB[3];
  if (active) {
    F[0] = *(const V *)&phi[om1 + row + x];
    F[1] = *(const V *)&phi[o0 + row + x];
    F[2] = *(const V *)&phi[o1 + row + x];
    Q[1] = *(const V *)&pi[o0 + row + x];
    *(LV *)&lphi[PL + me] = F[1];
  } else {
    F[0] = F[1] = F[2] = Q[1] = V{};
  }
  Q[0] = Q[2] = RA[0] = RA[1] = RA[2] = RB[0] = RB[1] = RB[2] = F[1];
  __syncthreads();
  // one iteration at plane za; S = (za - z0 + 2) % 3
  auto iter = [&](auto sc, int za) {
    constexpr int S = decltype(sc)::value;
    constexpr int c = (S + 1) % 3, m = S % 3, n = (S + 2) % 3;  // za; za-1, za+2; za+1, za-2
    // stage A, plane za, all rows: Pi of za in Q[c], result in RA[c]
    if (active) {
      *(LV *)&lphi[n * PL + me] = F[n];
      const LT *lp = lphi + c * PL;
      const T vxp = lp[mxp], vxm = lp[mxm];
      const V vyp = ty + 1 < RT ? *(const LV *)&lp[myp] : *(const V *)&phi[o0 + rowp + x];
      const V vym = ty > 0 ? *(const LV *)&lp[mym] : *(const V *)&phi[o0 + rowm + x];
      auto stage_a = [&](auto div) {
#pragma unroll
        for (int j = 0; j < W; j++) {
          const T right = (j == W - 1) ? vxp : F[c][(j + 1) % W];
          const T left = (j == 0) ? vxm : F[c][(j + W - 1) % W];
          T pj = Q[c][j];
          if constexpr (sizeof(T) == 8)
            site_update_f64<POW2, decltype(div)::value>(k, right, left, vyp[j], vym[j], F[n][j], F[m][j],
                                                        F[c][j], pj, dk0);
          else
            site_update_f32(k, right, left, vyp[j], vym[j], F[n][j], F[m][j], F[c][j], pj, dk0);
          Q[c][j] = pj;
          RA[c][j] = F[c][j] + pj * dd0;
        }
      };
      if constexpr (sizeof(T) == 8) {
        bool ok = true;
#pragma unroll
        for (int j = 0; j < W; j++) ok = ok && fast_div_site(F[c][j]);
        with_div_mode<FAST>(k, ok, stage_a);
      } else {
        stage_a(std::integral_constant<int, DIV_IEEE>());
      }
      *(LV *)&bufA[c * PL + me] = RA[c];
    }
    // stage B, plane za-1, rows 1..RT-2: Pi of za-1 in Q[m], result in RB[m]
    if (za >= z0 && brow) {
      const LT *ba = bufA + m * PL;
      const T bxp = ba[mxp], bxm = ba[mxm];
      const V byp = *(const LV *)&ba[myp], bym = *(const LV *)&ba[mym];
      auto stage_b = [&](auto div) {
#pragma unroll
        for (int j = 0; j < W; j++) {
          const T right = (j == W - 1) ? bxp : RA[m][(j + 1) % W];
          const T left = (j == 0) ? bxm : RA[m][(j + W - 1) % W];
          T qj = Q[m][j];
          if constexpr (sizeof(T) == 8)
            site_update_f64<POW2, decltype(div)::value>(k, right, left, byp[j], bym[j], RA[c][j], RA[n][j],
                                                        RA[m][j], qj, dk1);
          else
            site_update_f32(k, right, left, byp[j], bym[j], RA[c][j], RA[n][j], RA[m][j], qj, dk1);
          Q[m][j] = qj;
          RB[m][j] = RA[m][j] + qj * dd1;
        }
      };
      if constexpr (sizeof(T) == 8) {
        bool ok = true;
#pragma unroll
        for (int j = 0; j < W; j++) ok = ok && fast_div_site(RA[m][j]);
        with_div_mode<FAST>(k, ok, stage_b);
      } else {
        stage_b(std::integral_constant<int, DIV_IEEE>());
      }
      *(LV *)&bufB[m * PL + me] = RB[m];
    }
    // stage C, plane za-2, output rows: Pi of za-2 in Q[n]
    if (za >= z0 + 2 && crow) {
      const LT *bb = bufB + n * PL;
      const T cxp = bb[mxp], cxm = bb[mxm];
      const V cyp = *(const LV *)&bb[myp], cym = *(const LV *)&bb[mym];
      V f3;
      auto stage_c = [&](auto div) {
#pragma unroll
        for (int j = 0; j < W; j++) {
          const T right = (j == W - 1) ? cxp : RB[n][(j + 1) % W];
          const T left = (j == 0) ? cxm : RB[n][(j + W - 1) % W];
          T rj = Q[n][j];
          if constexpr (sizeof(T) == 8)
            site_update_f64<POW2, decltype(div)::value>(k, right, left, cyp[j], cym[j], RB[m][j], RB[c][j],
                                                        RB[n][j], rj, dk2);
          else
            site_update_f32(k, right, left, cyp[j], cym[j], RB[m][j], RB[c][j], RB[n][j], rj, dk2);
          Q[n][j] = rj;
          f3[j] = RB[n][j] + rj * dd2;
        }
      };
      if constexpr (sizeof(T) == 8) {
        bool ok = true;
#pragma unroll
        for (int j = 0; j < W; j++) ok = ok && fast_div_site(RB[n][j]);
        with_div_mode<FAST>(k, ok, stage_c);
      } else {
        stage_c(std::integral_constant<int, DIV_IEEE>());
      }
      __builtin_nontemporal_store(Q[n], (V *)&pio[om2 + row + x]);
      __builtin_nontemporal_store(f3, (V *)&phin[om2 + row + x]);
    }
    if (active && za <= z1) {  // Phi of plane za+2 (over za-1's) and Pi of plane za+1 (over za-2's)
      F[m] = *(const V *)&phi[o2 + row + x];
      Q[n] = *(const V *)&pi[o1 + row + x];
    }
    om2 = om1;
    om1 = o0;
    o0 = o1;
    o1 = o2;
    o2 = o2 + P == NP ? 0 : o2 + P;
    __syncthreads();
  };
  for (int za = z0 - 2; za <= z1 + 1; za += 3) {
    iter(std::integral_constant<int, 0>(), za);
    if (za + 1 > z1 + 1) break;
    iter(std::integral_constant<int, 1>(), za + 1);
    if (za + 2 > z1 + 1) break;
    iter(std::integral_constant<int, 2>(), za + 2);
  }
}
