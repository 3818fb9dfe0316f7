// pn_core.hip -- context, field transfer, the fused leapfrog stencil step and
// the volume reductions of PhiNum's lattice hot path on gfx950.
//
// Device layout: Phi and Pi are dense structure-of-arrays fields of N^3
// values, x fastest (the reference's pos() order without halo planes).  Phi
// is double-buffered: the fused step reads Phi[cur] and writes Phi[cur^1],
// so a site's drift never races with a neighbour's kick; Pi is updated in
// place (each site reads and writes only its own Pi).  The periodic wrap that
// the reference implements with halo planes and Communication::exchange
// (comm.cpp:63-78) and with mod() (util.cpp:12-18) is index arithmetic here.
//
// FP64 arithmetic keeps the reference's expression order and its correctly
// rounded divisions (dynamics.cpp:80,84) -- by a run-constant divisor they are
// computed in four FMA-pipe operations (pn_fastdiv.h) where the operands allow,
// with the same result bits -- and is compiled with -ffp-contract=off, so each
// site's update is bit-identical to a strict build of the reference.
#include <hip/hip_runtime.h>
#include <hip/hip_version.h>
#include <rocfft/rocfft-version.h>
#include <rocfft/rocfft.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <new>
#include <set>
#include <vector>

#include "pn_internal.h"
#include "pn_kernels.h"

// ---------------------------------------------------------------------------
// errors
// ---------------------------------------------------------------------------
static thread_local std::string g_thread_err;

void pn_set_thread_error(const char *msg) { g_thread_err = msg ? msg : ""; }

int pn_fail(pn_ctx *c, int code, const char *fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  if (c) c->err = buf;
  else g_thread_err = buf;
  return code;
}

extern "C" const char *pn_last_error(const pn_ctx *c) {
  return c ? c->err.c_str() : g_thread_err.c_str();
}

extern "C" int pn_abi_version(void) { return PN_ABI_VERSION; }

// ---------------------------------------------------------------------------
// launch geometry
// ---------------------------------------------------------------------------
struct StepGeom {
  int N, XT, YT, ZT;
};

static inline StepGeom step_geom(int N, int by, int zb) {
  StepGeom g;
  g.N = N;
  g.XT = (N + 63) / 64;
  g.YT = (N + by - 1) / by;
  g.ZT = (N + zb - 1) / zb;
  return g;
}

// ---------------------------------------------------------------------------
// K1: fused kick + drift (updateMomenta, updateFields, exchange)
// ---------------------------------------------------------------------------
// One thread per (x,y) column of a 64 x BY tile, streaming ZB planes in z and
// carrying phi(z-1), phi(z), phi(z+1) in registers.  In-plane neighbours are
// read from the same or adjacent cache lines (x) or from rows the tile's other
// waves have just brought into L1/L2 (y).  MODE 0: fused step (writes phin);
// MODE 1: kick only (Pi in place); MODE 2: kick into phin (Pi unchanged).
template <typename T, typename C, bool POW2, int BY, int ZB, int MODE>
__global__ __launch_bounds__(64 * BY) void k_step(const T *__restrict__ phi, T *__restrict__ pi,
                                                  T *__restrict__ phin, StepGeom g, C k, T dtk,
                                                  T dtd) {
  const int nblocks = g.XT * g.YT * g.ZT;
  const int t = xcd_tile(blockIdx.x, nblocks);
  const int xt = t % g.XT;
  const int yt = (t / g.XT) % g.YT;
  const int zt = t / (g.XT * g.YT);
  const int N = g.N;
  const int x = xt * 64 + (int)threadIdx.x;
  const int y = yt * BY + (int)threadIdx.y;
  if (x >= N || y >= N) return;
  const int z0 = zt * ZB;
  const int z1 = min(z0 + ZB, N);
  const int xp = (x + 1 == N) ? 0 : x + 1;
  const int xm = (x == 0) ? N - 1 : x - 1;
  const int yp = (y + 1 == N) ? 0 : y + 1;
  const int ym = (y == 0) ? N - 1 : y - 1;
  const long long P = (long long)N * N;
  const int row = y * N, rowp = yp * N, rowm = ym * N;
  const int zm0 = (z0 == 0) ? N - 1 : z0 - 1;
  T fm = phi[zm0 * P + row + x];
  T fc = phi[z0 * P + row + x];
  for (int z = z0; z < z1; z++) {
    const int zp = (z + 1 == N) ? 0 : z + 1;
    const long long base = z * P;
    const T fp = phi[zp * P + row + x];
    const T vxp = phi[base + row + xp];
    const T vxm = phi[base + row + xm];
    const T vyp = phi[base + rowp + x];
    const T vym = phi[base + rowm + x];
    T p = __builtin_nontemporal_load(&pi[base + row + x]);
    if constexpr (sizeof(T) == 8) {
      site_update_f64<POW2, DIV_PER_SITE>(k, vxp, vxm, vyp, vym, fp, fm, fc, p, dtk);
    } else {
      site_update_f32(k, vxp, vxm, vyp, vym, fp, fm, fc, p, dtk);
    }
    if constexpr (MODE == 2) {
      __builtin_nontemporal_store(p, &phin[base + row + x]);  // the kicked Pi, elsewhere
    } else {
      __builtin_nontemporal_store(p, &pi[base + row + x]);
    }
    if constexpr (MODE == 0) {
      const T fnew = fc + p * dtd;
      __builtin_nontemporal_store(fnew, &phin[base + row + x]);
    }
    fm = fc;
    fc = fp;
  }
}

// Vector form: each thread owns W consecutive x sites of one row (16-byte
// loads and stores: W = 2 doubles or 4 floats), a wave covers 64*W sites of
// the row.  Neighbours inside the vector come from registers; only the two
// outer x neighbours are separate loads.  Needs N % W == 0.
template <typename T, typename C, bool POW2, int W, int BY, int ZB, int MODE>
__global__ __launch_bounds__(64 * BY) void k_step_vec(const T *__restrict__ phi, T *__restrict__ pi,
                                                      T *__restrict__ phin, StepGeom g, C k,
                                                      T dtk, T dtd) {
  typedef typename VecT<T, W>::type V;
  const int nblocks = g.XT * g.YT * g.ZT;
  const int t = xcd_tile(blockIdx.x, nblocks);
  const int xt = t % g.XT;
  const int yt = (t / g.XT) % g.YT;
  const int zt = t / (g.XT * g.YT);
  const int N = g.N;
  const int x = xt * 64 * W + W * (int)threadIdx.x;
  const int y = yt * BY + (int)threadIdx.y;
  if (x >= N || y >= N) return;
  const int z0 = zt * ZB;
  const int z1 = min(z0 + ZB, N);
  const int xp = (x + W == N) ? 0 : x + W;
  const int xm = (x == 0) ? N - 1 : x - 1;
  const int yp = (y + 1 == N) ? 0 : y + 1;
  const int ym = (y == 0) ? N - 1 : y - 1;
  const long long P = (long long)N * N;
  const int row = y * N, rowp = yp * N, rowm = ym * N;
  const int zm0 = (z0 == 0) ? N - 1 : z0 - 1;
  V fm = *(const V *)&phi[zm0 * P + row + x];
  V fc = *(const V *)&phi[z0 * P + row + x];
  for (int z = z0; z < z1; z++) {
    const int zp = (z + 1 == N) ? 0 : z + 1;
    const long long base = z * P;
    const V fp = *(const V *)&phi[zp * P + row + x];
    const T vxp = phi[base + row + xp];
    const T vxm = phi[base + row + xm];
    const V vyp = *(const V *)&phi[base + rowp + x];
    const V vym = *(const V *)&phi[base + rowm + x];
    V p = __builtin_nontemporal_load((const V *)&pi[base + row + x]);
    if constexpr (sizeof(T) == 8) {
      bool ok = true;
#pragma unroll
      for (int j = 0; j < W; j++) ok = ok && fast_div_site(fc[j]);
      with_div_mode(k, ok, [&](auto div) {
#pragma unroll
        for (int j = 0; j < W; j++) {
          const T right = (j == W - 1) ? vxp : fc[(j + 1) % W];
          const T left = (j == 0) ? vxm : fc[(j + W - 1) % W];
          T pj = p[j];
          site_update_f64<POW2, decltype(div)::value>(k, right, left, vyp[j], vym[j], fp[j], fm[j],
                                                      fc[j], pj, dtk);
          p[j] = pj;
        }
      });
    } else {
#pragma unroll
      for (int j = 0; j < W; j++) {
        const T right = (j == W - 1) ? vxp : fc[(j + 1) % W];
        const T left = (j == 0) ? vxm : fc[(j + W - 1) % W];
        T pj = p[j];
        site_update_f32(k, right, left, vyp[j], vym[j], fp[j], fm[j], fc[j], pj, dtk);
        p[j] = pj;
      }
    }
    if constexpr (MODE == 2) {
      __builtin_nontemporal_store(p, (V *)&phin[base + row + x]);  // the kicked Pi, elsewhere
    } else {
      __builtin_nontemporal_store(p, (V *)&pi[base + row + x]);
    }
    if constexpr (MODE == 0) {
      V fnew;
#pragma unroll
      for (int j = 0; j < W; j++) fnew[j] = fc[j] + p[j] * dtd;
      __builtin_nontemporal_store(fnew, (V *)&phin[base + row + x]);
    }
    fm = fc;
    fc = fp;
  }
}

// Two leapfrog steps per pass over HBM (temporal blocking): one block per
// tile of RT full x rows (N/W threads of W sites each) and ZB planes; the
// tile's work is step2_tile.  Needs N % W == 0, (N/W)*RT <= 1024 and
// 4*RT*N values of LDS (2*RT*N without the Phi planes, LP false).
template <typename T, typename C, bool POW2, int W, int RT, bool LP = true, int FAST = DIV_FAST>
__global__ __launch_bounds__(1024) void k_step2(const T *__restrict__ phi, const T *__restrict__ pi,
                                                T *__restrict__ pio, T *__restrict__ phin, int N,
                                                int ZB, C k, T dk0, T dd0, T dk1, T dd1) {
  extern __shared__ unsigned char lds_raw[];
  const int YT = (N + RT - 3) / (RT - 2);
  // y-neighbouring tiles (which re-read each other's edge rows) on one XCD
  const int tile = xcd_tile((int)blockIdx.x, (int)gridDim.x);
  const int yt = tile % YT, zt = tile / YT;
  const int z0 = zt * ZB;
  step2_tile<T, C, POW2, W, true, LP, FAST>(phi, pi, pio, phin, (T *)lds_raw, N, RT, (int)threadIdx.x,
                            (int)threadIdx.y, true, yt * (RT - 2), z0, min(z0 + ZB, N), k, dk0, dd0,
                            dk1, dd1);
}

// LDS planes of the three-step tile: the unrolled tile (a ring of three
// planes per buffer) where it keeps the blocks per CU of the two-plane tile;
// the FP32 12-row tile fits two blocks per CU only with two-plane buffers
// (74 KB and 78 registers against 110 KB and 111), and keeps them.
template <typename T, int RT>
constexpr int step3_planes() {
  return (sizeof(T) == 4 && RT == 12) ? 6 : 9;
}

// Three leapfrog steps per pass (step3_tile_r3 or step3_tile): tiles of RT
// rows with RT-4 output rows and ZB planes; step3_planes() planes of RT*N
// values in LDS.
template <typename T, typename C, bool POW2, int W, int RT, int FAST = DIV_FAST>
__global__ __launch_bounds__(1024) void k_step3(const T *__restrict__ phi, const T *__restrict__ pi,
                                                T *__restrict__ pio, T *__restrict__ phin, int N,
                                                int ZB, C k, T dk0, T dd0, T dk1, T dd1, T dk2, T dd2) {
  extern __shared__ unsigned char lds_raw[];
  const int TY = RT - 4, YT = (N + TY - 1) / TY;
  const int tile = xcd_tile((int)blockIdx.x, (int)gridDim.x);
  const int yt = tile % YT, zt = tile / YT;
  const int z0 = zt * ZB;
  if constexpr (step3_planes<T, RT>() == 9)
    step3_tile_r3<T, C, POW2, W, FAST>(phi, pi, pio, phin, (T *)lds_raw, N, RT, (int)threadIdx.x,
                                       (int)threadIdx.y, true, yt * TY, z0, min(z0 + ZB, N), k, dk0,
                                       dd0, dk1, dd1, dk2, dd2);
  else
    step3_tile<T, C, POW2, W, FAST>(phi, pi, pio, phin, (T *)lds_raw, N, RT, (int)threadIdx.x,
                                    (int)threadIdx.y, yt * TY, z0, min(z0 + ZB, N), k, dk0, dd0, dk1,
                                    dd1, dk2, dd2);
}

// updateFields alone (dynamics.cpp:182-183): Phi += Pi*dtd in place.
template <typename T>
__global__ void k_drift(T *__restrict__ phi, const T *__restrict__ pi, long long n, T dtd) {
  long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const long long stride = (long long)gridDim.x * blockDim.x;
  for (; i < n; i += stride) phi[i] = phi[i] + pi[i] * dtd;
}

// tile shape used by the single-lattice path (the batched path autotunes)
static constexpr int kBY = 4;
static constexpr int kZB = 16;

template <typename T, typename C, bool POW2, int MODE>
static void launch_step_t(pn_ctx *c, const C &k, T dtk, T dtd) {
  constexpr int W = 16 / sizeof(T);
  const int N = c->p.N;
  const T *phi = (const T *)c->phi[c->cur];
  T *phin = (T *)c->phi[c->cur ^ 1];
  if (N % W == 0) {
    // z blocks of 32 planes below 512^3 keep enough tiles in flight; 16 above
    if (N < 512) {
      StepGeom g = step_geom(N, kBY, 32);
      g.XT = (N + 64 * W - 1) / (64 * W);
      const int nblocks = g.XT * g.YT * g.ZT;
      hipLaunchKernelGGL((k_step_vec<T, C, POW2, W, kBY, 32, MODE>), dim3(nblocks), dim3(64, kBY),
                         0, c->stream, phi, (T *)c->pi, phin, g, k, dtk, dtd);
    } else {
      StepGeom g = step_geom(N, kBY, kZB);
      g.XT = (N + 64 * W - 1) / (64 * W);
      const int nblocks = g.XT * g.YT * g.ZT;
      hipLaunchKernelGGL((k_step_vec<T, C, POW2, W, kBY, kZB, MODE>), dim3(nblocks), dim3(64, kBY),
                         0, c->stream, phi, (T *)c->pi, phin, g, k, dtk, dtd);
    }
    return;
  }
  StepGeom g = step_geom(N, kBY, kZB);
  int nblocks = g.XT * g.YT * g.ZT;
  hipLaunchKernelGGL((k_step<T, C, POW2, kBY, kZB, MODE>), dim3(nblocks), dim3(64, kBY), 0,
                     c->stream, phi, (T *)c->pi, phin, g, k, dtk, dtd);
}

static CoefF64 coef64(const pn_ctx *c) {
  CoefF64 k;
  k.a2 = c->k.a2;
  k.inv_a2 = c->k.inv_a2;
  k.c2 = c->k.c2;
  k.lmode = c->k.lmode;
  k.kappa = c->k.kappa;
  coef64_recips(k);
  return k;
}

static CoefF32 coef32(const pn_ctx *c) {
  CoefF32 k;
  k.inv_a2 = (float)(1.0 / c->k.a2);
  k.c2 = (float)c->k.c2;
  k.c4 = (float)(c->k.lmode / 6.0);
  k.c6 = (float)(1.0 / (120.0 * c->k.kappa * c->k.kappa));
  return k;
}

// MODE 0 fused step, MODE 1 kick only
static int launch_step(pn_ctx *c, int mode, double dtk, double dtd) {
  if (c->p.precision == PN_FP64) {
    CoefF64 k = coef64(c);
    if (c->k.a2_pow2) {
      if (mode == 0) launch_step_t<double, CoefF64, true, 0>(c, k, dtk, dtd);
      else if (mode == 1) launch_step_t<double, CoefF64, true, 1>(c, k, dtk, dtd);
      else launch_step_t<double, CoefF64, true, 2>(c, k, dtk, dtd);
    } else {
      if (mode == 0) launch_step_t<double, CoefF64, false, 0>(c, k, dtk, dtd);
      else if (mode == 1) launch_step_t<double, CoefF64, false, 1>(c, k, dtk, dtd);
      else launch_step_t<double, CoefF64, false, 2>(c, k, dtk, dtd);
    }
  } else {
    CoefF32 k = coef32(c);
    if (mode == 0) launch_step_t<float, CoefF32, true, 0>(c, k, (float)dtk, (float)dtd);
    else if (mode == 1) launch_step_t<float, CoefF32, true, 1>(c, k, (float)dtk, (float)dtd);
    else launch_step_t<float, CoefF32, true, 2>(c, k, (float)dtk, (float)dtd);
  }
  PN_CHECK_LAUNCH(c);
  if (mode == 0) c->cur ^= 1;
  return PN_OK;
}

// Two-step tile shape: W sites per thread, RT thread rows, ZB planes per block.
// Full rows need (N/W)*RT <= 1024 threads; LDS holds four planes of RT*N
// values (two of stage A's results, two of Phi; LP 0: the first two only), up
// to the device's limit (160 KB a block on MI355X).
struct Step2Shape {
  int W, RT, ZB, LP;
};

static bool step2_shape(const pn_ctx *c, Step2Shape *s) {
  const int N = c->p.N;
  const bool dp = c->p.precision == PN_FP64;
  if (N < 4) return false;
  // measured on MI355X at 256^3 and 512^3 (ms per step, FP64: 0.059 / 0.477,
  // FP32: 0.030 / 0.207); neighbouring shapes are within a few percent
  if (dp) {
    if (N <= 256 && N % 2 == 0) *s = {2, 6, 32, 1};
    else if (N <= 512 && N % 4 == 0) *s = {4, 6, 64, 1};
    // at N = 1024, 256 threads a row leave room for four rows only (two output
    // rows per tile: stage A runs on twice the output rows)
    else if (N <= 1024 && N % 4 == 0) *s = {4, 4, 64, 1};
    else return false;
  } else {
    if (N <= 256 && N % 4 == 0) *s = {4, N >= 8 ? 10 : 8, 16, 1};
    else if (N <= 512 && N % 4 == 0) *s = {4, 6, 64, 1};
    else if (N <= 1024 && N % 8 == 0) *s = {8, 8, 64, 1};
    else return false;
  }
  if (s->ZB > N) s->ZB = N;
  // a tile's rows y0-1 .. y0+RT-2 must wrap into [0, N) with one fold
  const int TY = s->RT - 2, YT = (N + TY - 1) / TY;
  if ((YT - 1) * TY + s->RT - 2 >= 2 * N) return false;
  const size_t lds = (size_t)(s->LP ? 4 : 2) * s->RT * N * c->esize;
  return (N / s->W) * s->RT <= 1024 && lds <= (size_t)c->lds_max;
}

// Raises a kernel's dynamic LDS limit to the device's maximum, once per
// (kernel, device): the attribute is per device, and the launches of a
// context run on its device, which is current.
static void raise_lds_limit(const pn_ctx *c, const void *kernel) {
  static std::mutex mu;
  static std::set<std::pair<const void *, int>> done;
  std::lock_guard<std::mutex> lock(mu);
  if (done.insert({kernel, c->p.device}).second)
    (void)hipFuncSetAttribute(kernel, hipFuncAttributeMaxDynamicSharedMemorySize, c->lds_max);
}

template <typename T, typename C, bool POW2, int W, int RT, bool LP = true, int FAST = DIV_FAST>
static void launch_step2_t(pn_ctx *c, const C &k, int ZB, T dk0, T dd0, T dk1, T dd1) {
  const int N = c->p.N;
  const int YT = (N + RT - 3) / (RT - 2), ZT = (N + ZB - 1) / ZB;
  const size_t lds = (size_t)(LP ? 4 : 2) * RT * N * sizeof(T);
  if (lds > 65536)  // beyond 64 KB the kernel's dynamic LDS limit is raised (once per device)
    raise_lds_limit(c, (const void *)&k_step2<T, C, POW2, W, RT, LP, FAST>);
  hipLaunchKernelGGL((k_step2<T, C, POW2, W, RT, LP, FAST>), dim3(YT * ZT), dim3(N / W, RT), lds, c->stream,
                     (const T *)c->phi[c->cur], (const T *)c->pi, (T *)c->pi_alt,
                     (T *)c->phi[c->cur ^ 1], N, ZB, k, dk0, dd0, dk1, dd1);
}

// Three-step tile shape (W, RT, ZB): step3_planes() LDS planes of RT*N values must fit.
static bool step3_shape(const pn_ctx *c, Step2Shape *s) {
  const int N = c->p.N;
  const bool dp = c->p.precision == PN_FP64;
  if (N < 8) return false;
  if (dp) {
    if (N <= 256 && N % 2 == 0) *s = {2, 8, 32, 1};
    else return false;
  } else {
    if (N <= 256 && N % 4 == 0) *s = {4, 12, 32, 1};
    else if (N <= 512 && N % 4 == 0) *s = {4, 8, 32, 1};
    else return false;
  }
  if (s->ZB > N) s->ZB = N;
  const int TY = s->RT - 4, YT = (N + TY - 1) / TY;
  if ((YT - 1) * TY + s->RT - 3 >= 2 * N) return false;
  const int planes = dp ? step3_planes<double, 8>()
                        : (s->RT == 12 ? step3_planes<float, 12>() : step3_planes<float, 8>());
  const size_t lds = (size_t)planes * s->RT * N * c->esize;
  return (N / s->W) * s->RT <= 1024 && lds <= (size_t)c->lds_max;
}

template <typename T, typename C, bool POW2, int W, int RT, int FAST = DIV_FAST>
static void launch_step3_t(pn_ctx *c, const C &k, int ZB, const T *d) {
  const int N = c->p.N;
  const int YT = (N + RT - 5) / (RT - 4), ZT = (N + ZB - 1) / ZB;
  const size_t lds = (size_t)step3_planes<T, RT>() * RT * N * sizeof(T);
  if (lds > 65536) raise_lds_limit(c, (const void *)&k_step3<T, C, POW2, W, RT, FAST>);
  hipLaunchKernelGGL((k_step3<T, C, POW2, W, RT, FAST>), dim3(YT * ZT), dim3(N / W, RT), lds, c->stream,
                     (const T *)c->phi[c->cur], (const T *)c->pi, (T *)c->pi_alt,
                     (T *)c->phi[c->cur ^ 1], N, ZB, k, d[0], d[1], d[2], d[3], d[4], d[5]);
}

// steps n, n+1, n+2 in one pass; Pi moves to the second buffer
static int launch_step3(pn_ctx *c, const Step2Shape &s, const double *dtk, const double *dtd) {
  if (c->p.precision == PN_FP64) {
    CoefF64 k = coef64(c);
    const double d[6] = {dtk[0], dtd[0], dtk[1], dtd[1], dtk[2], dtd[2]};
    if (k.rk_bmr) {
      if (c->k.a2_pow2) launch_step3_t<double, CoefF64, true, 2, 8, DIV_FAST_2OP>(c, k, s.ZB, d);
      else launch_step3_t<double, CoefF64, false, 2, 8, DIV_FAST_2OP>(c, k, s.ZB, d);
    } else {
      if (c->k.a2_pow2) launch_step3_t<double, CoefF64, true, 2, 8, DIV_FAST_KC>(c, k, s.ZB, d);
      else launch_step3_t<double, CoefF64, false, 2, 8, DIV_FAST_KC>(c, k, s.ZB, d);
    }
  } else {
    CoefF32 k = coef32(c);
    const float d[6] = {(float)dtk[0], (float)dtd[0], (float)dtk[1], (float)dtd[1], (float)dtk[2], (float)dtd[2]};
    if (s.RT == 12) launch_step3_t<float, CoefF32, true, 4, 12>(c, k, s.ZB, d);
    else launch_step3_t<float, CoefF32, true, 4, 8>(c, k, s.ZB, d);
  }
  PN_CHECK_LAUNCH(c);
  c->cur ^= 1;
  std::swap(c->pi, c->pi_alt);
  return PN_OK;
}

template <int FAST>
static void launch_step2_f64(pn_ctx *c, const CoefF64 &k, const Step2Shape &s, const double *dtk,
                             const double *dtd) {
  const double a0 = dtk[0], b0 = dtd[0], a1 = dtk[1], b1 = dtd[1];
  if (s.W == 2) {
    if (c->k.a2_pow2) launch_step2_t<double, CoefF64, true, 2, 6, true, FAST>(c, k, s.ZB, a0, b0, a1, b1);
    else launch_step2_t<double, CoefF64, false, 2, 6, true, FAST>(c, k, s.ZB, a0, b0, a1, b1);
  } else if (s.RT == 6) {
    if (c->k.a2_pow2) launch_step2_t<double, CoefF64, true, 4, 6, true, FAST>(c, k, s.ZB, a0, b0, a1, b1);
    else launch_step2_t<double, CoefF64, false, 4, 6, true, FAST>(c, k, s.ZB, a0, b0, a1, b1);
  } else {
    if (c->k.a2_pow2) launch_step2_t<double, CoefF64, true, 4, 4, true, FAST>(c, k, s.ZB, a0, b0, a1, b1);
    else launch_step2_t<double, CoefF64, false, 4, 4, true, FAST>(c, k, s.ZB, a0, b0, a1, b1);
  }
}

// steps n and n+1 in one pass; Pi moves to the second buffer
static int launch_step2(pn_ctx *c, const Step2Shape &s, const double *dtk, const double *dtd) {
  if (c->p.precision == PN_FP64) {
    CoefF64 k = coef64(c);
    // q/kappa in two operations, or with the correction step: a separate
    // kernel each, so neither carries the other's instructions
    if (k.rk_bmr) launch_step2_f64<DIV_FAST_2OP>(c, k, s, dtk, dtd);
    else launch_step2_f64<DIV_FAST_KC>(c, k, s, dtk, dtd);
  } else {
    CoefF32 k = coef32(c);
    const float a0 = (float)dtk[0], b0 = (float)dtd[0], a1 = (float)dtk[1], b1 = (float)dtd[1];
    if (s.W == 4 && s.RT == 10) launch_step2_t<float, CoefF32, true, 4, 10>(c, k, s.ZB, a0, b0, a1, b1);
    else if (s.W == 4 && s.RT == 6) launch_step2_t<float, CoefF32, true, 4, 6>(c, k, s.ZB, a0, b0, a1, b1);
    else if (s.W == 4) launch_step2_t<float, CoefF32, true, 4, 8>(c, k, s.ZB, a0, b0, a1, b1);
    else launch_step2_t<float, CoefF32, true, 8, 8>(c, k, s.ZB, a0, b0, a1, b1);
  }
  PN_CHECK_LAUNCH(c);
  c->cur ^= 1;
  std::swap(c->pi, c->pi_alt);
  return PN_OK;
}

// ---------------------------------------------------------------------------
// context
// ---------------------------------------------------------------------------
static bool is_pow2_double(double v) {
  if (!(v > 0) || !std::isfinite(v)) return false;
  int e;
  double m = std::frexp(v, &e);
  return m == 0.5 && e > -1000 && e < 1000;
}

int pn_reduce_blocks(const pn_ctx *c) { return c->red_blocks; }

static constexpr int kRedThreads = 256;
static constexpr int kNSums = 9;   // phi, pi, phi2, pi2, phipi, phi4, phi6, grad, inter
// block partials of the volume sums per call (and their grid's size cap), and
// tiles per call (N <= 2048 has at most 8 x 2048 columns of tiles)
static constexpr int kVolMaxBlocks = 2048;
static constexpr int kVolMaxTiles = 16384;
// devices a process may use with the per-device caches below
static constexpr int kScratchMaxDevices = 64;

// A process that loaded another HIP runtime or rocFFT before this library (a
// framework that bundles its own ROCm libraries under the same sonames) binds
// the library's calls to those: rocFFT plans described through this build's
// headers are then executed by another version's implementation, and code
// objects built by this compiler are loaded by another runtime.  Refused
// before any device work.
int pn_runtime_check() {
  static std::once_flag once;
  static int rc = PN_OK;
  static char msg[256];
  std::call_once(once, [] {
    int v = 0;
    if (hipRuntimeGetVersion(&v) != hipSuccess) {
      rc = PN_ERR_HIP;
      std::snprintf(msg, sizeof msg, "hipRuntimeGetVersion failed");
      return;
    }
    if (v / 100000 != HIP_VERSION / 100000) {
      rc = PN_ERR_STATE;
      std::snprintf(msg, sizeof msg,
                    "the process loaded HIP runtime %d.%d, this library is built for HIP %d.%d: load it "
                    "before any other ROCm library, in a process of its own",
                    v / 10000000, (v / 100000) % 100, HIP_VERSION_MAJOR, HIP_VERSION_MINOR);
      return;
    }
    char got[128] = {0}, want[64];
    std::snprintf(want, sizeof want, "%d.%d.%d.", rocfft_version_major, rocfft_version_minor,
                  rocfft_version_patch);
    if (rocfft_get_version_string(got, sizeof got - 1) != rocfft_status_success ||
        std::strncmp(got, want, std::strlen(want)) != 0) {
      rc = PN_ERR_STATE;
      std::snprintf(msg, sizeof msg,
                    "the process loaded rocFFT %s, this library is built for rocFFT %s: load it before "
                    "any other ROCm library, in a process of its own",
                    got, want);
    }
  });
  if (rc != PN_OK) return pn_fail(nullptr, rc, "%s", msg);
  return PN_OK;
}

extern "C" int pn_host_alloc(size_t bytes, void **out) {
  if (!out) return pn_fail(nullptr, PN_ERR_ARG, "pn_host_alloc: out is NULL");
  *out = nullptr;
  if (bytes == 0) return pn_fail(nullptr, PN_ERR_ARG, "pn_host_alloc: zero bytes");
  if (const int rc = pn_runtime_check()) return rc;
  if (hipHostMalloc(out, bytes, hipHostMallocDefault) != hipSuccess) {
    *out = nullptr;
    return pn_fail(nullptr, PN_ERR_NOMEM, "pn_host_alloc: %zu bytes of page-locked memory", bytes);
  }
  return PN_OK;
}

extern "C" int pn_host_free(void *p) {
  if (p && hipHostFree(p) != hipSuccess) return pn_fail(nullptr, PN_ERR_HIP, "pn_host_free failed");
  return PN_OK;
}

extern "C" int pn_create(const pn_params *p, pn_ctx **out) {
  if (!out) return pn_fail(nullptr, PN_ERR_ARG, "pn_create: out is NULL");
  *out = nullptr;
  if (!p) return pn_fail(nullptr, PN_ERR_ARG, "pn_create: params is NULL");
  if (p->N < 4 || p->N > 2048)
    return pn_fail(nullptr, PN_ERR_ARG, "pn_create: N=%d outside [4, 2048]", p->N);
  if (!(p->a > 0) || !std::isfinite(p->a))
    return pn_fail(nullptr, PN_ERR_ARG, "pn_create: lattice spacing a must be > 0");
  if (p->precision != PN_FP64 && p->precision != PN_FP32)
    return pn_fail(nullptr, PN_ERR_ARG, "pn_create: unknown precision %d", p->precision);
  if (!(p->kappa != 0) || !std::isfinite(p->kappa))
    return pn_fail(nullptr, PN_ERR_ARG, "pn_create: Kappa must be finite and nonzero");
  if (const int rc = pn_runtime_check()) return rc;
  int ndev = 0;
  hipError_t e = hipGetDeviceCount(&ndev);
  if (e != hipSuccess || ndev == 0)
    return pn_fail(nullptr, PN_ERR_HIP, "pn_create: no HIP device (%s)", hipGetErrorString(e));
  if (p->device < 0 || p->device >= ndev)
    return pn_fail(nullptr, PN_ERR_ARG, "pn_create: device %d of %d", p->device, ndev);

  pn_ctx *c = new (std::nothrow) pn_ctx();
  if (!c) return pn_fail(nullptr, PN_ERR_NOMEM, "pn_create: out of host memory");
  c->p = *p;
  c->n3 = (long long)p->N * p->N * p->N;
  c->esize = p->precision == PN_FP64 ? 8 : 4;
  c->k.a2 = p->a * p->a;
  c->k.a2_pow2 = is_pow2_double(c->k.a2);
  c->k.inv_a2 = 1.0 / c->k.a2;
  c->k.c2 = p->mass_mode * p->m2;
  c->k.lmode = (double)p->lambda_mode;
  c->k.kappa = p->kappa;
  c->red_blocks = kVolMaxBlocks;

  auto cleanup = [&](int code) {
    pn_destroy(c);
    return code;
  };
  if (hipSetDevice(p->device) != hipSuccess)
    return cleanup(pn_fail(nullptr, PN_ERR_HIP, "pn_create: hipSetDevice failed"));
  if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess)
    return cleanup(pn_fail(nullptr, PN_ERR_HIP, "pn_create: stream creation failed"));
  if (hipDeviceGetAttribute(&c->lds_max, hipDeviceAttributeMaxSharedMemoryPerBlock, p->device) != hipSuccess)
    c->lds_max = 65536;
  size_t bytes = (size_t)c->n3 * c->esize;
  for (int i = 0; i < 2; i++)
    if (hipMalloc(&c->phi[i], bytes) != hipSuccess)
      return cleanup(pn_fail(nullptr, PN_ERR_NOMEM, "pn_create: cannot allocate %zu bytes", bytes));
  if (hipMalloc(&c->pi, bytes) != hipSuccess)
    return cleanup(pn_fail(nullptr, PN_ERR_NOMEM, "pn_create: cannot allocate %zu bytes", bytes));
  Step2Shape s2;
  if (step2_shape(c, &s2)) {
    // the two-step kernel needs a second Pi buffer; without the room, step singly
    if (hipMalloc(&c->pi_alt, bytes) == hipSuccess) c->step2 = 1;
    else (void)hipGetLastError();
  }
  if (hipMalloc(&c->red_part, (size_t)c->red_blocks * 2 * kNSums * sizeof(double)) != hipSuccess ||
      hipMalloc(&c->ext_part, (size_t)c->red_blocks * 6 * sizeof(long long)) != hipSuccess ||
      hipMalloc(&c->tile_flag, (size_t)(kVolMaxTiles + kVolMaxBlocks) * sizeof(int)) != hipSuccess ||
      hipMalloc(&c->red_out, 64 * sizeof(double)) != hipSuccess)
    return cleanup(pn_fail(nullptr, PN_ERR_NOMEM, "pn_create: cannot allocate reduction scratch"));
  if (hipHostMalloc(&c->host_pin, 64 * sizeof(double), hipHostMallocDefault) != hipSuccess)
    return cleanup(pn_fail(nullptr, PN_ERR_NOMEM, "pn_create: cannot allocate pinned staging"));
  if (hipMemsetAsync(c->phi[0], 0, bytes, c->stream) != hipSuccess ||
      hipMemsetAsync(c->phi[1], 0, bytes, c->stream) != hipSuccess ||
      hipMemsetAsync(c->pi, 0, bytes, c->stream) != hipSuccess ||
      hipStreamSynchronize(c->stream) != hipSuccess)
    return cleanup(pn_fail(nullptr, PN_ERR_HIP, "pn_create: zero fill failed"));
  c->cur = 0;
  *out = c;
  return PN_OK;
}

int pn_make_view(const pn_params *p, void *phi0, void *phi1, void *pi, int cur, hipStream_t s,
                 pn_ctx **out) {
  pn_ctx *c = new (std::nothrow) pn_ctx();
  if (!c) return pn_fail(nullptr, PN_ERR_NOMEM, "pn_make_view: out of host memory");
  c->p = *p;
  c->is_view = 1;
  c->n3 = (long long)p->N * p->N * p->N;
  c->esize = p->precision == PN_FP64 ? 8 : 4;
  c->k.a2 = p->a * p->a;
  c->k.a2_pow2 = is_pow2_double(c->k.a2);
  c->k.inv_a2 = 1.0 / c->k.a2;
  c->k.c2 = p->mass_mode * p->m2;
  c->k.lmode = (double)p->lambda_mode;
  c->k.kappa = p->kappa;
  c->red_blocks = kVolMaxBlocks;
  c->phi[0] = phi0;
  c->phi[1] = phi1;
  c->pi = pi;
  c->cur = cur;
  c->stream = s;
  if (hipMalloc(&c->red_part, (size_t)c->red_blocks * 2 * kNSums * sizeof(double)) != hipSuccess ||
      hipMalloc(&c->ext_part, (size_t)c->red_blocks * 6 * sizeof(long long)) != hipSuccess ||
      hipMalloc(&c->tile_flag, (size_t)(kVolMaxTiles + kVolMaxBlocks) * sizeof(int)) != hipSuccess ||
      hipMalloc(&c->red_out, 64 * sizeof(double)) != hipSuccess ||
      hipHostMalloc(&c->host_pin, 64 * sizeof(double), hipHostMallocDefault) != hipSuccess) {
    pn_destroy(c);
    return pn_fail(nullptr, PN_ERR_NOMEM, "pn_make_view: cannot allocate reduction scratch");
  }
  *out = c;
  return PN_OK;
}

extern "C" int pn_set_couplings(pn_ctx *c, int mass_mode, double m2, int lambda_mode, double kappa) {
  if (!c) return pn_fail(nullptr, PN_ERR_ARG, "pn_set_couplings: NULL context");
  if (!(kappa != 0) || !std::isfinite(kappa))
    return pn_fail(c, PN_ERR_ARG, "pn_set_couplings: Kappa must be finite and nonzero");
  c->p.mass_mode = mass_mode;
  c->p.m2 = m2;
  c->p.lambda_mode = lambda_mode;
  c->p.kappa = kappa;
  c->k.c2 = mass_mode * m2;
  c->k.lmode = (double)lambda_mode;
  c->k.kappa = kappa;
  return PN_OK;
}

extern "C" int pn_destroy(pn_ctx *c) {
  if (!c) return PN_OK;
  hipSetDevice(c->p.device);
  if (c->stream) hipStreamSynchronize(c->stream);
  pn_free_spectrum(c);
  pn_free_islands(c);
  if (!c->is_view) {
    for (int i = 0; i < 2; i++)
      if (c->phi[i]) (void)hipFree(c->phi[i]);
    if (c->pi) (void)hipFree(c->pi);
    if (c->pi_alt) (void)hipFree(c->pi_alt);
  }
  if (c->red_part) (void)hipFree(c->red_part);
  if (c->ext_part) (void)hipFree(c->ext_part);
  if (c->tile_flag) (void)hipFree(c->tile_flag);
  if (c->red_out) (void)hipFree(c->red_out);
  if (c->host_pin) (void)hipHostFree(c->host_pin);
  if (c->stream && !c->is_view) (void)hipStreamDestroy(c->stream);
  delete c;
  return PN_OK;
}

// ---------------------------------------------------------------------------
// fields
// ---------------------------------------------------------------------------
static int put_field(pn_ctx *c, void *dst, const double *src, int layout) {
  const double *interior = layout == PN_LAYOUT_REF ? src + (long long)c->p.N * c->p.N : src;
  size_t n = (size_t)c->n3;
  if (c->p.precision == PN_FP64) {
    PN_HIP(c, hipMemcpyAsync(dst, interior, n * 8, hipMemcpyHostToDevice, c->stream));
    PN_HIP(c, hipStreamSynchronize(c->stream));
  } else {
    std::vector<float> tmp(n);
    for (size_t i = 0; i < n; i++) tmp[i] = (float)interior[i];
    PN_HIP(c, hipMemcpyAsync(dst, tmp.data(), n * 4, hipMemcpyHostToDevice, c->stream));
    PN_HIP(c, hipStreamSynchronize(c->stream));
  }
  return PN_OK;
}

static int get_field(pn_ctx *c, const void *src, double *dst, int layout) {
  const long long plane = (long long)c->p.N * c->p.N;
  double *interior = layout == PN_LAYOUT_REF ? dst + plane : dst;
  size_t n = (size_t)c->n3;
  if (c->p.precision == PN_FP64) {
    PN_HIP(c, hipMemcpyAsync(interior, src, n * 8, hipMemcpyDeviceToHost, c->stream));
    PN_HIP(c, hipStreamSynchronize(c->stream));
  } else {
    std::vector<float> tmp(n);
    PN_HIP(c, hipMemcpyAsync(tmp.data(), src, n * 4, hipMemcpyDeviceToHost, c->stream));
    PN_HIP(c, hipStreamSynchronize(c->stream));
    for (size_t i = 0; i < n; i++) interior[i] = (double)tmp[i];
  }
  if (layout == PN_LAYOUT_REF) {
    const int N = c->p.N;
    std::memcpy(dst + plane * (N + 1), dst + plane, plane * sizeof(double));
    std::memcpy(dst, dst + plane * N, plane * sizeof(double));
  }
  return PN_OK;
}

extern "C" int pn_set_fields(pn_ctx *c, const double *phi, const double *pi, int layout) {
  if (!c) return pn_fail(nullptr, PN_ERR_ARG, "pn_set_fields: NULL context");
  if (layout != PN_LAYOUT_REF && layout != PN_LAYOUT_DENSE)
    return pn_fail(c, PN_ERR_ARG, "pn_set_fields: unknown layout %d", layout);
  PN_HIP(c, hipSetDevice(c->p.device));
  int rc;
  if (phi && (rc = put_field(c, c->phi[c->cur], phi, layout)) != PN_OK) return rc;
  if (pi && (rc = put_field(c, c->pi, pi, layout)) != PN_OK) return rc;
  return PN_OK;
}

extern "C" int pn_get_fields(pn_ctx *c, double *phi, double *pi, int layout) {
  if (!c) return pn_fail(nullptr, PN_ERR_ARG, "pn_get_fields: NULL context");
  if (layout != PN_LAYOUT_REF && layout != PN_LAYOUT_DENSE)
    return pn_fail(c, PN_ERR_ARG, "pn_get_fields: unknown layout %d", layout);
  PN_HIP(c, hipSetDevice(c->p.device));
  int rc;
  if (phi && (rc = get_field(c, c->phi[c->cur], phi, layout)) != PN_OK) return rc;
  if (pi && (rc = get_field(c, c->pi, pi, layout)) != PN_OK) return rc;
  return PN_OK;
}

// ---------------------------------------------------------------------------
// dynamics entry points
// ---------------------------------------------------------------------------
extern "C" int pn_kick(pn_ctx *c, double dtk) {
  if (!c) return pn_fail(nullptr, PN_ERR_ARG, "pn_kick: NULL context");
  if (c->is_view) return pn_fail(c, PN_ERR_STATE, "pn_kick: a batch view is stepped by pn_batch_evolve");
  PN_HIP(c, hipSetDevice(c->p.device));
  int rc = launch_step(c, 1, dtk, 0.0);
  if (rc != PN_OK) return rc;
  PN_HIP(c, hipStreamSynchronize(c->stream));
  return PN_OK;
}

extern "C" int pn_drift(pn_ctx *c, double dtd) {
  if (!c) return pn_fail(nullptr, PN_ERR_ARG, "pn_drift: NULL context");
  if (c->is_view) return pn_fail(c, PN_ERR_STATE, "pn_drift: a batch view is stepped by pn_batch_evolve");
  PN_HIP(c, hipSetDevice(c->p.device));
  const int threads = 256;
  const int blocks = (int)std::min<long long>(4096, (c->n3 + threads - 1) / threads);
  if (c->p.precision == PN_FP64)
    hipLaunchKernelGGL(k_drift<double>, dim3(blocks), dim3(threads), 0, c->stream,
                       (double *)c->phi[c->cur], (const double *)c->pi, c->n3, dtd);
  else
    hipLaunchKernelGGL(k_drift<float>, dim3(blocks), dim3(threads), 0, c->stream,
                       (float *)c->phi[c->cur], (const float *)c->pi, c->n3, (float)dtd);
  PN_CHECK_LAUNCH(c);
  PN_HIP(c, hipStreamSynchronize(c->stream));
  return PN_OK;
}

// Online choice of the multi-step kernel and its z block.  The candidates are
// the two-step kernel with z blocks 8 ... 64 and, where it fits, the
// three-step kernel with z blocks 16 ... 64 (all give the same bits).  Each
// is timed with events on the caller's own steps: it runs twice in a row and
// the second launch is timed (so that the caches hold what that candidate
// leaves, not what the previous one did), over two interleaved rounds; each
// keeps its fastest time per step.  The timing goes on across evolve calls,
// as many candidates per call as its steps allow (calls of at least
// kTuneMin steps), and once every candidate has run both rounds the fastest
// is kept; the two-step default wins near-ties (3 %).  The number of tiles
// against the device's CUs, the recomputed halo and the occupancy are traded
// off by measurement.  Choices the caller fixed are kept.
static constexpr int kTuneMin = 12;
static constexpr int kTuneRounds = 2;

static int eff_zb2(const pn_ctx *c, const Step2Shape &s) {
  const int z = c->step2_zb_user ? c->step2_zb_user : (c->step2_zb_auto ? c->step2_zb_auto : s.ZB);
  return std::min(z, c->p.N);
}
static int eff_zb3(const pn_ctx *c, const Step2Shape &s) {
  const int z = c->step3_zb_user ? c->step3_zb_user : (c->step3_zb_auto ? c->step3_zb_auto : s.ZB);
  return std::min(z, c->p.N);
}
static bool use_step3(const pn_ctx *c) {
  Step2Shape s3;
  return c->step2 && c->pi_alt && (c->step3_mode == 1 || (c->step3_mode < 0 && c->step3_auto)) &&
         step3_shape(c, &s3);
}

static void tune_reset(pn_ctx *c) {
  c->tuned = 0;
  c->tune_pos = 0;
  c->tune_fam.clear();
  c->tune_zb.clear();
  c->tune_ms.clear();
}

static int tune_multistep(pn_ctx *c, const double *dtk, const double *dtd, int *n, int nsteps) {
  static const int z2[] = {8, 12, 16, 24, 32, 48, 64};
  static const int z3[] = {16, 24, 32, 48, 64};
  Step2Shape s2, s3;
  const bool have2 = step2_shape(c, &s2) && c->step3_mode != 1;
  const bool have3 = c->step3_mode != 0 && step3_shape(c, &s3);
  if (c->tune_fam.empty()) {
    auto add = [c](int fam, int zb) {
      c->tune_fam.push_back(fam);
      c->tune_zb.push_back(zb);
    };
    if (have2) {
      if (c->step2_zb_user) add(2, c->step2_zb_user);
      else
        for (int z : z2)
          if (z <= c->p.N) add(2, z);
    }
    if (have3) {
      if (c->step3_zb_user) add(3, c->step3_zb_user);
      else
        for (int z : z3)
          if (z <= c->p.N) add(3, z);
    }
    if (c->tune_fam.size() < 2) {
      tune_reset(c);
      c->tuned = 1;
      return PN_OK;
    }
    c->tune_ms.assign(c->tune_fam.size(), 1e30f);
    c->tune_pos = 0;
  }
  const int nc = (int)c->tune_fam.size();
  hipEvent_t e0, e1;
  PN_HIP(c, hipEventCreate(&e0));
  hipError_t ee = hipEventCreate(&e1);
  if (ee != hipSuccess) {
    (void)hipEventDestroy(e0);
    return pn_fail(c, PN_ERR_HIP, "pn_evolve: %s", hipGetErrorString(ee));
  }
  // one launch of candidate i on the next steps; *ms = its time per step
  auto run = [&](int i, float *ms) -> int {
    const int fam = c->tune_fam[i];
    Step2Shape t = fam == 2 ? s2 : s3;
    t.ZB = std::min(c->tune_zb[i], c->p.N);
    hipError_t e = hipEventRecord(e0, c->stream);
    int rc = e != hipSuccess ? PN_ERR_HIP
             : fam == 2       ? launch_step2(c, t, dtk + *n, dtd + *n)
                              : launch_step3(c, t, dtk + *n, dtd + *n);
    *n += fam;
    if (rc == PN_OK && e == hipSuccess) e = hipEventRecord(e1, c->stream);
    if (rc == PN_OK && e == hipSuccess) e = hipEventSynchronize(e1);
    if (rc == PN_OK && e == hipSuccess) e = hipEventElapsedTime(ms, e0, e1);
    if (rc == PN_OK && e != hipSuccess) rc = pn_fail(c, PN_ERR_HIP, "pn_evolve: %s", hipGetErrorString(e));
    *ms /= (float)fam;
    return rc;
  };
  int rc = PN_OK;
  float ms = 0.f;
  while (rc == PN_OK && c->tune_pos < kTuneRounds * nc) {
    const int i = c->tune_pos % nc;
    if (2 * c->tune_fam[i] > nsteps - *n) break;  // the next call goes on
    rc = run(i, &ms);  // primes the caches for this candidate
    if (rc == PN_OK) rc = run(i, &ms);
    if (rc == PN_OK) {
      c->tune_ms[i] = std::min(c->tune_ms[i], ms);
      c->tune_pos++;
    }
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  if (rc != PN_OK || c->tune_pos < kTuneRounds * nc) return rc;
  // the default (the shape's own or the caller's z block) wins near-ties
  const int dfam = have2 ? 2 : 3, dzb = have2 ? eff_zb2(c, s2) : eff_zb3(c, s3);
  int best = -1;
  float best_ms = 0.f;
  for (int i = 0; i < nc; i++) {
    const bool def = c->tune_fam[i] == dfam && std::min(c->tune_zb[i], c->p.N) == dzb;
    const float ti = def ? c->tune_ms[i] * 0.97f : c->tune_ms[i];
    if (best < 0 || ti < best_ms) {
      best_ms = ti;
      best = i;
    }
  }
  c->step3_auto = c->tune_fam[best] == 3;
  if (c->tune_fam[best] == 3) c->step3_zb_auto = c->tune_zb[best];
  else c->step2_zb_auto = c->tune_zb[best];
  tune_reset(c);
  c->tuned = 1;
  return PN_OK;
}

extern "C" int pn_evolve(pn_ctx *c, int nsteps, const double *dtk, const double *dtd) {
  if (!c) return pn_fail(nullptr, PN_ERR_ARG, "pn_evolve: NULL context");
  if (nsteps < 0 || (nsteps > 0 && (!dtk || !dtd)))
    return pn_fail(c, PN_ERR_ARG, "pn_evolve: bad step arguments");
  if (c->is_view) return pn_fail(c, PN_ERR_STATE, "pn_evolve: a batch view is stepped by pn_batch_evolve");
  PN_HIP(c, hipSetDevice(c->p.device));
  int n = 0;
  if (c->step2 && c->pi_alt) {
    if (!c->tuned && nsteps >= kTuneMin) {
      int rc = tune_multistep(c, dtk, dtd, &n, nsteps);
      if (rc != PN_OK) return rc;
    }
    Step2Shape s2, s3;
    if (use_step3(c) && step3_shape(c, &s3)) {
      s3.ZB = eff_zb3(c, s3);
      for (; n + 3 <= nsteps; n += 3) {
        int rc = launch_step3(c, s3, dtk + n, dtd + n);
        if (rc != PN_OK) return rc;
      }
    }
    if (step2_shape(c, &s2)) {
      s2.ZB = eff_zb2(c, s2);
      for (; n + 2 <= nsteps; n += 2) {
        int rc = launch_step2(c, s2, dtk + n, dtd + n);
        if (rc != PN_OK) return rc;
      }
    }
  }
  for (; n < nsteps; n++) {
    int rc = launch_step(c, 0, dtk[n], dtd[n]);
    if (rc != PN_OK) return rc;
  }
  return PN_OK;
}

extern "C" int pn_set_option(pn_ctx *c, int option, int value) {
  if (!c) return pn_fail(nullptr, PN_ERR_ARG, "pn_set_option: NULL context");
  if (option == PN_OPT_TWO_STEP) {
    if (value && !c->pi_alt) return pn_fail(c, PN_ERR_STATE, "pn_set_option: no two-step kernel for this lattice");
    c->step2 = value ? 1 : 0;
    return PN_OK;
  }
  if (option == PN_OPT_IC_ON_DEVICE) {
    c->ic_device = value ? 1 : 0;
    return PN_OK;
  }
  if (option == PN_OPT_THREE_STEP) {
    Step2Shape s3;
    if (value > 0 && (!c->pi_alt || !step3_shape(c, &s3)))
      return pn_fail(c, PN_ERR_STATE, "pn_set_option: no three-step kernel for this lattice");
    c->step3_mode = value > 0 ? 1 : (value == 0 ? 0 : -1);
    tune_reset(c);  // the next evolve calls time the remaining choices
    return PN_OK;
  }
  if (option == PN_OPT_KERNEL_TIMED)
    return pn_fail(c, PN_ERR_ARG, "pn_set_option: PN_OPT_KERNEL_TIMED is read only");
  if (option == PN_OPT_STEP3_ZBLOCK || option == PN_OPT_STEP2_ZBLOCK) {
    if (value < 0) return pn_fail(c, PN_ERR_ARG, "pn_set_option: z block %d", value);
    (option == PN_OPT_STEP3_ZBLOCK ? c->step3_zb_user : c->step2_zb_user) = value;
    tune_reset(c);
    return PN_OK;
  }
  return pn_fail(c, PN_ERR_ARG, "pn_set_option: unknown option %d", option);
}

extern "C" int pn_get_option(pn_ctx *c, int option, int *value) {
  if (!c || !value) return pn_fail(c, PN_ERR_ARG, "pn_get_option: NULL argument");
  if (option == PN_OPT_KERNEL_TIMED) {
    *value = c->tuned;
    return PN_OK;
  }
  if (option == PN_OPT_TWO_STEP) {
    *value = c->step2;
    return PN_OK;
  }
  if (option == PN_OPT_IC_ON_DEVICE) {
    *value = c->ic_device;
    return PN_OK;
  }
  if (option == PN_OPT_THREE_STEP) {
    *value = use_step3(c) ? 1 : 0;
    return PN_OK;
  }
  if (option == PN_OPT_STEP3_ZBLOCK) {
    Step2Shape s3;
    *value = step3_shape(c, &s3) ? eff_zb3(c, s3) : 0;
    return PN_OK;
  }
  if (option == PN_OPT_STEP2_ZBLOCK) {
    Step2Shape s2;
    *value = step2_shape(c, &s2) ? eff_zb2(c, s2) : 0;
    return PN_OK;
  }
  return pn_fail(c, PN_ERR_ARG, "pn_get_option: unknown option %d", option);
}

extern "C" int pn_evolve_timed(pn_ctx *c, int nsteps, const double *dtk, const double *dtd,
                               double *ms) {
  if (!c || !ms) return pn_fail(c, PN_ERR_ARG, "pn_evolve_timed: NULL argument");
  PN_HIP(c, hipSetDevice(c->p.device));
  hipEvent_t e0, e1;
  PN_HIP(c, hipEventCreate(&e0));
  PN_HIP(c, hipEventCreate(&e1));
  int rc = PN_OK;
  hipError_t he = hipEventRecord(e0, c->stream);
  if (he == hipSuccess) rc = pn_evolve(c, nsteps, dtk, dtd);
  if (he == hipSuccess && rc == PN_OK) he = hipEventRecord(e1, c->stream);
  if (he == hipSuccess && rc == PN_OK) he = hipEventSynchronize(e1);
  float f = 0.f;
  if (he == hipSuccess && rc == PN_OK) he = hipEventElapsedTime(&f, e0, e1);
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  if (rc != PN_OK) return rc;
  if (he != hipSuccess) return pn_fail(c, PN_ERR_HIP, "pn_evolve_timed: %s", hipGetErrorString(he));
  *ms = f;
  return PN_OK;
}

extern "C" int pn_sync(pn_ctx *c) {
  if (!c) return pn_fail(nullptr, PN_ERR_ARG, "pn_sync: NULL context");
  PN_HIP(c, hipStreamSynchronize(c->stream));
  return PN_OK;
}

// ---------------------------------------------------------------------------
// K3/K4: volume sums and extrema (fieldanalysis.cpp:319-699)
// ---------------------------------------------------------------------------
// Sums are FP64 and compensated: each thread keeps (hi, lo) pairs updated
// with an exact two-sum, partial pairs are merged in a fixed tree, so the
// result is deterministic and within a few ulp of the exact sum.
// extrema candidates: strict comparisons in raster order, ties -> lowest index
struct Ext {
  double mn, mx, ab;
  long long imn, imx, iab;
};

__device__ __forceinline__ void ext_merge(Ext &a, const Ext &b) {
  if (b.mn < a.mn || (b.mn == a.mn && b.imn >= 0 && (a.imn < 0 || b.imn < a.imn))) {
    a.mn = b.mn; a.imn = b.imn;
  }
  if (b.mx > a.mx || (b.mx == a.mx && b.imx >= 0 && (a.imx < 0 || b.imx < a.imx))) {
    a.mx = b.mx; a.imx = b.imx;
  }
  if (b.ab > a.ab || (b.ab == a.ab && b.iab >= 0 && (a.iab < 0 || b.iab < a.iab))) {
    a.ab = b.ab; a.iab = b.iab;
  }
}

__device__ __forceinline__ double shfl_d(double v, int src) { return __shfl(v, src, 64); }
__device__ __forceinline__ long long shfl_ll(long long v, int src) { return __shfl(v, src, 64); }

// The divisions of the potential sum (fieldanalysis.cpp:571) are by run
// constants, 24 and 720*Nc*Nc*Kappa*Kappa: two-term reciprocals
// (pn_fastdiv.h) give the IEEE quotient in two or four FMA-pipe operations.
struct VolCoef {
  RedCoef r;
  pn_recip r24, rd6;
  int fdiv;    // both divisors lie where pn_div_fast is exact
  int d6_bmr;  // the division by d6 needs no correction step
};

// Tiles of RB rows x XC vectors of W consecutive x sites, ZB planes deep; a
// block of 256 threads takes one tile at a time and streams its planes in z,
// so the z neighbour of plane z is the own field of plane z+1.
struct VolShape {
  int W, NV, XC, RB, nxc, nyt, nzb, ZB, tiles;
};
// The main pass runs as many blocks as are resident at once, each taking tiles
// in turn; the z blocks are cut to balance the blocks' work (kVolTileStart:
// a tile's start-up cost in planes).
constexpr int kVolTileStart = 3;
// planes of a tile held in registers at once (the loads' look-ahead)
constexpr int kVolRing = 4;

static VolShape vol_shape(int N, int W, int grid) {
  VolShape s;
  s.W = W;
  s.NV = N / W;
  s.XC = s.NV < kRedThreads ? s.NV : kRedThreads;
  s.nxc = (s.NV + kRedThreads - 1) / kRedThreads;
  s.RB = std::min(kRedThreads / s.XC, N);
  s.nyt = (N + s.RB - 1) / s.RB;
  const long long cols = (long long)s.nxc * s.nyt;  // at most kVolMaxTiles for N <= 2048
  // planes per tile: the busiest block's planes, plus a few planes' worth of
  // start-up per tile, as small as it gets (z blocks of at least 4 planes)
  long long best = -1, best_nzb = 1;
  const long long nz_max = std::max<long long>(1, std::min<long long>(N / 4, kVolMaxTiles / cols));
  for (long long nzb = 1; nzb <= nz_max; nzb++) {
    const long long zb = (N + nzb - 1) / nzb;
    if ((N + zb - 1) / zb != nzb) continue;  // the same z blocks as a smaller count
    const long long per_block = (cols * nzb + grid - 1) / grid;
    const long long cost = per_block * (zb + kVolTileStart);
    if (best < 0 || cost < best) { best = cost; best_nzb = nzb; }
  }
  s.ZB = (int)((N + best_nzb - 1) / best_nzb);
  s.nzb = (N + s.ZB - 1) / s.ZB;
  s.tiles = (int)(cols * s.nzb);
  return s;
}

// Device scratch of one volume analysis: the main pass's block partials
// (kVolMaxBlocks * 2 * kNSums doubles and kVolMaxBlocks * 6 extrema words) and
// its tile flags (kVolMaxTiles).
struct VolScratch {
  double *part;
  long long *epart;
  int *tflag;
  int *bflag;  // per block: one of its tiles has a flag (kVolMaxBlocks)
};

// One vector's terms (fieldanalysis.cpp:331-355,437-571), sites in raster
// order.  The sums of phi, pi and phi*pi, whose terms can have either sign,
// take every term through an exact two-sum; so does the interaction, with the
// two terms of a site (:457 and :571) added first.  The terms of phi^2, pi^2,
// phi^4, phi^6 and the gradient are never negative: a run of the terms of two
// planes' vectors (at most 2W) is added first (relative error below 2W ulp
// of the run, which is at most the total) and the run takes one two-sum.
// FAST: the wave's sites all have |phi| in [2^-100, 2^100), so
//  * phi^4 and phi^6 lie where the two-term divisions are exact;
//  * phi*phi is a normal number, for which sqrt(RN(phi*phi)) == |phi| in
//    binary arithmetic (S. Boldo, "Stupid is as stupid does: taking the
//    square root of the square of a floating-point number", 2015): MaxAbs
//    takes |phi| without the square root;
//  * every nonzero difference of neighbours is at least 2^-152, so with a2 a
//    power of two (1/a2 within [2^-200, 2^200]) no product or sum of the
//    gradient under- or overflows, and ((dx^2/a2 + dy^2/a2) + dz^2/a2) equals
//    ((dx^2 + dy^2) + dz^2) * (1/a2) bit for bit (scaling by a power of two
//    commutes with rounding).
template <bool POW2, bool FAST>
__device__ __forceinline__ void vol_site(const VolCoef &k, double f, double p, double fx, double fy,
                                         double fz, long long rp, bool first, double (&r)[5],
                                         double (&hi)[kNSums], double (&lo)[kNSums], Ext &e) {
  const double q = f * f;  // the reference's 0.0 + phi*phi: never -0
  two_acc(hi[0], lo[0], f);
  two_acc(hi[1], lo[1], p);
  two_acc(hi[4], lo[4], f * p);
  const double q2 = q * q, q3 = q2 * q;
  double g, v, ab;
  if (FAST) {
    const double dx = fx - f, dy = fy - f, dz = fz - f;
    g = POW2 ? ((dx * dx + dy * dy) + dz * dz) * k.r.inv_a2
             : (dx * dx / k.r.a2 + dy * dy / k.r.a2) + dz * dz / k.r.a2;  // :445
    double v6 = pn_div_bmr(q3, k.rd6);
    if (!k.d6_bmr) v6 = pn_div_correct(q3, k.rd6, v6);
    v = k.r.lmode * pn_div_bmr(q2, k.r24) + v6;  // :571
    ab = fabs(f);
  } else {
    const double dx = fx - f, dy = fy - f, dz = fz - f;
    const double gx = POW2 ? dx * dx * k.r.inv_a2 : dx * dx / k.r.a2;
    const double gy = POW2 ? dy * dy * k.r.inv_a2 : dy * dy / k.r.a2;
    const double gz = POW2 ? dz * dz * k.r.inv_a2 : dz * dz / k.r.a2;
    g = (gx + gy) + gz;
    v = k.r.lmode * (q2 / 24.0) + q3 / k.r.d6;
    ab = sqrt(0.0 + f * f);
  }
  two_acc(hi[8], lo[8], k.r.cm * q + v);  // :457, :571
  const double t5[5] = {q, p * p, q2, q3, g};
#pragma unroll
  for (int u = 0; u < 5; u++) r[u] = first ? t5[u] : r[u] + t5[u];
  // a thread visits its sites in raster order, so strict comparisons keep
  // the first index
  if (ab > e.ab) { e.ab = ab; e.iab = rp; }
  if (f > e.mx) { e.mx = f; e.imx = rp; }
  if (f < e.mn) { e.mn = f; e.imn = rp; }
}

__device__ __forceinline__ void vol_runs(const double (&r)[5], double (&hi)[kNSums],
                                         double (&lo)[kNSums]) {
  two_acc(hi[2], lo[2], r[0]);
  two_acc(hi[3], lo[3], r[1]);
  two_acc(hi[5], lo[5], r[2]);
  two_acc(hi[6], lo[6], r[3]);
  two_acc(hi[7], lo[7], r[4]);
}

// a vector of W sites; start: the vector begins the runs of the never
// negative terms (vol_runs adds them to the sums)
template <int W, bool POW2, bool FAST, typename V>
__device__ __forceinline__ void vol_vec(const VolCoef &k, const V &fv, const V &pv, const V &fy,
                                        const V &fz, double fxw, long long rp, bool start,
                                        double (&r)[5], double (&hi)[kNSums],
                                        double (&lo)[kNSums], Ext &e) {
#pragma unroll
  for (int j = 0; j < W; j++)
    vol_site<POW2, FAST>(k, (double)fv[j], (double)pv[j],
                         (j + 1 < W) ? (double)fv[(j + 1) % W] : fxw, (double)fy[j],
                         (double)fz[j], rp + j, start && j == 0, r, hi, lo, e);
}

// One plane of a tile's vector: through the fast form when the whole wave
// allows it (PASS 0) or through the IEEE form when it does not (PASS 1).
// Returns 1 when the vector was summed, 2 when PASS 0 skipped it.
template <int W, bool POW2, int PASS, bool FASTOK, typename V>
__device__ __forceinline__ int vol_plane(const VolCoef &k, const V &f, const V &p, const V &fy,
                                         const V &fz, double fxw, long long rp, bool start,
                                         double (&r)[5], double (&hi)[kNSums],
                                         double (&lo)[kNSums], Ext &e) {
  bool ok = true;
#pragma unroll
  for (int j = 0; j < W; j++) ok = ok && fast_div_site((double)f[j]);
  const bool fast = FASTOK && k.fdiv && __all(ok);
  if (PASS == 0 && !FASTOK) {
    vol_vec<W, POW2, false>(k, f, p, fy, fz, fxw, rp, start, r, hi, lo, e);
    return 1;
  }
  if (PASS == 0) {
    if (!fast) return 2;
    vol_vec<W, POW2, true>(k, f, p, fy, fz, fxw, rp, start, r, hi, lo, e);
    return 1;
  }
  if (fast) return 0;
  vol_vec<W, POW2, false>(k, f, p, fy, fz, fxw, rp, start, r, hi, lo, e);
  return 1;
}

// One tile of the volume sums.  PASS 0 (k_volume): a vector whose wave has
// every site in the fast range is summed, any other is skipped and the call
// returns true (FASTOK false, for divisors outside the fast form's range: every
// vector is summed with the IEEE divisions and none is skipped).  PASS 1 (k_volume_final, over the tiles PASS 0 flagged): only
// the skipped vectors are summed, with the IEEE divisions and square root.
// Both passes map threads to sites and group them into waves the same way, so
// every vector is summed by exactly one of them.  Keeping the IEEE sequences
// out of the main kernel keeps its register count, and its occupancy, that
// of the fast form.
template <typename T, int W, bool POW2, int PASS, bool FASTOK = true>
__device__ __forceinline__ bool vol_tile(const T *__restrict__ phi, const T *__restrict__ pi,
                                         int N, const VolShape &s, const VolCoef &k, int t, int lt,
                                         double (&hi)[kNSums], double (&lo)[kNSums], Ext &e) {
  typedef typename VecT<T, W>::type V;
  const long long P = (long long)N * N;
  const int tr = lt / s.XC, tx = lt - tr * s.XC;  // lt: the thread's index in its tile group
  // tile = (z block, row block, x chunk), x chunk fastest
  const int xc = t % s.nxc;
  t /= s.nxc;
  const int yt = t % s.nyt, zb = t / s.nyt;
  const int xv = xc * kRedThreads + tx, y = yt * s.RB + tr;
  if (tr >= s.RB || xv >= s.NV || y >= N) return false;
  const int z0 = zb * s.ZB, z1 = min(z0 + s.ZB, N);
  const int x = xv * W;
  const long long row = (long long)y * N;
  const long long dyp = (long long)((y + 1 == N ? 0 : y + 1) - y) * N;
  const long long xw = row + (x + W == N ? 0 : x + W);  // x neighbour of the last site
  bool skipped = false;
  // A ring of kVolRing planes in registers: slot u holds plane z of the tile
  // with z - z0 = u (mod kVolRing): its field, momentum, y+1 row and x
  // neighbour.  A slot is refilled with the plane kVolRing further on right
  // after its sums, so each plane's loads are in flight while the planes
  // before it are summed.  Plane z1 (periodic: plane 0 after plane N-1) holds
  // the z neighbour of the tile's last plane; refills past it load plane z1
  // again (cached), so every iteration issues the same loads and the waits
  // for them stay exact.
  V F[kVolRing], Pm[kVolRing], Y[kVolRing];
  double X[kVolRing];
  auto load = [&](int u, int zz) {
    zz = zz < z1 ? zz : z1;
    const long long iz = (long long)(zz == N ? 0 : zz) * P + row;
    F[u] = *(const V *)&phi[iz + x];
    Pm[u] = __builtin_nontemporal_load((const V *)&pi[iz + x]);
    Y[u] = *(const V *)&phi[iz + dyp + x];
    X[u] = (double)phi[iz - row + xw];
  };
#pragma unroll
  for (int u = 0; u < kVolRing; u++) load(u, z0 + u);
  for (int zg = z0; zg < z1; zg += kVolRing) {
    double r[5];
    bool started = false;
#pragma unroll
    for (int u = 0; u < kVolRing; u++) {
      const int z = zg + u;
      if (z < z1) {
        const int res = vol_plane<W, POW2, PASS, FASTOK>(k, F[u], Pm[u], Y[u], F[(u + 1) % kVolRing],
                                                         X[u], (long long)z * P + row + x + P,
                                                         !started, r, hi, lo, e);
        started = started || (res & 1);
        skipped = skipped || (res & 2);
        // runs of two planes' terms at most
        if ((u & 1) || z + 1 == z1) {
          if (started) vol_runs(r, hi, lo);
          started = false;
        }
      }
      load(u, z + kVolRing);
    }
  }
  return skipped;
}

__device__ __forceinline__ void vol_init(double (&hi)[kNSums], double (&lo)[kNSums], Ext &e) {
#pragma unroll
  for (int q = 0; q < kNSums; q++) { hi[q] = 0.0; lo[q] = 0.0; }
  e.mn = 10000000.0; e.mx = -10000000.0; e.ab = 0.0;  // fieldanalysis.cpp:324
  e.imn = e.imx = e.iab = -1;
}

// merges the block's threads in a fixed order (wave trees, then the waves in
// order) into sh[0] and se[0]
template <int NT = kRedThreads>
__device__ __forceinline__ void vol_block_merge(double (&hi)[kNSums], double (&lo)[kNSums], Ext &e,
                                                double (*sh)[2 * kNSums], Ext *se) {
  // one sum at a time through the shuffle tree (fewer registers live than
  // all sums level by level)
#pragma unroll
  for (int q = 0; q < kNSums; q++)
    for (int off = 32; off > 0; off >>= 1) {
      const double h2 = __shfl_down(hi[q], off, 64), l2 = __shfl_down(lo[q], off, 64);
      two_merge(hi[q], lo[q], h2, l2);
    }
  for (int off = 32; off > 0; off >>= 1) {
    Ext o;
    o.mn = __shfl_down(e.mn, off, 64); o.imn = __shfl_down(e.imn, off, 64);
    o.mx = __shfl_down(e.mx, off, 64); o.imx = __shfl_down(e.imx, off, 64);
    o.ab = __shfl_down(e.ab, off, 64); o.iab = __shfl_down(e.iab, off, 64);
    ext_merge(e, o);
  }
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  if (lane == 0) {
#pragma unroll
    for (int q = 0; q < kNSums; q++) { sh[w][2 * q] = hi[q]; sh[w][2 * q + 1] = lo[q]; }
    se[w] = e;
  }
  __syncthreads();
  if (threadIdx.x == 0)
    for (int ww = 1; ww < NT / 64; ww++) {
      for (int q = 0; q < kNSums; q++) two_merge(sh[0][2 * q], sh[0][2 * q + 1], sh[ww][2 * q], sh[ww][2 * q + 1]);
      ext_merge(se[0], se[ww]);
    }
}

// Main pass: a grid of as many blocks as are resident at once, each taking
// tiles slot, slot + gridDim, ... (in each round of gridDim tiles, XCD k takes
// a contiguous run: tiles that share halo rows); per block one partial, per
// tile a flag saying whether it skipped vectors for the final pass.
template <typename T, int W, bool POW2, bool FASTOK>
__global__ __launch_bounds__(kRedThreads) void k_volume(const T *__restrict__ phi,
                                                        const T *__restrict__ pi, int N,
                                                        VolShape s, VolCoef k,
                                                        double *__restrict__ part,
                                                        long long *__restrict__ epart,
                                                        int *__restrict__ tflag,
                                                        int *__restrict__ bflag) {
  double hi[kNSums], lo[kNSums];
  Ext e;
  vol_init(hi, lo, e);
  const int slot = xcd_tile((int)blockIdx.x, (int)gridDim.x);
  int block_any = 0;
  for (int base = 0; base + slot < s.tiles; base += (int)gridDim.x) {
    const bool skipped =
        vol_tile<T, W, POW2, 0, FASTOK>(phi, pi, N, s, k, base + slot, (int)threadIdx.x, hi, lo, e);
    const int any = __syncthreads_or(skipped);
    if (threadIdx.x == 0) tflag[base + slot] = any;
    block_any |= any;
  }
  if (threadIdx.x == 0) bflag[blockIdx.x] = block_any;
  __shared__ double sh[kRedThreads / 64][2 * kNSums];
  __shared__ Ext se[kRedThreads / 64];
  vol_block_merge(hi, lo, e, sh, se);
  if (threadIdx.x == 0) {
    for (int q = 0; q < 2 * kNSums; q++) part[(long long)blockIdx.x * 2 * kNSums + q] = sh[0][q];
    long long *ep = epart + (long long)blockIdx.x * 6;
    ep[0] = __double_as_longlong(se[0].mn); ep[1] = se[0].imn;
    ep[2] = __double_as_longlong(se[0].mx); ep[3] = se[0].imx;
    ep[4] = __double_as_longlong(se[0].ab); ep[5] = se[0].iab;
  }
}

// Values written by k_volume_final: the ten averages in pn_averages order,
// then min, minpos, max, maxpos, maxabs, maxabspos (positions as doubles).
constexpr int kVolOut = 16;

// One block: the vectors the main pass skipped (none in the usual case;
// flagged tiles in order), then the main pass's block partials (thread t takes
// partials t, t + kVolFinalThreads, ...), merged in a fixed order (wave trees,
// then the waves in order); then the averages as fieldanalysis.cpp:619-638
// forms them (one rank: /nodes == /1).  One tile group of threads, so the
// IEEE path keeps its registers without spilling.
constexpr int kVolFinalThreads = kRedThreads;
template <typename T, int W, bool POW2>
__global__ __launch_bounds__(kVolFinalThreads) void k_volume_final(
    const T *__restrict__ phi, const T *__restrict__ pi, int N, VolShape s, VolCoef k,
    const double *__restrict__ part, const long long *__restrict__ epart, int nb,
    const int *__restrict__ tflag, const int *__restrict__ bflag, double vol,
    double *__restrict__ out) {
  double hi[kNSums], lo[kNSums];
  Ext e;
  vol_init(hi, lo, e);
  int mine = 0;
  for (int b = threadIdx.x; b < nb; b += kVolFinalThreads) mine |= bflag[b];
  if (__syncthreads_or(mine))
    for (int t = 0; t < s.tiles; t++)
      if (tflag[t]) vol_tile<T, W, POW2, 1>(phi, pi, N, s, k, t, (int)threadIdx.x, hi, lo, e);
  for (int b = threadIdx.x; b < nb; b += kVolFinalThreads) {
    const double *pb = part + (long long)b * 2 * kNSums;
    double v[2 * kNSums];
#pragma unroll
    for (int q = 0; q < 2 * kNSums; q++) v[q] = pb[q];
    const long long *ep = epart + (long long)b * 6;
    Ext o;
    o.mn = __longlong_as_double(ep[0]); o.imn = ep[1];
    o.mx = __longlong_as_double(ep[2]); o.imx = ep[3];
    o.ab = __longlong_as_double(ep[4]); o.iab = ep[5];
#pragma unroll
    for (int q = 0; q < kNSums; q++) two_merge(hi[q], lo[q], v[2 * q], v[2 * q + 1]);
    ext_merge(e, o);
  }
  __shared__ double sh[kVolFinalThreads / 64][2 * kNSums];
  __shared__ Ext se[kVolFinalThreads / 64];
  vol_block_merge<kVolFinalThreads>(hi, lo, e, sh, se);
  if (threadIdx.x == 0) {
    e = se[0];
    double S[kNSums];
    for (int q = 0; q < kNSums; q++) S[q] = sh[0][2 * q] + sh[0][2 * q + 1];
    double phi1 = S[0], pi1 = S[1], phi2 = S[2], pi2 = S[3], phipi = S[4], phi4 = S[5], phi6 = S[6];
    double G = S[7], I = S[8];
    phi1 /= (vol); pi1 /= (vol);
    double Charge = (0.5 * pi2 - G / 6. - I);
    double Energy = (0.5 * pi2 + 0.5 * G + I);
    Energy /= (vol); I /= (vol); Charge /= (vol);
    phi2 /= (vol); pi2 /= (vol); phi4 /= (vol); phi6 /= (vol); phipi /= (vol);
    out[0] = phi1; out[1] = pi1; out[2] = phi2; out[3] = pi2; out[4] = phipi;
    out[5] = phi4; out[6] = phi6; out[7] = Energy; out[8] = I; out[9] = Charge;
    out[10] = e.mn; out[11] = (double)e.imn;
    out[12] = e.mx; out[13] = (double)e.imx;
    out[14] = e.ab; out[15] = (double)e.iab;
  }
}

static RedCoef red_coef(const pn_params &p, int a2_pow2) {
  RedCoef k;
  const int Nc = 1;
  k.a2 = p.a * p.a;
  k.cm = (p.mass_mode) * 0.5 * p.m2;
  k.lmode = (double)p.lambda_mode;
  k.d6 = 720.0 * Nc * Nc * p.kappa * p.kappa;
  k.inv_a2 = 1.0 / k.a2;
  k.pow2 = a2_pow2;
  return k;
}

static RedCoef red_coef(const pn_ctx *c) { return red_coef(c->p, c->k.a2_pow2); }

static VolCoef vol_coef(const pn_params &p, int a2_pow2) {
  VolCoef k;
  k.r = red_coef(p, a2_pow2);
  // the fused gradient scaling needs 1/a2 in [2^-200, 2^200]; otherwise divide
  if (!(k.r.inv_a2 >= 0x1p-200 && k.r.inv_a2 <= 0x1p+200)) k.r.pow2 = 0;
  int f24 = 0, fd6 = 0;
  k.r24 = pn_make_recip(24.0, &f24);
  k.rd6 = pn_make_recip(k.r.d6, &fd6);
  k.fdiv = f24 && fd6 && pn_bmr_exact(k.r24);
  k.d6_bmr = k.fdiv && pn_bmr_exact(k.rd6);
  return k;
}

// blocks of k_volume<T, W, POW2> resident at once on device dev (capped at
// the partials' room), found once per device
template <typename T, int W, bool POW2, bool FASTOK>
static int vol_grid(int dev) {
  static std::mutex mu;
  static int cached[kScratchMaxDevices];
  std::lock_guard<std::mutex> lk(mu);
  if (dev < 0 || dev >= kScratchMaxDevices) return 256;
  if (!cached[dev]) {
    int per_cu = 0, cus = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, (const void *)k_volume<T, W, POW2, FASTOK>,
                                                     kRedThreads, 0) != hipSuccess || per_cu < 1)
      per_cu = 1;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cus < 1)
      cus = 256;
    cached[dev] = std::max(1, std::min(kVolMaxBlocks, per_cu * cus));
  }
  return cached[dev];
}

template <typename T, int W, bool POW2, bool FASTOK>
static hipError_t launch_volume_t(const pn_params &p, const VolCoef &k, const void *phi,
                                  const void *pi, hipStream_t st, const VolScratch &sc, double *out) {
  const int grid = vol_grid<T, W, POW2, FASTOK>(p.device);
  const VolShape s = vol_shape(p.N, W, grid);
  const int nb = std::min(s.tiles, grid);
  hipLaunchKernelGGL((k_volume<T, W, POW2, FASTOK>), dim3(nb), dim3(kRedThreads), 0, st,
                     (const T *)phi, (const T *)pi, p.N, s, k, sc.part, sc.epart, sc.tflag, sc.bflag);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  const double vol = (double)((long long)p.N * p.N * p.N);  // N^3 overflows int above 1290
  hipLaunchKernelGGL((k_volume_final<T, W, POW2>), dim3(1), dim3(kVolFinalThreads), 0, st,
                     (const T *)phi, (const T *)pi, p.N, s, k, sc.part, sc.epart, nb, sc.tflag,
                     sc.bflag, vol, out);
  return hipGetLastError();
}

template <typename T, int W>
static hipError_t launch_volume_w(const pn_params &p, const VolCoef &k, const void *phi,
                                  const void *pi, hipStream_t st, const VolScratch &sc, double *out) {
  if (!k.fdiv)
    return k.r.pow2 ? launch_volume_t<T, W, true, false>(p, k, phi, pi, st, sc, out)
                    : launch_volume_t<T, W, false, false>(p, k, phi, pi, st, sc, out);
  return k.r.pow2 ? launch_volume_t<T, W, true, true>(p, k, phi, pi, st, sc, out)
                  : launch_volume_t<T, W, false, true>(p, k, phi, pi, st, sc, out);
}

// Enqueues K3/K4 on stream st over dense fields phi, pi of p.N^3 values
// (p.precision's type); the kVolOut results go to out.  Vector loads of W
// sites need N divisible by W and W-aligned pointers.
static hipError_t launch_volume(const pn_params &p, int a2_pow2, const void *phi, const void *pi,
                                hipStream_t st, const VolScratch &sc, double *out) {
  const VolCoef k = vol_coef(p, a2_pow2);
  const size_t es = p.precision == PN_FP64 ? 8 : 4;
  const uintptr_t al = (uintptr_t)phi | (uintptr_t)pi;
  int W = p.precision == PN_FP64 ? 2 : 4;
  while (W > 1 && (p.N % W != 0 || al % (W * es) != 0)) W >>= 1;
  if (p.precision == PN_FP64)
    return W == 2 ? launch_volume_w<double, 2>(p, k, phi, pi, st, sc, out)
                  : launch_volume_w<double, 1>(p, k, phi, pi, st, sc, out);
  if (W == 4) return launch_volume_w<float, 4>(p, k, phi, pi, st, sc, out);
  if (W == 2) return launch_volume_w<float, 2>(p, k, phi, pi, st, sc, out);
  return launch_volume_w<float, 1>(p, k, phi, pi, st, sc, out);
}

static VolScratch vol_scratch(const pn_ctx *c) {
  VolScratch sc;
  sc.part = c->red_part;
  sc.epart = c->ext_part;
  sc.tflag = c->tile_flag;
  sc.bflag = c->tile_flag + kVolMaxTiles;
  return sc;
}

// runs K3/K4 and leaves the kVolOut results in c->host_pin
static int run_volume(pn_ctx *c) {
  PN_HIP(c, hipSetDevice(c->p.device));
  PN_HIP(c, launch_volume(c->p, c->k.a2_pow2, c->phi[c->cur], c->pi, c->stream, vol_scratch(c),
                          c->red_out));
  PN_HIP(c, hipMemcpyAsync(c->host_pin, c->red_out, kVolOut * sizeof(double),
                           hipMemcpyDeviceToHost, c->stream));
  PN_HIP(c, hipStreamSynchronize(c->stream));
  return PN_OK;
}

static void vol_unpack(const double *s, pn_averages *avg, pn_extrema *ext) {
  if (avg) {
    avg->phi = s[0]; avg->pi = s[1]; avg->phi2 = s[2]; avg->pi2 = s[3]; avg->phipi = s[4];
    avg->phi4 = s[5]; avg->phi6 = s[6]; avg->energy = s[7]; avg->interaction = s[8];
    avg->charge = s[9];
  }
  if (ext) {
    ext->min = s[10]; ext->minpos = (long long)s[11];
    ext->max = s[12]; ext->maxpos = (long long)s[13];
    ext->maxabs = s[14]; ext->maxabspos = (long long)s[15];
  }
}

extern "C" int pn_volume_averages(pn_ctx *c, pn_averages *out) {
  if (!c || !out) return pn_fail(c, PN_ERR_ARG, "pn_volume_averages: NULL argument");
  int rc = run_volume(c);
  if (rc != PN_OK) return rc;
  vol_unpack(c->host_pin, out, nullptr);
  return PN_OK;
}

extern "C" int pn_volume_analysis(pn_ctx *c, pn_averages *avg, pn_extrema *ext) {
  if (!c) return pn_fail(c, PN_ERR_ARG, "pn_volume_analysis: NULL context");
  if (!avg && !ext) return PN_OK;
  int rc = run_volume(c);  // one k_volume run for both
  if (rc != PN_OK) return rc;
  vol_unpack(c->host_pin, avg, ext);
  return PN_OK;
}

extern "C" int pn_extrema_get(pn_ctx *c, pn_extrema *out) {
  if (!c || !out) return pn_fail(c, PN_ERR_ARG, "pn_extrema_get: NULL argument");
  int rc = run_volume(c);
  if (rc != PN_OK) return rc;
  vol_unpack(c->host_pin, nullptr, out);
  return PN_OK;
}

// Averages with Pi brought to Phi's time: Pi lives half a step behind Phi
// between leapfrog steps (the first kick is a half kick, dynamics.cpp:196-213),
// so Energy mixes phi(t) with pi(t - dt/2).  The kick of updateMomenta
// (dynamics.cpp:52-175) with factor half_dt, written into the idle Phi
// buffer, gives pi(t) to the leapfrog's order, and the volume sums run on
// (Phi, that buffer).  No reference counterpart: the precision form's
// time-synchronised energy.  The fields are unchanged.
extern "C" int pn_volume_averages_synced(pn_ctx *c, double half_dt, pn_averages *out) {
  if (!c || !out) return pn_fail(c, PN_ERR_ARG, "pn_volume_averages_synced: NULL argument");
  if (c->is_view)
    return pn_fail(c, PN_ERR_STATE, "pn_volume_averages_synced: not available on a batch view");
  PN_HIP(c, hipSetDevice(c->p.device));
  int rc = launch_step(c, 2, half_dt, 0.0);  // phi[cur ^ 1] = Pi kicked by half_dt
  if (rc != PN_OK) return rc;
  PN_HIP(c, launch_volume(c->p, c->k.a2_pow2, c->phi[c->cur], c->phi[c->cur ^ 1], c->stream,
                          vol_scratch(c), c->red_out));
  PN_HIP(c, hipMemcpyAsync(c->host_pin, c->red_out, kVolOut * sizeof(double),
                           hipMemcpyDeviceToHost, c->stream));
  PN_HIP(c, hipStreamSynchronize(c->stream));
  vol_unpack(c->host_pin, out, nullptr);
  return PN_OK;
}

// ---------------------------------------------------------------------------
// volume analysis that allocates nothing (no reference counterpart)
// ---------------------------------------------------------------------------
// For callers that cannot allocate (code under HIP graph capture, serving
// loops): the block partials of K3/K4 live in a file-scope device array
// instead of context memory.  Each device holds its own copy of the array at
// its own address, looked up once per device with that device current; the
// launches take that address, never the array's host-side name.  Every call on
// a device shares its copy, so calls that overlap on one device race.
// Nothing in it needs clearing: the main pass writes every partial and every
// tile flag the final pass reads, earlier in the same stream.  Layout: the
// block partials (2 * kNSums doubles and 6 extrema words per block), then the
// tile and block flags (two per double): 456 KiB
constexpr size_t kScratchDoubles =
    (size_t)kVolMaxBlocks * (2 * kNSums + 6) + (kVolMaxTiles + kVolMaxBlocks) / 2;
static __device__ __align__(256) double scratch[kScratchDoubles];

static std::mutex g_scratch_mu;
static double *g_scratch_dev[kScratchMaxDevices];  // device address of `scratch` per device

// the device address of this device's copy (the device must be current)
static int scratch_on(int dev, double **out) {
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (!g_scratch_dev[dev]) {
    void *ptr = nullptr;
    const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !ptr)
      return pn_fail(nullptr, PN_ERR_HIP, "pn_volume_analysis_noalloc: hipGetSymbolAddress on device %d: %s",
                     dev, hipGetErrorString(e));
    g_scratch_dev[dev] = (double *)ptr;
  }
  *out = g_scratch_dev[dev];
  return PN_OK;
}

extern "C" int pn_volume_analysis_noalloc(const pn_params *p, const void *phi, const void *pi,
                                          void *stream, double *out) {
  if (!p || !phi || !pi || !out)
    return pn_fail(nullptr, PN_ERR_ARG, "pn_volume_analysis_noalloc: NULL argument");
  if (p->N < 4 || p->N > PN_NOALLOC_MAX_N)
    return pn_fail(nullptr, PN_ERR_ARG,
                   "pn_volume_analysis_noalloc: N=%d outside [4, %d] (the static scratch holds the "
                   "partials of at most %d blocks)", p->N, PN_NOALLOC_MAX_N, kVolMaxBlocks);
  if (p->precision != PN_FP64 && p->precision != PN_FP32)
    return pn_fail(nullptr, PN_ERR_ARG, "pn_volume_analysis_noalloc: unknown precision %d", p->precision);
  if (!(p->a > 0) || !std::isfinite(p->a) || !(p->kappa != 0) || !std::isfinite(p->kappa))
    return pn_fail(nullptr, PN_ERR_ARG, "pn_volume_analysis_noalloc: a must be > 0 and Kappa finite and nonzero");
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || p->device < 0 || p->device >= ndev ||
      p->device >= kScratchMaxDevices)
    return pn_fail(nullptr, PN_ERR_ARG, "pn_volume_analysis_noalloc: device %d of %d", p->device, ndev);
  int prev = 0;
  if (hipGetDevice(&prev) != hipSuccess || hipSetDevice(p->device) != hipSuccess)
    return pn_fail(nullptr, PN_ERR_HIP, "pn_volume_analysis_noalloc: cannot select device %d", p->device);
  double *sc = nullptr;
  int rc = scratch_on(p->device, &sc);
  if (rc == PN_OK) {
    const double a2 = p->a * p->a;
    VolScratch vs;
    vs.part = sc;
    vs.epart = (long long *)(sc + (size_t)kVolMaxBlocks * 2 * kNSums);
    vs.tflag = (int *)(sc + (size_t)kVolMaxBlocks * (2 * kNSums + 6));
    vs.bflag = vs.tflag + kVolMaxTiles;
    const hipError_t e = launch_volume(*p, is_pow2_double(a2), phi, pi, (hipStream_t)stream, vs, out);
    if (e != hipSuccess)
      rc = pn_fail(nullptr, PN_ERR_HIP, "pn_volume_analysis_noalloc: launch failed: %s", hipGetErrorString(e));
  }
  (void)hipSetDevice(prev);
  return rc;
}

// ---------------------------------------------------------------------------
// per-site energy density (the summand of Energy, fieldanalysis.cpp:625)
// ---------------------------------------------------------------------------
template <typename T>
__global__ void k_rho(const T *__restrict__ phi, const T *__restrict__ pi, int N, RedCoef k,
                      double *__restrict__ rho) {
  const long long n3 = (long long)N * N * N, P = (long long)N * N;
  const long long stride = (long long)gridDim.x * blockDim.x;
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n3; i += stride) {
    const int x = (int)(i % N), y = (int)((i / N) % N), z = (int)(i / P);
    const double f = (double)phi[i], p = (double)pi[i];
    const double dx = (double)phi[i - x + (x + 1 == N ? 0 : x + 1)] - f;
    const double dy = (double)phi[i + (long long)((y + 1 == N ? 0 : y + 1) - y) * N] - f;
    const double dz = (double)phi[i + (long long)((z + 1 == N ? 0 : z + 1) - z) * P] - f;
    const double q = 0.0 + f * f;
    const double g = (dx * dx / k.a2 + dy * dy / k.a2) + dz * dz / k.a2;
    const double v = k.cm * (f * f) + (k.lmode * ((q * q) / 24.0) + ((q * q) * q) / k.d6);
    rho[i] = (0.5 * (p * p) + 0.5 * g) + v;
  }
}

extern "C" int pn_energy_density(pn_ctx *c, double *rho) {
  if (!c || !rho) return pn_fail(c, PN_ERR_ARG, "pn_energy_density: NULL argument");
  PN_HIP(c, hipSetDevice(c->p.device));
  double *d = nullptr;
  PN_HIP(c, hipMalloc(&d, (size_t)c->n3 * sizeof(double)));
  RedCoef k = red_coef(c);
  const int blocks = (int)std::min<long long>(8192, (c->n3 + 255) / 256);
  if (c->p.precision == PN_FP64)
    hipLaunchKernelGGL(k_rho<double>, dim3(blocks), dim3(256), 0, c->stream,
                       (const double *)c->phi[c->cur], (const double *)c->pi, c->p.N, k, d);
  else
    hipLaunchKernelGGL(k_rho<float>, dim3(blocks), dim3(256), 0, c->stream,
                       (const float *)c->phi[c->cur], (const float *)c->pi, c->p.N, k, d);
  hipError_t e = hipGetLastError();
  if (e == hipSuccess) e = hipMemcpyAsync(rho, d, (size_t)c->n3 * 8, hipMemcpyDeviceToHost, c->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
  (void)hipFree(d);
  if (e != hipSuccess) return pn_fail(c, PN_ERR_HIP, "pn_energy_density: %s", hipGetErrorString(e));
  return PN_OK;
}

// ---------------------------------------------------------------------------
// field difference of two lattices: the precision form's field error against
// an FP64 run (no reference counterpart)
// ---------------------------------------------------------------------------
constexpr int kDiffVals = 12;  // 4 maxima, 4 compensated sums (hi, lo)

__device__ __forceinline__ double nan_max(double m, double v) { return (v > m || v != v) ? v : m; }

template <typename TA, typename TB>
__global__ __launch_bounds__(256) void k_field_diff(const TA *__restrict__ pa, const TA *__restrict__ qa,
                                                    const TB *__restrict__ pb, const TB *__restrict__ qb,
                                                    long long n3, double *__restrict__ part) {
  // mx: max|dphi|, max|dpi|, max|phi_b|, max|pi_b|;
  // sums: (dphi)^2, phi_b^2, (dpi)^2, pi_b^2
  double mx[4] = {0.0, 0.0, 0.0, 0.0}, hi[4] = {0.0, 0.0, 0.0, 0.0}, lo[4] = {0.0, 0.0, 0.0, 0.0};
  const long long stride = (long long)gridDim.x * blockDim.x;
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n3; i += stride) {
    const double fa = (double)pa[i], fb = (double)pb[i], ga = (double)qa[i], gb = (double)qb[i];
    const double df = fa - fb, dg = ga - gb;
    mx[0] = nan_max(mx[0], fabs(df));
    mx[1] = nan_max(mx[1], fabs(dg));
    mx[2] = nan_max(mx[2], fabs(fb));
    mx[3] = nan_max(mx[3], fabs(gb));
    two_acc(hi[0], lo[0], df * df);
    two_acc(hi[1], lo[1], fb * fb);
    two_acc(hi[2], lo[2], dg * dg);
    two_acc(hi[3], lo[3], gb * gb);
  }
  for (int off = 32; off > 0; off >>= 1) {
#pragma unroll
    for (int s = 0; s < 4; s++) {
      mx[s] = nan_max(mx[s], __shfl_down(mx[s], off, 64));
      const double h2 = __shfl_down(hi[s], off, 64), l2 = __shfl_down(lo[s], off, 64);
      two_merge(hi[s], lo[s], h2, l2);
    }
  }
  __shared__ double sh[4][kDiffVals];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  if (lane == 0)
    for (int s = 0; s < 4; s++) {
      sh[w][s] = mx[s];
      sh[w][4 + 2 * s] = hi[s];
      sh[w][5 + 2 * s] = lo[s];
    }
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int ww = 1; ww < (int)(blockDim.x >> 6); ww++)
      for (int s = 0; s < 4; s++) {
        sh[0][s] = nan_max(sh[0][s], sh[ww][s]);
        two_merge(sh[0][4 + 2 * s], sh[0][5 + 2 * s], sh[ww][4 + 2 * s], sh[ww][5 + 2 * s]);
      }
    for (int s = 0; s < kDiffVals; s++) part[(long long)blockIdx.x * kDiffVals + s] = sh[0][s];
  }
}

__global__ __launch_bounds__(64) void k_field_diff_final(const double *__restrict__ part, int nb,
                                                         double *__restrict__ out) {
  double mx[4] = {0.0, 0.0, 0.0, 0.0}, hi[4] = {0.0, 0.0, 0.0, 0.0}, lo[4] = {0.0, 0.0, 0.0, 0.0};
  for (int b = threadIdx.x; b < nb; b += 64)
    for (int s = 0; s < 4; s++) {
      mx[s] = nan_max(mx[s], part[(long long)b * kDiffVals + s]);
      two_merge(hi[s], lo[s], part[(long long)b * kDiffVals + 4 + 2 * s], part[(long long)b * kDiffVals + 5 + 2 * s]);
    }
  for (int off = 32; off > 0; off >>= 1)
    for (int s = 0; s < 4; s++) {
      mx[s] = nan_max(mx[s], __shfl_down(mx[s], off, 64));
      const double h2 = __shfl_down(hi[s], off, 64), l2 = __shfl_down(lo[s], off, 64);
      two_merge(hi[s], lo[s], h2, l2);
    }
  if (threadIdx.x == 0)
    for (int s = 0; s < 4; s++) {
      out[s] = mx[s];
      out[4 + s] = hi[s] + lo[s];
    }
}

template <typename TA, typename TB>
static void launch_field_diff(pn_ctx *a, pn_ctx *b, int nb) {
  hipLaunchKernelGGL((k_field_diff<TA, TB>), dim3(nb), dim3(kRedThreads), 0, a->stream,
                     (const TA *)a->phi[a->cur], (const TA *)a->pi, (const TB *)b->phi[b->cur],
                     (const TB *)b->pi, a->n3, a->red_part);
}

extern "C" int pn_field_diff_get(pn_ctx *a, pn_ctx *b, pn_field_diff *out) {
  if (!a || !b || !out) return pn_fail(a, PN_ERR_ARG, "pn_field_diff_get: NULL argument");
  if (a->p.N != b->p.N) return pn_fail(a, PN_ERR_ARG, "pn_field_diff_get: N %d vs %d", a->p.N, b->p.N);
  if (a->p.device != b->p.device)
    return pn_fail(a, PN_ERR_ARG, "pn_field_diff_get: lattices on devices %d and %d", a->p.device, b->p.device);
  PN_HIP(a, hipSetDevice(a->p.device));
  PN_HIP(a, hipStreamSynchronize(b->stream));  // b's pending steps
  const int nb = a->red_blocks;  // red_part holds 2 * kNSums >= kDiffVals values per block
  static_assert(2 * kNSums >= kDiffVals, "partials fit the reduction scratch");
  const bool a64 = a->p.precision == PN_FP64, b64 = b->p.precision == PN_FP64;
  if (a64 && b64) launch_field_diff<double, double>(a, b, nb);
  else if (!a64 && b64) launch_field_diff<float, double>(a, b, nb);
  else if (a64 && !b64) launch_field_diff<double, float>(a, b, nb);
  else launch_field_diff<float, float>(a, b, nb);
  PN_CHECK_LAUNCH(a);
  hipLaunchKernelGGL(k_field_diff_final, dim3(1), dim3(64), 0, a->stream, a->red_part, nb, a->red_out);
  PN_CHECK_LAUNCH(a);
  PN_HIP(a, hipMemcpyAsync(a->host_pin, a->red_out, 8 * sizeof(double), hipMemcpyDeviceToHost, a->stream));
  PN_HIP(a, hipStreamSynchronize(a->stream));
  const double *s = a->host_pin;
  out->max_dphi = s[0];
  out->max_dpi = s[1];
  out->max_phi = s[2];
  out->max_pi = s[3];
  out->sum_dphi2 = s[4];
  out->sum_phi2 = s[5];
  out->sum_dpi2 = s[6];
  out->sum_pi2 = s[7];
  return PN_OK;
}
