// pn_batch.hip -- many independent lattices (seeds, couplings and sizes may
// differ) advanced together.
//
// Layout: the realizations' fields are packed one after the other in three
// device buffers (Phi ping-pong and Pi), each realization dense x-fastest as
// in the single-lattice path.  A descriptor per realization carries its
// offset, size, couplings, tile shape and a decayed flag.
//
// Step: one launch per leapfrog step for all realizations.  The launch walks
// a tile list that concatenates every realization's (x, y, z) tiles; blocks
// loop over it with a grid stride, so the grid stays a fixed multiple of the
// CU count whatever the mix of sizes.  The per-site arithmetic is the same
// device function as the single-lattice kernel (bit-identical results).
// Sequences of steps are captured once into a HIP graph and replayed.
//
// Two-step mode (the default when every realization fits): one launch per
// pair of steps, each tile running the single-lattice two-step tile code
// (step2_tile) on full x rows; a block of kB2 threads covers kB2*W/N rows of
// an N^3 realization.  Pi then alternates between two buffers.
//
// Tile shape (rows per tile, planes per z block) is chosen per size class by
// an autotuner that times the candidates on the live batch and restores it.
//
// Analysis (averages, extrema, spectra, islands, IC) goes through views: a
// pn_ctx that aliases one realization's slice of the buffers.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <map>
#include <new>
#include <string>
#include <vector>

#include "pn_internal.h"
#include "pn_kernels.h"

struct BatchDesc {
  long long off;   // element offset of this realization
  int N, XT, YT, ZT, TY, ZB;
  int tile_start;  // first tile of this realization in the tile list
  int pow2, decayed;
  int RT2, YT2, ZB2, ZT2;  // two-step tiles: RT2 thread rows (RT2-2 output rows), ZB2 planes
  int tile_start2;         // first tile in the two-step tile list
  int RT3, YT3, ZB3, ZT3;  // three-step tiles: RT3 thread rows (RT3-4 output rows), ZB3 planes
  int tile_start3;         // first tile in the three-step tile list
  CoefF64 k64;
  CoefF32 k32;
};

struct pn_batch {
  int R = 0, precision = PN_FP64, device = 0;
  size_t esize = 8;
  std::vector<pn_realization> re;
  std::vector<long long> off;
  long long total = 0;
  void *phi[2] = {nullptr, nullptr};
  void *pib[2] = {nullptr, nullptr};  // Pi; the second buffer only in two-step mode
  int pcur = 0;                        // pib[pcur] holds the current Pi
  int cur = 0;
  int step2_ok = 0;                    // every realization fits a two-step tile
  int step2 = 0;                       // two-step mode on
  std::map<int, int> zb2;              // N -> planes per two-step tile
  std::vector<int> htile2;
  int *dtile2 = nullptr;
  int total_tiles2 = 0;
  int step3_ok = 0;                    // every realization fits a three-step tile
  int step3 = 0;                       // three-step mode on (needs the two-step buffers)
  std::map<int, int> zb3;              // N -> planes per three-step tile
  std::vector<int> htile3;
  int *dtile3 = nullptr;
  int total_tiles3 = 0;
  hipStream_t stream = nullptr;
  std::vector<BatchDesc> hdesc;
  BatchDesc *ddesc = nullptr;
  int *dtile = nullptr;  // R+1 tile starts
  int total_tiles = 0;
  std::vector<int> htile;
  double *dtk = nullptr, *dtd = nullptr;  // [R][cap]
  int table_cap = 0;
  int *dbase = nullptr;  // step base for graph replays
  std::map<int, std::pair<int, int>> tiles;  // N -> (TY, ZB)
  std::vector<pn_ctx *> views;
  hipGraphExec_t gexec[2][2] = {{nullptr, nullptr}, {nullptr, nullptr}};  // [cur][pcur]
  int graph_len = 0;
  std::string err;
};

static int bfail(pn_batch *b, int code, const char *msg, const char *extra = "") {
  if (b) b->err = std::string(msg) + extra;
  pn_set_thread_error((std::string(msg) + extra).c_str());
  return code;
}

#define PB_HIP(b, call)                                                              \
  do {                                                                               \
    hipError_t e_ = (call);                                                          \
    if (e_ != hipSuccess) return bfail((b), PN_ERR_HIP, #call ": ", hipGetErrorString(e_)); \
  } while (0)

// ---------------------------------------------------------------------------
// kernel
// ---------------------------------------------------------------------------
__global__ void k_set_int(int *p, int v) { *p = v; }

template <typename T, int W>
__global__ __launch_bounds__(256) void k_batch_step(const T *__restrict__ phi_all,
                                                    T *__restrict__ pi_all,
                                                    T *__restrict__ phin_all,
                                                    const BatchDesc *__restrict__ desc,
                                                    const int *__restrict__ tile_start, int R,
                                                    int total_tiles, const double *__restrict__ dtk_tab,
                                                    const double *__restrict__ dtd_tab, int cap,
                                                    const int *__restrict__ base_ptr, int j) {
  typedef typename VecT<T, W>::type V;
  const int step = *base_ptr + j;
  for (int b = blockIdx.x; b < total_tiles; b += gridDim.x) {
    // realization owning tile b: last r with tile_start[r] <= b
    int lo = 0, hi = R;
    while (hi - lo > 1) {
      const int mid = (lo + hi) >> 1;
      if (tile_start[mid] <= b) lo = mid;
      else hi = mid;
    }
    const BatchDesc d = desc[lo];
    const int t = b - d.tile_start;
    const int xt = t % d.XT, yt = (t / d.XT) % d.YT, zt = t / (d.XT * d.YT);
    const int N = d.N;
    const T *phi = phi_all + d.off;
    T *pi = pi_all + d.off;
    T *phin = phin_all + d.off;
    const int x = xt * 64 * W + W * (int)threadIdx.x;
    const int z0 = zt * d.ZB, z1 = min(z0 + d.ZB, N);
    const T dk = (T)dtk_tab[(long long)lo * cap + step];
    const T dd = (T)dtd_tab[(long long)lo * cap + step];
    if (x >= N) continue;
    for (int yy = (int)threadIdx.y; yy < d.TY; yy += blockDim.y) {
      const int y = yt * d.TY + yy;
      if (y >= N) break;
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
        if (d.decayed) {
          // a decayed realization is frozen (main.cpp:85-89 ends its run);
          // the copy keeps both ping-pong buffers holding its last field
          __builtin_nontemporal_store(fc, (V *)&phin[base + row + x]);
          fm = fc;
          fc = fp;
          continue;
        }
        const T vxp = phi[base + row + xp];
        const T vxm = phi[base + row + xm];
        const V vyp = *(const V *)&phi[base + rowp + x];
        const V vym = *(const V *)&phi[base + rowm + x];
        V p = __builtin_nontemporal_load((const V *)&pi[base + row + x]);
#pragma unroll
        for (int jj = 0; jj < W; jj++) {
          const T right = (jj == W - 1) ? vxp : fc[(jj + 1) % W];
          const T left = (jj == 0) ? vxm : fc[(jj + W - 1) % W];
          T pj = p[jj];
          if constexpr (sizeof(T) == 8) {
            if (d.pow2) site_update_f64<true, DIV_PER_SITE>(d.k64, right, left, vyp[jj], vym[jj], fp[jj], fm[jj], fc[jj], pj, dk);
            else site_update_f64<false, DIV_PER_SITE>(d.k64, right, left, vyp[jj], vym[jj], fp[jj], fm[jj], fc[jj], pj, dk);
          } else {
            site_update_f32(d.k32, right, left, vyp[jj], vym[jj], fp[jj], fm[jj], fc[jj], pj, dk);
          }
          p[jj] = pj;
        }
        __builtin_nontemporal_store(p, (V *)&pi[base + row + x]);
        V fnew;
#pragma unroll
        for (int jj = 0; jj < W; jj++) fnew[jj] = fc[jj] + p[jj] * dd;
        __builtin_nontemporal_store(fnew, (V *)&phin[base + row + x]);
        fm = fc;
        fc = fp;
      }
    }
  }
}

// Two steps per launch for all realizations: a grid-stride walk over the
// two-step tile list.  Block = kB2 threads, thread t -> (t % (N/W), t / (N/W))
// of the tile's realization; threads beyond its RT2 rows only join barriers.
// Two blocks per CU are 6 waves per SIMD, which needs at most 80 VGPRs: the
// launch bound asks for it, and the tiles skip the early HBM reads of the
// single-lattice kernel (with them the kernel needs 87 VGPRs, spills under the
// bound, and the batch runs 12 % slower).
static constexpr int kB2 = 768;

template <typename T, int W>
__global__ __launch_bounds__(kB2, 6) void k_batch_step2(const T *__restrict__ phi_all,
                                                     const T *__restrict__ pi_all,
                                                     T *__restrict__ pio_all,
                                                     T *__restrict__ phin_all,
                                                     const BatchDesc *__restrict__ desc,
                                                     const int *__restrict__ tile_start, int R,
                                                     int total_tiles,
                                                     const double *__restrict__ dtk_tab,
                                                     const double *__restrict__ dtd_tab, int cap,
                                                     const int *__restrict__ base_ptr, int j) {
  typedef typename VecT<T, W>::type V;
  extern __shared__ unsigned char lds_raw[];
  const int step = *base_ptr + j;  // this launch advances steps `step` and `step + 1`
  for (int b = blockIdx.x; b < total_tiles; b += gridDim.x) {
    int lo = 0, hi = R;
    while (hi - lo > 1) {
      const int mid = (lo + hi) >> 1;
      if (tile_start[mid] <= b) lo = mid;
      else hi = mid;
    }
    const BatchDesc d = desc[lo];
    const int t = b - d.tile_start2;
    const int yt = t % d.YT2, zt = t / d.YT2;
    const int N = d.N, RT = d.RT2, tpr = N / W;
    const int tx = (int)threadIdx.x % tpr, ty = (int)threadIdx.x / tpr;
    const bool active = ty < RT;
    const int y0 = yt * (RT - 2), z0 = zt * d.ZB2, z1 = min(z0 + d.ZB2, N);
    const T *phi = phi_all + d.off;
    const T *pi = pi_all + d.off;
    T *pio = pio_all + d.off;
    T *phin = phin_all + d.off;
    if (d.decayed) {
      // frozen (main.cpp:85-89 ends a decayed run): the tile's output rows
      // keep their Phi and Pi in the buffers the next launch reads
      const int y = y0 + ty - 1;
      if (active && ty >= 1 && ty <= RT - 2 && y < N) {
        const long long P = (long long)N * N;
        for (int z = z0; z < z1; z++) {
          const long long i = z * P + (long long)y * N + W * tx;
          *(V *)&phin[i] = *(const V *)&phi[i];
          *(V *)&pio[i] = *(const V *)&pi[i];
        }
      }
      continue;
    }
    const long long tb = (long long)lo * cap + step;
    const T dk0 = (T)dtk_tab[tb], dd0 = (T)dtd_tab[tb], dk1 = (T)dtk_tab[tb + 1], dd1 = (T)dtd_tab[tb + 1];
    if constexpr (sizeof(T) == 8) {
      if (d.pow2)
        step2_tile<T, CoefF64, true, W, false>(phi, pi, pio, phin, (T *)lds_raw, N, RT, tx, ty, active, y0,
                                        z0, z1, d.k64, dk0, dd0, dk1, dd1);
      else
        step2_tile<T, CoefF64, false, W, false>(phi, pi, pio, phin, (T *)lds_raw, N, RT, tx, ty, active,
                                         y0, z0, z1, d.k64, dk0, dd0, dk1, dd1);
    } else {
      step2_tile<T, CoefF32, true, W, false>(phi, pi, pio, phin, (T *)lds_raw, N, RT, tx, ty, active, y0,
                                      z0, z1, d.k32, dk0, dd0, dk1, dd1);
    }
  }
}

// Three steps per launch for all realizations: a grid-stride walk over the
// three-step tile list (interleaved, so that the size classes' tiles, which
// differ in cost, spread over the XCDs), each tile step3_tile_r3 (nine LDS
// planes of RT3*N <= kB3*W values: one block per CU).
// Block = kB3 = batch_b3<T>() threads, thread t -> (t % (N/W), t / (N/W)) of
// the tile's realization; threads beyond its RT3 rows only join barriers.
// FP64 blocks are 768 threads: with 1024 the tile spills at 128 registers;
// 768 allows 168.
template <typename T>
constexpr int batch_b3() {
  return sizeof(T) == 8 ? 768 : 1024;
}

template <typename T, int W>
__global__ __launch_bounds__(batch_b3<T>()) void k_batch_step3(const T *__restrict__ phi_all,
                                                     const T *__restrict__ pi_all,
                                                     T *__restrict__ pio_all,
                                                     T *__restrict__ phin_all,
                                                     const BatchDesc *__restrict__ desc,
                                                     const int *__restrict__ tile_start, int R,
                                                     int total_tiles,
                                                     const double *__restrict__ dtk_tab,
                                                     const double *__restrict__ dtd_tab, int cap,
                                                     const int *__restrict__ base_ptr, int j) {
  typedef typename VecT<T, W>::type V;
  extern __shared__ unsigned char lds_raw[];
  const int step = *base_ptr + j;  // this launch advances steps `step` .. `step + 2`
  for (int b = blockIdx.x; b < total_tiles; b += gridDim.x) {
    int lo = 0, hi = R;
    while (hi - lo > 1) {
      const int mid = (lo + hi) >> 1;
      if (tile_start[mid] <= b) lo = mid;
      else hi = mid;
    }
    const BatchDesc d = desc[lo];
    const int t = b - d.tile_start3;
    const int yt = t % d.YT3, zt = t / d.YT3;
    const int N = d.N, RT = d.RT3, tpr = N / W;
    const int tx = (int)threadIdx.x % tpr, ty = (int)threadIdx.x / tpr;
    const bool active = ty < RT;
    const int y0 = yt * (RT - 4), z0 = zt * d.ZB3, z1 = min(z0 + d.ZB3, N);
    const T *phi = phi_all + d.off;
    const T *pi = pi_all + d.off;
    T *pio = pio_all + d.off;
    T *phin = phin_all + d.off;
    if (d.decayed) {
      // frozen (main.cpp:85-89 ends a decayed run): the tile's output rows
      // keep their Phi and Pi in the buffers the next launch reads
      const int y = y0 + ty - 2;
      if (active && ty >= 2 && ty <= RT - 3 && y < N) {
        const long long P = (long long)N * N;
        for (int z = z0; z < z1; z++) {
          const long long i = z * P + (long long)y * N + W * tx;
          *(V *)&phin[i] = *(const V *)&phi[i];
          *(V *)&pio[i] = *(const V *)&pi[i];
        }
      }
      continue;
    }
    const long long tb = (long long)lo * cap + step;
    const T dk0 = (T)dtk_tab[tb], dd0 = (T)dtd_tab[tb], dk1 = (T)dtk_tab[tb + 1],
            dd1 = (T)dtd_tab[tb + 1], dk2 = (T)dtk_tab[tb + 2], dd2 = (T)dtd_tab[tb + 2];
    if constexpr (sizeof(T) == 8) {
      if (d.pow2)
        step3_tile_r3<T, CoefF64, true, W>(phi, pi, pio, phin, (T *)lds_raw, N, RT, tx, ty, active, y0,
                                           z0, z1, d.k64, dk0, dd0, dk1, dd1, dk2, dd2);
      else
        step3_tile_r3<T, CoefF64, false, W>(phi, pi, pio, phin, (T *)lds_raw, N, RT, tx, ty, active,
                                            y0, z0, z1, d.k64, dk0, dd0, dk1, dd1, dk2, dd2);
    } else {
      step3_tile_r3<T, CoefF32, true, W>(phi, pi, pio, phin, (T *)lds_raw, N, RT, tx, ty, active, y0,
                                         z0, z1, d.k32, dk0, dd0, dk1, dd1, dk2, dd2);
    }
    __syncthreads();  // the next tile's prologue rewrites the LDS planes
  }
}

// ---------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------
static bool pow2d(double v) {
  if (!(v > 0) || !std::isfinite(v)) return false;
  int e;
  return std::frexp(v, &e) == 0.5;
}

static int vec_width(const pn_batch *b) { return b->precision == PN_FP64 ? 2 : 4; }

static std::pair<int, int> default_tiles(int N) { return {4, N >= 256 ? 16 : 32}; }

// two-step rows per tile for an N^3 realization (0 when it does not fit)
static int step2_rows(const pn_batch *b, int N) {
  const int W = vec_width(b);
  if (N % W != 0 || N / W > kB2) return 0;
  const int rt = std::min(kB2 / (N / W), N + 2);
  return rt >= 3 ? rt : 0;
}

// four planes of the tile in LDS: two of stage A results, two of Phi
static size_t step2_lds(const pn_batch *b) { return (size_t)4 * kB2 * vec_width(b) * b->esize; }

// three-step rows per tile for an N^3 realization (0 when it does not fit):
// at least one output row, and the tile's rows y0-2 .. y0+RT-3 wrap into
// [0, N) with one fold
static int b3_threads(const pn_batch *b) {
  return b->precision == PN_FP64 ? batch_b3<double>() : batch_b3<float>();
}

static int step3_rows(const pn_batch *b, int N) {
  const int W = vec_width(b), kB3 = b3_threads(b);
  if (N % W != 0 || N / W > kB3 || N < 8) return 0;
  const int rt = std::min(kB3 / (N / W), N + 4);
  if (rt < 5) return 0;
  const int TY = rt - 4, YT = (N + TY - 1) / TY;
  return (YT - 1) * TY + rt - 3 < 2 * N ? rt : 0;
}

// nine planes of RT3*N <= kB3*W values (step3_tile_r3)
static size_t step3_lds(const pn_batch *b) { return (size_t)9 * b3_threads(b) * vec_width(b) * b->esize; }

// rebuild descriptors and the tile list; mask (optional) selects the
// realizations that get tiles
static int upload_desc(pn_batch *b, const std::vector<int> *mask = nullptr) {
  const int W = vec_width(b);
  b->hdesc.resize(b->R);
  b->htile.assign(b->R + 1, 0);
  b->htile2.assign(b->R + 1, 0);
  b->htile3.assign(b->R + 1, 0);
  int tiles = 0, tiles2 = 0, tiles3 = 0;
  for (int r = 0; r < b->R; r++) {
    const pn_realization &q = b->re[r];
    BatchDesc &d = b->hdesc[r];
    auto it = b->tiles.find(q.N);
    const std::pair<int, int> ts = it != b->tiles.end() ? it->second : default_tiles(q.N);
    const int decayed = d.decayed;
    d = BatchDesc();
    d.decayed = decayed;
    d.off = b->off[r];
    d.N = q.N;
    d.TY = ts.first;
    d.ZB = ts.second;
    d.XT = (q.N + 64 * W - 1) / (64 * W);
    d.YT = (q.N + d.TY - 1) / d.TY;
    d.ZT = (q.N + d.ZB - 1) / d.ZB;
    d.k64.a2 = q.a * q.a;
    d.k64.inv_a2 = 1.0 / d.k64.a2;
    d.pow2 = pow2d(d.k64.a2);
    d.k64.c2 = q.mass_mode * q.m2;
    d.k64.lmode = (double)q.lambda_mode;
    d.k64.kappa = q.kappa;
    coef64_recips(d.k64);
    d.k32.inv_a2 = (float)(1.0 / d.k64.a2);
    d.k32.c2 = (float)d.k64.c2;
    d.k32.c4 = (float)(d.k64.lmode / 6.0);
    d.k32.c6 = (float)(1.0 / (120.0 * q.kappa * q.kappa));
    d.tile_start = tiles;
    b->htile[r] = tiles;
    if (!mask || (*mask)[r]) tiles += d.XT * d.YT * d.ZT;
    d.RT2 = step2_rows(b, q.N);
    if (d.RT2) {
      auto z2 = b->zb2.find(q.N);
      d.ZB2 = std::min(z2 != b->zb2.end() ? z2->second : 16, q.N);
      d.YT2 = (q.N + d.RT2 - 3) / (d.RT2 - 2);
      d.ZT2 = (q.N + d.ZB2 - 1) / d.ZB2;
    }
    d.tile_start2 = tiles2;
    b->htile2[r] = tiles2;
    if (d.RT2 && (!mask || (*mask)[r])) tiles2 += d.YT2 * d.ZT2;
    d.RT3 = step3_rows(b, q.N);
    if (d.RT3) {
      auto z3 = b->zb3.find(q.N);
      d.ZB3 = std::min(z3 != b->zb3.end() ? z3->second : 32, q.N);
      d.YT3 = (q.N + d.RT3 - 5) / (d.RT3 - 4);
      d.ZT3 = (q.N + d.ZB3 - 1) / d.ZB3;
    }
    d.tile_start3 = tiles3;
    b->htile3[r] = tiles3;
    if (d.RT3 && (!mask || (*mask)[r])) tiles3 += d.YT3 * d.ZT3;
  }
  b->htile[b->R] = tiles;
  b->total_tiles = tiles;
  b->htile2[b->R] = tiles2;
  b->total_tiles2 = tiles2;
  b->htile3[b->R] = tiles3;
  b->total_tiles3 = tiles3;
  PB_HIP(b, hipMemcpyAsync(b->ddesc, b->hdesc.data(), sizeof(BatchDesc) * b->R, hipMemcpyHostToDevice, b->stream));
  PB_HIP(b, hipMemcpyAsync(b->dtile, b->htile.data(), sizeof(int) * (b->R + 1), hipMemcpyHostToDevice, b->stream));
  PB_HIP(b, hipMemcpyAsync(b->dtile2, b->htile2.data(), sizeof(int) * (b->R + 1), hipMemcpyHostToDevice, b->stream));
  if (b->dtile3)
    PB_HIP(b, hipMemcpyAsync(b->dtile3, b->htile3.data(), sizeof(int) * (b->R + 1), hipMemcpyHostToDevice, b->stream));
  PB_HIP(b, hipStreamSynchronize(b->stream));
  return PN_OK;
}

static void drop_graphs(pn_batch *b) {
  for (int i = 0; i < 2; i++)
    for (int k = 0; k < 2; k++)
      if (b->gexec[i][k]) {
        (void)hipGraphExecDestroy(b->gexec[i][k]);
        b->gexec[i][k] = nullptr;
      }
}

// views follow the current Phi parity and the current Pi buffer
static void sync_views(pn_batch *b) {
  for (int r = 0; r < b->R; r++) {
    b->views[r]->cur = b->cur;
    b->views[r]->pi = (char *)b->pib[b->pcur] + b->off[r] * b->esize;
  }
}

extern "C" int pn_batch_destroy(pn_batch *b) {
  if (!b) return PN_OK;
  (void)hipSetDevice(b->device);
  if (b->stream) (void)hipStreamSynchronize(b->stream);
  drop_graphs(b);
  for (pn_ctx *v : b->views) pn_destroy(v);
  for (int i = 0; i < 2; i++) (void)hipFree(b->phi[i]);
  for (int i = 0; i < 2; i++) (void)hipFree(b->pib[i]);
  (void)hipFree(b->ddesc);
  (void)hipFree(b->dtile);
  (void)hipFree(b->dtile2);
  (void)hipFree(b->dtile3);
  (void)hipFree(b->dtk);
  (void)hipFree(b->dtd);
  (void)hipFree(b->dbase);
  if (b->stream) (void)hipStreamDestroy(b->stream);
  delete b;
  return PN_OK;
}

extern "C" const char *pn_batch_last_error(const pn_batch *b) {
  return b ? b->err.c_str() : pn_last_error(nullptr);
}

extern "C" int pn_batch_create(int R, const pn_realization *reals, int precision, int device,
                               pn_batch **out) {
  if (!out || !reals || R < 1) return bfail(nullptr, PN_ERR_ARG, "pn_batch_create: bad argument");
  *out = nullptr;
  if (const int rc = pn_runtime_check()) return rc;
  if (precision != PN_FP64 && precision != PN_FP32)
    return bfail(nullptr, PN_ERR_ARG, "pn_batch_create: unknown precision");
  pn_batch *b = new (std::nothrow) pn_batch();
  if (!b) return bfail(nullptr, PN_ERR_NOMEM, "pn_batch_create: out of host memory");
  b->R = R;
  b->precision = precision;
  b->device = device;
  b->esize = precision == PN_FP64 ? 8 : 4;
  b->re.assign(reals, reals + R);
  const int W = vec_width(b);
  long long total = 0;
  for (int r = 0; r < R; r++) {
    const pn_realization &q = b->re[r];
    if (q.N < 4 || q.N > 2048 || q.N % W != 0 || !(q.a > 0) || !(q.kappa != 0)) {
      pn_batch_destroy(b);
      return bfail(nullptr, PN_ERR_ARG, "pn_batch_create: realization has N < 4, N % vector width != 0, "
                                        "a <= 0 or Kappa == 0");
    }
    b->off.push_back(total);
    total += (long long)q.N * q.N * q.N;
    total = (total + 63) / 64 * 64;  // keep every realization 256/512-byte aligned
  }
  b->total = total;
  auto fail = [&](int code, const char *m, const char *x = "") {
    int rc = bfail(nullptr, code, m, x);
    pn_batch_destroy(b);
    return rc;
  };
  if (hipSetDevice(device) != hipSuccess) return fail(PN_ERR_HIP, "pn_batch_create: hipSetDevice failed");
  if (hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking) != hipSuccess)
    return fail(PN_ERR_HIP, "pn_batch_create: stream creation failed");
  const size_t bytes = (size_t)total * b->esize;
  for (int i = 0; i < 2; i++)
    if (hipMalloc(&b->phi[i], bytes) != hipSuccess) return fail(PN_ERR_NOMEM, "pn_batch_create: field allocation failed");
  if (hipMalloc(&b->pib[0], bytes) != hipSuccess) return fail(PN_ERR_NOMEM, "pn_batch_create: field allocation failed");
  b->step2_ok = 1;
  for (int r = 0; r < R; r++)
    if (!step2_rows(b, b->re[r].N)) b->step2_ok = 0;
  // the second Pi buffer of two-step mode; without room the batch runs single steps
  if (b->step2_ok && hipMalloc(&b->pib[1], bytes) != hipSuccess) {
    (void)hipGetLastError();
    b->pib[1] = nullptr;
    b->step2_ok = 0;
  }
  b->step2 = b->step2_ok;
  // three-step mode: every realization fits a tile and the device gives a
  // block the nine LDS planes
  int lds_max = 0;
  if (hipDeviceGetAttribute(&lds_max, hipDeviceAttributeMaxSharedMemoryPerBlock, device) != hipSuccess)
    lds_max = 0;
  b->step3_ok = b->step2_ok && step3_lds(b) <= (size_t)lds_max;
  for (int r = 0; r < R; r++)
    if (!step3_rows(b, b->re[r].N)) b->step3_ok = 0;
  if (b->step3_ok) {
    const void *kern = precision == PN_FP64 ? (const void *)&k_batch_step3<double, 2>
                                            : (const void *)&k_batch_step3<float, 4>;
    if (hipFuncSetAttribute(kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)step3_lds(b)) != hipSuccess) {
      (void)hipGetLastError();
      b->step3_ok = 0;
    }
  }
  b->step3 = b->step3_ok;
  if (hipMalloc(&b->ddesc, sizeof(BatchDesc) * R) != hipSuccess || hipMalloc(&b->dtile, sizeof(int) * (R + 1)) != hipSuccess ||
      hipMalloc(&b->dtile2, sizeof(int) * (R + 1)) != hipSuccess || hipMalloc(&b->dtile3, sizeof(int) * (R + 1)) != hipSuccess ||
      hipMalloc(&b->dbase, sizeof(int)) != hipSuccess)
    return fail(PN_ERR_NOMEM, "pn_batch_create: descriptor allocation failed");
  for (int i = 0; i < 2; i++)
    if (hipMemsetAsync(b->phi[i], 0, bytes, b->stream) != hipSuccess) return fail(PN_ERR_HIP, "pn_batch_create: zero fill failed");
  if (hipMemsetAsync(b->pib[0], 0, bytes, b->stream) != hipSuccess) return fail(PN_ERR_HIP, "pn_batch_create: zero fill failed");
  if (b->pib[1] && hipMemsetAsync(b->pib[1], 0, bytes, b->stream) != hipSuccess)
    return fail(PN_ERR_HIP, "pn_batch_create: zero fill failed");
  b->hdesc.assign(R, BatchDesc());
  if (upload_desc(b) != PN_OK) return fail(PN_ERR_HIP, "pn_batch_create: descriptor upload failed");
  for (int r = 0; r < R; r++) {
    const pn_realization &q = b->re[r];
    pn_params p{q.N, q.a, q.mass_mode, q.m2, q.lambda_mode, q.kappa, precision, device};
    char *p0 = (char *)b->phi[0] + b->off[r] * b->esize;
    char *p1 = (char *)b->phi[1] + b->off[r] * b->esize;
    char *pp = (char *)b->pib[0] + b->off[r] * b->esize;
    pn_ctx *v = nullptr;
    if (pn_make_view(&p, p0, p1, pp, 0, b->stream, &v) != PN_OK)
      return fail(PN_ERR_NOMEM, "pn_batch_create: view creation failed");
    b->views.push_back(v);
  }
  *out = b;
  return PN_OK;
}

extern "C" int pn_batch_view(pn_batch *b, int r, pn_ctx **view) {
  if (!b || !view || r < 0 || r >= b->R) return bfail(b, PN_ERR_ARG, "pn_batch_view: bad argument");
  sync_views(b);
  *view = b->views[r];
  return PN_OK;
}

extern "C" int pn_batch_set_decayed(pn_batch *b, const int *flags) {
  if (!b || !flags) return bfail(b, PN_ERR_ARG, "pn_batch_set_decayed: NULL argument");
  for (int r = 0; r < b->R; r++) b->hdesc[r].decayed = flags[r] ? 1 : 0;
  PB_HIP(b, hipSetDevice(b->device));
  PB_HIP(b, hipMemcpyAsync(b->ddesc, b->hdesc.data(), sizeof(BatchDesc) * b->R, hipMemcpyHostToDevice, b->stream));
  PB_HIP(b, hipStreamSynchronize(b->stream));
  return PN_OK;
}

extern "C" int pn_batch_set_tiles(pn_batch *b, int N, int rows, int zblock) {
  if (!b || rows < 1 || rows > 64 || zblock < 1 || zblock > 2048)
    return bfail(b, PN_ERR_ARG, "pn_batch_set_tiles: bad argument");
  b->tiles[N] = {rows, zblock};
  drop_graphs(b);
  return upload_desc(b);
}

static int grid_blocks(const pn_batch *b) {
  return std::max(1, std::min(b->total_tiles, 256 * 8));
}

static void launch_batch_step(pn_batch *b, int parity, int pparity, int j) {
  const int grid = grid_blocks(b);
  if (b->precision == PN_FP64)
    hipLaunchKernelGGL((k_batch_step<double, 2>), dim3(grid), dim3(64, 4), 0, b->stream,
                       (const double *)b->phi[parity], (double *)b->pib[pparity],
                       (double *)b->phi[parity ^ 1], b->ddesc, b->dtile, b->R, b->total_tiles,
                       b->dtk, b->dtd, b->table_cap, b->dbase, j);
  else
    hipLaunchKernelGGL((k_batch_step<float, 4>), dim3(grid), dim3(64, 4), 0, b->stream,
                       (const float *)b->phi[parity], (float *)b->pib[pparity],
                       (float *)b->phi[parity ^ 1], b->ddesc, b->dtile, b->R, b->total_tiles,
                       b->dtk, b->dtd, b->table_cap, b->dbase, j);
}

// steps j and j+1 of the tables: Phi parity -> parity^1, Pi pib[pparity] -> pib[pparity^1]
static void launch_batch_step2(pn_batch *b, int parity, int pparity, int j) {
  // two blocks of kB2 threads per CU
  const int grid = std::max(1, std::min(b->total_tiles2, 256 * 2));
  const size_t lds = step2_lds(b);
  if (b->precision == PN_FP64)
    hipLaunchKernelGGL((k_batch_step2<double, 2>), dim3(grid), dim3(kB2), lds, b->stream,
                       (const double *)b->phi[parity], (const double *)b->pib[pparity],
                       (double *)b->pib[pparity ^ 1], (double *)b->phi[parity ^ 1], b->ddesc,
                       b->dtile2, b->R, b->total_tiles2, b->dtk, b->dtd, b->table_cap, b->dbase, j);
  else
    hipLaunchKernelGGL((k_batch_step2<float, 4>), dim3(grid), dim3(kB2), lds, b->stream,
                       (const float *)b->phi[parity], (const float *)b->pib[pparity],
                       (float *)b->pib[pparity ^ 1], (float *)b->phi[parity ^ 1], b->ddesc,
                       b->dtile2, b->R, b->total_tiles2, b->dtk, b->dtd, b->table_cap, b->dbase, j);
}

// steps j .. j+2 of the tables: Phi parity -> parity^1, Pi pib[pparity] -> pib[pparity^1]
static void launch_batch_step3(pn_batch *b, int parity, int pparity, int j) {
  // one block per CU (nine LDS planes)
  const int grid = std::max(1, std::min(b->total_tiles3, 256));
  const size_t lds = step3_lds(b);
  if (b->precision == PN_FP64)
    hipLaunchKernelGGL((k_batch_step3<double, 2>), dim3(grid), dim3(batch_b3<double>()), lds, b->stream,
                       (const double *)b->phi[parity], (const double *)b->pib[pparity],
                       (double *)b->pib[pparity ^ 1], (double *)b->phi[parity ^ 1], b->ddesc,
                       b->dtile3, b->R, b->total_tiles3, b->dtk, b->dtd, b->table_cap, b->dbase, j);
  else
    hipLaunchKernelGGL((k_batch_step3<float, 4>), dim3(grid), dim3(batch_b3<float>()), lds, b->stream,
                       (const float *)b->phi[parity], (const float *)b->pib[pparity],
                       (float *)b->pib[pparity ^ 1], (float *)b->phi[parity ^ 1], b->ddesc,
                       b->dtile3, b->R, b->total_tiles3, b->dtk, b->dtd, b->table_cap, b->dbase, j);
}

// steps per captured graph: 100 single steps, 50 two-step launches or 32
// three-step launches (96 steps), all even numbers of launches, so a replay
// leaves the Phi parity and the Pi buffer where it found them
static constexpr int kGraphLen = 100;
static constexpr int kGraphLen3 = 96;
static int graph_len(const pn_batch *b) { return b->step3 ? kGraphLen3 : kGraphLen; }

static int ensure_graph(pn_batch *b, int parity, int pparity) {
  const int glen = graph_len(b);
  if (b->gexec[parity][pparity] && b->graph_len == glen) return PN_OK;
  hipGraph_t g = nullptr;
  PB_HIP(b, hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
  if (b->step3)
    for (int j = 0; j < kGraphLen3; j += 3) {
      const int flip = (j / 3) & 1;
      launch_batch_step3(b, parity ^ flip, pparity ^ flip, j);
    }
  else if (b->step2)
    for (int j = 0; j < kGraphLen; j += 2) {
      const int flip = (j / 2) & 1;
      launch_batch_step2(b, parity ^ flip, pparity ^ flip, j);
    }
  else
    for (int j = 0; j < kGraphLen; j++) launch_batch_step(b, parity ^ (j & 1), pparity, j);
  hipError_t e = hipStreamEndCapture(b->stream, &g);
  if (e != hipSuccess) return bfail(b, PN_ERR_HIP, "pn_batch_evolve: graph capture: ", hipGetErrorString(e));
  e = hipGraphInstantiate(&b->gexec[parity][pparity], g, nullptr, nullptr, 0);
  (void)hipGraphDestroy(g);
  if (e != hipSuccess) return bfail(b, PN_ERR_HIP, "pn_batch_evolve: graph instantiate: ", hipGetErrorString(e));
  b->graph_len = glen;
  return PN_OK;
}

static int upload_tables(pn_batch *b, int nsteps, const double *dtk, const double *dtd) {
  if (nsteps > b->table_cap) {
    PB_HIP(b, hipStreamSynchronize(b->stream));
    (void)hipFree(b->dtk);
    (void)hipFree(b->dtd);
    b->dtk = b->dtd = nullptr;
    b->table_cap = 0;
    const int cap = std::max(nsteps, 256);
    PB_HIP(b, hipMalloc(&b->dtk, sizeof(double) * (size_t)b->R * cap));
    PB_HIP(b, hipMalloc(&b->dtd, sizeof(double) * (size_t)b->R * cap));
    b->table_cap = cap;
    drop_graphs(b);  // kernel arguments changed
  }
  PB_HIP(b, hipStreamSynchronize(b->stream));
  for (int r = 0; r < b->R; r++) {
    PB_HIP(b, hipMemcpyAsync(b->dtk + (size_t)r * b->table_cap, dtk + (size_t)r * nsteps,
                             sizeof(double) * nsteps, hipMemcpyHostToDevice, b->stream));
    PB_HIP(b, hipMemcpyAsync(b->dtd + (size_t)r * b->table_cap, dtd + (size_t)r * nsteps,
                             sizeof(double) * nsteps, hipMemcpyHostToDevice, b->stream));
  }
  PB_HIP(b, hipStreamSynchronize(b->stream));
  return PN_OK;
}

// enqueues nsteps steps that read the uploaded tables
static int run_steps(pn_batch *b, int nsteps, int use_graph) {
  int s = 0;
  while (s < nsteps) {
    hipLaunchKernelGGL(k_set_int, dim3(1), dim3(1), 0, b->stream, b->dbase, s);
    if (use_graph && nsteps - s >= graph_len(b)) {
      int rc = ensure_graph(b, b->cur, b->pcur);
      if (rc != PN_OK) return rc;
      PB_HIP(b, hipGraphLaunch(b->gexec[b->cur][b->pcur], b->stream));
      s += graph_len(b);  // even number of launches: parity and Pi buffer unchanged
    } else {
      const int n = nsteps - s;
      int j = 0;
      if (b->step3)
        for (; j + 2 < n; j += 3) {
          launch_batch_step3(b, b->cur, b->pcur, j);
          b->cur ^= 1;
          b->pcur ^= 1;
        }
      if (b->step2)
        for (; j + 1 < n; j += 2) {
          launch_batch_step2(b, b->cur, b->pcur, j);
          b->cur ^= 1;
          b->pcur ^= 1;
        }
      for (; j < n; j++) {
        launch_batch_step(b, b->cur, b->pcur, j);
        b->cur ^= 1;
      }
      s += n;
    }
    PB_HIP(b, hipGetLastError());
  }
  sync_views(b);
  return PN_OK;
}

// dtk/dtd: [R][nsteps] host tables (per-realization time replay)
extern "C" int pn_batch_evolve(pn_batch *b, int nsteps, const double *dtk, const double *dtd,
                               int use_graph) {
  if (!b || nsteps < 0 || (nsteps > 0 && (!dtk || !dtd)))
    return bfail(b, PN_ERR_ARG, "pn_batch_evolve: bad argument");
  if (nsteps == 0) return PN_OK;
  PB_HIP(b, hipSetDevice(b->device));
  int rc = upload_tables(b, nsteps, dtk, dtd);
  if (rc != PN_OK) return rc;
  return run_steps(b, nsteps, use_graph);
}

extern "C" int pn_batch_sync(pn_batch *b) {
  if (!b) return bfail(b, PN_ERR_ARG, "pn_batch_sync: NULL batch");
  PB_HIP(b, hipStreamSynchronize(b->stream));
  return PN_OK;
}

extern "C" int pn_batch_evolve_timed(pn_batch *b, int nsteps, const double *dtk, const double *dtd,
                                     int use_graph, double *ms) {
  if (!b || !ms) return bfail(b, PN_ERR_ARG, "pn_batch_evolve_timed: NULL argument");
  if (nsteps < 1 || !dtk || !dtd) return bfail(b, PN_ERR_ARG, "pn_batch_evolve_timed: bad argument");
  PB_HIP(b, hipSetDevice(b->device));
  // tables are uploaded (and graphs captured) before the timed region
  int rc = upload_tables(b, nsteps, dtk, dtd);
  if (rc != PN_OK) return rc;
  if (use_graph && nsteps >= graph_len(b) && (rc = ensure_graph(b, b->cur, b->pcur)) != PN_OK) return rc;
  hipEvent_t e0, e1;
  PB_HIP(b, hipEventCreate(&e0));
  PB_HIP(b, hipEventCreate(&e1));
  PB_HIP(b, hipStreamSynchronize(b->stream));
  PB_HIP(b, hipEventRecord(e0, b->stream));
  rc = run_steps(b, nsteps, use_graph);
  if (rc == PN_OK) {
    PB_HIP(b, hipEventRecord(e1, b->stream));
    PB_HIP(b, hipEventSynchronize(e1));
    float f = 0;
    PB_HIP(b, hipEventElapsedTime(&f, e0, e1));
    *ms = f;
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  return rc;
}

// Autotuner: for every size class, time each candidate tile shape on that
// class's realizations only, keep the fastest, then restore the batch.
extern "C" int pn_batch_autotune(pn_batch *b, int trial_steps, int *out, int *nclasses) {
  if (!b || trial_steps < 1) return bfail(b, PN_ERR_ARG, "pn_batch_autotune: bad argument");
  PB_HIP(b, hipSetDevice(b->device));
  PB_HIP(b, hipStreamSynchronize(b->stream));
  const size_t bytes = (size_t)b->total * b->esize;
  void *save_phi = nullptr, *save_pi = nullptr;
  PB_HIP(b, hipMalloc(&save_phi, bytes));
  hipError_t ea = hipMalloc(&save_pi, bytes);
  if (ea != hipSuccess) {
    (void)hipFree(save_phi);
    return bfail(b, PN_ERR_NOMEM, "pn_batch_autotune: no room for the state backup");
  }
  const int saved_cur = b->cur, saved_pcur = b->pcur;
  PB_HIP(b, hipMemcpyAsync(save_phi, b->phi[saved_cur], bytes, hipMemcpyDeviceToDevice, b->stream));
  PB_HIP(b, hipMemcpyAsync(save_pi, b->pib[saved_pcur], bytes, hipMemcpyDeviceToDevice, b->stream));
  trial_steps = (trial_steps + 5) / 6 * 6;  // whole two- and three-step launches
  std::vector<double> tk((size_t)b->R * trial_steps, 1e-3), td((size_t)b->R * trial_steps, 1e-3);
  std::vector<int> classes;
  for (auto &q : b->re)
    if (std::find(classes.begin(), classes.end(), q.N) == classes.end()) classes.push_back(q.N);
  std::sort(classes.begin(), classes.end());
  const int rows_c[] = {2, 4, 8, 16};
  const int zb_c[] = {8, 16, 32, 64};
  int rc = PN_OK;
  // tunes every size class in the current mode; returns the sum over the
  // classes of their fastest time (the batch's time per trial) and their choices
  auto tune_mode = [&](std::vector<int> &results, double &total) {
    results.clear();
    total = 0;
    for (int N : classes) {
      // every class is timed on the caller's fields, not on what the previous
      // class's trials left in the buffers
      b->cur = saved_cur;
      b->pcur = saved_pcur;
      hipError_t er = hipMemcpyAsync(b->phi[saved_cur], save_phi, bytes, hipMemcpyDeviceToDevice, b->stream);
      if (er == hipSuccess)
        er = hipMemcpyAsync(b->pib[saved_pcur], save_pi, bytes, hipMemcpyDeviceToDevice, b->stream);
      if (er != hipSuccess) {
        rc = bfail(b, PN_ERR_HIP, "pn_batch_autotune: ", hipGetErrorString(er));
        return;
      }
      std::vector<int> mask(b->R, 0);
      for (int r = 0; r < b->R; r++) mask[r] = b->re[r].N == N;
      double best = 1e30;
      std::pair<int, int> bestp = b->step3   ? std::make_pair(step3_rows(b, N), 32)
                                  : b->step2 ? std::make_pair(step2_rows(b, N), 16)
                                             : default_tiles(N);
      // multi-step modes tune the planes per tile (their rows follow from N);
      // single steps tune rows and planes
      const bool multi = b->step3 || b->step2;
      const int nrows = multi ? 1 : 4;
      for (int ri = 0; ri < nrows; ri++)
        for (int zb : zb_c) {
          const int ty = b->step3 ? step3_rows(b, N) : b->step2 ? step2_rows(b, N) : rows_c[ri];
          if ((!multi && ty > N) || zb > N) continue;
          if (b->step3) b->zb3[N] = zb;
          else if (b->step2) b->zb2[N] = zb;
          else b->tiles[N] = {ty, zb};
          if ((rc = upload_desc(b, &mask)) != PN_OK) return;
          double ms = 0;
          if ((rc = pn_batch_evolve(b, trial_steps, tk.data(), td.data(), 0)) != PN_OK) return;  // warm
          if ((rc = pn_batch_evolve_timed(b, trial_steps, tk.data(), td.data(), 0, &ms)) != PN_OK) return;
          if (ms < best) {
            best = ms;
            bestp = {ty, zb};
          }
        }
      if (b->step3) b->zb3[N] = bestp.second;
      else if (b->step2) b->zb2[N] = bestp.second;
      else b->tiles[N] = bestp;
      total += best;
      results.push_back(N);
      results.push_back(bestp.first);
      results.push_back(bestp.second);
    }
  };
  std::vector<int> results;
  double total = 0;
  if (b->step2 && b->step3_ok) {
    // both multi-step modes: the faster one stays on (three steps wins ties within 2 %)
    std::vector<int> res3;
    double t3 = 0;
    b->step3 = 1;
    tune_mode(res3, t3);
    if (rc == PN_OK) {
      b->step3 = 0;
      tune_mode(results, total);
    }
    if (rc == PN_OK && t3 <= total * 1.02) {
      b->step3 = 1;
      results = res3;
    }
  } else {
    tune_mode(results, total);
  }
  // restore the state and the full tile list
  hipError_t e = hipStreamSynchronize(b->stream);
  b->cur = saved_cur;
  b->pcur = saved_pcur;
  if (e == hipSuccess) e = hipMemcpyAsync(b->phi[saved_cur], save_phi, bytes, hipMemcpyDeviceToDevice, b->stream);
  if (e == hipSuccess) e = hipMemcpyAsync(b->pib[saved_pcur], save_pi, bytes, hipMemcpyDeviceToDevice, b->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(b->stream);
  (void)hipFree(save_phi);
  (void)hipFree(save_pi);
  drop_graphs(b);
  sync_views(b);
  if (rc != PN_OK) return rc;
  if (e != hipSuccess) return bfail(b, PN_ERR_HIP, "pn_batch_autotune: restore: ", hipGetErrorString(e));
  if ((rc = upload_desc(b)) != PN_OK) return rc;
  if (nclasses) *nclasses = (int)classes.size();
  if (out) std::copy(results.begin(), results.end(), out);
  return PN_OK;
}

extern "C" int pn_batch_set_option(pn_batch *b, int option, int value) {
  if (!b) return bfail(b, PN_ERR_ARG, "pn_batch_set_option: NULL batch");
  if (option == PN_OPT_THREE_STEP) {
    if (value > 0 && !(b->step3_ok && b->step2))
      return bfail(b, PN_ERR_STATE, "pn_batch_set_option: a realization does not fit a three-step tile, "
                                    "or two-step mode is off");
    if ((value > 0) != (b->step3 != 0)) {
      PB_HIP(b, hipSetDevice(b->device));
      PB_HIP(b, hipStreamSynchronize(b->stream));
      b->step3 = value > 0;
      drop_graphs(b);
    }
    return PN_OK;
  }
  if (option != PN_OPT_TWO_STEP) return bfail(b, PN_ERR_ARG, "pn_batch_set_option: unknown option");
  if (value && !b->step2_ok)
    return bfail(b, PN_ERR_STATE, "pn_batch_set_option: a realization does not fit a two-step tile "
                                  "(N % vector width, or N/width above the block size)");
  if ((value != 0) != (b->step2 != 0)) {
    PB_HIP(b, hipSetDevice(b->device));
    PB_HIP(b, hipStreamSynchronize(b->stream));
    b->step2 = value != 0;
    if (!b->step2) b->step3 = 0;  // three-step mode uses the two-step buffers
    drop_graphs(b);
  }
  return PN_OK;
}

extern "C" int pn_batch_get_option(pn_batch *b, int option, int *value) {
  if (!b || !value) return bfail(b, PN_ERR_ARG, "pn_batch_get_option: NULL argument");
  if (option == PN_OPT_THREE_STEP) {
    *value = b->step3;
    return PN_OK;
  }
  if (option != PN_OPT_TWO_STEP) return bfail(b, PN_ERR_ARG, "pn_batch_get_option: unknown option");
  *value = b->step2;
  return PN_OK;
}
