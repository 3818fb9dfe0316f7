// pn_islands.hip -- island catalog as GPU connected-component labelling.
//
// Replaces the queue-based flood fill of the reference (BFS, islands.cpp:90-279,
// driven by the raster scan of Full_Analysis, :827-835) with a lock-free
// union-find over the thresholded lattice:
//   1. mask + init : label[i] = i for island sites, -1 otherwise (is_island,
//                    islands.cpp:72-87)
//   2. merge       : every island site unites with its x-1, y-1 (periodic) and
//                    z-1 (open: the reference never links plane 1 with plane N
//                    at one rank, islands.cpp:224-236) neighbours with atomicMin
//                    on roots, so every root ends up being the smallest index
//                    of its component
//   3. flatten     : label[i] = root
//   4. numbering   : roots in increasing index order get 0,1,2,... -- exactly
//                    the order in which the raster scan meets the islands
//   5. statistics  : over each island's sites in raster order (a stable sort
//                    groups them): sizes and offset sums exactly, the peak as
//                    the largest key (phi^2 or Delta) with ties broken towards
//                    the smallest index, and the FP64 sums (phi, potential,
//                    kinetic and gradient energy) compensated and merged in a
//                    fixed order: the same bits on every call.  One wave per
//                    island up to kBig sites; larger islands in chunks of kBig
//                    sites, one atomic per chunk for the integers and an
//                    ordered merge of the chunks' partial sums.
#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <cmath>
#include <new>
#include <vector>

#include "pn_internal.h"
#include "pn_kernels.h"

struct IslCoef {
  double a2, cm, lmode, d6, thr, thr2h;
  int mode;
};

struct pn_island_state {
  int *label = nullptr;      // N^3
  int *cid = nullptr;        // N^3: island number stored at root positions
  int *blk = nullptr;        // per-block root counts / offsets
  long long nblk = 0;
  int cap = 0;               // allocated catalog entries
  unsigned long long *ikey = nullptr;  // per island: max key (ordered bits)
  long long *iint = nullptr;           // per island: nn, first, peak, ties, sdx, sdy, sdz
  double *idbl = nullptr;              // per island: sum phi, pot, kin, grad, peak value
  int *count = nullptr;
  // the island sites grouped by island for the FP sums: raster-order
  // compaction, then a stable sort by island number
  int *mblk = nullptr;                 // per-block island-site counts / offsets
  long long mcap = 0;                  // allocated island-site entries
  int *mkey = nullptr, *mkey2 = nullptr, *msite = nullptr, *msite2 = nullptr;
  int *ist = nullptr;                  // per island: first entry in the sorted sites (+ total)
  void *sort_tmp = nullptr;
  size_t sort_bytes = 0;
  // islands above kBig sites, in chunks of kBig sorted entries
  int *counters = nullptr;             // chunks, large islands
  long long ccap = 0;                  // allocated chunks (large islands: half as many)
  int *chunk_isl = nullptr, *chunk_rank = nullptr, *big_isl = nullptr, *big_c0 = nullptr;
  double *cpart = nullptr;             // per chunk: (hi, lo) of the four sums
  // catalog rows assembled on the device, and the label array handed to the
  // caller, kept across calls (grown on demand)
  pn_island *rows = nullptr;
  long long rcap = 0;
  int *dlab = nullptr;
  // pn_islands_lookup's positions and answers on the device and their
  // page-locked host staging (grown on demand)
  long long *lpos = nullptr, *hpos = nullptr;
  int *lout = nullptr, *hout = nullptr;
  long long lcap = 0;
  // one island's keys and peak values for the breadth-first order of tied peaks
  double *kv = nullptr;
  long long kvcap = 0;
  // the last call: its coefficients, its island count and the starts of the
  // islands among the sorted sites (host copy, read on demand)
  IslCoef last{};
  int nisl = 0;
  int lgN = -1;
};

void pn_free_islands(pn_ctx *c) {
  pn_island_state *s = c->isl;
  if (!s) return;
  hipFree(s->label);
  hipFree(s->cid);
  hipFree(s->blk);
  hipFree(s->ikey);
  hipFree(s->iint);
  hipFree(s->idbl);
  hipFree(s->count);
  hipFree(s->mblk);
  hipFree(s->mkey);
  hipFree(s->mkey2);
  hipFree(s->msite);
  hipFree(s->msite2);
  hipFree(s->ist);
  hipFree(s->sort_tmp);
  hipFree(s->counters);
  hipFree(s->chunk_isl);
  hipFree(s->chunk_rank);
  hipFree(s->big_isl);
  hipFree(s->big_c0);
  hipFree(s->cpart);
  hipFree(s->rows);
  hipFree(s->dlab);
  hipFree(s->lpos);
  hipFree(s->lout);
  hipHostFree(s->hpos);
  hipHostFree(s->hout);
  hipFree(s->kv);
  delete s;
  c->isl = nullptr;
}

static constexpr int kT = 256;          // threads per block
static constexpr int kChunk = kT * 16;  // sites per block in the numbering pass
static constexpr int kNI = 7;           // integer stats per island
static constexpr int kND = 5;           // double stats per island (+ peak value)


// site energy density Delta (fieldanalysis.cpp:447,514-515)
template <typename T>
__device__ __forceinline__ double site_delta(const T *phi, const T *pi, long long i, int x, int y,
                                             int z, int N, const IslCoef &k) {
  const long long P = (long long)N * N;
  const double f = (double)phi[i], p = (double)pi[i];
  const double dx = (double)phi[i - x + (x + 1 == N ? 0 : x + 1)] - f;
  const double dy = (double)phi[i + (long long)((y + 1 == N ? 0 : y + 1) - y) * N] - f;
  const double dz = (double)phi[i + (long long)((z + 1 == N ? 0 : z + 1) - z) * P] - f;
  double d = (0.5 * (p * p) + (dx * dx) / k.a2 + (dy * dy) / k.a2 + (dz * dz) / k.a2);
  d += k.cm * (f * f);
  const double q2 = (f * f) * (f * f);
  d += k.lmode * (q2 / 24.0) + (q2 * (f * f) / k.d6);
  return d;
}

template <typename T>
__device__ __forceinline__ bool site_mask(const T *phi, const T *pi, long long i, int x, int y,
                                          int z, int N, const IslCoef &k, double *key) {
  if (k.mode == 0) {
    const double f = (double)phi[i], p = (double)pi[i];
    if (key) *key = f * f;
    return (f * f * 0.5 + p * p * 0.5 > k.thr2h);  // islands.cpp:76
  }
  const double d = site_delta(phi, pi, i, x, y, z, N, k);
  if (key) *key = d;
  return d > k.thr;  // islands.cpp:83
}

// x, y, z of site i (< 2^31): shifts when N is a power of two (lgN >= 0)
__device__ __forceinline__ void site_xyz(int i, int N, int lgN, int &x, int &y, int &z) {
  if (lgN >= 0) {
    x = i & (N - 1);
    y = (i >> lgN) & (N - 1);
    z = i >> (2 * lgN);
  } else {
    const int r = i / N;
    x = i - r * N;
    y = r % N;
    z = r / N;
  }
}

// mode 0 reads only the site itself; mode 1 needs its position (32-bit, and
// shifts for power-of-two N)
template <typename T>
__global__ __launch_bounds__(kT) void k_isl_init(const T *__restrict__ phi, const T *__restrict__ pi,
                                                 int N, int lgN, IslCoef k, int *__restrict__ label) {
  const int n3 = N * N * N;  // below 2^31 (pn_islands checks)
  for (int i = (int)blockIdx.x * kT + (int)threadIdx.x; i < n3; i += (int)gridDim.x * kT) {
    bool in;
    if (k.mode == 0) {
      in = site_mask(phi, pi, i, 0, 0, 0, N, k, nullptr);
    } else {
      int x, y, z;
      site_xyz(i, N, lgN, x, y, z);
      in = site_mask(phi, pi, i, x, y, z, N, k, nullptr);
    }
    label[i] = in ? i : -1;
  }
}

// parent reads bypass the non-coherent per-CU cache so that roots written
// by other CUs' atomics are seen; parents only ever decrease, so any value
// read is an ancestor and the walk terminates
__device__ __forceinline__ int uf_load(const int *label, int x) {
  return __hip_atomic_load(const_cast<int *>(label) + x, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ __forceinline__ int uf_find(const int *label, int x) {
  int p = uf_load(label, x);
  while (p != x) {
    x = p;
    p = uf_load(label, x);
  }
  return x;
}

__device__ __forceinline__ void uf_union(int *label, int a, int b) {
  bool done;
  do {
    a = uf_find(label, a);
    b = uf_find(label, b);
    if (a < b) {
      const int old = atomicMin(&label[b], a);
      done = (old == b);
      b = old;
    } else if (b < a) {
      const int old = atomicMin(&label[a], b);
      done = (old == a);
      a = old;
    } else {
      done = true;
    }
  } while (!done);
}

__global__ __launch_bounds__(kT) void k_isl_merge(int *label, int N) {
  const long long n3 = (long long)N * N * N, P = (long long)N * N;
  for (long long i = (long long)blockIdx.x * kT + threadIdx.x; i < n3;
       i += (long long)gridDim.x * kT) {
    if (label[i] < 0) continue;
    const int x = (int)(i % N), y = (int)((i / N) % N), z = (int)(i / P);
    const long long jx = i - x + (x == 0 ? N - 1 : x - 1);
    const long long jy = i + (long long)((y == 0 ? N - 1 : y - 1) - y) * N;
    if (label[jx] >= 0) uf_union(label, (int)i, (int)jx);
    if (label[jy] >= 0) uf_union(label, (int)i, (int)jy);
    if (z > 0 && label[i - P] >= 0) uf_union(label, (int)i, (int)(i - P));
  }
}

__global__ __launch_bounds__(kT) void k_isl_flatten(int *__restrict__ label, long long n3) {
  for (long long i = (long long)blockIdx.x * kT + threadIdx.x; i < n3;
       i += (long long)gridDim.x * kT) {
    if (label[i] >= 0) label[i] = uf_find(label, (int)i);
  }
}

// numbering: count roots per chunk, scan chunk counts, assign in order
__global__ __launch_bounds__(kT) void k_isl_count(const int *__restrict__ label, long long n3,
                                                  int *__restrict__ blk) {
  __shared__ int s[kT];
  const long long base = (long long)blockIdx.x * kChunk;
  int cnt = 0;
  for (int j = threadIdx.x; j < kChunk; j += kT) {
    const long long i = base + j;
    if (i < n3 && label[i] == (int)i) cnt++;
  }
  s[threadIdx.x] = cnt;
  __syncthreads();
  for (int o = kT / 2; o > 0; o >>= 1) {
    if ((int)threadIdx.x < o) s[threadIdx.x] += s[threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x == 0) blk[blockIdx.x] = s[0];
}

// exclusive scan of nb block counts by one block; total in blk[nb]
__global__ __launch_bounds__(1024) void k_isl_scan(int *__restrict__ blk, long long nb) {
  __shared__ int s[1024];
  __shared__ int carry;
  if (threadIdx.x == 0) carry = 0;
  __syncthreads();
  for (long long base = 0; base < nb; base += 1024) {
    const long long i = base + threadIdx.x;
    const int v = i < nb ? blk[i] : 0;
    s[threadIdx.x] = v;
    __syncthreads();
    for (int o = 1; o < 1024; o <<= 1) {
      const int t = (int)threadIdx.x >= o ? s[threadIdx.x - o] : 0;
      __syncthreads();
      s[threadIdx.x] += t;
      __syncthreads();
    }
    if (i < nb) blk[i] = carry + s[threadIdx.x] - v;
    __syncthreads();
    if (threadIdx.x == 1023) carry += s[1023];
    __syncthreads();
  }
  if (threadIdx.x == 0) blk[nb] = carry;
}

__global__ __launch_bounds__(kT) void k_isl_assign(const int *__restrict__ label, long long n3,
                                                   const int *__restrict__ blk,
                                                   int *__restrict__ cid) {
  __shared__ int s[kT];
  __shared__ int carry;
  const long long base = (long long)blockIdx.x * kChunk;
  if (threadIdx.x == 0) carry = blk[blockIdx.x];
  __syncthreads();
  for (int j0 = 0; j0 < kChunk; j0 += kT) {
    const long long i = base + j0 + threadIdx.x;
    const int r = (i < n3 && label[i] == (int)i) ? 1 : 0;
    s[threadIdx.x] = r;
    __syncthreads();
    for (int o = 1; o < kT; o <<= 1) {
      const int t = (int)threadIdx.x >= o ? s[threadIdx.x - o] : 0;
      __syncthreads();
      s[threadIdx.x] += t;
      __syncthreads();
    }
    if (r) cid[i] = carry + s[threadIdx.x] - 1;
    __syncthreads();
    if (threadIdx.x == kT - 1) carry += s[kT - 1];
    __syncthreads();
  }
}

// order-preserving map of a double onto unsigned 64-bit integers
__device__ __forceinline__ unsigned long long ord_key(double v) {
  const unsigned long long b = (unsigned long long)__double_as_longlong(v);
  return (b >> 63) ? ~b : (b | 0x8000000000000000ULL);
}

// island sites per chunk, for the raster-order compaction below
__global__ __launch_bounds__(kT) void k_isl_mcount(const int *__restrict__ label, long long n3,
                                                   int *__restrict__ mblk) {
  __shared__ int s[kT];
  const long long base = (long long)blockIdx.x * kChunk;
  int cnt = 0;
  for (int j = threadIdx.x; j < kChunk; j += kT) {
    const long long i = base + j;
    if (i < n3 && label[i] >= 0) cnt++;
  }
  s[threadIdx.x] = cnt;
  __syncthreads();
  for (int o = kT / 2; o > 0; o >>= 1) {
    if ((int)threadIdx.x < o) s[threadIdx.x] += s[threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x == 0) mblk[blockIdx.x] = s[0];
}

// island sites in raster order, each with its island number
__global__ __launch_bounds__(kT) void k_isl_mscatter(const int *__restrict__ label, long long n3,
                                                     const int *__restrict__ mblk,
                                                     const int *__restrict__ cid,
                                                     int *__restrict__ mkey, int *__restrict__ msite) {
  __shared__ int s[kT];
  __shared__ int carry;
  const long long base = (long long)blockIdx.x * kChunk;
  if (threadIdx.x == 0) carry = mblk[blockIdx.x];
  __syncthreads();
  for (int j0 = 0; j0 < kChunk; j0 += kT) {
    const long long i = base + j0 + threadIdx.x;
    const int r = i < n3 ? label[i] : -1;
    const int f = r >= 0 ? 1 : 0;
    s[threadIdx.x] = f;
    __syncthreads();
    for (int o = 1; o < kT; o <<= 1) {
      const int t = (int)threadIdx.x >= o ? s[threadIdx.x - o] : 0;
      __syncthreads();
      s[threadIdx.x] += t;
      __syncthreads();
    }
    if (f) {
      const int pos = carry + s[threadIdx.x] - 1;
      mkey[pos] = cid[r];
      msite[pos] = (int)i;
    }
    __syncthreads();
    if (threadIdx.x == kT - 1) carry += s[kT - 1];
    __syncthreads();
  }
}

// first entry of each island among the sorted sites; ist[nisl] = total
__global__ void k_isl_starts(const int *__restrict__ key, int m, int nisl, int *__restrict__ ist) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= m) return;
  if (p == 0 || key[p - 1] != key[p]) ist[key[p]] = p;
  if (p == m - 1) ist[nisl] = m;
}

// the peak key of a site: phi^2 (mode 0) or Delta (mode 1)
template <typename T>
__device__ __forceinline__ double isl_key(const T *phi, const T *pi, int i, int N, int lgN,
                                          const IslCoef &k) {
  if (k.mode == 0) {
    const double f = (double)phi[i];
    return f * f;
  }
  int x, y, z;
  site_xyz(i, N, lgN, x, y, z);
  return site_delta(phi, pi, (long long)i, x, y, z, N, k);
}

// the summands of one site (islands.cpp:131-142): phi, potential, kinetic and
// gradient energy density
template <typename T>
__device__ __forceinline__ void isl_terms(const T *phi, const T *pi, int i, int N, int lgN,
                                          const IslCoef &k, double t[4]) {
  int x, y, z;
  site_xyz(i, N, lgN, x, y, z);
  const long long P = (long long)N * N, ii = i;
  const double f = (double)phi[ii], p = (double)pi[ii];
  const double fx = (double)phi[ii - x + (x + 1 == N ? 0 : x + 1)];
  const double fy = (double)phi[ii + (long long)((y + 1 == N ? 0 : y + 1) - y) * N];
  const double fz = (double)phi[ii + (long long)((z + 1 == N ? 0 : z + 1) - z) * P];
  t[3] = 0.5 * ((fx - f) * (fx - f) + (fy - f) * (fy - f) + (fz - f) * (fz - f)) / k.a2;
  t[2] = 0.5 * (p * p);
  const double q2 = (f * f) * (f * f);
  t[1] = k.cm * (f * f) + k.lmode * (q2 / 24.0) + (q2 * (f * f) / k.d6);
  t[0] = f;
}

// offset of site i from the peak (px, py, pz): x and y minimum image, z plain
__device__ __forceinline__ void isl_offset(int i, int N, int lgN, int px, int py, int pz,
                                           long long &sx, long long &sy, long long &sz) {
  int x, y, z;
  site_xyz(i, N, lgN, x, y, z);
  int dx = x - px, dy = y - py;
  if (dx >= N - N / 2) dx -= N;
  if (dx < -(N / 2)) dx += N;
  if (dy >= N - N / 2) dy -= N;
  if (dy < -(N / 2)) dy += N;
  sx += dx;
  sy += dy;
  sz += z - pz;
}

// a wave's largest key, its count and its first entry (entries visited in
// increasing order per lane; kc == 0: the lane saw no entry): merged over the
// lanes
__device__ __forceinline__ void wave_peak(unsigned long long &km, int &kc, int &kp) {
  for (int off = 32; off > 0; off >>= 1) {
    const unsigned long long k2 = __shfl_down(km, off, 64);
    const int c2 = __shfl_down(kc, off, 64), p2 = __shfl_down(kp, off, 64);
    if (c2 > 0 && (kc == 0 || k2 > km)) {
      km = k2; kc = c2; kp = p2;
    } else if (c2 > 0 && k2 == km) {
      kc += c2;
      kp = min(kp, p2);
    }
  }
}

__device__ __forceinline__ long long wave_sum_ll(long long v) {
  for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
  return v;
}

__device__ __forceinline__ void wave_sum4(double hi[4], double lo[4]) {
  for (int off = 32; off > 0; off >>= 1)
#pragma unroll
    for (int j = 0; j < 4; j++) {
      const double h2 = __shfl_down(hi[j], off, 64), l2 = __shfl_down(lo[j], off, 64);
      two_merge(hi[j], lo[j], h2, l2);
    }
}

// Islands of up to kBig sites, one wave each, over their sites in raster
// order (the stable sort): size, first site, peak (largest key, smallest
// index among equal keys) and its tie count, the FP64 sums in (hi, lo) form
// merged in a fixed tree, and the offset sums from the peak.  No atomics.
static constexpr int kBig = 4096;

template <typename T>
__global__ __launch_bounds__(kT) void k_isl_small(const T *__restrict__ phi, const T *__restrict__ pi,
                                                  int N, int lgN, IslCoef k, const int *__restrict__ ist,
                                                  const int *__restrict__ site, int nisl,
                                                  unsigned long long *__restrict__ ikey,
                                                  long long *__restrict__ iint,
                                                  double *__restrict__ idbl) {
  const int lane = (int)(threadIdx.x & 63), wpb = (int)(blockDim.x >> 6);
  for (int id = (int)blockIdx.x * wpb + (int)(threadIdx.x >> 6); id < nisl; id += (int)gridDim.x * wpb) {
    const int e0 = ist[id], e1 = ist[id + 1];
    if (e1 - e0 > kBig) continue;  // chunked: k_isl_big*
    unsigned long long km = 0;
    int kc = 0, kp = 0x7fffffff;
    double hi[4] = {0.0, 0.0, 0.0, 0.0}, lo[4] = {0.0, 0.0, 0.0, 0.0};
    for (int e = e0 + lane; e < e1; e += 64) {
      const int i = site[e];
      const unsigned long long kk = ord_key(isl_key(phi, pi, i, N, lgN, k));
      if (kc == 0 || kk > km) {
        km = kk; kc = 1; kp = e;
      } else if (kk == km) {
        kc++;
      }
      double t[4];
      isl_terms(phi, pi, i, N, lgN, k, t);
#pragma unroll
      for (int j = 0; j < 4; j++) two_acc(hi[j], lo[j], t[j]);
    }
    wave_peak(km, kc, kp);
    wave_sum4(hi, lo);
    const int pk = site[__shfl(kp, 0, 64)];
    int px, py, pz;
    site_xyz(pk, N, lgN, px, py, pz);
    long long sx = 0, sy = 0, sz = 0;
    for (int e = e0 + lane; e < e1; e += 64) isl_offset(site[e], N, lgN, px, py, pz, sx, sy, sz);
    sx = wave_sum_ll(sx);
    sy = wave_sum_ll(sy);
    sz = wave_sum_ll(sz);
    if (lane == 0) {
      long long *q = iint + (long long)id * kNI;
      q[0] = e1 - e0;
      q[1] = site[e0];
      q[2] = pk;
      q[3] = kc;
      q[4] = sx;
      q[5] = sy;
      q[6] = sz;
      ikey[id] = km;
#pragma unroll
      for (int j = 0; j < 4; j++) idbl[(long long)id * kND + j] = hi[j] + lo[j];
    }
  }
}

// Islands above kBig sites, in chunks of kBig consecutive sorted entries
// (counters *nchunk, *nbig; the chunks of an island are consecutive).
// plan: one thread per island reserves its chunks
__global__ void k_isl_bigplan(const int *__restrict__ ist, int nisl, int *__restrict__ counters,
                              int *__restrict__ chunk_isl, int *__restrict__ chunk_rank,
                              int *__restrict__ big_isl, int *__restrict__ big_c0) {
  const int id = blockIdx.x * blockDim.x + threadIdx.x;
  if (id >= nisl) return;
  const int n = ist[id + 1] - ist[id];
  if (n <= kBig) return;
  const int nch = (n + kBig - 1) / kBig;
  const int c0 = atomicAdd(&counters[0], nch);
  const int b = atomicAdd(&counters[1], 1);
  for (int c = 0; c < nch; c++) {
    chunk_isl[c0 + c] = id;
    chunk_rank[c0 + c] = c;
  }
  big_isl[b] = id;
  big_c0[b] = c0;
}

// pass 1 per chunk: its largest key (one atomicMax per chunk) and its FP64
// partial sums (merged in chunk order by k_isl_bigmerge)
template <typename T>
__global__ __launch_bounds__(kT) void k_isl_big1(const T *__restrict__ phi, const T *__restrict__ pi,
                                                 int N, int lgN, IslCoef k, const int *__restrict__ ist,
                                                 const int *__restrict__ site,
                                                 const int *__restrict__ counters,
                                                 const int *__restrict__ chunk_isl,
                                                 const int *__restrict__ chunk_rank,
                                                 unsigned long long *__restrict__ ikey,
                                                 double *__restrict__ cpart) {
  const int lane = (int)(threadIdx.x & 63), wpb = (int)(blockDim.x >> 6);
  const int nc = counters[0];
  for (int c = (int)blockIdx.x * wpb + (int)(threadIdx.x >> 6); c < nc; c += (int)gridDim.x * wpb) {
    const int id = chunk_isl[c];
    const int e0 = ist[id] + chunk_rank[c] * kBig, e1 = min(e0 + kBig, ist[id + 1]);
    unsigned long long km = 0;
    double hi[4] = {0.0, 0.0, 0.0, 0.0}, lo[4] = {0.0, 0.0, 0.0, 0.0};
    for (int e = e0 + lane; e < e1; e += 64) {
      const int i = site[e];
      const unsigned long long kk = ord_key(isl_key(phi, pi, i, N, lgN, k));
      km = kk > km ? kk : km;
      double t[4];
      isl_terms(phi, pi, i, N, lgN, k, t);
#pragma unroll
      for (int j = 0; j < 4; j++) two_acc(hi[j], lo[j], t[j]);
    }
    for (int off = 32; off > 0; off >>= 1) {
      const unsigned long long k2 = __shfl_down(km, off, 64);
      km = k2 > km ? k2 : km;
    }
    wave_sum4(hi, lo);
    if (lane == 0) {
      atomicMax(&ikey[id], km);
#pragma unroll
      for (int j = 0; j < 4; j++) {
        cpart[(long long)c * 8 + 2 * j] = hi[j];
        cpart[(long long)c * 8 + 2 * j + 1] = lo[j];
      }
    }
  }
}

// pass 2 per chunk: entries holding the island's largest key: their count
// and the first of them (one atomicAdd and one atomicMin per chunk)
template <typename T>
__global__ __launch_bounds__(kT) void k_isl_big2(const T *__restrict__ phi, const T *__restrict__ pi,
                                                 int N, int lgN, IslCoef k, const int *__restrict__ ist,
                                                 const int *__restrict__ site,
                                                 const int *__restrict__ counters,
                                                 const int *__restrict__ chunk_isl,
                                                 const int *__restrict__ chunk_rank,
                                                 const unsigned long long *__restrict__ ikey,
                                                 long long *__restrict__ iint) {
  const int lane = (int)(threadIdx.x & 63), wpb = (int)(blockDim.x >> 6);
  const int nc = counters[0];
  for (int c = (int)blockIdx.x * wpb + (int)(threadIdx.x >> 6); c < nc; c += (int)gridDim.x * wpb) {
    const int id = chunk_isl[c];
    const int e0 = ist[id] + chunk_rank[c] * kBig, e1 = min(e0 + kBig, ist[id + 1]);
    const unsigned long long km = ikey[id];
    int kc = 0, kp = 0x7fffffff;
    for (int e = e0 + lane; e < e1; e += 64) {
      if (ord_key(isl_key(phi, pi, site[e], N, lgN, k)) == km) {
        kc++;
        kp = min(kp, e);
      }
    }
    for (int off = 32; off > 0; off >>= 1) {
      kc += __shfl_down(kc, off, 64);
      kp = min(kp, __shfl_down(kp, off, 64));
    }
    if (lane == 0 && kc > 0) {
      atomicAdd((unsigned long long *)&iint[(long long)id * kNI + 3], (unsigned long long)kc);
      atomicMin(&iint[(long long)id * kNI + 2], (long long)site[kp]);
    }
  }
}

// pass 3 per chunk: offset sums from the peak (one atomicAdd per sum and chunk)
__global__ __launch_bounds__(kT) void k_isl_big3(int N, int lgN, const int *__restrict__ ist,
                                                 const int *__restrict__ site,
                                                 const int *__restrict__ counters,
                                                 const int *__restrict__ chunk_isl,
                                                 const int *__restrict__ chunk_rank,
                                                 long long *__restrict__ iint) {
  const int lane = (int)(threadIdx.x & 63), wpb = (int)(blockDim.x >> 6);
  const int nc = counters[0];
  for (int c = (int)blockIdx.x * wpb + (int)(threadIdx.x >> 6); c < nc; c += (int)gridDim.x * wpb) {
    const int id = chunk_isl[c];
    const int e0 = ist[id] + chunk_rank[c] * kBig, e1 = min(e0 + kBig, ist[id + 1]);
    int px, py, pz;
    site_xyz((int)iint[(long long)id * kNI + 2], N, lgN, px, py, pz);
    long long sx = 0, sy = 0, sz = 0;
    for (int e = e0 + lane; e < e1; e += 64) isl_offset(site[e], N, lgN, px, py, pz, sx, sy, sz);
    sx = wave_sum_ll(sx);
    sy = wave_sum_ll(sy);
    sz = wave_sum_ll(sz);
    if (lane == 0) {
      long long *q = iint + (long long)id * kNI;
      atomicAdd((unsigned long long *)&q[4], (unsigned long long)sx);
      atomicAdd((unsigned long long *)&q[5], (unsigned long long)sy);
      atomicAdd((unsigned long long *)&q[6], (unsigned long long)sz);
    }
  }
}

// per large island (one wave): size, first site, and the FP64 sums merged
// from its chunks' partials in chunk order (lanes take chunks strided, then
// a fixed tree)
__global__ __launch_bounds__(kT) void k_isl_bigmerge(const int *__restrict__ ist,
                                                     const int *__restrict__ site,
                                                     const int *__restrict__ counters,
                                                     const int *__restrict__ big_isl,
                                                     const int *__restrict__ big_c0,
                                                     const double *__restrict__ cpart,
                                                     long long *__restrict__ iint,
                                                     double *__restrict__ idbl) {
  const int lane = (int)(threadIdx.x & 63), wpb = (int)(blockDim.x >> 6);
  const int nb = counters[1];
  for (int b = (int)blockIdx.x * wpb + (int)(threadIdx.x >> 6); b < nb; b += (int)gridDim.x * wpb) {
    const int id = big_isl[b], c0 = big_c0[b];
    const int n = ist[id + 1] - ist[id], nch = (n + kBig - 1) / kBig;
    double hi[4] = {0.0, 0.0, 0.0, 0.0}, lo[4] = {0.0, 0.0, 0.0, 0.0};
    for (int c = lane; c < nch; c += 64)
#pragma unroll
      for (int j = 0; j < 4; j++)
        two_merge(hi[j], lo[j], cpart[(long long)(c0 + c) * 8 + 2 * j], cpart[(long long)(c0 + c) * 8 + 2 * j + 1]);
    wave_sum4(hi, lo);
    if (lane == 0) {
      iint[(long long)id * kNI + 0] = n;
      iint[(long long)id * kNI + 1] = site[ist[id]];
#pragma unroll
      for (int j = 0; j < 4; j++) idbl[(long long)id * kND + j] = hi[j] + lo[j];
    }
  }
}

// the label array the caller asked for: island number per site, -1 elsewhere
__global__ __launch_bounds__(kT) void k_isl_labels(const int *__restrict__ label, long long n3,
                                                   const int *__restrict__ cid,
                                                   int *__restrict__ out) {
  for (long long i = (long long)blockIdx.x * kT + threadIdx.x; i < n3; i += (long long)gridDim.x * kT) {
    const int r = label[i];
    out[i] = r < 0 ? -1 : cid[r];
  }
}
__global__ void k_isl_prep(int n, unsigned long long *__restrict__ ikey, long long *__restrict__ iint,
                           double *__restrict__ idbl) {
  const int id = blockIdx.x * blockDim.x + threadIdx.x;
  if (id >= n) return;
  ikey[id] = 0ULL;
  iint[(long long)id * kNI + 0] = 0;
  iint[(long long)id * kNI + 1] = 0;
  iint[(long long)id * kNI + 2] = 0x7fffffffffffffffLL;
  for (int j = 3; j < kNI; j++) iint[(long long)id * kNI + j] = 0;
  for (int j = 0; j < kND; j++) idbl[(long long)id * kND + j] = 0.0;
}

// peak value: Phi (mode 0, islands.cpp:257) or Delta (mode 1, :260) at the peak
template <typename T>
__global__ void k_isl_peakval(const T *__restrict__ phi, const T *__restrict__ pi, int N, int lgN,
                              IslCoef k, int n, const long long *__restrict__ iint,
                              double *__restrict__ idbl) {
  const int id = blockIdx.x * blockDim.x + threadIdx.x;
  if (id >= n) return;
  const int pk = (int)iint[(long long)id * kNI + 2];
  int x, y, z;
  site_xyz(pk, N, lgN, x, y, z);
  idbl[(long long)id * kND + 4] =
      k.mode == 0 ? (double)phi[pk] : site_delta(phi, pi, (long long)pk, x, y, z, N, k);
}

// centroid of a row from its peak and offset sums: x and y wrapped onto the
// lattice, z plain (the same expressions on the device and on the host)
__host__ __device__ inline void isl_centroid(pn_island &I, int N) {
  const long long P = (long long)N * N;
  const long long pk = I.peak_pos - P;
  const double px = (double)(pk % N), py = (double)((pk / N) % N), pz = (double)(pk / P);
  const double NN = (double)I.nn;
  double cx = px + (double)I.sum_dx / NN, cy = py + (double)I.sum_dy / NN;
  cx = fmod(cx + N, (double)N);
  cy = fmod(cy + N, (double)N);
  I.cx = cx;
  I.cy = cy;
  I.cz = pz + (double)I.sum_dz / NN;
}

// catalog rows as BFS finishes them (islands.cpp:240-245): volume, averages
// and energies times a^3, positions in the reference's pos() numbering
__global__ void k_isl_rows(int n, const long long *__restrict__ iint, const double *__restrict__ idbl,
                           int N, int lgN, double a, pn_island *__restrict__ rows) {
  const int id = blockIdx.x * blockDim.x + threadIdx.x;
  if (id >= n) return;
  (void)lgN;
  const long long *q = iint + (long long)id * kNI;
  const double *d = idbl + (long long)id * kND;
  const long long P = (long long)N * N;
  const double a3 = (a * a) * a;
  pn_island I;
  const long long NN = q[0];
  I.index = id;
  I.nn = NN;
  I.first_pos = q[1] + P;
  I.peak_pos = q[2] + P;
  I.peak_ties = q[3] - 1;
  I.volume = NN * a * a * a;
  I.avg_phi = d[0] / (double)(NN);
  I.pot_en = d[1] * a3;
  I.kin_en = d[2] * a3;
  I.grad_en = d[3] * a3;
  I.sum_dx = q[4];
  I.sum_dy = q[5];
  I.sum_dz = q[6];
  I.peak_value = d[4];
  isl_centroid(I, N);
  rows[id] = I;
}

// the peak key (phi^2 or Delta) and the peak value (phi or Delta) of the
// given sites, as the catalog kernels compute them (isl_key, k_isl_peakval)
template <typename T>
__global__ void k_isl_keyvals(const T *__restrict__ phi, const T *__restrict__ pi, int N, int lgN,
                              IslCoef k, const int *__restrict__ site, int n, double *__restrict__ kv) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= n) return;
  const int i = site[j];
  if (k.mode == 0) {
    const double f = (double)phi[i];
    kv[2 * j] = f * f;
    kv[2 * j + 1] = f;
  } else {
    int x, y, z;
    site_xyz(i, N, lgN, x, y, z);
    const double d = site_delta(phi, pi, (long long)i, x, y, z, N, k);
    kv[2 * j] = d;
    kv[2 * j + 1] = d;
  }
}

// The peak of island `id` of the last pn_islands call as the reference's BFS
// finds it from site `seed` (dense index): the queue order of islands.cpp:
// 90-279 (x+1, x-1, y+1, y-1 periodic, then z+1 and z-1 inside the lattice)
// and the first strict maximum of the key.  Only islands with equal maxima
// need it (a unique maximum does not depend on the order); it runs on the
// host over the island's sites (sorted, so membership is a binary search) and
// stops at the first site holding the island's largest key.  Fills
// peak_pos, peak_value, the offset sums and the centroid of *row.
template <typename T>
static int isl_bfs_peak(pn_ctx *c, int id, long long seed, pn_island *row) {
  pn_island_state *s = c->isl;
  const int N = c->p.N;
  const long long P = (long long)N * N;
  if (!s || id < 0 || id >= s->nisl) return pn_fail(c, PN_ERR_STATE, "pn_islands: no island %d", id);
  int se[2];
  PN_HIP(c, hipMemcpyAsync(se, s->ist + id, 2 * sizeof(int), hipMemcpyDeviceToHost, c->stream));
  PN_HIP(c, hipStreamSynchronize(c->stream));
  const int n = se[1] - se[0];
  std::vector<int> sites;
  std::vector<double> kv;
  std::vector<int> queue;
  std::vector<char> seen;
  try {
    sites.resize(n);
    kv.resize(2 * (size_t)n);
    queue.reserve(n);
    seen.assign(n, 0);
  } catch (const std::bad_alloc &) {
    return pn_fail(c, PN_ERR_NOMEM, "pn_islands: no host memory for an island of %d sites", n);
  }
  if (2LL * n > s->kvcap) {
    hipFree(s->kv);
    s->kv = nullptr;
    s->kvcap = 0;
    PN_HIP(c, hipMalloc(&s->kv, sizeof(double) * 2 * (size_t)n));
    s->kvcap = 2LL * n;
  }
  hipLaunchKernelGGL(k_isl_keyvals<T>, dim3((n + 255) / 256), dim3(256), 0, c->stream,
                     (const T *)c->phi[c->cur], (const T *)c->pi, N, s->lgN, s->last, s->msite2 + se[0],
                     n, s->kv);
  PN_CHECK_LAUNCH(c);
  PN_HIP(c, hipMemcpyAsync(sites.data(), s->msite2 + se[0], sizeof(int) * (size_t)n, hipMemcpyDeviceToHost,
                           c->stream));
  PN_HIP(c, hipMemcpyAsync(kv.data(), s->kv, sizeof(double) * 2 * (size_t)n, hipMemcpyDeviceToHost, c->stream));
  PN_HIP(c, hipStreamSynchronize(c->stream));
  double M = kv[0];
  for (int j = 1; j < n; j++) M = kv[2 * j] > M ? kv[2 * j] : M;
  auto find = [&](long long site) -> int {
    auto it = std::lower_bound(sites.begin(), sites.end(), (int)site);
    return (it != sites.end() && *it == (int)site) ? (int)(it - sites.begin()) : -1;
  };
  const int s0 = find(seed);
  if (s0 < 0) return pn_fail(c, PN_ERR_STATE, "pn_islands: site %lld is not in island %d", seed, id);
  int best = s0;
  queue.push_back(s0);
  seen[s0] = 1;
  for (size_t head = 0; head < queue.size() && !(kv[2 * best] == M); head++) {
    const int u = queue[head];
    if (kv[2 * u] > kv[2 * best]) best = u;  // islands.cpp:110-116
    if (kv[2 * best] == M) break;
    const long long i = sites[u];
    const int x = (int)(i % N), y = (int)((i / N) % N), z = (int)(i / P);
    const long long row0 = i - x, plane0 = i - (long long)y * N - x;
    const long long nb[6] = {row0 + (x + 1 == N ? 0 : x + 1), row0 + (x == 0 ? N - 1 : x - 1),
                             plane0 + (long long)(y + 1 == N ? 0 : y + 1) * N + x,
                             plane0 + (long long)(y == 0 ? N - 1 : y - 1) * N + x,
                             z + 1 < N ? i + P : -1, z > 0 ? i - P : -1};
    for (int j = 0; j < 6; j++) {
      if (nb[j] < 0) continue;
      const int v = find(nb[j]);
      if (v >= 0 && !seen[v]) {
        seen[v] = 1;
        queue.push_back(v);
      }
    }
  }
  const long long pk = sites[best];
  const int px = (int)(pk % N), py = (int)((pk / N) % N), pz = (int)(pk / P);
  long long sx = 0, sy = 0, sz = 0;
  for (int j = 0; j < n; j++) {
    const long long i = sites[j];
    int dx = (int)(i % N) - px, dy = (int)((i / N) % N) - py;
    if (dx >= N - N / 2) dx -= N;
    if (dx < -(N / 2)) dx += N;
    if (dy >= N - N / 2) dy -= N;
    if (dy < -(N / 2)) dy += N;
    sx += dx;
    sy += dy;
    sz += (int)(i / P) - pz;
  }
  row->peak_pos = pk + P;
  row->peak_value = kv[2 * best + 1];
  row->sum_dx = sx;
  row->sum_dy = sy;
  row->sum_dz = sz;
  isl_centroid(*row, N);
  return PN_OK;
}

int pn_islands_seeded_peak(pn_ctx *c, int id, long long seed_pos, long long *peak_pos, double *peak_value) {
  if (!c || !c->isl) return pn_fail(c, PN_ERR_STATE, "pn_islands_seeded_peak: no pn_islands call yet");
  pn_island row{};
  row.nn = 1;
  const long long seed = seed_pos - (long long)c->p.N * c->p.N;
  const int rc = c->p.precision == PN_FP64 ? isl_bfs_peak<double>(c, id, seed, &row)
                                           : isl_bfs_peak<float>(c, id, seed, &row);
  if (rc != PN_OK) return rc;
  *peak_pos = row.peak_pos;
  *peak_value = row.peak_value;
  return PN_OK;
}

template <typename T>
static int run_islands(pn_ctx *c, const IslCoef &k, pn_catalog *out) {
  pn_island_state *s = c->isl;
  const int N = c->p.N;
  const long long n3 = c->n3;
  const int grid = (int)std::min<long long>(8192, (n3 + kT - 1) / kT);
  const T *phi = (const T *)c->phi[c->cur];
  const T *pi = (const T *)c->pi;
  int lgN = -1;
  if ((N & (N - 1)) == 0) {
    lgN = 0;
    while ((1 << lgN) < N) lgN++;
  }
  s->last = k;
  s->lgN = lgN;
  s->nisl = 0;
  hipLaunchKernelGGL(k_isl_init<T>, dim3(grid), dim3(kT), 0, c->stream, phi, pi, N, lgN, k, s->label);
  hipLaunchKernelGGL(k_isl_merge, dim3(grid), dim3(kT), 0, c->stream, s->label, N);
  hipLaunchKernelGGL(k_isl_flatten, dim3(grid), dim3(kT), 0, c->stream, s->label, n3);
  hipLaunchKernelGGL(k_isl_count, dim3((unsigned)s->nblk), dim3(kT), 0, c->stream, s->label, n3, s->blk);
  hipLaunchKernelGGL(k_isl_scan, dim3(1), dim3(1024), 0, c->stream, s->blk, s->nblk);
  hipLaunchKernelGGL(k_isl_assign, dim3((unsigned)s->nblk), dim3(kT), 0, c->stream, s->label, n3,
                     s->blk, s->cid);
  hipLaunchKernelGGL(k_isl_mcount, dim3((unsigned)s->nblk), dim3(kT), 0, c->stream, s->label, n3, s->mblk);
  hipLaunchKernelGGL(k_isl_scan, dim3(1), dim3(1024), 0, c->stream, s->mblk, s->nblk);
  PN_CHECK_LAUNCH(c);
  int nisl = 0, nsite = 0;
  PN_HIP(c, hipMemcpyAsync(&nisl, s->blk + s->nblk, sizeof(int), hipMemcpyDeviceToHost, c->stream));
  PN_HIP(c, hipMemcpyAsync(&nsite, s->mblk + s->nblk, sizeof(int), hipMemcpyDeviceToHost, c->stream));
  PN_HIP(c, hipStreamSynchronize(c->stream));
  out->count = nisl;
  if (nisl > s->cap) {
    hipFree(s->ikey);
    hipFree(s->iint);
    hipFree(s->idbl);
    s->ikey = nullptr; s->iint = nullptr; s->idbl = nullptr;
    hipFree(s->ist);
    s->ist = nullptr;
    s->cap = 0;
    const int cap = std::max(nisl, 1024);
    PN_HIP(c, hipMalloc(&s->ikey, (size_t)cap * sizeof(unsigned long long)));
    PN_HIP(c, hipMalloc(&s->iint, (size_t)cap * kNI * sizeof(long long)));
    PN_HIP(c, hipMalloc(&s->idbl, (size_t)cap * kND * sizeof(double)));
    PN_HIP(c, hipMalloc(&s->ist, ((size_t)cap + 1) * sizeof(int)));
    s->cap = cap;
  }
  // island sites grouped by island: compaction in raster order, stable sort
  // by island number (radix sort over the bits island numbers use)
  int end_bit = 1;
  while (end_bit < 31 && (1LL << end_bit) <= (long long)nisl) end_bit++;
  if (nisl > 0) {
    if (nsite > s->mcap) {
      hipFree(s->mkey); hipFree(s->mkey2); hipFree(s->msite); hipFree(s->msite2);
      s->mkey = s->mkey2 = s->msite = s->msite2 = nullptr;
      s->mcap = 0;
      const size_t m = (size_t)nsite;
      PN_HIP(c, hipMalloc(&s->mkey, m * sizeof(int)));
      PN_HIP(c, hipMalloc(&s->mkey2, m * sizeof(int)));
      PN_HIP(c, hipMalloc(&s->msite, m * sizeof(int)));
      PN_HIP(c, hipMalloc(&s->msite2, m * sizeof(int)));
      s->mcap = nsite;
    }
    size_t bytes = 0;
    PN_HIP(c, hipcub::DeviceRadixSort::SortPairs(nullptr, bytes, s->mkey, s->mkey2, s->msite, s->msite2,
                                                 nsite, 0, end_bit, c->stream));
    if (bytes > s->sort_bytes) {
      hipFree(s->sort_tmp);
      s->sort_tmp = nullptr;
      s->sort_bytes = 0;
      PN_HIP(c, hipMalloc(&s->sort_tmp, bytes));
      s->sort_bytes = bytes;
    }
  }
  if (nisl > 0) {
    // chunk tables for islands above kBig sites: fewer than 2 nsite / kBig chunks
    const long long ccap = 2LL * nsite / kBig + 2;
    if (ccap > s->ccap || !s->counters) {
      hipFree(s->chunk_isl); hipFree(s->chunk_rank); hipFree(s->big_isl); hipFree(s->big_c0);
      hipFree(s->cpart);
      s->chunk_isl = s->chunk_rank = s->big_isl = s->big_c0 = nullptr;
      s->cpart = nullptr;
      s->ccap = 0;
      if (!s->counters) PN_HIP(c, hipMalloc(&s->counters, 2 * sizeof(int)));
      PN_HIP(c, hipMalloc(&s->chunk_isl, (size_t)ccap * sizeof(int)));
      PN_HIP(c, hipMalloc(&s->chunk_rank, (size_t)ccap * sizeof(int)));
      PN_HIP(c, hipMalloc(&s->big_isl, (size_t)ccap * sizeof(int)));
      PN_HIP(c, hipMalloc(&s->big_c0, (size_t)ccap * sizeof(int)));
      PN_HIP(c, hipMalloc(&s->cpart, (size_t)ccap * 8 * sizeof(double)));
      s->ccap = ccap;
    }
    hipLaunchKernelGGL(k_isl_prep, dim3((nisl + 255) / 256), dim3(256), 0, c->stream, nisl,
                       s->ikey, s->iint, s->idbl);
    hipLaunchKernelGGL(k_isl_mscatter, dim3((unsigned)s->nblk), dim3(kT), 0, c->stream, s->label, n3,
                       s->mblk, s->cid, s->mkey, s->msite);
    size_t bytes = s->sort_bytes;
    PN_HIP(c, hipcub::DeviceRadixSort::SortPairs(s->sort_tmp, bytes, s->mkey, s->mkey2, s->msite,
                                                 s->msite2, nsite, 0, end_bit, c->stream));
    hipLaunchKernelGGL(k_isl_starts, dim3((nsite + 255) / 256), dim3(256), 0, c->stream, s->mkey2, nsite,
                       nisl, s->ist);
    PN_HIP(c, hipMemsetAsync(s->counters, 0, 2 * sizeof(int), c->stream));
    hipLaunchKernelGGL(k_isl_small<T>, dim3((unsigned)std::min(8192, (nisl + 3) / 4)), dim3(kT), 0,
                       c->stream, phi, pi, N, lgN, k, s->ist, s->msite2, nisl, s->ikey, s->iint, s->idbl);
    hipLaunchKernelGGL(k_isl_bigplan, dim3((nisl + 255) / 256), dim3(256), 0, c->stream, s->ist, nisl,
                       s->counters, s->chunk_isl, s->chunk_rank, s->big_isl, s->big_c0);
    const int gbig = (int)std::max<long long>(1, std::min<long long>(2048, (ccap + 3) / 4));
    hipLaunchKernelGGL(k_isl_big1<T>, dim3(gbig), dim3(kT), 0, c->stream, phi, pi, N, lgN, k, s->ist,
                       s->msite2, s->counters, s->chunk_isl, s->chunk_rank, s->ikey, s->cpart);
    hipLaunchKernelGGL(k_isl_big2<T>, dim3(gbig), dim3(kT), 0, c->stream, phi, pi, N, lgN, k, s->ist,
                       s->msite2, s->counters, s->chunk_isl, s->chunk_rank, s->ikey, s->iint);
    hipLaunchKernelGGL(k_isl_big3, dim3(gbig), dim3(kT), 0, c->stream, N, lgN, s->ist, s->msite2,
                       s->counters, s->chunk_isl, s->chunk_rank, s->iint);
    hipLaunchKernelGGL(k_isl_bigmerge, dim3(gbig), dim3(kT), 0, c->stream, s->ist, s->msite2, s->counters,
                       s->big_isl, s->big_c0, s->cpart, s->iint, s->idbl);
    hipLaunchKernelGGL(k_isl_peakval<T>, dim3((nisl + 255) / 256), dim3(256), 0, c->stream, phi, pi,
                       N, lgN, k, nisl, s->iint, s->idbl);
  }
  s->nisl = nisl;
  // the rows, assembled on the device (isl_row) and copied once into the
  // caller's array
  const int ncopy = std::min(nisl, out->capacity);
  if (ncopy > s->rcap) {
    hipFree(s->rows);
    s->rows = nullptr;
    s->rcap = 0;
    const long long rc = std::max<long long>(ncopy + ncopy / 4, 1024);
    PN_HIP(c, hipMalloc(&s->rows, (size_t)rc * sizeof(pn_island)));
    s->rcap = rc;
  }
  if (out->labels && !s->dlab) PN_HIP(c, hipMalloc(&s->dlab, (size_t)n3 * sizeof(int)));
  if (out->labels)
    hipLaunchKernelGGL(k_isl_labels, dim3(grid), dim3(kT), 0, c->stream, s->label, n3, s->cid, s->dlab);
  if (ncopy > 0)
    hipLaunchKernelGGL(k_isl_rows, dim3((ncopy + 255) / 256), dim3(256), 0, c->stream, ncopy, s->iint,
                       s->idbl, N, lgN, c->p.a, s->rows);
  hipError_t e = hipGetLastError();
  if (e == hipSuccess && ncopy > 0)
    e = hipMemcpyAsync(out->islands, s->rows, (size_t)ncopy * sizeof(pn_island), hipMemcpyDeviceToHost,
                       c->stream);
  if (e == hipSuccess && out->labels)
    e = hipMemcpyAsync(out->labels, s->dlab, (size_t)n3 * sizeof(int), hipMemcpyDeviceToHost, c->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
  if (e != hipSuccess) return pn_fail(c, PN_ERR_HIP, "pn_islands: %s", hipGetErrorString(e));
  // equal maxima: the reference keeps the first of them in its breadth-first
  // order from the island's first site (islands.cpp:104-116,200-236)
  for (int id = 0; id < ncopy; id++) {
    if (out->islands[id].peak_ties == 0) continue;
    int rc = isl_bfs_peak<T>(c, id, out->islands[id].first_pos - (long long)N * N, &out->islands[id]);
    if (rc != PN_OK) return rc;
  }
  if (nisl > out->capacity)
    return pn_fail(c, PN_ERR_CAPACITY, "pn_islands: %d islands, capacity %d", nisl, out->capacity);
  return PN_OK;
}

extern "C" int pn_islands(pn_ctx *c, int mode, double thr, pn_catalog *out) {
  if (!c || !out || (out->capacity > 0 && !out->islands))
    return pn_fail(c, PN_ERR_ARG, "pn_islands: NULL argument");
  if (mode != 0 && mode != 1) return pn_fail(c, PN_ERR_ARG, "pn_islands: mode %d", mode);
  if (c->n3 > 0x7fffffffLL) return pn_fail(c, PN_ERR_ARG, "pn_islands: lattice too large for int labels");
  PN_HIP(c, hipSetDevice(c->p.device));
  if (!c->isl) {
    c->isl = new (std::nothrow) pn_island_state();
    if (!c->isl) return pn_fail(c, PN_ERR_NOMEM, "pn_islands: out of host memory");
    pn_island_state *s = c->isl;
    s->nblk = (c->n3 + kChunk - 1) / kChunk;
    PN_HIP(c, hipMalloc(&s->label, (size_t)c->n3 * sizeof(int)));
    PN_HIP(c, hipMalloc(&s->cid, (size_t)c->n3 * sizeof(int)));
    PN_HIP(c, hipMalloc(&s->blk, (size_t)(s->nblk + 1) * sizeof(int)));
    PN_HIP(c, hipMalloc(&s->mblk, (size_t)(s->nblk + 1) * sizeof(int)));
  }
  IslCoef k;
  const int Nc = 1;
  k.a2 = c->p.a * c->p.a;
  k.cm = (c->p.mass_mode) * 0.5 * c->p.m2;
  k.lmode = (double)c->p.lambda_mode;
  k.d6 = 720.0 * Nc * Nc * c->p.kappa * c->p.kappa;
  k.thr = thr;
  k.thr2h = thr * thr * 0.5;
  k.mode = mode;
  out->threshold = thr;
  if (c->p.precision == PN_FP64) return run_islands<double>(c, k, out);
  return run_islands<float>(c, k, out);
}

// island number of the last pn_islands call at given sites (or -1 outside the
// mask): label[] holds each site's root, cid[] the number stored at roots
__global__ void k_isl_lookup(const int *__restrict__ label, const int *__restrict__ cid,
                             const long long *__restrict__ dense, int n, int *__restrict__ out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int r = label[dense[i]];
  out[i] = r >= 0 ? cid[r] : -1;
}

extern "C" int pn_islands_lookup(pn_ctx *c, int n, const long long *pos, int *island) {
  if (!c || n < 0 || (n > 0 && (!pos || !island)))
    return pn_fail(c, PN_ERR_ARG, "pn_islands_lookup: bad argument");
  if (!c->isl) return pn_fail(c, PN_ERR_STATE, "pn_islands_lookup: no pn_islands call yet");
  if (n == 0) return PN_OK;
  pn_island_state *s = c->isl;
  const long long P = (long long)c->p.N * c->p.N;
  for (int i = 0; i < n; i++)
    if (pos[i] - P < 0 || pos[i] - P >= c->n3)
      return pn_fail(c, PN_ERR_ARG, "pn_islands_lookup: pos %lld is not an interior site", pos[i]);
  PN_HIP(c, hipSetDevice(c->p.device));
  if (n > s->lcap) {  // kept across calls, grown on demand
    hipFree(s->lpos);
    hipFree(s->lout);
    hipHostFree(s->hpos);
    hipHostFree(s->hout);
    s->lpos = s->hpos = nullptr;
    s->lout = s->hout = nullptr;
    s->lcap = 0;
    const long long cap = std::max<long long>(n + n / 4, 1024);
    PN_HIP(c, hipMalloc(&s->lpos, sizeof(long long) * (size_t)cap));
    PN_HIP(c, hipMalloc(&s->lout, sizeof(int) * (size_t)cap));
    PN_HIP(c, hipHostMalloc(&s->hpos, sizeof(long long) * (size_t)cap, hipHostMallocDefault));
    PN_HIP(c, hipHostMalloc(&s->hout, sizeof(int) * (size_t)cap, hipHostMallocDefault));
    s->lcap = cap;
  }
  for (int i = 0; i < n; i++) s->hpos[i] = pos[i] - P;  // reference pos() (one z halo plane below) -> dense
  hipError_t e = hipMemcpyAsync(s->lpos, s->hpos, sizeof(long long) * n, hipMemcpyHostToDevice, c->stream);
  if (e == hipSuccess) {
    hipLaunchKernelGGL(k_isl_lookup, dim3((n + 255) / 256), dim3(256), 0, c->stream, s->label, s->cid,
                       s->lpos, n, s->lout);
    e = hipGetLastError();
  }
  if (e == hipSuccess) e = hipMemcpyAsync(s->hout, s->lout, sizeof(int) * n, hipMemcpyDeviceToHost, c->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
  if (e != hipSuccess) return pn_fail(c, PN_ERR_HIP, "pn_islands_lookup: %s", hipGetErrorString(e));
  std::copy(s->hout, s->hout + n, island);
  return PN_OK;
}
