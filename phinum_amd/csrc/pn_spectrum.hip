// pn_spectrum.hip -- |phi_k|^2 shell spectrum: rocFFT r2c for the 3D
// transform, hand-written kernels for the layout change and the binning.
//
// The reference transforms with FFTW in the order (z, x, y) with y as the
// halved, fastest dimension (clusterfft.cpp:40,193-199) and bins modes in the
// order posLOC = py + px*(N/2+1) + pz*N*(N/2+1) (momentumspace.cpp:222).  Here
// a tiled transpose writes the field as [z][x][y] into a scratch buffer, so
// rocFFT's r2c (fastest dimension halved) yields exactly that mode order, and
// every mode gets the reference's bin from the same per-axis momentum tables
// (built on the host with the reference's expressions, momentumspace.cpp:102-125).
#include <hip/hip_runtime.h>
#include <rocfft/rocfft.h>

#include <cmath>
#include <mutex>
#include <vector>

#include "pn_internal.h"
#include "pn_kernels.h"

struct pn_fft_state {
  rocfft_plan plan = nullptr;
  rocfft_execution_info info = nullptr;
  void *work = nullptr;
  void *scratch = nullptr;  // real [z][x][y]
  void *cphi = nullptr;     // complex [pz][px][py], py <= N/2
  void *cpi = nullptr;
  double *tab = nullptr;    // sqr of lattice momenta: N (x) + (N/2+1) (y) + N (z)
  double *bin_part = nullptr;
  int *cnt_part = nullptr;
  double *bin_out = nullptr;
  int *cnt_out = nullptr;
  int nbins = 0;
  double binsize = 0;
  int nblocks = 0;
};

static std::once_flag g_rocfft_once;
static rocfft_status g_rocfft_setup = rocfft_status_success;

void pn_free_spectrum(pn_ctx *c) {
  pn_fft_state *f = c->fft;
  if (!f) return;
  if (f->plan) rocfft_plan_destroy(f->plan);
  if (f->info) rocfft_execution_info_destroy(f->info);
  hipFree(f->work);
  hipFree(f->scratch);
  hipFree(f->cphi);
  hipFree(f->cpi);
  hipFree(f->tab);
  hipFree(f->bin_part);
  hipFree(f->cnt_part);
  hipFree(f->bin_out);
  hipFree(f->cnt_out);
  delete f;
  c->fft = nullptr;
}

static double sqrd(double x) { return x * x; }

extern "C" int pn_spectrum_nbins(int N, double a) {
  if (N < 2 || !(a > 0)) return PN_ERR_ARG;
  double pTmin = (2.0 / (a)) * sin(acos(-1.) / N);   // momentumspace.cpp:40
  double pTmax = sqrt(3.0 * (4.0 / sqrd(a)));        // :41
  return (int)(pTmax / pTmin + 0.5) + 2;             // :47
}

static constexpr int kBinBlocks = 512;

// Waves per binning block.  Each wave keeps its own compensated histogram,
// (hi, lo) of the three sums per bin (48 bytes), and the block one int count
// per bin; up to four waves within 64 KB of LDS.
static int bin_waves(int nbins) {
  const long long room = (65536LL - 4LL * nbins) / (48LL * nbins);
  return room < 1 ? 1 : (room > 4 ? 4 : (int)room);
}

static int ensure_fft(pn_ctx *c) {
  if (c->fft) return PN_OK;
  std::call_once(g_rocfft_once, [] { g_rocfft_setup = rocfft_setup(); });
  if (g_rocfft_setup != rocfft_status_success)
    return pn_fail(c, PN_ERR_FFT, "rocfft_setup failed (%d)", (int)g_rocfft_setup);
  pn_fft_state *f = new pn_fft_state();
  c->fft = f;
  const int N = c->p.N, H = N / 2 + 1;
  const double a = c->p.a;
  const bool dp = c->p.precision == PN_FP64;
  size_t lengths[3] = {(size_t)N, (size_t)N, (size_t)N};
  rocfft_status s = rocfft_plan_create(&f->plan, rocfft_placement_notinplace,
                                       rocfft_transform_type_real_forward,
                                       dp ? rocfft_precision_double : rocfft_precision_single, 3,
                                       lengths, 1, nullptr);
  if (s != rocfft_status_success) return pn_fail(c, PN_ERR_FFT, "rocfft_plan_create failed (%d)", (int)s);
  size_t wsize = 0;
  rocfft_plan_get_work_buffer_size(f->plan, &wsize);
  if (rocfft_execution_info_create(&f->info) != rocfft_status_success)
    return pn_fail(c, PN_ERR_FFT, "rocfft_execution_info_create failed");
  if (wsize) {
    PN_HIP(c, hipMalloc(&f->work, wsize));
    rocfft_execution_info_set_work_buffer(f->info, f->work, wsize);
  }
  rocfft_execution_info_set_stream(f->info, c->stream);
  const size_t rsz = (size_t)c->n3 * c->esize;
  const size_t csz = (size_t)N * N * H * 2 * c->esize;
  PN_HIP(c, hipMalloc(&f->scratch, rsz));
  PN_HIP(c, hipMalloc(&f->cphi, csz));
  PN_HIP(c, hipMalloc(&f->cpi, csz));

  // per-axis squared lattice momenta, reference expressions
  // (momentumspace.cpp:102-125): x and z use the signed branches, y the
  // positive one; sqr() removes the sign exactly.
  std::vector<double> tab((size_t)N + H + N);
  const double pi_ = acos(-1.);
  for (int px = 0; px < N; px++) {
    double ppx = (px < N / 2 + 1) ? 2.0 / (a)*sin(px * pi_ / N) : -2.0 / (a)*sin(px * pi_ / N);
    tab[px] = sqrd(ppx);
  }
  for (int py = 0; py < H; py++) tab[N + py] = sqrd(2.0 / (a)*sin(py * pi_ / N));
  for (int pz = 0; pz < N; pz++) {
    double nu;
    if (pz == 0) nu = 0;
    else nu = (pz < N / 2 + 1) ? 2.0 / (a)*sin((pz)*pi_ / N) : -2.0 / (a)*sin((pz)*pi_ / N);
    tab[N + H + pz] = sqrd(nu);
  }
  PN_HIP(c, hipMalloc(&f->tab, tab.size() * sizeof(double)));
  PN_HIP(c, hipMemcpy(f->tab, tab.data(), tab.size() * sizeof(double), hipMemcpyHostToDevice));
  f->nbins = pn_spectrum_nbins(N, a);
  f->binsize = (2.0 / (a)) * sin(acos(-1.) / N);
  f->nblocks = kBinBlocks;
  PN_HIP(c, hipMalloc(&f->bin_part, (size_t)f->nblocks * f->nbins * 6 * sizeof(double)));
  PN_HIP(c, hipMalloc(&f->cnt_part, (size_t)f->nblocks * f->nbins * sizeof(int)));
  PN_HIP(c, hipMalloc(&f->bin_out, (size_t)f->nbins * 3 * sizeof(double)));
  PN_HIP(c, hipMalloc(&f->cnt_out, (size_t)f->nbins * sizeof(int)));
  return PN_OK;
}

// [z][y][x] -> [z][x][y], one 32x32 tile per block through LDS
template <typename T>
__global__ __launch_bounds__(256) void k_transpose_xy(const T *__restrict__ in, T *__restrict__ out,
                                                      int N) {
  __shared__ T tile[32][33];
  const int bx = blockIdx.x * 32, by = blockIdx.y * 32, z = blockIdx.z;
  const long long base = (long long)z * N * N;
  for (int j = threadIdx.y; j < 32; j += 8) {
    const int x = bx + threadIdx.x, y = by + j;
    if (x < N && y < N) tile[j][threadIdx.x] = in[base + (long long)y * N + x];
  }
  __syncthreads();
  for (int j = threadIdx.y; j < 32; j += 8) {
    const int y = by + threadIdx.x, x = bx + j;
    if (x < N && y < N) out[base + (long long)x * N + y] = tile[threadIdx.x][j];
  }
}

// Shell binning of momentumspace.cpp:196-247 with compensated, deterministic
// FP64 sums.  A wave takes one (pz, px) row of modes at a time, py = lane,
// lane + 64, ...: mode m = py + H*(pz*N + px).  Along a row the bin index
// never decreases (psqr grows with py and every operation on it is monotone),
// so equal bins are contiguous lanes.  A segmented scan over shuffles sums
// each run in (hi, lo) form, and the run's last lane adds it to the wave's own
// LDS histogram: no two lanes touch one bin at once and the order of the
// additions is fixed by the launch.  The waves' histograms are merged in wave
// order, the blocks' by k_bin_final in a fixed tree.  Counts are integers
// (atomic adds are exact).
template <typename T2>
__global__ __launch_bounds__(256) void k_bin(const T2 *__restrict__ cphi, const T2 *__restrict__ cpi,
                                             const double *__restrict__ tab, int N, double fac,
                                             double binsize, int nbins, double *__restrict__ part,
                                             int *__restrict__ cpart) {
  extern __shared__ double lds[];
  const int nw = (int)(blockDim.x >> 6), w = (int)(threadIdx.x >> 6), lane = (int)(threadIdx.x & 63);
  double *hist = lds + (size_t)w * nbins * 6;  // per bin: F hi, F lo, K hi, K lo, Q hi, Q lo
  int *sc = (int *)(lds + (size_t)nw * nbins * 6);
  for (int i = threadIdx.x; i < nw * nbins * 6; i += blockDim.x) lds[i] = 0.0;
  for (int b = threadIdx.x; b < nbins; b += blockDim.x) sc[b] = 0;
  __syncthreads();
  const int H = N / 2 + 1;
  const double *tx = tab, *ty = tab + N, *tz = tab + N + H;
  const long long rows = (long long)N * N;
  for (long long r = (long long)blockIdx.x * nw + w; r < rows; r += (long long)gridDim.x * nw) {
    const int px = (int)(r % N), pz = (int)(r / N);
    const long long m0 = r * H;
    for (int p0 = 0; p0 < H; p0 += 64) {
      const int py = p0 + lane;
      int b = nbins;  // lanes past the row end: a run of their own, never added
      double hi[3] = {0.0, 0.0, 0.0}, lo[3] = {0.0, 0.0, 0.0};
      if (py < H) {
        const double psqr = tx[px] + ty[py] + tz[pz];
        b = (int)(sqrt(psqr) / binsize + 0.5);
        if (b > nbins) b = nbins;  // cannot happen: nbins covers sqrt(3)*pTmax
        const T2 u = cphi[m0 + py], v = cpi[m0 + py];
        const double pr = fac * (double)u.x, pim = fac * (double)u.y;
        const double kr = fac * (double)v.x, kim = fac * (double)v.y;
        const double ff = pr * pr + pim * pim, kk = kr * kr + kim * kim, qq = pr * kr + pim * kim;
        const bool twice = py != 0 && py != N / 2;  // the py -> N-py half not stored
        hi[0] = twice ? ff + ff : ff;
        hi[1] = twice ? kk + kk : kk;
        hi[2] = twice ? qq + qq : qq;
        if (b < nbins) atomicAdd(&sc[b], twice ? 2 : 1);
      }
#pragma unroll
      for (int d = 1; d < 64; d <<= 1) {
        const int bo = __shfl_up(b, d, 64);
        double h2[3], l2[3];
#pragma unroll
        for (int j = 0; j < 3; j++) {
          h2[j] = __shfl_up(hi[j], d, 64);
          l2[j] = __shfl_up(lo[j], d, 64);
        }
        if (lane >= d && bo == b)
#pragma unroll
          for (int j = 0; j < 3; j++) two_merge(hi[j], lo[j], h2[j], l2[j]);
      }
      const int bn = __shfl_down(b, 1, 64);
      if (b < nbins && (lane == 63 || bn != b))
#pragma unroll
        for (int j = 0; j < 3; j++) two_merge(hist[6 * b + 2 * j], hist[6 * b + 2 * j + 1], hi[j], lo[j]);
    }
  }
  __syncthreads();
  for (int i = threadIdx.x; i < nbins * 3; i += blockDim.x) {
    const int b = i / 3, j = i % 3;
    double h = lds[6 * b + 2 * j], l = lds[6 * b + 2 * j + 1];
    for (int ww = 1; ww < nw; ww++)
      two_merge(h, l, lds[(size_t)ww * nbins * 6 + 6 * b + 2 * j], lds[(size_t)ww * nbins * 6 + 6 * b + 2 * j + 1]);
    const long long o = ((long long)blockIdx.x * nbins + b) * 6 + 2 * j;
    part[o] = h;
    part[o + 1] = l;
  }
  for (int b = threadIdx.x; b < nbins; b += blockDim.x) cpart[(long long)blockIdx.x * nbins + b] = sc[b];
}

// one wave per shell: lane l merges the partials of blocks l, l + 64, ...,
// then a fixed shuffle tree merges the lanes (the same order on every call)
__global__ __launch_bounds__(256) void k_bin_final(const double *__restrict__ part,
                                                   const int *__restrict__ cpart, int nb, int nbins,
                                                   double *__restrict__ out, int *__restrict__ cout) {
  const int lane = (int)(threadIdx.x & 63);
  const int b = (int)blockIdx.x * (int)(blockDim.x >> 6) + (int)(threadIdx.x >> 6);
  if (b >= nbins) return;  // whole waves
  double hi[3] = {0, 0, 0}, lo[3] = {0, 0, 0};
  int cnt = 0;
  for (int k = lane; k < nb; k += 64) {
    const long long o = ((long long)k * nbins + b) * 6;
    for (int j = 0; j < 3; j++) two_merge(hi[j], lo[j], part[o + 2 * j], part[o + 2 * j + 1]);
    cnt += cpart[(long long)k * nbins + b];
  }
  for (int off = 32; off > 0; off >>= 1) {
    for (int j = 0; j < 3; j++) {
      const double h2 = __shfl_down(hi[j], off, 64), l2 = __shfl_down(lo[j], off, 64);
      two_merge(hi[j], lo[j], h2, l2);
    }
    cnt += __shfl_down(cnt, off, 64);
  }
  if (lane == 0) {
    for (int j = 0; j < 3; j++) out[3 * b + j] = hi[j] + lo[j];
    cout[b] = cnt;
  }
}

// one field through transpose + r2c into dst
static int transform_field(pn_ctx *c, const void *field, void *dst) {
  pn_fft_state *f = c->fft;
  const int N = c->p.N;
  dim3 grid((N + 31) / 32, (N + 31) / 32, N), block(32, 8);
  if (c->p.precision == PN_FP64)
    hipLaunchKernelGGL(k_transpose_xy<double>, grid, block, 0, c->stream, (const double *)field,
                       (double *)f->scratch, N);
  else
    hipLaunchKernelGGL(k_transpose_xy<float>, grid, block, 0, c->stream, (const float *)field,
                       (float *)f->scratch, N);
  PN_CHECK_LAUNCH(c);
  void *in[1] = {f->scratch};
  void *out[1] = {dst};
  rocfft_status s = rocfft_execute(f->plan, in, out, f->info);
  if (s != rocfft_status_success) return pn_fail(c, PN_ERR_FFT, "rocfft_execute failed (%d)", (int)s);
  return PN_OK;
}

static int run_transforms(pn_ctx *c) {
  int rc = ensure_fft(c);
  if (rc != PN_OK) return rc;
  if ((rc = transform_field(c, c->phi[c->cur], c->fft->cphi)) != PN_OK) return rc;
  if ((rc = transform_field(c, c->pi, c->fft->cpi)) != PN_OK) return rc;
  return PN_OK;
}

extern "C" int pn_spectrum_get(pn_ctx *c, pn_spectrum *out) {
  if (!c || !out || !out->counts || !out->corr_f || !out->corr_k || !out->corr_q)
    return pn_fail(c, PN_ERR_ARG, "pn_spectrum_get: NULL argument");
  PN_HIP(c, hipSetDevice(c->p.device));
  int rc = run_transforms(c);
  if (rc != PN_OK) return rc;
  pn_fft_state *f = c->fft;
  if (out->nbins < f->nbins)
    return pn_fail(c, PN_ERR_CAPACITY, "pn_spectrum_get: nbins %d < %d", out->nbins, f->nbins);
  const double a = c->p.a;
  const double fac = (a * a) * a;  // Factor = sqr(a_t)*a_eta (clusterfft.cpp:187)
  const int nw = bin_waves(f->nbins);
  const size_t lds = (size_t)nw * f->nbins * 6 * sizeof(double) + (size_t)f->nbins * sizeof(int);
  if (lds > 65536) return pn_fail(c, PN_ERR_ARG, "pn_spectrum_get: %d bins exceed the binning kernel's LDS", f->nbins);
  if (c->p.precision == PN_FP64)
    hipLaunchKernelGGL(k_bin<double2>, dim3(f->nblocks), dim3(64 * nw), lds, c->stream,
                       (const double2 *)f->cphi, (const double2 *)f->cpi, f->tab, c->p.N, fac,
                       f->binsize, f->nbins, f->bin_part, f->cnt_part);
  else
    hipLaunchKernelGGL(k_bin<float2>, dim3(f->nblocks), dim3(64 * nw), lds, c->stream,
                       (const float2 *)f->cphi, (const float2 *)f->cpi, f->tab, c->p.N, fac,
                       f->binsize, f->nbins, f->bin_part, f->cnt_part);
  PN_CHECK_LAUNCH(c);
  hipLaunchKernelGGL(k_bin_final, dim3((f->nbins + 3) / 4), dim3(256), 0, c->stream,
                     f->bin_part, f->cnt_part, f->nblocks, f->nbins, f->bin_out, f->cnt_out);
  PN_CHECK_LAUNCH(c);
  std::vector<double> sums((size_t)f->nbins * 3);
  PN_HIP(c, hipMemcpyAsync(sums.data(), f->bin_out, sums.size() * sizeof(double),
                           hipMemcpyDeviceToHost, c->stream));
  PN_HIP(c, hipMemcpyAsync(out->counts, f->cnt_out, (size_t)f->nbins * sizeof(int),
                           hipMemcpyDeviceToHost, c->stream));
  PN_HIP(c, hipStreamSynchronize(c->stream));
  for (int b = 0; b < f->nbins; b++) {
    out->corr_f[b] = sums[3 * b + 0];
    out->corr_k[b] = sums[3 * b + 1];
    out->corr_q[b] = sums[3 * b + 2];
  }
  out->nbins = f->nbins;
  out->binsize = f->binsize;
  return PN_OK;
}

// fftXtoP_direct (clusterfft.cpp:185-225): a^3-scaled transform of Phi
// (field 0) or Pi (field 1), interleaved complex in reference p-space order.
extern "C" int pn_fft_modes(pn_ctx *c, int field, double *out) {
  if (!c || !out || (field != 0 && field != 1))
    return pn_fail(c, PN_ERR_ARG, "pn_fft_modes: bad argument");
  PN_HIP(c, hipSetDevice(c->p.device));
  int rc = run_transforms(c);
  if (rc != PN_OK) return rc;
  const int N = c->p.N, H = N / 2 + 1;
  const size_t nm = (size_t)N * N * H;
  const void *src = field == 0 ? c->fft->cphi : c->fft->cpi;
  const double a = c->p.a, fac = (a * a) * a;
  if (c->p.precision == PN_FP64) {
    PN_HIP(c, hipMemcpyAsync(out, src, nm * 16, hipMemcpyDeviceToHost, c->stream));
    PN_HIP(c, hipStreamSynchronize(c->stream));
    for (size_t i = 0; i < 2 * nm; i++) out[i] = fac * out[i];
  } else {
    std::vector<float> tmp(2 * nm);
    PN_HIP(c, hipMemcpyAsync(tmp.data(), src, nm * 8, hipMemcpyDeviceToHost, c->stream));
    PN_HIP(c, hipStreamSynchronize(c->stream));
    for (size_t i = 0; i < 2 * nm; i++) out[i] = fac * (double)tmp[i];
  }
  return PN_OK;
}
