// pn_icgen.hip -- the initialize==3 mode generator on the device.
//
// It makes the same draws as the host generator pn_ic_modes
// (initialconditions.cpp:83-594 with randGen, :12-42): one POSIX drand48
// stream; per mode, two uniform phases, then one polar Box-Muller pair whose
// two values are the mode's two |Gauss| amplitudes.  The cached second value
// of the pair is consumed by the same mode, so a mode uses 2 + 2k draws (k >= 1
// polar attempts) and every mode starts at an even stream position.  The
// start of each mode is found in parallel:
//
//  1. The stream is cut into chunks of kC draws.  A block computes its chunk's
//     draws (drand48 jump-ahead), the next mode start after each even position
//     (one polar loop each), and by pointer jumping in LDS, for each of the
//     kWin possible entry offsets of the chunk, where the chain of modes leaves
//     it and how many modes it started inside.
//  2. Three short passes chain the chunks into each chunk's entry offset and
//     first mode number: within groups of kG chunks, across the groups, then
//     back down into the chunks.
//  3. A block per chunk recomputes its draws and the 2^r-step jump tables and
//     writes its modes; a thread finds its mode's start with at most kLev
//     table lookups.
//
// ensureRealResult (clusterfft.cpp:55-145) follows as an elementwise pass:
// in the py = 0 and py = N/2 planes, of each pair (pz,px) <-> (-pz,-px) the
// one the reference visits first keeps its value and the other becomes its
// conjugate; a self-conjugate mode becomes +-|F| (sign of its real part).
//
// The draws, the polar acceptance test (x*x + y*y, a correctly rounded sqrt,
// < 1) and so the mode positions are exact.  log, sincos and hypot are the
// device library's and may differ from the host C library in the last bit.
#include <hip/hip_runtime.h>

#include <cmath>
#include <vector>

#include "pn_internal.h"

namespace {

constexpr int kC = 2048;          // draws per chunk
constexpr int kM = 256;           // look-ahead draws past the chunk
constexpr int kU = kC + kM;       // draws a block holds in LDS
constexpr int kNodes = kC / 2;    // even positions (possible mode starts) per chunk
constexpr int kWin = 64;          // entry offsets tracked per chunk (even positions 0..126)
constexpr int kLev = 10;          // 2^10 >= the longest chain in a chunk (kNodes / 2 modes)
constexpr int kT = 256;           // threads per block
constexpr long long kG = 1024;    // chunks per group in the chaining passes
static_assert(kU % kT == 0, "each thread fills kU / kT draws");
constexpr unsigned long long kMask48 = (1ULL << 48) - 1;
constexpr unsigned long long kLcgA = 0x5DEECE66DULL, kLcgC = 0xBULL;

// J.a[j], J.c[j]: the drand48 map applied 2^j times, X -> a X + c (mod 2^48)
struct IcJump {
  unsigned long long a[48], c[48];
};

IcJump make_jump() {
  IcJump J;
  unsigned long long a = kLcgA, c = kLcgC;
  for (int j = 0; j < 48; j++) {
    J.a[j] = a;
    J.c[j] = c;
    c = (a * c + c) & kMask48;  // (a x + c) twice: a^2 x + (a c + c)
    a = (a * a) & kMask48;
  }
  return J;
}

struct IcModeArgs {
  int N, H;
  double two_pi, QS, scale2, amplitude, noise, m2_long;
  int justnoise;
  const double *latp;  // lat_p(k) for k < N (host-computed, momentumspace.cpp:102-118)
  const double *ppy;   // 2/a sin(py pi / N) for py < H (host-computed)
};

__device__ __forceinline__ unsigned long long lcg_jump(const IcJump &J, unsigned long long x,
                                                      unsigned long long n) {
  for (int j = 0; n; j++, n >>= 1)
    if (n & 1) x = (J.a[j] * x + J.c[j]) & kMask48;
  return x;
}

// u[i] = draw number S + i of the stream (X_{S+i+1} / 2^48)
__device__ void fill_draws(const IcJump &J, unsigned long long x0, long long S, double *u,
                           unsigned long long *xs) {
  if (threadIdx.x == 0) *xs = lcg_jump(J, x0, (unsigned long long)S);
  __syncthreads();
  constexpr int per = kU / kT;
  unsigned long long x = lcg_jump(J, *xs, (unsigned long long)(per * threadIdx.x));
  for (int k = 0; k < per; k++) {
    x = (kLcgA * x + kLcgC) & kMask48;
    u[per * threadIdx.x + k] = (double)x * 0x1p-48;
  }
  __syncthreads();
}

// node i = even draw 2i of the chunk; returns the node where the next mode
// starts (two phases, then polar attempts until x^2 + y^2 < 1), or a node past
// the entry window when the attempts run past the look-ahead (flagged)
__device__ __forceinline__ int next_node(const double *u, int i, bool &bad) {
  int q = 2 * i + 2;
  for (;;) {
    if (q + 1 >= kU) {
      bad = true;
      return kNodes + kWin;
    }
    const double x = 2.0 * u[q] - 1.0, y = 2.0 * u[q + 1] - 1.0;
    q += 2;
    if (sqrt(x * x + y * y) < 1.0) return q >> 1;
  }
}

// pass 1: per chunk and entry offset e < kWin, (exit offset into the next
// chunk << 16) | modes started in this chunk
__global__ __launch_bounds__(kT) void k_icg_chunks(IcJump J, unsigned long long x0,
                                                   unsigned *__restrict__ table, int *__restrict__ err) {
  __shared__ double u[kU];
  __shared__ unsigned long long xs;
  __shared__ unsigned short nx[2][kNodes], cn[2][kNodes];
  const long long ch = blockIdx.x;
  fill_draws(J, x0, ch * kC, u, &xs);
  bool bad = false;
  for (int i = threadIdx.x; i < kNodes; i += kT) {
    nx[0][i] = (unsigned short)next_node(u, i, bad);
    cn[0][i] = 1;
  }
  __syncthreads();
  int s = 0;
  for (int r = 0; r < kLev; r++) {
    for (int i = threadIdx.x; i < kNodes; i += kT) {
      const int n = nx[s][i];
      if (n < kNodes) {
        nx[s ^ 1][i] = nx[s][n];
        cn[s ^ 1][i] = (unsigned short)(cn[s][i] + cn[s][n]);
      } else {
        nx[s ^ 1][i] = (unsigned short)n;
        cn[s ^ 1][i] = cn[s][i];
      }
    }
    __syncthreads();
    s ^= 1;
  }
  if (threadIdx.x < kWin) {
    int ex = (int)nx[s][threadIdx.x] - kNodes;
    if (ex < 0 || ex >= kWin) {
      bad = true;
      ex = 0;
    }
    table[ch * kWin + threadIdx.x] = ((unsigned)ex << 16) | cn[s][threadIdx.x];
  }
  if (bad) err[0] = 1;
}

// pass 2a: compose the chunks of each group, one thread per entry offset
__global__ __launch_bounds__(kWin) void k_icg_group(const unsigned *__restrict__ table, long long nchunks,
                                                    int *__restrict__ gexit,
                                                    unsigned long long *__restrict__ gcnt) {
  const long long g = blockIdx.x;
  const long long c0 = g * kG, c1 = c0 + kG < nchunks ? c0 + kG : nchunks;
  int e = threadIdx.x;
  unsigned long long cnt = 0;
  for (long long ch = c0; ch < c1; ch++) {
    const unsigned t = table[ch * kWin + e];
    cnt += t & 0xFFFFu;
    e = (int)(t >> 16);
  }
  gexit[g * kWin + threadIdx.x] = e;
  gcnt[g * kWin + threadIdx.x] = cnt;
}

// pass 2b: walk the groups from the stream start (entry 0 of chunk 0)
__global__ __launch_bounds__(64) void k_icg_top(const int *__restrict__ gexit,
                                                const unsigned long long *__restrict__ gcnt, long long ngroups,
                                                unsigned long long nm, int *__restrict__ gentry,
                                                unsigned long long *__restrict__ gbase, int *__restrict__ err) {
  if (threadIdx.x != 0) return;
  int e = 0;
  unsigned long long base = 0;
  for (long long g = 0; g < ngroups; g++) {
    gentry[g] = e;
    gbase[g] = base;
    base += gcnt[g * kWin + e];
    e = gexit[g * kWin + e];
  }
  if (base < nm) err[1] = 1;  // the provisioned stream held too few modes
}

// pass 2c: entry offset and first mode number of every chunk
__global__ __launch_bounds__(64) void k_icg_down(const unsigned *__restrict__ table, long long nchunks,
                                                 long long ngroups, const int *__restrict__ gentry,
                                                 const unsigned long long *__restrict__ gbase,
                                                 unsigned char *__restrict__ centry,
                                                 unsigned long long *__restrict__ cbase) {
  const long long g = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= ngroups) return;
  const long long c0 = g * kG, c1 = c0 + kG < nchunks ? c0 + kG : nchunks;
  int e = gentry[g];
  unsigned long long base = gbase[g];
  for (long long ch = c0; ch < c1; ch++) {
    centry[ch] = (unsigned char)e;
    cbase[ch] = base;
    const unsigned t = table[ch * kWin + e];
    base += t & 0xFFFFu;
    e = (int)(t >> 16);
  }
}

// one mode from its first draw q (initialconditions.cpp:83-594, as pn_ic_modes)
__device__ void gen_mode(const double *u, int q, unsigned long long m, const IcModeArgs &g,
                         double2 *__restrict__ A, double2 *__restrict__ E) {
  const double phase1 = g.two_pi * u[q];
  const double phase2 = g.two_pi * u[q + 1];
  int r = q + 2;
  double x, y, rr;
  do {
    if (r + 1 >= kU) return;  // flagged by the chunk pass; the call fails
    x = 2.0 * u[r] - 1.0;
    y = 2.0 * u[r + 1] - 1.0;
    rr = sqrt(x * x + y * y);
    r += 2;
  } while (!(rr < 1.0));
  const double c = x / rr, s = y / rr;
  const double gr1 = c * sqrt(-2.0 * log(rr * rr));
  const double gr2 = (s / c) * gr1;
  const double g1 = fabs(gr1), g2 = fabs(gr2);
  const int py = (int)(m % (unsigned long long)g.H);
  const unsigned long long t = m / (unsigned long long)g.H;
  const int px = (int)(t % (unsigned long long)g.N);
  const int pz = (int)(t / (unsigned long long)g.N);
  const double nu = (pz == 0) ? 0 : g.latp[pz];
  const double ppx = g.latp[px];
  const double ppy = g.ppy[py];
  const double pt = sqrt(ppx * ppx + ppy * ppy);
  const double psqr = pt * pt + nu * nu;
  const double pchisqr = pt * pt + nu * nu;
  double occ;
  if (g.justnoise == 0 && pchisqr < g.QS * g.QS) occ = g.amplitude + g.noise;
  else if (g.scale2 == 0) occ = g.noise;
  else occ = (sqrt(pchisqr) < g.scale2) ? g.noise : 0.0;
  const double omega = sqrt(psqr + g.m2_long);
  const double s1 = sqrt(occ) * g1, s2 = sqrt(occ) * g2;
  double sn1, cs1, sn2, cs2;
  sincos(phase1, &sn1, &cs1);
  sincos(phase2, &sn2, &cs2);
  if (fabs(omega) < 1e-8) {
    A[m] = make_double2(0.0, 0.0);
    E[m] = make_double2(0.0, 0.0);
  } else {
    const double so = sqrt(omega);
    A[m] = make_double2((s1 * cs1) / so, (s1 * sn1) / so);
    E[m] = make_double2((s2 * cs2) * so, (s2 * sn2) * so);
  }
}

// pass 3: the modes of each chunk
__global__ __launch_bounds__(kT) void k_icg_modes(IcJump J, unsigned long long x0,
                                                  const unsigned *__restrict__ table,
                                                  const unsigned char *__restrict__ centry,
                                                  const unsigned long long *__restrict__ cbase,
                                                  unsigned long long nm, IcModeArgs g,
                                                  double2 *__restrict__ A, double2 *__restrict__ E) {
  const long long ch = blockIdx.x;
  const unsigned long long base = cbase[ch];
  if (base >= nm) return;  // past the last mode (uniform over the block)
  __shared__ double u[kU];
  __shared__ unsigned long long xs;
  __shared__ unsigned short jt[kLev][kNodes];
  fill_draws(J, x0, ch * kC, u, &xs);
  bool bad = false;
  for (int i = threadIdx.x; i < kNodes; i += kT) jt[0][i] = (unsigned short)next_node(u, i, bad);
  __syncthreads();
  for (int r = 1; r < kLev; r++) {
    for (int i = threadIdx.x; i < kNodes; i += kT) {
      const int n = jt[r - 1][i];
      jt[r][i] = n < kNodes ? jt[r - 1][n] : (unsigned short)n;
    }
    __syncthreads();
  }
  const int e = centry[ch];
  const int n = (int)(table[ch * kWin + e] & 0xFFFFu);
  for (int k = threadIdx.x; k < n; k += kT) {
    const unsigned long long m = base + (unsigned long long)k;
    if (m >= nm) break;
    int p = e;
    for (int b = 0; b < kLev && p < kNodes; b++)
      if ((k >> b) & 1) p = jt[b][p];
    if (p >= kNodes) break;  // only after a flagged node
    gen_mode(u, 2 * p, m, g, A, E);
  }
}

// ensureRealResult for the plane py (py = 0, and py = N/2 for even N)
__global__ __launch_bounds__(256) void k_icg_hermitian(double2 *__restrict__ F, int N, int H, int py) {
  const long long id = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (id >= (long long)N * N) return;
  const int pz = (int)(id / N), px = (int)(id % N);
  const int mpz = (N - pz) % N, mpx = (N - px) % N;
  const long long m = py + (long long)px * H + (long long)pz * N * H;
  if (pz == mpz && px == mpx) {
    const double2 v = F[m];
    F[m] = make_double2(hypot(v.x, v.y) * (v.x > 0 ? 1 : -1), 0.0);
  } else if (pz > mpz || (pz == mpz && px > mpx)) {
    // the mirror mode came first in the pz -> px order and keeps its value
    const double2 v = F[py + (long long)mpx * H + (long long)mpz * N * H];
    F[m] = make_double2(v.x, -v.y);
  }
}

double lat_p_host(int k, int N, double a) {  // momentumspace.cpp:102-118
  const double pi_ = std::acos(-1.);
  return (k < N / 2 + 1) ? 2.0 / (a)*std::sin(k * pi_ / N) : -2.0 / (a)*std::sin(k * pi_ / N);
}

}  // namespace

// Modes A and E of pn_ic_modes for the context's N and a, generated on its
// device into dA and dE (N*N*(N/2+1) double2 each), on its stream.
int pn_icgen_modes(pn_ctx *c, const pn_ic_params *ic, void *dA, void *dE) {
  const int N = c->p.N, H = N / 2 + 1;
  const double a = c->p.a;
  const unsigned long long nm = (unsigned long long)N * N * H;
  // draws: 2 + 2k per mode with k geometric (acceptance pi/4): mean 2 + 8/pi,
  // standard deviation 1.18 per mode; provision 64 standard deviations more
  const double draws = (2.0 + 8.0 / std::acos(-1.)) * (double)nm + 64.0 * 1.18 * std::sqrt((double)nm);
  const long long nchunks = (long long)(draws / kC) + 2;
  const long long ngroups = (nchunks + kG - 1) / kG;
  if (nchunks > 0x7fffffffLL) return pn_fail(c, PN_ERR_ARG, "pn_icgen: lattice too large");

  std::vector<double> tabs((size_t)N + H);
  for (int k = 0; k < N; k++) tabs[k] = lat_p_host(k, N, a);
  for (int py = 0; py < H; py++) tabs[(size_t)N + py] = 2.0 / (a)*std::sin(py * std::acos(-1.) / N);

  void *buf = nullptr;
  const size_t o_table = 0;
  const size_t o_gexit = o_table + (size_t)nchunks * kWin * 4;
  const size_t o_gcnt = (o_gexit + (size_t)ngroups * kWin * 4 + 255) & ~(size_t)255;
  const size_t o_gentry = o_gcnt + (size_t)ngroups * kWin * 8;
  const size_t o_gbase = (o_gentry + (size_t)ngroups * 4 + 255) & ~(size_t)255;
  const size_t o_cbase = o_gbase + (size_t)ngroups * 8;
  const size_t o_centry = o_cbase + (size_t)nchunks * 8;
  const size_t o_tabs = (o_centry + (size_t)nchunks + 255) & ~(size_t)255;
  const size_t o_err = (o_tabs + tabs.size() * 8 + 255) & ~(size_t)255;
  const size_t total = o_err + 256;
  PN_HIP(c, hipMalloc(&buf, total));
  char *B = (char *)buf;
  auto done = [&](int code) {
    (void)hipFree(buf);
    return code;
  };
  hipError_t e = hipMemcpyAsync(B + o_tabs, tabs.data(), tabs.size() * 8, hipMemcpyHostToDevice, c->stream);
  if (e == hipSuccess) e = hipMemsetAsync(B + o_err, 0, 8, c->stream);
  if (e != hipSuccess) return done(pn_fail(c, PN_ERR_HIP, "pn_icgen: %s", hipGetErrorString(e)));

  const IcJump J = make_jump();
  const unsigned long long x0 = ((unsigned long long)(unsigned)ic->seed << 16) | 0x330Eu;  // srand48
  unsigned *table = (unsigned *)(B + o_table);
  int *err = (int *)(B + o_err);
  hipLaunchKernelGGL(k_icg_chunks, dim3((unsigned)nchunks), dim3(kT), 0, c->stream, J, x0, table, err);
  hipLaunchKernelGGL(k_icg_group, dim3((unsigned)ngroups), dim3(kWin), 0, c->stream, (const unsigned *)table,
                     nchunks, (int *)(B + o_gexit), (unsigned long long *)(B + o_gcnt));
  hipLaunchKernelGGL(k_icg_top, dim3(1), dim3(64), 0, c->stream, (const int *)(B + o_gexit),
                     (const unsigned long long *)(B + o_gcnt), ngroups, nm, (int *)(B + o_gentry),
                     (unsigned long long *)(B + o_gbase), err);
  hipLaunchKernelGGL(k_icg_down, dim3((unsigned)((ngroups + 63) / 64)), dim3(64), 0, c->stream,
                     (const unsigned *)table, nchunks, ngroups, (const int *)(B + o_gentry),
                     (const unsigned long long *)(B + o_gbase), (unsigned char *)(B + o_centry),
                     (unsigned long long *)(B + o_cbase));
  IcModeArgs g;
  g.N = N;
  g.H = H;
  g.two_pi = 2.0 * std::acos(-1.);
  g.QS = ic->QS;
  g.scale2 = ic->scale2;
  g.amplitude = ic->amplitude;
  g.noise = ic->noise;
  g.m2_long = ic->m2_long;
  g.justnoise = ic->justnoise;
  g.latp = (const double *)(B + o_tabs);
  g.ppy = (const double *)(B + o_tabs) + N;
  hipLaunchKernelGGL(k_icg_modes, dim3((unsigned)nchunks), dim3(kT), 0, c->stream, J, x0,
                     (const unsigned *)table, (const unsigned char *)(B + o_centry),
                     (const unsigned long long *)(B + o_cbase), nm, g, (double2 *)dA, (double2 *)dE);
  const long long plane = (long long)N * N;
  const unsigned hb = (unsigned)((plane + 255) / 256);
  for (double2 *F : {(double2 *)dA, (double2 *)dE}) {
    hipLaunchKernelGGL(k_icg_hermitian, dim3(hb), dim3(256), 0, c->stream, F, N, H, 0);
    if (N % 2 == 0) hipLaunchKernelGGL(k_icg_hermitian, dim3(hb), dim3(256), 0, c->stream, F, N, H, N / 2);
  }
  e = hipGetLastError();
  int herr[2] = {0, 0};
  if (e == hipSuccess) e = hipMemcpyAsync(herr, err, 8, hipMemcpyDeviceToHost, c->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
  if (e != hipSuccess) return done(pn_fail(c, PN_ERR_HIP, "pn_icgen: %s", hipGetErrorString(e)));
  if (herr[0] || herr[1])
    return done(pn_fail(c, PN_ERR_STATE,
                        "pn_icgen: the draw stream left the generator's window (%d, %d); "
                        "use the host modes (PN_OPT_IC_ON_DEVICE = 0)", herr[0], herr[1]));
  return done(PN_OK);
}

extern "C" int pn_ic_modes_device(pn_ctx *c, const pn_ic_params *ic, double *A, double *E) {
  if (!c || !ic || !A || !E) return pn_fail(c, PN_ERR_ARG, "pn_ic_modes_device: NULL argument");
  PN_HIP(c, hipSetDevice(c->p.device));
  const int N = c->p.N, H = N / 2 + 1;
  const size_t bytes = (size_t)N * N * H * 16;
  void *d = nullptr;
  PN_HIP(c, hipMalloc(&d, 2 * bytes));
  int rc = pn_icgen_modes(c, ic, d, (char *)d + bytes);
  hipError_t e = hipSuccess;
  if (rc == PN_OK) e = hipMemcpy(A, d, bytes, hipMemcpyDeviceToHost);
  if (rc == PN_OK && e == hipSuccess) e = hipMemcpy(E, (char *)d + bytes, bytes, hipMemcpyDeviceToHost);
  (void)hipFree(d);
  if (rc != PN_OK) return rc;
  if (e != hipSuccess) return pn_fail(c, PN_ERR_HIP, "pn_ic_modes_device: %s", hipGetErrorString(e));
  return PN_OK;
}
