// pn_adjoint.hip -- reverse-mode gradient of a smooth oscillon energy fraction
// through the leapfrog (no reference counterpart).
//
// Forward model: the theory-0 leapfrog of dynamics.cpp:52-237 with the force
// written as a polynomial with continuous couplings,
//   pi  <- pi + (Lap phi - ((c2 + c4 q) + c6 q^2) phi) * dtk[n],   q = phi^2
//   phi <- phi + pi * dtd[n]
// (c2 = Mass_Mode*M^2, c4 = Lambda_Mode/6, c6 = 1/(120 Kappa^2) in reference
// units; c4 = -lambda, c6 = g for V = phi^2/2 - lambda phi^4/4 + g phi^6/6).
//
// Observable after S steps: F = sum rho s(u) / sum rho, u = (rho - t rhobar) /
// (w rhobar), s logistic, rho the energy density summand of Energy
// (fieldanalysis.cpp:625) with V = c2 q/2 + c4 q^2/4 + c6 q^3/6.
//
// Reverse pass, per step n from S-1 down to 0, with pb = pibar + dtd[n]*phibar:
//   phibar <- phibar + dtk[n] * (Lap pb - f'(phi_n) pb),   pibar <- pb,
//   cbar   -= dtk[n] * sum pb * (phi, q phi, q^2 phi)
// pb is formed once per site: the reverse step for n also writes the next
// step's pb = pibar + dtd[n-1]*phibar, so the stencil reads one array.
// phi_n comes from checkpoints of (phi, pi) every K steps: each segment is
// recomputed forward into a K-plane store, then walked backwards.  With
// K >= S every phi_n stays in HBM from the forward pass (S planes, 67 GB for
// 4000 steps at 128^3) and nothing is recomputed.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <vector>

#include "pn_internal.h"
#include "pn_kernels.h"

namespace {

struct AdjCoef {
  int N;
  double a2, c2, c4, c6, t, w;
};

constexpr int kT = 256;
constexpr int kBlocks = 1024;
// planes per block of the z-streaming step kernels (both directions); 2, 4
// and 8 planes ran at the same rate at 128^3
constexpr int kZB = 16;

__device__ __forceinline__ long long nb_idx(long long i, int x, int y, int z, int N, int d, int s) {
  // neighbour of site i in direction d (0 x, 1 y, 2 z) with step s = +-1, periodic
  const long long P = (long long)N * N;
  if (d == 0) {
    int xx = x + s;
    xx = xx < 0 ? N - 1 : (xx == N ? 0 : xx);
    return i - x + xx;
  }
  if (d == 1) {
    int yy = y + s;
    yy = yy < 0 ? N - 1 : (yy == N ? 0 : yy);
    return i + (long long)(yy - y) * N;
  }
  int zz = z + s;
  zz = zz < 0 ? N - 1 : (zz == N ? 0 : zz);
  return i + (long long)(zz - z) * P;
}

// forward step, z-streaming columns of kZB planes, polynomial force
__global__ __launch_bounds__(256) void k_poly_step(const double *__restrict__ phi,
                                                   double *__restrict__ pi,
                                                   double *__restrict__ phin, AdjCoef k, double dtk,
                                                   double dtd) {
  const int N = k.N;
  const int XT = (N + 63) / 64, YT = (N + 3) / 4, ZB = kZB, ZT = (N + ZB - 1) / ZB;
  const int b = blockIdx.x;
  if (b >= XT * YT * ZT) return;
  const int xt = b % XT, yt = (b / XT) % YT, zt = b / (XT * YT);
  const int x = xt * 64 + threadIdx.x, y = yt * 4 + threadIdx.y;
  if (x >= N || y >= N) return;
  const int z0 = zt * ZB, z1 = min(z0 + ZB, N);
  const int xp = x + 1 == N ? 0 : x + 1, xm = x == 0 ? N - 1 : x - 1;
  const int yp = y + 1 == N ? 0 : y + 1, ym = y == 0 ? N - 1 : y - 1;
  const long long P = (long long)N * N;
  const int row = y * N;
  double fm = phi[(long long)(z0 == 0 ? N - 1 : z0 - 1) * P + row + x];
  double fc = phi[(long long)z0 * P + row + x];
  for (int z = z0; z < z1; z++) {
    const long long base = (long long)z * P;
    const double fp = phi[(long long)(z + 1 == N ? 0 : z + 1) * P + row + x];
    double s = phi[base + row + xp] + phi[base + row + xm];
    s = s + phi[base + yp * N + x];
    s = s + phi[base + ym * N + x];
    s = s + fp;
    s = s + fm;
    const double lap = (s - 6.0 * fc) / k.a2;
    const double q = fc * fc;
    const double force = ((k.c2 + k.c4 * q) + k.c6 * (q * q)) * fc;
    const double p = pi[base + row + x] + (lap - force) * dtk;
    pi[base + row + x] = p;
    phin[base + row + x] = fc + p * dtd;
    fm = fc;
    fc = fp;
  }
}

static int poly_blocks(int N) { return ((N + 63) / 64) * ((N + 3) / 4) * ((N + kZB - 1) / kZB); }

__device__ __forceinline__ double site_rho(const double *phi, const double *pi, long long i, int x,
                                           int y, int z, const AdjCoef &k) {
  const double f = phi[i];
  const double dx = phi[nb_idx(i, x, y, z, k.N, 0, 1)] - f;
  const double dy = phi[nb_idx(i, x, y, z, k.N, 1, 1)] - f;
  const double dz = phi[nb_idx(i, x, y, z, k.N, 2, 1)] - f;
  const double q = f * f;
  const double V = ((0.5 * k.c2) * q + (0.25 * k.c4) * (q * q)) + (k.c6 / 6.0) * (q * q * q);
  return 0.5 * (pi[i] * pi[i]) + 0.5 * ((dx * dx + dy * dy + dz * dz) / k.a2) + V;
}

// per-block (hi, lo) partials of three per-site quantities
template <typename Fn>
__device__ void block_sum3(Fn fn, long long n3, double *part) {
  double h[3] = {0, 0, 0}, l[3] = {0, 0, 0};
  for (long long i = (long long)blockIdx.x * kT + threadIdx.x; i < n3; i += (long long)gridDim.x * kT) {
    double v[3];
    fn(i, v);
    for (int j = 0; j < 3; j++) two_acc(h[j], l[j], v[j]);
  }
  __shared__ double sh[kT][6];
  for (int j = 0; j < 3; j++) { sh[threadIdx.x][2 * j] = h[j]; sh[threadIdx.x][2 * j + 1] = l[j]; }
  __syncthreads();
  for (int o = kT / 2; o > 0; o >>= 1) {
    if ((int)threadIdx.x < o)
      for (int j = 0; j < 3; j++)
        two_merge(sh[threadIdx.x][2 * j], sh[threadIdx.x][2 * j + 1], sh[threadIdx.x + o][2 * j],
                  sh[threadIdx.x + o][2 * j + 1]);
    __syncthreads();
  }
  if (threadIdx.x == 0)
    for (int j = 0; j < 6; j++) part[(long long)blockIdx.x * 6 + j] = sh[0][j];
}

// mode 0: (rho, 0, 0), writes rho;  mode 1: (rho s, rho^2 s(1-s), 0) given R;
// mode 2: direct coupling terms of V weighted by g: (g q/2, g q^2/4, g q^3/6)
__global__ __launch_bounds__(kT) void k_obs_sums(const double *__restrict__ phi,
                                                 const double *__restrict__ pi,
                                                 double *__restrict__ rho,
                                                 const double *__restrict__ g, AdjCoef k, int mode,
                                                 double R, double *__restrict__ part) {
  const int N = k.N;
  const long long n3 = (long long)N * N * N, P = (long long)N * N;
  const double rbar = R / (double)n3;
  block_sum3(
      [&](long long i, double *v) {
        if (mode == 0) {
          const int x = (int)(i % N), y = (int)((i / N) % N), z = (int)(i / P);
          const double r = site_rho(phi, pi, i, x, y, z, k);
          rho[i] = r;
          v[0] = r; v[1] = 0; v[2] = 0;
        } else if (mode == 1) {
          const double r = rho[i];
          const double u = (r - k.t * rbar) / (k.w * rbar);
          const double s = 1.0 / (1.0 + exp(-u));
          v[0] = r * s; v[1] = r * r * s * (1.0 - s); v[2] = 0;
        } else {
          const double f = phi[i], q = f * f;
          v[0] = g[i] * q / 2.0; v[1] = g[i] * q * q / 4.0; v[2] = g[i] * q * q * q / 6.0;
        }
      },
      n3, part);
}

__global__ void k_final3(const double *__restrict__ part, int nb, double *__restrict__ out) {
  if (threadIdx.x != 0) return;
  double h[3] = {0, 0, 0}, l[3] = {0, 0, 0};
  for (int b = 0; b < nb; b++)
    for (int j = 0; j < 3; j++) two_merge(h[j], l[j], part[(long long)b * 6 + 2 * j], part[(long long)b * 6 + 2 * j + 1]);
  for (int j = 0; j < 3; j++) out[j] = h[j] + l[j];
}

// g = dF/drho
__global__ void k_obs_seed(const double *__restrict__ rho, AdjCoef k, double R, double F, double A,
                           double *__restrict__ g) {
  const long long n3 = (long long)k.N * k.N * k.N;
  const double rbar = R / (double)n3;
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n3; i += (long long)gridDim.x * blockDim.x) {
    const double r = rho[i];
    const double u = (r - k.t * rbar) / (k.w * rbar);
    const double s = 1.0 / (1.0 + exp(-u));
    g[i] = (s - F) / R + (r * s * (1.0 - s) - A / R) / (k.w * rbar * R);
  }
}

// phibar, pibar at the final step from g (chain rule through rho)
__global__ void k_obs_back(const double *__restrict__ phi, const double *__restrict__ pi,
                           const double *__restrict__ g, AdjCoef k, double *__restrict__ bphi,
                           double *__restrict__ bpi) {
  const int N = k.N;
  const long long n3 = (long long)N * N * N, P = (long long)N * N;
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n3; i += (long long)gridDim.x * blockDim.x) {
    const int x = (int)(i % N), y = (int)((i / N) % N), z = (int)(i / P);
    const double f = phi[i], q = f * f;
    double v = (k.c2 + k.c4 * q + k.c6 * q * q) * f * g[i];
    for (int d = 0; d < 3; d++) {
      const long long up = nb_idx(i, x, y, z, N, d, 1), dn = nb_idx(i, x, y, z, N, d, -1);
      v += ((f - phi[up]) * g[i] + (f - phi[dn]) * g[dn]) / k.a2;
    }
    bphi[i] = v;
    bpi[i] = pi[i] * g[i];
  }
}

// One reverse step as a z-streaming stencil on pb, in the forward step's
// geometry (64 x 4 columns of kZB planes per block, periodic in x, y and z):
//   bphi_o = bphi + dtk (Lap pb - f'(phi_n) pb),   (pibar after the step = pb)
//   pbm    = pb + dtd_prev bphi_o                  (the next step's pb; skipped when pbm is null)
// and the block's coupling partials into its own slot of gacc (one per
// block of the grid).  Per site: the six neighbours of pb summed left to
// right in the order x+, x-, y+, y-, z+, z-, then (s - 6 pb) / a^2.
__global__ __launch_bounds__(256) void k_adj_zstep(const double *__restrict__ phin,
                                                   const double *__restrict__ bphi,
                                                   const double *__restrict__ pb,
                                                   double *__restrict__ bphi_o,
                                                   double *__restrict__ pbm, AdjCoef k, double dtk,
                                                   double dtd_prev, double *__restrict__ gacc) {
  const int N = k.N;
  const int XT = (N + 63) / 64, YT = (N + 3) / 4, ZB = kZB, ZT = (N + ZB - 1) / ZB;
  const int b = blockIdx.x;
  if (b >= XT * YT * ZT) return;  // whole blocks only
  const int xt = b % XT, yt = (b / XT) % YT, zt = b / (XT * YT);
  const int x = xt * 64 + threadIdx.x, y = yt * 4 + threadIdx.y;
  const bool active = x < N && y < N;
  double c0 = 0, c1 = 0, c2 = 0;
  if (active) {
    const int z0 = zt * ZB, z1 = min(z0 + ZB, N);
    const int xp = x + 1 == N ? 0 : x + 1, xm = x == 0 ? N - 1 : x - 1;
    const int yp = y + 1 == N ? 0 : y + 1, ym = y == 0 ? N - 1 : y - 1;
    const long long P = (long long)N * N;
    const int row = y * N;
    double bm = pb[(long long)(z0 == 0 ? N - 1 : z0 - 1) * P + row + x];
    double bc = pb[(long long)z0 * P + row + x];
    for (int z = z0; z < z1; z++) {
      const long long base = (long long)z * P;
      const double bp = pb[(long long)(z + 1 == N ? 0 : z + 1) * P + row + x];
      double s = pb[base + row + xp] + pb[base + row + xm];
      s += pb[base + yp * N + x];
      s += pb[base + ym * N + x];
      s += bp;
      s += bm;
      const double f = phin[base + row + x], q = f * f;
      const double fd = k.c2 + 3.0 * k.c4 * q + 5.0 * k.c6 * q * q;
      const double o = bphi[base + row + x] + dtk * ((s - 6.0 * bc) / k.a2 - fd * bc);
      bphi_o[base + row + x] = o;
      if (pbm) pbm[base + row + x] = bc + dtd_prev * o;
      c0 -= dtk * bc * f;
      c1 -= dtk * bc * q * f;
      c2 -= dtk * bc * q * q * f;
      bm = bc;
      bc = bp;
    }
  }
  __shared__ double sh[256][3];
  const int t = threadIdx.y * 64 + threadIdx.x;
  sh[t][0] = c0;
  sh[t][1] = c1;
  sh[t][2] = c2;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if (t < o)
      for (int j = 0; j < 3; j++) sh[t][j] += sh[t + o][j];
    __syncthreads();
  }
  if (t == 0)
    for (int j = 0; j < 3; j++) gacc[(long long)b * 3 + j] += sh[0][j];
}

// pb for the last step: bpi + dtd bphi
__global__ void k_adj_pb(const double *__restrict__ bphi, const double *__restrict__ bpi, double dtd,
                         long long n3, double *__restrict__ pb) {
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n3; i += (long long)gridDim.x * blockDim.x)
    pb[i] = bpi[i] + dtd * bphi[i];
}

__global__ void k_zero(double *p, long long n) {
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x) p[i] = 0.0;
}

struct AdjRun {
  int device = 0;
  hipStream_t s = nullptr;
  std::vector<void *> allocs;
  long long bytes = 0;
  std::string err;
  ~AdjRun() {
    for (void *p : allocs) (void)hipFree(p);
    if (s) (void)hipStreamDestroy(s);
  }
  double *alloc(size_t n) {
    void *p = nullptr;
    if (hipMalloc(&p, n * sizeof(double)) != hipSuccess) return nullptr;
    allocs.push_back(p);
    bytes += (long long)(n * sizeof(double));
    return (double *)p;
  }
};

int afail(int code, const char *m, hipError_t e = hipSuccess) {
  std::string s = m;
  if (e != hipSuccess) s += std::string(": ") + hipGetErrorString(e);
  return pn_fail(nullptr, code, "%s", s.c_str());
}

#define AD_HIP(call)                                            \
  do {                                                          \
    hipError_t e_ = (call);                                     \
    if (e_ != hipSuccess) return afail(PN_ERR_HIP, #call, e_);  \
  } while (0)

int check_params(const pn_adj_params *p, int nsteps) {
  if (!p || p->N < 4 || p->N > 1024 || !(p->a > 0) || nsteps < 0 || !(p->width > 0))
    return afail(PN_ERR_ARG, "adjoint: bad parameters");
  if (p->checkpoint_every < 1) return afail(PN_ERR_ARG, "adjoint: checkpoint_every must be >= 1");
  return pn_runtime_check();
}

AdjCoef coef(const pn_adj_params *p) {
  AdjCoef k;
  k.N = p->N;
  k.a2 = p->a * p->a;
  k.c2 = p->c2;
  k.c4 = p->c4;
  k.c6 = p->c6;
  k.t = p->threshold;
  k.w = p->width;
  return k;
}

// F (and, if want_seed, g, bphi, bpi and the direct coupling terms) at (phi, pi)
int observable(AdjRun &r, const AdjCoef &k, const double *phi, const double *pi, double *rho,
               double *part, double *out3, double *F, double *g, double *bphi, double *bpi,
               double *cdirect) {
  double h[3];
  hipLaunchKernelGGL(k_obs_sums, dim3(kBlocks), dim3(kT), 0, r.s, phi, pi, rho, (const double *)nullptr, k, 0, 0.0, part);
  hipLaunchKernelGGL(k_final3, dim3(1), dim3(64), 0, r.s, part, kBlocks, out3);
  AD_HIP(hipMemcpyAsync(h, out3, 3 * sizeof(double), hipMemcpyDeviceToHost, r.s));
  AD_HIP(hipStreamSynchronize(r.s));
  const double R = h[0];
  hipLaunchKernelGGL(k_obs_sums, dim3(kBlocks), dim3(kT), 0, r.s, phi, pi, rho, (const double *)nullptr, k, 1, R, part);
  hipLaunchKernelGGL(k_final3, dim3(1), dim3(64), 0, r.s, part, kBlocks, out3);
  AD_HIP(hipMemcpyAsync(h, out3, 3 * sizeof(double), hipMemcpyDeviceToHost, r.s));
  AD_HIP(hipStreamSynchronize(r.s));
  *F = h[0] / R;
  if (!g) return PN_OK;
  const double A = h[1];
  hipLaunchKernelGGL(k_obs_seed, dim3(kBlocks), dim3(kT), 0, r.s, rho, k, R, *F, A, g);
  hipLaunchKernelGGL(k_obs_back, dim3(kBlocks), dim3(kT), 0, r.s, phi, pi, g, k, bphi, bpi);
  hipLaunchKernelGGL(k_obs_sums, dim3(kBlocks), dim3(kT), 0, r.s, phi, pi, rho, (const double *)g, k, 2, R, part);
  hipLaunchKernelGGL(k_final3, dim3(1), dim3(64), 0, r.s, part, kBlocks, out3);
  AD_HIP(hipMemcpyAsync(cdirect, out3, 3 * sizeof(double), hipMemcpyDeviceToHost, r.s));
  AD_HIP(hipStreamSynchronize(r.s));
  return PN_OK;
}

}  // namespace

extern "C" int pn_forward_observable(const pn_adj_params *p, const double *phi0, const double *pi0,
                                     int nsteps, const double *dtk, const double *dtd, double *F) {
  int rc = check_params(p, nsteps);
  if (rc != PN_OK) return rc;
  if (!phi0 || !pi0 || !F || (nsteps > 0 && (!dtk || !dtd))) return afail(PN_ERR_ARG, "pn_forward_observable: NULL argument");
  AdjRun r;
  AD_HIP(hipSetDevice(p->device));
  AD_HIP(hipStreamCreateWithFlags(&r.s, hipStreamNonBlocking));
  const AdjCoef k = coef(p);
  const size_t n3 = (size_t)p->N * p->N * p->N;
  double *phi[2] = {r.alloc(n3), r.alloc(n3)}, *pi = r.alloc(n3), *rho = r.alloc(n3);
  double *part = r.alloc(kBlocks * 6), *out3 = r.alloc(3);
  if (!phi[0] || !phi[1] || !pi || !rho || !part || !out3) return afail(PN_ERR_NOMEM, "pn_forward_observable: allocation");
  AD_HIP(hipMemcpyAsync(phi[0], phi0, n3 * 8, hipMemcpyHostToDevice, r.s));
  AD_HIP(hipMemcpyAsync(pi, pi0, n3 * 8, hipMemcpyHostToDevice, r.s));
  int cur = 0;
  for (int n = 0; n < nsteps; n++) {
    hipLaunchKernelGGL(k_poly_step, dim3(poly_blocks(p->N)), dim3(64, 4), 0, r.s, phi[cur], pi, phi[cur ^ 1], k, dtk[n], dtd[n]);
    cur ^= 1;
  }
  AD_HIP(hipGetLastError());
  return observable(r, k, phi[cur], pi, rho, part, out3, F, nullptr, nullptr, nullptr, nullptr);
}

extern "C" int pn_adjoint(const pn_adj_params *p, const double *phi0, const double *pi0, int nsteps,
                          const double *dtk, const double *dtd, double *grad_phi0, double *grad_pi0,
                          pn_adj_result *res) {
  int rc = check_params(p, nsteps);
  if (rc != PN_OK) return rc;
  if (!phi0 || !pi0 || !res || (nsteps > 0 && (!dtk || !dtd))) return afail(PN_ERR_ARG, "pn_adjoint: NULL argument");
  AdjRun r;
  AD_HIP(hipSetDevice(p->device));
  AD_HIP(hipStreamCreateWithFlags(&r.s, hipStreamNonBlocking));
  const AdjCoef k = coef(p);
  const int N = p->N, K = p->checkpoint_every;
  const size_t n3 = (size_t)N * N * N;
  const int nseg = (nsteps + K - 1) / K;
  // K >= nsteps: every phi_n (n < nsteps) kept from the forward pass, no
  // checkpoints and no recompute
  const bool resident = nsteps > 0 && K >= nsteps;
  // checkpoints of (phi, pi) at steps 0, K, 2K, ...; a K-plane store of phi_n
  double *ck = resident ? nullptr : r.alloc(n3 * 2 * std::max(nseg, 1));
  double *store = r.alloc(n3 * (size_t)std::min(K, std::max(nsteps, 1)));
  double *phi[2] = {r.alloc(n3), r.alloc(n3)}, *pi = r.alloc(n3);
  double *rho = r.alloc(n3), *g = r.alloc(n3);
  double *bphi[2] = {r.alloc(n3), r.alloc(n3)}, *bpi[2] = {r.alloc(n3), r.alloc(n3)};
  const int pb = poly_blocks(N);  // blocks of the step kernels, one coupling slot each
  double *part = r.alloc(kBlocks * 6), *out3 = r.alloc(3), *gacc = r.alloc((size_t)pb * 3);
  if ((!resident && !ck) || !store || !phi[0] || !phi[1] || !pi || !rho || !g || !bphi[0] || !bphi[1] ||
      !bpi[0] || !bpi[1] || !part || !out3 || !gacc)
    return afail(PN_ERR_NOMEM, "pn_adjoint: device allocation");
  hipEvent_t e0, e1, e2;
  AD_HIP(hipEventCreate(&e0));
  AD_HIP(hipEventCreate(&e1));
  AD_HIP(hipEventCreate(&e2));
  AD_HIP(hipMemcpyAsync(resident ? store : phi[0], phi0, n3 * 8, hipMemcpyHostToDevice, r.s));
  AD_HIP(hipMemcpyAsync(pi, pi0, n3 * 8, hipMemcpyHostToDevice, r.s));
  AD_HIP(hipStreamSynchronize(r.s));
  // forward with checkpoints, or through the store
  AD_HIP(hipEventRecord(e0, r.s));
  int cur = 0;
  if (resident) {
    for (int n = 0; n < nsteps; n++) {
      double *dst = n + 1 < nsteps ? store + (size_t)(n + 1) * n3 : phi[0];
      hipLaunchKernelGGL(k_poly_step, dim3(pb), dim3(64, 4), 0, r.s, store + (size_t)n * n3, pi, dst, k, dtk[n],
                         dtd[n]);
    }
  } else {
    for (int n = 0; n < nsteps; n++) {
      if (n % K == 0) {
        double *c = ck + (size_t)(n / K) * 2 * n3;
        AD_HIP(hipMemcpyAsync(c, phi[cur], n3 * 8, hipMemcpyDeviceToDevice, r.s));
        AD_HIP(hipMemcpyAsync(c + n3, pi, n3 * 8, hipMemcpyDeviceToDevice, r.s));
      }
      hipLaunchKernelGGL(k_poly_step, dim3(pb), dim3(64, 4), 0, r.s, phi[cur], pi, phi[cur ^ 1], k, dtk[n], dtd[n]);
      cur ^= 1;
    }
  }
  AD_HIP(hipGetLastError());
  double F = 0, cdir[3] = {0, 0, 0};
  rc = observable(r, k, phi[cur], pi, rho, part, out3, &F, g, bphi[0], bpi[0], cdir);
  if (rc != PN_OK) return rc;
  AD_HIP(hipEventRecord(e1, r.s));
  // backward: segments in reverse.  bphi[bc] holds phibar; bpi[pc] holds pb
  // of the step about to run (after the last one: pibar of step 0)
  hipLaunchKernelGGL(k_zero, dim3(16), dim3(256), 0, r.s, gacc, (long long)pb * 3);
  int bc = 0, pc = 0;
  if (nsteps > 0) {
    hipLaunchKernelGGL(k_adj_pb, dim3(kBlocks), dim3(kT), 0, r.s, bphi[0], bpi[0], dtd[nsteps - 1], (long long)n3,
                       bpi[1]);
    pc = 1;
  }
  for (int sg = nseg - 1; sg >= 0; sg--) {
    const int n0 = sg * K, n1 = std::min(nsteps, n0 + K);
    if (!resident) {
      // recompute phi_{n0..n1-1} into the store; pi evolves in the scratch pi
      const double *c = ck + (size_t)sg * 2 * n3;
      AD_HIP(hipMemcpyAsync(store, c, n3 * 8, hipMemcpyDeviceToDevice, r.s));
      AD_HIP(hipMemcpyAsync(pi, c + n3, n3 * 8, hipMemcpyDeviceToDevice, r.s));
      for (int n = n0; n < n1 - 1; n++) {
        double *src = store + (size_t)(n - n0) * n3, *dst = store + (size_t)(n - n0 + 1) * n3;
        hipLaunchKernelGGL(k_poly_step, dim3(pb), dim3(64, 4), 0, r.s, src, pi, dst, k, dtk[n], dtd[n]);
      }
    }
    for (int n = n1 - 1; n >= n0; n--) {
      double *next = n > 0 ? bpi[pc ^ 1] : nullptr;
      hipLaunchKernelGGL(k_adj_zstep, dim3(pb), dim3(64, 4), 0, r.s, store + (size_t)(n - n0) * n3, bphi[bc],
                         bpi[pc], bphi[bc ^ 1], next, k, dtk[n], n > 0 ? dtd[n - 1] : 0.0, gacc);
      bc ^= 1;
      if (n > 0) pc ^= 1;
    }
    AD_HIP(hipGetLastError());
  }
  AD_HIP(hipEventRecord(e2, r.s));
  std::vector<double> ga((size_t)pb * 3);
  AD_HIP(hipMemcpyAsync(ga.data(), gacc, ga.size() * 8, hipMemcpyDeviceToHost, r.s));
  if (grad_phi0) AD_HIP(hipMemcpyAsync(grad_phi0, bphi[bc], n3 * 8, hipMemcpyDeviceToHost, r.s));
  if (grad_pi0) AD_HIP(hipMemcpyAsync(grad_pi0, bpi[pc], n3 * 8, hipMemcpyDeviceToHost, r.s));
  AD_HIP(hipStreamSynchronize(r.s));
  float fms = 0, bms = 0;
  AD_HIP(hipEventElapsedTime(&fms, e0, e1));
  AD_HIP(hipEventElapsedTime(&bms, e1, e2));
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  (void)hipEventDestroy(e2);
  for (int j = 0; j < 3; j++) {
    double hi = cdir[j], lo = 0;
    for (int b = 0; b < pb; b++) {
      const double x = ga[(size_t)b * 3 + j], s = hi + x, bp = s - hi;
      lo += (hi - (s - bp)) + (x - bp);
      hi = s;
    }
    res->grad_c[j] = hi + lo;
  }
  res->F = F;
  res->forward_ms = fms;
  res->backward_ms = bms;
  res->peak_bytes = r.bytes;
  return PN_OK;
}
