// pn_ic.hip -- initial conditions, initialize==3 (initialconditions.cpp:45-723).
//
// The random draws stay on the host so that the stream is the reference's
// exactly: POSIX drand48 seeded by srand48(seed) (the reference seeds with
// time(NULL)*(ID+1), initialconditions.cpp:16; here the seed is a parameter),
// two uniform phases and two |Gauss| amplitudes per mode in the order
// pz -> px -> py (:83-594), then the Hermitian completion of ensureRealResult
// (clusterfft.cpp:55-145).  The c2r transforms run on the device with rocFFT;
// a transpose kernel applies the /Norm, *NormFactor and the background value
// while moving the result into the dense x-fastest layout
// (clusterfft.cpp:148-182, initialconditions.cpp:636-703).
#include <hip/hip_runtime.h>
#include <rocfft/rocfft.h>

#include <cmath>
#include <cstdint>
#include <mutex>
#include <vector>

#include "pn_internal.h"

namespace {

// POSIX drand48: X <- (a X + c) mod 2^48 with a = 0x5DEECE66D, c = 0xB;
// the value is X / 2^48, exact in binary64.  srand48(s) sets X to the low 32
// bits of s followed by 0x330E.
struct Drand48 {
  uint64_t x = 0x330E;
  void seed(long s) { x = ((uint64_t)(uint32_t)s << 16) | 0x330Eu; }
  double next() {
    x = (0x5DEECE66DULL * x + 0xBULL) & 0xFFFFFFFFFFFFULL;
    return std::ldexp((double)x, -48);
  }
};

// polar Box-Muller with the second value cached (initialconditions.cpp:27-41)
struct Gauss {
  Drand48 *r;
  bool have = false;
  double second = 0;
  double next() {
    if (have) {
      have = false;
      return second;
    }
    double x, y, rr;
    do {
      x = 2.0 * r->next() - 1.0;
      y = 2.0 * r->next() - 1.0;
      rr = std::sqrt(x * x + y * y);
    } while (!(rr < 1.0));
    const double c = x / rr, s = y / rr;
    const double g1 = c * std::sqrt(-2.0 * std::log(rr * rr));
    second = (s / c) * g1;
    have = true;
    return g1;
  }
};

double sq(double v) { return v * v; }

// signed lattice momentum of index k (momentumspace.cpp:102-118 and the
// identical expressions in initialconditions.cpp:88-111)
double lat_p(int k, int N, double a) {
  const double pi_ = std::acos(-1.);
  return (k < N / 2 + 1) ? 2.0 / (a)*std::sin(k * pi_ / N) : -2.0 / (a)*std::sin(k * pi_ / N);
}

}  // namespace

// Host part: modes A (field) and E (momentum) in reference p-space order,
// interleaved complex, N*N*(N/2+1) each, Hermitian-completed.
extern "C" int pn_ic_modes(int N, double a, const pn_ic_params *ic, double *A, double *E) {
  if (N < 2 || !(a > 0) || !ic || !A || !E)
    return pn_fail(nullptr, PN_ERR_ARG, "pn_ic_modes: bad argument");
  const int H = N / 2 + 1;
  Drand48 rng;
  rng.seed(ic->seed);
  Gauss gauss{&rng};
  const double two_pi = 2.0 * std::acos(-1.);
  for (int pz = 0; pz < N; pz++) {
    const double nu = (pz == 0) ? 0 : lat_p(pz, N, a);
    for (int px = 0; px < N; px++) {
      const double ppx = lat_p(px, N, a);
      for (int py = 0; py < H; py++) {
        const double ppy = 2.0 / (a)*std::sin(py * std::acos(-1.) / N);
        const double pt = std::sqrt(sq(ppx) + sq(ppy));
        const double psqr = sq(pt) + sq(nu);
        const size_t m = (size_t)py + (size_t)px * H + (size_t)pz * N * H;
        const double phase1 = two_pi * rng.next();
        const double phase2 = two_pi * rng.next();
        const double g1 = std::fabs(gauss.next());
        const double g2 = std::fabs(gauss.next());
        const double pchisqr = sq(pt) + sq(nu);
        double occ;
        if (ic->justnoise == 0 && pchisqr < ic->QS * ic->QS) occ = ic->amplitude + ic->noise;
        else if (ic->scale2 == 0) occ = ic->noise;
        else occ = (std::sqrt(pchisqr) < ic->scale2) ? ic->noise : 0.0;
        const double omega = std::sqrt(psqr + ic->m2_long);
        const double s1 = std::sqrt(occ) * g1, s2 = std::sqrt(occ) * g2;
        // complex(cos(phase), sin(phase)): g++ turns the pair into one sincos
        // call when it builds the reference, and glibc's sincos may differ from
        // separate sin/cos in the last bit, so call it explicitly
        double sn1, cs1, sn2, cs2;
        ::sincos(phase1, &sn1, &cs1);
        ::sincos(phase2, &sn2, &cs2);
        if (std::fabs(omega) < 1e-8) {
          A[2 * m] = A[2 * m + 1] = E[2 * m] = E[2 * m + 1] = 0.0;
        } else {
          const double so = std::sqrt(omega);
          A[2 * m] = (s1 * cs1) / so;
          A[2 * m + 1] = (s1 * sn1) / so;
          E[2 * m] = (s2 * cs2) * so;
          E[2 * m + 1] = (s2 * sn2) * so;
        }
      }
    }
  }
  // ensureRealResult at one rank, first-visited mode of each pair wins
  for (double *F : {A, E}) {
    for (int pz = 0; pz < N; pz++)
      for (int px = 0; px < N; px++)
        for (int py = 0; py < H; py++) {
          const int mpy = (N - py) % N;
          if (mpy != py) continue;
          const int mpz = (N - pz) % N, mpx = (N - px) % N;
          const size_t m = (size_t)py + (size_t)px * H + (size_t)pz * N * H;
          if (px == mpx && pz == mpz) {
            const double v = std::hypot(F[2 * m], F[2 * m + 1]) * (F[2 * m] > 0 ? 1 : -1);
            F[2 * m] = v;
            F[2 * m + 1] = 0.0;
          } else {
            const size_t mm = (size_t)mpy + (size_t)mpx * H + (size_t)mpz * N * H;
            F[2 * mm] = F[2 * m];
            F[2 * mm + 1] = -F[2 * m + 1];
          }
        }
  }
  return PN_OK;
}

// X[z][x][y] (rocFFT c2r output) -> dense [z][y][x]:  (X/Norm)*NormFactor + bg
template <typename T>
__global__ __launch_bounds__(256) void k_ic_fill(const double *__restrict__ X, T *__restrict__ out,
                                                 int N, double norm, double normfactor, double bg) {
  __shared__ double tile[32][33];
  const int bx = blockIdx.x * 32, by = blockIdx.y * 32, z = blockIdx.z;
  const long long base = (long long)z * N * N;
  for (int j = threadIdx.y; j < 32; j += 8) {
    const int y = bx + threadIdx.x, x = by + j;  // X is [z][x][y]: y fastest
    if (x < N && y < N) tile[j][threadIdx.x] = X[base + (long long)x * N + y];
  }
  __syncthreads();
  for (int j = threadIdx.y; j < 32; j += 8) {
    const int x = by + threadIdx.x, y = bx + j;
    if (x < N && y < N) out[base + (long long)y * N + x] = (T)((tile[threadIdx.x][j] / norm) * normfactor + bg);
  }
}

static std::once_flag g_fft_once;
static rocfft_status g_fft_setup = rocfft_status_success;

extern "C" int pn_initial_conditions(pn_ctx *c, const pn_ic_params *ic) {
  if (!c || !ic) return pn_fail(c, PN_ERR_ARG, "pn_initial_conditions: NULL argument");
  PN_HIP(c, hipSetDevice(c->p.device));
  std::call_once(g_fft_once, [] { g_fft_setup = rocfft_setup(); });
  if (g_fft_setup != rocfft_status_success) return pn_fail(c, PN_ERR_FFT, "rocfft_setup failed");
  const int N = c->p.N, H = N / 2 + 1;
  const double a = c->p.a;
  const size_t nm = (size_t)N * N * H;
  // modes: drawn on the device (pn_icgen.hip) into dmodes[0..1], or on the
  // host and uploaded one field at a time into dmodes[0]
  std::vector<double> A, E;
  if (!c->ic_device) {
    A.resize(2 * nm);
    E.resize(2 * nm);
    int rc = pn_ic_modes(N, a, ic, A.data(), E.data());
    if (rc != PN_OK) return pn_fail(c, rc, "%s", pn_last_error(nullptr));
  }

  rocfft_plan plan = nullptr;
  rocfft_execution_info info = nullptr;
  void *work = nullptr, *dmodes[2] = {nullptr, nullptr}, *dreal = nullptr;
  auto cleanup = [&](int code) {
    if (plan) rocfft_plan_destroy(plan);
    if (info) rocfft_execution_info_destroy(info);
    (void)hipFree(work);
    (void)hipFree(dmodes[0]);
    (void)hipFree(dmodes[1]);
    (void)hipFree(dreal);
    return code;
  };
  size_t lengths[3] = {(size_t)N, (size_t)N, (size_t)N};
  if (rocfft_plan_create(&plan, rocfft_placement_notinplace, rocfft_transform_type_real_inverse,
                         rocfft_precision_double, 3, lengths, 1, nullptr) != rocfft_status_success)
    return cleanup(pn_fail(c, PN_ERR_FFT, "pn_initial_conditions: c2r plan failed"));
  size_t wsize = 0;
  rocfft_plan_get_work_buffer_size(plan, &wsize);
  if (rocfft_execution_info_create(&info) != rocfft_status_success)
    return cleanup(pn_fail(c, PN_ERR_FFT, "pn_initial_conditions: execution info failed"));
  hipError_t e = hipSuccess;
  if (wsize) e = hipMalloc(&work, wsize);
  if (e == hipSuccess) e = hipMalloc(&dmodes[0], nm * 16);
  if (e == hipSuccess && c->ic_device) e = hipMalloc(&dmodes[1], nm * 16);
  if (e == hipSuccess) e = hipMalloc(&dreal, (size_t)c->n3 * 8);
  if (e != hipSuccess) return cleanup(pn_fail(c, PN_ERR_NOMEM, "pn_initial_conditions: %s", hipGetErrorString(e)));
  if (c->ic_device) {
    int rc = pn_icgen_modes(c, ic, dmodes[0], dmodes[1]);
    if (rc != PN_OK) return cleanup(rc);
  }
  if (wsize) rocfft_execution_info_set_work_buffer(info, work, wsize);
  rocfft_execution_info_set_stream(info, c->stream);
  const double norm = ((a * N) * (a * N)) * N * a;        // clusterfft.cpp:150
  const double normfactor = std::sqrt(((N * a) * (N * a)) * N * a);  // initialconditions.cpp:643
  dim3 grid((N + 31) / 32, (N + 31) / 32, N), block(32, 8);
  for (int f = 0; f < 2; f++) {
    void *dst = f == 0 ? c->phi[c->cur] : c->pi;
    const double bg = f == 0 ? ic->phi0 : ic->pi0;
    void *src = c->ic_device ? dmodes[f] : dmodes[0];
    if (!c->ic_device) {
      e = hipMemcpyAsync(src, (f == 0 ? A : E).data(), nm * 16, hipMemcpyHostToDevice, c->stream);
      if (e != hipSuccess) return cleanup(pn_fail(c, PN_ERR_HIP, "pn_initial_conditions: %s", hipGetErrorString(e)));
    }
    void *in[1] = {src}, *out[1] = {dreal};
    if (rocfft_execute(plan, in, out, info) != rocfft_status_success)
      return cleanup(pn_fail(c, PN_ERR_FFT, "pn_initial_conditions: c2r failed"));
    if (c->p.precision == PN_FP64)
      hipLaunchKernelGGL(k_ic_fill<double>, grid, block, 0, c->stream, (const double *)dreal,
                         (double *)dst, N, norm, normfactor, bg);
    else
      hipLaunchKernelGGL(k_ic_fill<float>, grid, block, 0, c->stream, (const double *)dreal,
                         (float *)dst, N, norm, normfactor, bg);
    e = hipGetLastError();
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    if (e != hipSuccess) return cleanup(pn_fail(c, PN_ERR_HIP, "pn_initial_conditions: %s", hipGetErrorString(e)));
  }
  return cleanup(PN_OK);
}
