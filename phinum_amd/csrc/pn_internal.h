// pn_internal.h -- context layout and helpers shared by the HIP sources.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdarg>
#include <cstdio>
#include <string>
#include <vector>

#include "../../include/phinum_gpu.h"

struct pn_fft_state;     // pn_spectrum.hip
struct pn_island_state;  // pn_islands.hip

// Coefficients of the theory-0 force, computed once on the host in the
// reference's own expression order (dynamics.cpp:84):
//   force = ((c2 + lmode*(q/6.0)) + ((q/kappa)*(q/kappa))/120.0) * phi
struct pn_coef {
  double a2;       // sqr(a_t)
  double inv_a2;   // 1/a2, used only when a2 is a power of two (exact)
  int a2_pow2;     // division by a2 may be replaced by the multiplication
  double c2;       // Mass_Mode * BareM_sqr_long
  double lmode;    // (double)Lambda_Mode
  double kappa;
};

struct pn_ctx {
  pn_params p;
  pn_coef k;
  long long n3;        // N^3 sites
  size_t esize;        // 8 (FP64) or 4 (FP32)
  void *phi[2];        // ping-pong field buffers, dense x-fastest
  void *pi;            // momenta, updated in place
  void *pi_alt;        // second Pi buffer for the two-step kernel (or NULL)
  int step2;           // pn_evolve advances pairs of steps in one pass
  int lds_max;         // LDS bytes a block may use on this device
  int ic_device = 1;   // initial-condition modes drawn on the device (else on the host)
  // multi-step kernels: the caller's choices (0 / -1: automatic) and the ones
  // the first long evolve call made by timing (pn_core.hip, tune_multistep)
  int step3_mode = -1;   // three-step kernel: -1 automatic, 0 off, 1 on
  int step2_zb_user = 0, step3_zb_user = 0;  // z blocks fixed by the caller
  int step2_zb_auto = 0, step3_zb_auto = 0;  // z blocks chosen by timing
  int step3_auto = 0;    // timing chose the three-step kernel
  int tuned = 0;         // the timing has run
  // timing in progress, carried across evolve calls: candidates as
  // (steps per pass, z block), each one's fastest time per step so far, and
  // the next (round, candidate) to time
  std::vector<int> tune_fam, tune_zb;
  std::vector<float> tune_ms;
  int tune_pos = 0;
  int cur;             // phi[cur] holds the current field
  hipStream_t stream;
  // reduction scratch
  double *red_part;    // per-block partial sums
  double *red_out;     // final sums (device)
  long long *ext_part; // per-block extrema
  int *tile_flag;      // per-tile flags of the volume sums' main pass
  double *host_pin;    // pinned host staging for small results
  int red_blocks;
  pn_fft_state *fft;
  pn_island_state *isl;
  std::string err;
  // a view aliases one realization of a batch: it owns neither the field
  // buffers nor the stream, and cannot step on its own (the batch steps it)
  int is_view;
};

// builds a view context over existing device buffers (pn_batch.hip)
int pn_make_view(const pn_params *p, void *phi0, void *phi1, void *pi, int cur, hipStream_t s,
                 pn_ctx **out);

// initialize==3 modes of pn_ic_modes generated on the device (pn_icgen.hip):
// dA, dE hold N*N*(N/2+1) double2 each; runs on c->stream and waits for it
int pn_icgen_modes(pn_ctx *c, const pn_ic_params *ic, void *dA, void *dE);

// The HIP runtime and rocFFT loaded into the process must be the versions the
// library was compiled against (PN_OK, or PN_ERR_STATE with the versions in
// the message); checked once per process, before any device work
int pn_runtime_check();

// error helpers ------------------------------------------------------------
int pn_fail(pn_ctx *c, int code, const char *fmt, ...);
void pn_set_thread_error(const char *msg);

#define PN_HIP(c, call)                                                                 \
  do {                                                                                  \
    hipError_t e_ = (call);                                                             \
    if (e_ != hipSuccess)                                                               \
      return pn_fail((c), PN_ERR_HIP, "%s failed: %s (%s:%d)", #call, hipGetErrorString(e_), \
                     __FILE__, __LINE__);                                               \
  } while (0)

#define PN_CHECK_LAUNCH(c) PN_HIP(c, hipGetLastError())

static inline long long pn_ref_pos(const pn_ctx *c, long long dense) {
  // dense interior index -> reference pos() index (one z halo plane below)
  return dense + (long long)c->p.N * c->p.N;
}

// The peak of island `id` of the context's last pn_islands call as the
// reference's breadth-first search finds it from site seed_pos (pos()
// numbering): the first site holding the island's largest key in that
// order (islands.cpp:90-279).  Positions in pos() numbering.
int pn_islands_seeded_peak(pn_ctx *c, int id, long long seed_pos, long long *peak_pos, double *peak_value);

// shared by several translation units
int pn_reduce_blocks(const pn_ctx *c);
void pn_free_spectrum(pn_ctx *c);
void pn_free_islands(pn_ctx *c);
