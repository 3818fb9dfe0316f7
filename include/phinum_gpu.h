/*
 * phinum_gpu.h -- C ABI of the MI355X (gfx950) lattice hot path of PhiNum.
 *
 * Plain C types only (no torch, no HIP types).  One context owns all device
 * memory, FFT plans and one HIP stream; the caller owns every host array and
 * every output struct.  Every call returns PN_OK (0) or a negative PN_ERR_*
 * code and leaves a message for pn_last_error().  Calls are synchronous at
 * return unless the name says otherwise.  One context per host thread.
 *
 * Lattice conventions follow the reference (one scalar component, Nc = 1):
 *   pos(x,y,z) = x + N*y + N*N*z                      (util.cpp:34-37)
 *   PN_LAYOUT_REF   : N*N*(N+2) values, interior planes z = 1..N plus the two
 *                     z halo planes 0 and N+1          (lattice.cpp:70-104)
 *   PN_LAYOUT_DENSE : N*N*N values, interior only, same x-fastest order
 * On the device the fields are stored dense (structure of arrays), and the
 * periodic wrap in x, y and z is index arithmetic inside the kernels.
 * Site indices returned by the library are reference pos() indices
 * (PN_LAYOUT_REF numbering) so they can be compared with the reference.
 */
#ifndef PHINUM_GPU_H
#define PHINUM_GPU_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#define PN_ABI_VERSION 1

#define PN_OK 0
#define PN_ERR_ARG (-1)      /* bad argument or size                        */
#define PN_ERR_HIP (-2)      /* HIP runtime error (message has the HIP text) */
#define PN_ERR_FFT (-3)      /* rocFFT error                                */
#define PN_ERR_STATE (-4)    /* call not valid in the context's state       */
#define PN_ERR_NOMEM (-5)    /* host or device allocation failed            */
#define PN_ERR_CAPACITY (-6) /* caller-provided output array too small      */

enum pn_precision { PN_FP64 = 0, PN_FP32 = 1 };
enum pn_layout { PN_LAYOUT_REF = 0, PN_LAYOUT_DENSE = 1 };

/* Theory 0 (phi^4-phi^6), no expansion.  Replaces the compile-time switches
 * and mutable globals of parameters.cpp:84-100,184-201 and the bare mass of
 * renormalization.cpp:3-31 (BareM_sqr_long = M_sqr). */
typedef struct pn_params {
  int N;           /* N_t = N_eta, lattice extent (>= 4)                     */
  double a;        /* a_t = a_eta, lattice spacing                           */
  int mass_mode;   /* Mass_Mode (+1, 0, -1)                                  */
  double m2;       /* BareM_sqr_long                                         */
  int lambda_mode; /* Lambda_Mode (+1, 0, -1)                                */
  double kappa;    /* Kappa                                                  */
  int precision;   /* enum pn_precision: FP64, or FP32 fields with FP64 sums */
  int device;      /* HIP device ordinal                                     */
} pn_params;

typedef struct pn_ctx pn_ctx;

/* Volume averages, fieldanalysis.cpp:399-699 (LocalLattice::avg*). */
typedef struct pn_averages {
  double phi, pi, phi2, pi2, phipi, phi4, phi6;
  double energy, interaction, charge;
} pn_averages;

/* Extrema, fieldanalysis.cpp:319-397 (GlminPhi, GlmaxPhi, GlmaxAbsPhi and the
 * first raster position of each, as pos() indices; -1 when never set). */
typedef struct pn_extrema {
  double min, max, maxabs;
  long long minpos, maxpos, maxabspos;
} pn_extrema;

/* Binned shell spectrum, momentumspace.cpp:140-278 (TotalCorrF/K/Q and
 * TotalBCounts).  Arrays are caller-allocated with pn_spectrum_nbins()
 * entries.  binsize is pTmin (momentumspace.cpp:40-42). */
typedef struct pn_spectrum {
  int nbins;
  double binsize;
  int *counts;
  double *corr_f; /* sum |phi_k|^2 per shell   */
  double *corr_k; /* sum |pi_k|^2 per shell    */
  double *corr_q; /* sum Re(phi_k pi_k^*)      */
} pn_spectrum;

/* One island (islands.cpp:13-28, filled as BFS does at :240-264).  Islands
 * are numbered in the order of their smallest site index, which is the order
 * in which the reference's raster scan (islands.cpp:827-835) discovers them
 * when no islands are tracked yet. */
typedef struct pn_island {
  long long index;      /* island number (0,1,2,... in discovery order)     */
  long long nn;         /* number of sites                                  */
  long long first_pos;  /* smallest pos() index of the island               */
  long long peak_pos;   /* pos() index of the peak (max phi^2, or Delta)    */
  double volume;        /* nn * a^3                                         */
  double peak_value;    /* Phi (mode 0) or Delta (mode 1) at the peak       */
  double avg_phi;       /* sum phi / nn                                     */
  double pot_en, kin_en, grad_en;  /* each times a^3                        */
  long long sum_dx, sum_dy, sum_dz;  /* sum of site offsets from the peak:
                           x, y minimum-image (periodic), z plain (open)     */
  double cx, cy, cz;    /* centroid = peak + sum_d / nn (x, y wrapped)      */
  long long peak_ties;  /* sites (besides the peak) with the same peak value */
} pn_island;

typedef struct pn_catalog {
  int capacity;        /* entries available in islands[]                    */
  int count;           /* islands found (may exceed capacity: PN_ERR_CAPACITY) */
  double threshold;    /* threshold used                                    */
  pn_island *islands;
  int *labels;         /* optional, N^3 dense: island index or -1           */
} pn_catalog;

/* Initial conditions, initialize==3 with withcharge==0 and no expansion
 * (parameters.cpp:279-303, initialconditions.cpp:45-723). */
typedef struct pn_ic_params {
  long long seed;   /* srand48 seed (the reference uses time(NULL)*(ID+1))     */
  double QS;        /* QS                                                      */
  double scale2;    /* scale2 (-Q2)                                            */
  double amplitude; /* Amplitude                                               */
  double noise;     /* Noise after CompleteParameters (times Lambda)           */
  double m2_long;   /* M2_long (= M_sqr, renormalization.cpp:17)               */
  int justnoise;    /* justnoise                                               */
  double phi0;      /* phi0_0, added after the transform                       */
  double pi0;       /* pi0_0                                                   */
} pn_ic_params;

/* ---- context ---------------------------------------------------------- */

/* LocalLattice::init (lattice.cpp:70-104). */
int pn_create(const pn_params *p, pn_ctx **out);
/* LocalLattice::fin (lattice.cpp:108-117). */
int pn_destroy(pn_ctx *c);
/* Replaces the theory-0 couplings after creation, as fix_tadpole does for
 * BareM_sqr_long (renormalization.cpp:3-31) and GetParameters for the rest
 * (parameters.cpp:316-321). */
int pn_set_couplings(pn_ctx *c, int mass_mode, double m2, int lambda_mode, double kappa);
/* Text of the last error on this context (or of the last failed pn_create
 * in this thread when c is NULL). */
const char *pn_last_error(const pn_ctx *c);
int pn_abi_version(void);

/* ---- fields ----------------------------------------------------------- */

/* Host -> device copy of Phi and Pi (either may be NULL to keep the device
 * copy).  Replaces direct writes to LocalLattice::Phi/Pi followed by
 * Communication::exchange (initialconditions.cpp:636-703).  FP32 contexts
 * round to nearest. */
int pn_set_fields(pn_ctx *c, const double *phi, const double *pi, int layout);
/* Device -> host copy.  PN_LAYOUT_REF also fills the halo planes as
 * Communication::exchange leaves them (comm.cpp:63-78). */
int pn_get_fields(pn_ctx *c, double *phi, double *pi, int layout);

/* Host part of SetInitialConditions (initialconditions.cpp:45-614): the seeded
 * drand48 draws per mode in the order pz -> px -> py and the Hermitian
 * completion of ensureRealResult (clusterfft.cpp:55-145).  A (field) and E
 * (momentum) receive N*N*(N/2+1) interleaved complex modes in the reference
 * order py + px*(N/2+1) + pz*N*(N/2+1).  No device work. */
int pn_ic_modes(int N, double a, const pn_ic_params *ic, double *A, double *E);
/* SetInitialConditions (initialconditions.cpp:45-723): pn_ic_modes on the
 * host, rocFFT c2r on the device, /Norm (clusterfft.cpp:150) times NormFactor
 * (initialconditions.cpp:643) plus the background, into Phi and Pi. */
int pn_initial_conditions(pn_ctx *c, const pn_ic_params *ic);
/* The modes of pn_ic_modes for the context's N and a, drawn on its device and
 * copied into the caller's A and E (no reference counterpart: the same
 * drand48 stream, polar Box-Muller draws and ensureRealResult completion,
 * found in parallel; the draws and mode positions are exact, log, sincos and
 * hypot are the device's and may differ from the host's in the last bit). */
int pn_ic_modes_device(pn_ctx *c, const pn_ic_params *ic, double *A, double *E);

/* ---- dynamics (dynamics.cpp) ------------------------------------------ */

/* updateMomenta (dynamics.cpp:52-175) with dtk = tau - Last_Momenta_update. */
int pn_kick(pn_ctx *c, double dtk);
/* updateFields (dynamics.cpp:177-194) with dtd = tau - Last_Field_update. */
int pn_drift(pn_ctx *c, double dtd);
/* evolve() body (dynamics.cpp:196-237): kick, drift and the periodic wrap of
 * Communication::exchange fused into one launch per step.  dtk[n], dtd[n]
 * are the host-replayed factors of step n.  Enqueued; pn_sync waits. */
int pn_evolve(pn_ctx *c, int nsteps, const double *dtk, const double *dtd);
int pn_sync(pn_ctx *c);

/* Options (no reference counterpart).
 * PN_OPT_TWO_STEP: 1 (default where the lattice fits: FP64 N <= 512, FP32
 * N <= 1024, N a multiple of the vector width) makes pn_evolve advance pairs
 * of steps in one pass over HBM (temporal blocking); 0 makes every step its
 * own launch.  Both give bit-identical fields. */
#define PN_OPT_TWO_STEP 1
/* PN_OPT_IC_ON_DEVICE: 1 (default) makes pn_initial_conditions draw the modes
 * on the device (pn_ic_modes_device); 0 draws them on the host (pn_ic_modes)
 * and uploads them. */
#define PN_OPT_IC_ON_DEVICE 2
/* Multi-step kernel choice.  pn_evolve calls of at least 12 steps time the
 * candidates on their own steps: the two-step kernel at z blocks 8 ... 64
 * and, where it fits (FP64 N <= 256 even, FP32 N <= 512 divisible by 4), the
 * three-step kernel (triples of steps per pass over HBM) at z blocks
 * 16 ... 64, as many per call as its steps allow (up to 118 steps in all at
 * 256^3), and the fastest is kept once every candidate has been timed;
 * every choice gives the same bits.  The options read the choice in use or
 * fix it (setting any of them times the rest again):
 * PN_OPT_STEP2_ZBLOCK: planes per two-step tile (0: chosen by timing).
 * PN_OPT_THREE_STEP: 1 on, 0 off, -1 chosen by timing (the default); reads 1
 *   when pn_evolve advances triples.
 * PN_OPT_STEP3_ZBLOCK: planes per three-step tile (0: chosen by timing).
 * PN_OPT_KERNEL_TIMED: read only; 1 once the timed choice has been made. */
#define PN_OPT_STEP2_ZBLOCK 3
#define PN_OPT_THREE_STEP 4
#define PN_OPT_STEP3_ZBLOCK 5
#define PN_OPT_KERNEL_TIMED 6
int pn_set_option(pn_ctx *c, int option, int value);
int pn_get_option(pn_ctx *c, int option, int *value);
/* pn_evolve bracketed by HIP events on the context's stream; waits for the
 * steps and returns their device time in *ms (measurement helper, no
 * reference counterpart). */
int pn_evolve_timed(pn_ctx *c, int nsteps, const double *dtk, const double *dtd, double *ms);

/* ---- analysis --------------------------------------------------------- */

/* CalculateVolumeAverages (fieldanalysis.cpp:399-699), FP64 compensated sums. */
int pn_volume_averages(pn_ctx *c, pn_averages *out);
/* CalculateExtrema (fieldanalysis.cpp:319-397). */
int pn_extrema_get(pn_ctx *c, pn_extrema *out);
/* Both of the above from one pass over the fields, as Print_AveragesExtrema
 * (fieldanalysis.cpp:748-799) calls them back to back; either output may be
 * NULL.  The values are those of the two separate calls. */
int pn_volume_analysis(pn_ctx *c, pn_averages *avg, pn_extrema *ext);
/* The volume averages with Pi brought to Phi's time by a kick of half_dt
 * (updateMomenta, dynamics.cpp:52-175, into scratch; the fields are
 * unchanged): between leapfrog steps Pi lags Phi by half a step, so Energy of
 * pn_volume_averages mixes phi(t) with pi(t - dt/2) (fieldanalysis.cpp:625).
 * half_dt = dtau/2 gives the time-synchronised energy.  No reference
 * counterpart; not available on a batch view. */
int pn_volume_averages_synced(pn_ctx *c, double half_dt, pn_averages *out);
/* pn_volume_analysis without allocating (no reference counterpart): for
 * callers that cannot call hipMalloc, such as code under HIP graph capture or
 * a serving loop.  phi and pi are the caller's dense device fields of N^3
 * values (double for PN_FP64, float for PN_FP32) on device p->device; the
 * launches are enqueued on `stream` (a hipStream_t; NULL: the null stream)
 * and nothing waits for them.  out is a device array of PN_VOLUME_OUT doubles
 * that receives the pn_averages fields in their order, then min, minpos, max,
 * maxpos, maxabs, maxabspos (positions as doubles).  The values are those of
 * pn_volume_analysis on the same fields.  The block partials and tile flags go
 * to a static device array of 448 KiB per device (at most 2048 blocks),
 * so N must lie in [4, PN_NOALLOC_MAX_N]; calls that overlap on one device
 * share that array and race.  Errors are reported by pn_last_error(NULL). */
#define PN_VOLUME_OUT 16
#define PN_NOALLOC_MAX_N 2048
int pn_volume_analysis_noalloc(const pn_params *p, const void *phi, const void *pi, void *stream,
                               double *out);
/* nTBins of momentumspace.cpp:40-47 for this N and a. */
int pn_spectrum_nbins(int N, double a);
/* Calculate_CorrelationFunctions (momentumspace.cpp:140-278) with
 * fftXtoP_direct (clusterfft.cpp:185-225): rocFFT r2c of Phi and Pi, times
 * a^3, binned by a hand-written kernel. */
int pn_spectrum_get(pn_ctx *c, pn_spectrum *out);
/* Field difference of lattice a against lattice b (same N and device, either
 * precision), both converted to double: max |phi_a - phi_b|, max |pi_a - pi_b|,
 * max |phi_b|, max |pi_b| and the compensated sums of (phi_a - phi_b)^2,
 * phi_b^2, (pi_a - pi_b)^2, pi_b^2.  The precision form's field error against
 * an FP64 run; no reference counterpart. */
typedef struct pn_field_diff {
  double max_dphi, max_dpi, max_phi, max_pi;
  double sum_dphi2, sum_phi2, sum_dpi2, sum_pi2;
} pn_field_diff;
int pn_field_diff_get(pn_ctx *a, pn_ctx *b, pn_field_diff *out);
/* fftXtoP_direct (clusterfft.cpp:185-225): the a^3-scaled r2c transform of
 * Phi (field 0) or Pi (field 1) as interleaved complex values in the
 * reference's p-space order py + px*(N/2+1) + pz*N*(N/2+1), N*N*(N/2+1)
 * modes (2x that many doubles). */
int pn_fft_modes(pn_ctx *c, int field, double *out);
/* Island catalog: is_island + BFS + the raster scan of Full_Analysis
 * (islands.cpp:72-279,788-835) as GPU connected-component labelling.
 * mode 0: phi^2/2 + pi^2/2 > thr^2/2 (islands.cpp:76); peak = max phi^2.
 * mode 1: Delta > thr (islands.cpp:83) with Delta of fieldanalysis.cpp:447,514-515;
 *         peak = max Delta.
 * Connectivity: 6 neighbours, periodic in x and y, open in z (islands.cpp:200-236). */
int pn_islands(pn_ctx *c, int mode, double thr, pn_catalog *out);
/* Page-locked host memory for output arrays that are filled on every call
 * (island catalogs: about 33 MB for 233k islands at 256^3), so that their copy
 * from the device runs at the link's rate instead of through pageable
 * staging.  No reference counterpart (the reference's catalog never leaves
 * host memory).  Free with pn_host_free. */
int pn_host_alloc(size_t bytes, void **out);
int pn_host_free(void *p);
/* Island number (row of the last pn_islands catalog) of each site pos[i]
 * (reference pos() index of an interior site), or -1 outside the mask.
 * Replaces the b[] array lookups of Full_Analysis (islands.cpp:798-817). */
int pn_islands_lookup(pn_ctx *c, int n, const long long *pos, int *island);

/* ---- island tracking over time (islands.cpp:61-70,651-736,788-835) ------
 * A tracker keeps the reference's found_islands list across Full_Analysis
 * calls: islands are re-grown from their last peak in order of decreasing
 * volume (merges drop the later one), retired when their peak leaves the
 * mask, and new components get new numbers.  Histories keep the first two
 * and the latest three entries (islands.cpp:255-267). */
typedef struct pn_tracker pn_tracker;
#define PN_TRACK_HIST 5
typedef struct pn_track_entry {
  double at_time;       /* tau of the analysis                             */
  long long center_pos; /* pos() index of the peak                         */
  double peak_value, avg_phi, volume, pot_en, kin_en, grad_en;
} pn_track_entry;
typedef struct pn_tracked_island {
  long long index;      /* number given at birth (number_of_islands)       */
  int status;           /* 0 island, 1 expanding bubble, 2 oscillon        */
  int disp;             /* 1: this process holds the island's peak         */
  int nhist;            /* entries in hist                                 */
  pn_track_entry hist[PN_TRACK_HIST];
} pn_tracked_island;
int pn_tracker_create(pn_tracker **out);
int pn_tracker_destroy(pn_tracker *t);
const char *pn_tracker_last_error(const pn_tracker *t);
/* One Full_Analysis call on the context's current fields at threshold thr
 * (mode as pn_islands; the caller supplies thr, e.g. 4<E> for mode 1 as in
 * islands.cpp:783-786) stamped with time tau.  outdir != NULL appends the
 * reference's non_critical_islands.txt, islands.txt and expanding_bubbles.txt
 * rows there (islands.cpp:651-736). */
int pn_tracker_update(pn_tracker *t, pn_ctx *c, int mode, double thr, double tau,
                      const char *outdir);
int pn_tracker_count(const pn_tracker *t, int *n);
/* The components of the last pn_tracker_update, as pn_islands numbers them
 * (the islands[] of Full_Analysis's raster scan, islands.cpp:827-835),
 * without running the analysis again: out->count receives their number and
 * up to out->capacity entries are copied (PN_ERR_CAPACITY when it is
 * smaller).  No labels. */
int pn_tracker_catalog(const pn_tracker *t, pn_catalog *out);
/* i-th tracked island in list order (after the update: the survivors in
 * decreasing volume of their previous entry, then the new ones). */
int pn_tracker_get(const pn_tracker *t, int i, pn_tracked_island *out);

/* Energy density per site rho = pi^2/2 + (1/2) sum_i (grad+_i phi)^2 + V(phi),
 * the summand of Energy (fieldanalysis.cpp:625), dense order. */
int pn_energy_density(pn_ctx *c, double *rho);

/* ---- batched realizations (no reference counterpart: the reference runs
 * `runs` realizations one after another, main.cpp:52-96) ------------------ */

/* One realization of a batch: size, spacing and theory-0 couplings. */
typedef struct pn_realization {
  int N;
  double a;
  int mass_mode;
  double m2;
  int lambda_mode;
  double kappa;
} pn_realization;

typedef struct pn_batch pn_batch;

/* R realizations, sizes may differ (N must be a multiple of 2 for FP64 and of
 * 4 for FP32).  Fields start at zero. */
int pn_batch_create(int R, const pn_realization *reals, int precision, int device, pn_batch **out);
int pn_batch_destroy(pn_batch *b);
const char *pn_batch_last_error(const pn_batch *b);
/* A pn_ctx aliasing realization r, owned by the batch (do not pn_destroy it).
 * Every analysis and field call of this header works on it (initial
 * conditions, fields, averages, extrema, spectrum, islands); stepping it
 * alone returns PN_ERR_STATE. */
int pn_batch_view(pn_batch *b, int r, pn_ctx **view);
/* Advances every realization by nsteps leapfrog steps in one launch per
 * step; dtk and dtd are [R][nsteps] tables (each realization's own replay of
 * dynamics.cpp:196-237).  use_graph replays captured 100-step HIP graphs. */
int pn_batch_evolve(pn_batch *b, int nsteps, const double *dtk, const double *dtd, int use_graph);
int pn_batch_evolve_timed(pn_batch *b, int nsteps, const double *dtk, const double *dtd,
                          int use_graph, double *ms);
int pn_batch_sync(pn_batch *b);
/* Freezes the flagged realizations (the false_vacuum_decayed break of
 * main.cpp:85-89): their fields stop changing. flags: R ints. */
int pn_batch_set_decayed(pn_batch *b, const int *flags);
/* Tile shape of the single-step kernel for every realization of size N: rows
 * per tile, planes per z block. */
int pn_batch_set_tiles(pn_batch *b, int N, int rows, int zblock);
/* Times candidate tile shapes per size class on the live batch (state is
 * restored afterwards) and keeps the fastest.  out receives nclasses triples
 * (N, rows, zblock); pass NULL to skip.  In the multi-step modes only the
 * planes per tile are tuned and rows is the tile's thread rows; with both
 * multi-step modes available it tunes each and keeps the faster mode (three
 * steps per launch on ties within 2 %). */
int pn_batch_autotune(pn_batch *b, int trial_steps, int *out, int *nclasses);
/* Options of a batch (no reference counterpart); all modes give bit-identical
 * fields.  PN_OPT_TWO_STEP: 1 (default when every realization fits: N a
 * multiple of the vector width, N/width <= 768) advances pairs of steps in
 * one launch; 0 runs single steps (and turns three-step mode off).
 * PN_OPT_THREE_STEP: 1 (default when every realization fits a three-step
 * tile: N >= 8, N/width <= 768 FP64 or 1024 FP32, and the device gives a
 * block the LDS) advances three steps in one launch; needs two-step mode. */
int pn_batch_set_option(pn_batch *b, int option, int value);
int pn_batch_get_option(pn_batch *b, int option, int *value);

/* ---- differentiable leapfrog (no reference counterpart) ----------------- */

/* The theory-0 leapfrog of dynamics.cpp:52-237 with the force as a
 * polynomial with continuous couplings: f = ((c2 + c4 q) + c6 q^2) phi,
 * q = phi^2 (reference units: c2 = Mass_Mode*M^2, c4 = Lambda_Mode/6,
 * c6 = 1/(120 Kappa^2); physical units of V = phi^2/2 - lambda phi^4/4 +
 * g phi^6/6: c2 = 1, c4 = -lambda, c6 = g).  Observable after the last step:
 * F = sum rho s(u) / sum rho, u = (rho - threshold*rhobar)/(width*rhobar),
 * s logistic, rho the per-site summand of Energy (fieldanalysis.cpp:625). */
typedef struct pn_adj_params {
  int N;
  double a;
  double c2, c4, c6;
  double threshold;      /* 4   */
  double width;          /* 0.1 */
  int checkpoint_every;  /* K: (phi, pi) saved every K steps */
  int device;
} pn_adj_params;

typedef struct pn_adj_result {
  double F;
  double grad_c[3];      /* dF/dc2, dF/dc4, dF/dc6 */
  double forward_ms;     /* device time of the checkpointed forward + F */
  double backward_ms;    /* device time of recompute + reverse sweep   */
  long long peak_bytes;  /* device memory held by the run              */
} pn_adj_result;

/* F after nsteps steps from (phi0, pi0), dense N^3 host arrays. */
int pn_forward_observable(const pn_adj_params *p, const double *phi0, const double *pi0, int nsteps,
                          const double *dtk, const double *dtd, double *F);
/* Reverse entry point: F, dF/dphi0 and dF/dpi0 (N^3 each, either may be
 * NULL) and dF/dc in res, by checkpointed recompute (every
 * checkpoint_every steps) and a hand-written adjoint stencil.  With
 * checkpoint_every >= nsteps every phi_n stays in device memory from the
 * forward pass (nsteps * N^3 doubles) and nothing is recomputed; the
 * results are the same bits either way. */
int pn_adjoint(const pn_adj_params *p, const double *phi0, const double *pi0, int nsteps,
               const double *dtk, const double *dtd, double *grad_phi0, double *grad_pi0,
               pn_adj_result *res);

#ifdef __cplusplus
}
#endif
#endif /* PHINUM_GPU_H */
