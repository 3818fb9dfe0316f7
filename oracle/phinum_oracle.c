/*
 * phinum_oracle.c -- CPU restatement of PhiNum's lattice hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker: only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it, and
 * only to compare against (or time beside) the HIP product library.  The
 * product (phinum_amd/) never links, loads or calls anything in oracle/.
 *
 * Built strict IEEE-754 binary64 (gcc -O2 -ffp-contract=off, no fast-math,
 * x86-64 SSE2 => no excess precision).  Every expression keeps the
 * reference's operand order so that results are bit-identical to a strict
 * build of the reference sources (checked against oracle/_ref in tests).
 *
 * Storage follows the reference exactly: one scalar component (Nc=1), the
 * field stored x-fastest with two z halo planes, index pos(x,y,z) =
 * x + N*y + N*N*z, interior planes z = 1..N (lattice.cpp:70-104,
 * util.cpp:34-37).
 *
 * Parity pins (what checks each function, besides the reference build):
 *   or_drand48, or_gauss      glibc drand48 stream; polar-method identities
 *   or_update_(momenta|fields), or_evolve oracle/_ref bitwise; free-field dispersion;
 *                            homogeneous-field ODE (second order); time reversibility
 *   or_evolve_f32            within binary32 rounding of or_evolve over 20 steps;
 *                            a dropped phi^6 term or flipped phi^4 sign is not
 *   or_volume_averages       oracle/_ref bitwise; math.fsum closed forms; its
 *                            Delta by sum Delta = vol (E + 1.5 (E + Charge - <pi^2>))
 *   or_extrema               oracle/_ref bitwise; numpy argmax/argmin
 *   or_bin_table             brute-force count identities (sum = N^3)
 *   or_spectrum_bins         Parseval identity; single plane-wave mode; every
 *                            mode in the shell nearest |p|/pTmin (half a bin)
 *   or_print_fp_rows         closed form on a single mode (omega = |B/A|, the
 *                            1/((N a)^3 count) norm)
 *   or_islands_bfs           oracle/_ref bitwise; brute-force flood fill
 *   or_ic_modes              Hermitian symmetry; drand48 stream; spectrum
 *                            of the IC against the occupation formula.
 *                            PARITY UNPINNED: the draw order (phase1, phase2,
 *                            |g1|, |g2|) and or_ensure_real's first-visited-wins
 *                            and |F| sgn(Re F) rules -- FFTW is absent, the
 *                            reference holds no fixtures, and no identity
 *                            separates them from their alternatives
 *   or_ic_fill               Parseval between modes and fields (the
 *                            /(aN)^3 * sqrt((aN)^3) normalisation); field
 *                            variances against the occupation closed form
 *   or_tlm_(...), or_adj_(...) dot-product test (<TLM v,w> = <v,ADJ w>);
 *                            central finite differences
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

typedef struct {
  int N;            /* N_t = N_eta (CompleteParameters, parameters.cpp:398-402) */
  double a;         /* a_t = a_eta */
  int mass_mode;    /* Mass_Mode */
  double m2;        /* BareM_sqr_long = M_sqr (renormalization.cpp:17) */
  int lambda_mode;  /* Lambda_Mode */
  double kappa;     /* Kappa */
} or_params;

static inline double sqr(double x) { return x * x; }            /* util.cpp:20-23 */
static inline int mod(int i, int N) { return i >= 0 ? i % N : (N + i) % N; } /* util.cpp:12-18 */
static inline long pos3(int N, int x, int y, int z) {            /* util.cpp:34-37 */
  return (long)x + (long)N * y + (long)N * N * z;
}

/* ---------------------------------------------------------------------------
 * Time bookkeeping: reset() dynamics.cpp:28-50 and evolve() :196-237.
 * tau advances by 0.5*dtau twice per step; the kick uses tau-Last_Momenta_update
 * and the drift tau-Last_Field_update, each as rounded by the host.
 * ------------------------------------------------------------------------- */
void or_time_replay(double T0, double dtbase, double a, int nsteps,
                    double *dtk, double *dtd, double *tau_after) {
  double tau = T0;
  double dtau = dtbase * a;                  /* dynamics.cpp:42 */
  double last_f = tau, last_m = tau;         /* dynamics.cpp:43-44 */
  for (int n = 0; n < nsteps; n++) {
    tau += 0.5 * dtau;                       /* :202 */
    dtk[n] = tau - last_m;                   /* :84 (tau-Last_Momenta_update) */
    last_m = tau;                            /* :171 */
    tau += 0.5 * dtau;                       /* :208 */
    dtd[n] = tau - last_f;                   /* :183 */
    last_f = tau;                            /* :190 */
    if (tau_after) tau_after[n] = tau;
  }
}

/* Communication::exchange at one rank (comm.cpp:63-78): plane z=1 goes to the
 * upper halo z=N+1, then plane z=N goes to the lower halo z=0. Only Phi. */
void or_exchange(int N, double *phi) {
  long plane = (long)N * N;
  memcpy(phi + plane * (N + 1), phi + plane * 1, plane * sizeof(double));
  memcpy(phi + 0, phi + plane * N, plane * sizeof(double));
}

/* updateMomenta, theory 0, no expansion, Nc=1 (dynamics.cpp:52-87).
 * LapPhi at :80, force at :84. */
void or_update_momenta(const or_params *p, const double *phi, double *pi, double dtk) {
  const int N = p->N;
  const double a = p->a;
  const int Nc = 1;
  long K = pos3(N, 0, 0, 1);
  for (int z = 1; z < N + 1; z++)
    for (int y = 0; y < N; y++)
      for (int x = 0; x < N; x++) {
        double q = 0.0;
        q += sqr(phi[K]);
        long Kuoo = pos3(N, mod(x + 1, N), y, z);
        long Kouo = pos3(N, x, mod(y + 1, N), z);
        long Koou = pos3(N, x, y, z + 1);
        long Kdoo = pos3(N, mod(x - 1, N), y, z);
        long Kodo = pos3(N, x, mod(y - 1, N), z);
        long Kood = pos3(N, x, y, z - 1);
        double lap = (phi[Kuoo] + phi[Kdoo] + phi[Kouo] + phi[Kodo] + phi[Koou] + phi[Kood]
                      - 6. * phi[K]) / sqr(a);
        pi[K] += (lap - (p->mass_mode * p->m2 + p->lambda_mode * (q / (6. * Nc))
                         + (sqr(q / p->kappa) / (120. * Nc * Nc))) * phi[K]) * dtk;
        K++;
      }
}

/* updateFields (dynamics.cpp:177-194): linear sweep of the interior. */
void or_update_fields(int N, double *phi, const double *pi, double dtd) {
  long end = (long)N * N * (N + 1);
  for (long K = pos3(N, 0, 0, 1); K < end; K++) phi[K] += pi[K] * dtd;
}

/* evolve() body for nsteps steps with host-replayed factors (dynamics.cpp:196-237). */
void or_evolve(const or_params *p, double *phi, double *pi, int nsteps,
               const double *dtk, const double *dtd) {
  for (int n = 0; n < nsteps; n++) {
    or_update_momenta(p, phi, pi, dtk[n]);
    or_update_fields(p->N, phi, pi, dtd[n]);
    or_exchange(p->N, phi);
  }
}

/* ---------------------------------------------------------------------------
 * CalculateVolumeAverages at one rank (fieldanalysis.cpp:399-699).
 * out[0..9] = avgphi, avgpi, avgphi2, avgpi2, avgphipi, avgphi4, avgphi6,
 *             avgEnergy, avgInteraction, avgCharge.
 * If delta != NULL it also fills the per-site energy density of
 * fieldanalysis.cpp:447,514-515 (withEnergyPS=1), using BareM_sqr_long where
 * the reference names the undeclared BareM_mono_sqr_long.
 * ------------------------------------------------------------------------- */
void or_volume_averages(const or_params *p, const double *phi, const double *pi,
                        double *out, double *delta) {
  const int N = p->N;
  const double a = p->a;
  const int Nc = 1;
  const double kap = p->kappa;
  int vol = N * N * N;
  double phi1 = 0.0, pi1 = 0.0, phi2 = 0., pi2 = 0., phi4 = 0., phi6 = 0., phipi = 0.;
  double Interaction = 0., Charge = 0.;
  double GradPhiSqr = 0.0;
  long K = pos3(N, 0, 0, 1);
  for (int z = 1; z < N + 1; z++)
    for (int y = 0; y < N; y++)
      for (int x = 0; x < N; x++) {
        long Kuoo = pos3(N, mod(x + 1, N), y, z);
        long Kouo = pos3(N, x, mod(y + 1, N), z);
        long Koou = pos3(N, x, y, z + 1);
        double q = 0.0;
        phi1 += phi[K]; pi1 += pi[K];                              /* :437 */
        phi2 += sqr(phi[K]); pi2 += sqr(pi[K]);                    /* :439 */
        phipi += phi[K] * pi[K];                                   /* :440 */
        q += sqr(phi[K]);                                          /* :442 */
        GradPhiSqr += (sqr(phi[Kuoo] - phi[K]) / sqr(a) + sqr(phi[Kouo] - phi[K]) / sqr(a)
                       + sqr(phi[Koou] - phi[K]) / sqr(a));        /* :445 */
        if (delta) {
          double d = (0.5 * sqr(pi[K]) + sqr(phi[Kuoo] - phi[K]) / sqr(a)
                      + sqr(phi[Kouo] - phi[K]) / sqr(a) + sqr(phi[Koou] - phi[K]) / sqr(a)); /* :447 */
          d += 0.5 * (p->mass_mode) * p->m2 * sqr(phi[K]);        /* :514 */
          d += (p->lambda_mode) * (sqr(sqr(phi[K])) / (24.0 * Nc))
               + (sqr(sqr(phi[K])) * sqr(phi[K]) / (720.0 * Nc * Nc * kap * kap)); /* :515 */
          delta[K] = d;
        }
        Interaction += (p->mass_mode) * 0.5 * p->m2 * sqr(phi[K]); /* :457 */
        K++;
        phi4 += sqr(q);                                            /* :565 */
        phi6 += sqr(q) * q;                                        /* :566 */
        Interaction += (p->lambda_mode) * (sqr(q) / (24.0 * Nc))
                       + (sqr(q) * q / (720.0 * Nc * Nc * kap * kap)); /* :571 */
      }
  phi1 /= (vol); pi1 /= (vol);                                     /* :621 */
  Charge = (0.5 * pi2 - GradPhiSqr / 6. - Interaction);            /* :624 */
  double Energy = (0.5 * pi2 + 0.5 * GradPhiSqr + Interaction);    /* :625 */
  Energy /= (vol); Interaction /= (vol); Charge /= (vol);          /* :631-633 */
  phi2 /= (vol); pi2 /= (vol); phi4 /= (vol); phi6 /= (vol); phipi /= (vol);
  /* one rank: MPI_SUM is the identity and the /nodes divisions are by 1 */
  out[0] = phi1; out[1] = pi1; out[2] = phi2; out[3] = pi2; out[4] = phipi;
  out[5] = phi4; out[6] = phi6; out[7] = Energy; out[8] = Interaction; out[9] = Charge;
}

/* CalculateExtrema at one rank (fieldanalysis.cpp:319-355).
 * vals = {min, max, maxabs}; idx = {minpos, maxpos, maxAbspos} as pos() indices. */
void or_extrema(int N, const double *phi, double *vals, int *idx) {
  const int Nc = 1;
  double maxPhi = -10000000.0, minPhi = 10000000.0, maxAbsPhi = 0.0;
  int maxpos = -1, minpos = -1, maxAbspos = -1;
  double AbsPhi = 0.;
  int end = Nc * N * N * (N + 1);
  for (int k = (int)pos3(N, 0, 0, 1); k < end; k++) {
    AbsPhi += ((phi[k]) * (phi[k]));
    if ((k + 1) % Nc == 0) {
      AbsPhi = sqrt(AbsPhi);
      if (AbsPhi > maxAbsPhi) { maxAbsPhi = AbsPhi; maxAbspos = k + 1 - Nc; }
      AbsPhi = 0;
    }
    if (phi[k] > maxPhi) { maxPhi = phi[k]; maxpos = k; }
    if (phi[k] < minPhi) { minPhi = phi[k]; minpos = k; }
  }
  vals[0] = minPhi; vals[1] = maxPhi; vals[2] = maxAbsPhi;
  idx[0] = minpos; idx[1] = maxpos; idx[2] = maxAbspos;
}

/* ---------------------------------------------------------------------------
 * Momentum-space shells (momentumspace.cpp:32-65 init, :89-138 InitBCounts).
 * Returns nTBins; fills binsize (pTmin) and counts[nTBins] if counts != NULL.
 * ------------------------------------------------------------------------- */
static double lat_mom(int k, int N, double a) {  /* momentumspace.cpp:102-118 */
  if (k < N / 2 + 1) return 2.0 / (a) * sin(k * acos(-1.) / N);
  return -2.0 / (a) * sin(k * acos(-1.) / N);
}

int or_bin_count(int N, double a, double *binsize) {
  double pTmin = (2.0 / (a)) * sin(acos(-1.) / N);    /* :40 */
  double pTmax = sqrt(3.0 * (4.0 / sqr(a)));           /* :41 */
  if (binsize) *binsize = pTmin;                       /* :42 (finebin never active) */
  return (int)(pTmax / pTmin + 0.5) + 2;               /* :47 */
}

int or_bin_table(int N, double a, int *counts) {
  double binsize;
  int nb = or_bin_count(N, a, &binsize);
  for (int k = 0; k < nb; k++) counts[k] = 0;
  for (int pz = 0; pz < N; pz++) {
    double nu = (pz == 0) ? 0 : lat_mom(pz, N, a);
    for (int px = 0; px < N; px++) {
      double ppx = lat_mom(px, N, a);
      for (int py = 0; py < N / 2 + 1; py++) {
        double ppy = 2.0 / (a) * sin(py * acos(-1.) / N);
        double psqr = sqr(ppx) + sqr(ppy) + sqr(nu);
        int b = (int)(sqrt(psqr) / binsize + 0.5);
        if (b >= nb) return -1;                        /* :128-129 */
        counts[b]++;
        if (py != 0 && py != N / 2) counts[b]++;
      }
    }
  }
  return nb;
}

/* Shell sums of Calculate_CorrelationFunctions (momentumspace.cpp:196-247)
 * given the transforms already scaled by a^3 (clusterfft.cpp:185-225):
 * phik/pik are interleaved complex arrays in the reference's p-space order
 * posLOC = py + px*(N/2+1) + pz*N*(N/2+1). */
void or_spectrum_bins(int N, double a, const double *phik, const double *pik,
                      double *CorrF, double *CorrK, double *CorrQ) {
  double binsize;
  int nb = or_bin_count(N, a, &binsize);
  for (int k = 0; k < nb; k++) { CorrF[k] = 0.; CorrK[k] = 0.; CorrQ[k] = 0.; }
  for (int pz = 0; pz < N; pz++) {
    double nu = (pz == 0) ? 0 : lat_mom(pz, N, a);
    for (int px = 0; px < N; px++) {
      double ppx = lat_mom(px, N, a);
      for (int py = 0; py < N / 2 + 1; py++) {
        double ppy = 2.0 / (a) * sin(py * acos(-1.) / N);
        double psqr = sqr(ppx) + sqr(ppy) + sqr(nu);
        long posLOC = py + (long)px * (N / 2 + 1) + (long)pz * N * (N / 2 + 1);
        int b = (int)(sqrt(psqr) / binsize + 0.5);
        double pr = phik[2 * posLOC], pim = phik[2 * posLOC + 1];
        double kr = pik[2 * posLOC], kim = pik[2 * posLOC + 1];
        double f = sqr(hypot(pr, pim));                 /* sqr(abs(PhiP)) */
        double g = sqr(hypot(kr, kim));
        /* 0.5*real(P*conj(K) + conj(P)*K) = pr*kr + pim*kim exactly */
        double h = pr * kr + pim * kim;
        CorrF[b] += f; CorrK[b] += g; CorrQ[b] += h;
        if (py != 0 && py != N / 2) { CorrF[b] += f; CorrK[b] += g; CorrQ[b] += h; }
      }
    }
  }
}

/* Print_fp row values (momentumspace.cpp:318-369).  rows[k*7 + c] with
 * c = p, count, omega, FF*norm, KK*norm, sqrt(FF*KK)*norm, FF*norm*p^3/(2 pi^2);
 * valid[k] = 1 where the reference writes a row. */
void or_print_fp_rows(int N, double a, int nb, const int *counts, const double *F,
                      const double *Kk, double *rows, int *valid) {
  double binsize;
  or_bin_count(N, a, &binsize);
  double NormCorr = 1. / (sqr(N * a) * N * a);
  for (int k = 0; k < nb; k++) {
    double norm2 = (counts[k] == 0 ? 0 : 1 / (double)(counts[k]));
    double pabs = binsize * k;
    valid[k] = (norm2 != 0 && F[k] != 0);
    double FF = F[k] / 1, KK = Kk[k] / 1;
    double *r = rows + 7 * k;
    r[0] = pabs; r[1] = counts[k];
    r[2] = sqrt(KK / FF);
    r[3] = FF * NormCorr * norm2;
    r[4] = KK * NormCorr * norm2;
    r[5] = sqrt(FF * KK) * NormCorr * norm2;
    r[6] = FF * NormCorr * norm2 * pabs * pabs * pabs / 2. / acos(-1.) / acos(-1.);
  }
}

/* ---------------------------------------------------------------------------
 * Initial conditions, initialize==3, withcharge==0, expansion==0
 * (initialconditions.cpp:12-42 randGen, :83-594 mode loop, :603-614
 * ensureRealResult = clusterfft.cpp:55-145 at one rank).
 * ------------------------------------------------------------------------- */
static uint64_t g_x48;
static int g_gauss_on = 0;
static double g_gauss2;

/* POSIX drand48: X <- (0x5DEECE66D*X + 0xB) mod 2^48, value X/2^48; srand48
 * puts the low 32 bits of the seed above 0x330E. */
void or_srand48(long seed) {
  g_x48 = (((uint64_t)(uint32_t)seed) << 16) | 0x330Eu;
  g_gauss_on = 0;
}
double or_drand48(void) {
  g_x48 = (0x5DEECE66DULL * g_x48 + 0xBULL) & ((1ULL << 48) - 1);
  return (double)g_x48 * (1.0 / 281474976710656.0);
}
double or_gauss(void) {                                   /* initialconditions.cpp:27-41 */
  if (g_gauss_on == 1) { g_gauss_on = 0; return g_gauss2; }
  double x, y, r, c, s, g1;
  int accept = 0;
  while (accept == 0) {
    x = 2.0 * or_drand48() - 1.0; y = 2.0 * or_drand48() - 1.0;
    r = sqrt(sqr(x) + sqr(y));
    if (r < 1.0) accept = 1;
  }
  c = x / r; s = y / r;
  g1 = c * sqrt(-2.0 * log(sqr(r)));
  g_gauss2 = (s / c) * g1;
  g_gauss_on = 1;
  return g1;
}

typedef struct {
  double QS, scale2, amplitude, noise; /* noise already multiplied by Lambda (parameters.cpp:408-410) */
  double m2_long;                      /* M2_long = M_sqr (renormalization.cpp:17) */
  int justnoise;
} or_ic_params;

static int sgn(double x) { return x > 0 ? 1 : -1; }      /* clusterfft.cpp:10-16 */

/* A, E: interleaved complex arrays of N*N*(N/2+1) modes in p-space order. */
void or_ic_modes(int N, double a, const or_ic_params *ic, long seed, double *A, double *E) {
  const int H = N / 2 + 1;
  or_srand48(seed);
  for (int pz = 0; pz < N; pz++) {
    double nu = (pz == 0) ? 0 : lat_mom(pz, N, a);
    for (int px = 0; px < N; px++) {
      double ppx = lat_mom(px, N, a);
      for (int py = 0; py < H; py++) {
        double ppy = 2.0 / (a) * sin(py * acos(-1.) / N);
        double pt = sqrt(sqr(ppx) + sqr(ppy));
        double psqr = sqr(pt) + sqr(nu);
        long m = py + (long)px * H + (long)pz * N * H;
        A[2 * m] = 0.0; A[2 * m + 1] = 0.0; E[2 * m] = 0.0; E[2 * m + 1] = 0.0;
        double phase1 = 2.0 * acos(-1.) * or_drand48();
        double phase2 = 2.0 * acos(-1.) * or_drand48();
        double gauss1 = fabs(or_gauss());
        double gauss2 = fabs(or_gauss());
        double pchisqr = sqr(pt) + sqr(nu);
        double Occ;
        if (ic->justnoise == 0) {
          if (pchisqr < ic->QS * ic->QS) Occ = ic->amplitude + ic->noise;
          else if (ic->scale2 == 0) Occ = ic->noise;
          else Occ = (sqrt(pchisqr) < ic->scale2) ? ic->noise : 0.0;
        } else {
          if (ic->scale2 == 0) Occ = ic->noise;
          else Occ = (sqrt(pchisqr) < ic->scale2) ? ic->noise : 0.0;
        }
        double Omega = sqrt(psqr + ic->m2_long);
        double s1 = sqrt(Occ) * gauss1, s2 = sqrt(Occ) * gauss2;
        double f1r = s1 * cos(phase1), f1i = s1 * sin(phase1);
        double f2r = s2 * cos(phase2), f2i = s2 * sin(phase2);
        if (fabs(Omega) < 1e-8) {
          A[2 * m] = 0; A[2 * m + 1] = 0; E[2 * m] = 0; E[2 * m + 1] = 0;
        } else {
          double so = sqrt(Omega);
          A[2 * m] = f1r / so; A[2 * m + 1] = f1i / so;
          E[2 * m] = f2r * so; E[2 * m + 1] = f2i * so;
        }
      }
    }
  }
}

/* ensureRealResult at one rank (clusterfft.cpp:55-145). */
void or_ensure_real(int N, double *F) {
  const int H = N / 2 + 1;
  for (int pz = 0; pz < N; pz++)
    for (int px = 0; px < N; px++)
      for (int py = 0; py < H; py++) {
        int mpz = mod(N - pz, N), mpx = mod(N - px, N), mpy = mod(N - py, N);
        if (mpy != py) continue;
        long m = py + (long)px * H + (long)pz * N * H;
        if (px == mpx && pz == mpz) {
          double v = hypot(F[2 * m], F[2 * m + 1]) * sgn(F[2 * m]);
          F[2 * m] = v; F[2 * m + 1] = 0.0;
        } else {
          long mm = mpy + (long)mpx * H + (long)mpz * N * H;
          F[2 * mm] = F[2 * m]; F[2 * mm + 1] = -F[2 * m + 1];
        }
      }
}

/* Lattice fill after the c2r transforms (initialconditions.cpp:636-703).
 * AX/EX hold the unnormalised c2r outputs in the reference order
 * XData[y + x*N + z*N*N] already mapped to Out[x + y*N + z*N*N]
 * (clusterfft.cpp:176-179, before the /Norm). */
void or_ic_fill(int N, double a, const double *AX, const double *EX, double phi0,
                double pi0, double *phi, double *pi) {
  double Norm = (sqr(a * N) * N * a);                      /* clusterfft.cpp:150 */
  double NormFactor = sqrt(sqr(N * a) * N * a);            /* initialconditions.cpp:643 */
  for (int z = 0; z < N; z++)
    for (int y = 0; y < N; y++)
      for (int x = 0; x < N; x++) {
        long posLOC = pos3(N, x, y, z + 1);
        long posBUFF = x + (long)y * N + (long)z * N * N;
        pi[posLOC] = (EX[posBUFF] / Norm) * NormFactor + pi0;
        phi[posLOC] = (AX[posBUFF] / Norm) * NormFactor + phi0;
      }
  or_exchange(N, phi);
}

/* ---------------------------------------------------------------------------
 * Islands: is_island (islands.cpp:72-87), BFS (:90-279) and the raster scan of
 * Full_Analysis (:827-835) for a call with no tracked islands (first call).
 * mode 0: phi^2/2 + pi^2/2 > thr^2/2 (withEnergyPS=0); peak = first strict max of phi^2.
 * mode 1: Delta > thr (withEnergyPS=1, EnergybasedID=1); peak = first strict max of Delta.
 * labels: pos()-indexed, size N*N*(N+2), -1 outside islands, else island order.
 * cat: per island 10 doubles: index, NN, V, center pos, Peak_value, avPhi,
 *      PotEn, KinEn, GradEn, 0.  Returns the island count (or -1 if cap hit).
 * ------------------------------------------------------------------------- */
static inline int is_island(int mode, const double *phi, const double *pi,
                            const double *delta, double thr, long K) {
  if (mode == 0)
    return (phi[K] * phi[K] * 0.5 + pi[K] * pi[K] * 0.5 > thr * thr * 0.5);
  return (delta[K] > thr);
}

int or_islands_bfs(const or_params *p, int mode, double thr, const double *phi,
                   const double *pi, const double *delta, int *b, double *cat, int cap) {
  const int N = p->N;
  const double a = p->a, a_eta = p->a;
  const int Nc = 1;
  const double kap = p->kappa;
  long total = (long)N * N * (N + 2);
  for (long k = 0; k < total; k++) b[k] = -1;                   /* :788-789 */
  int *q = (int *)malloc(sizeof(int) * (size_t)N * N * N);
  int nisl = 0;
  long end = (long)Nc * N * N * (N + 1);
  for (long K0 = pos3(N, 0, 0, 1); K0 < end; K0++) {
    if (!(is_island(mode, phi, pi, delta, thr, K0) && b[K0] == -1)) continue;
    int flag = nisl;
    int NN = 0;
    double cPotEn = 0., cGradEn = 0., cKinEn = 0., cavPhi = 0.;
    long cnt = K0;
    long qh = 0, qt = 0;
    b[K0] = flag;
    q[qt++] = (int)K0;
    while (qh < qt) {
      long Kooo = q[qh++];
      if (mode == 0) {
        if (phi[Kooo] * phi[Kooo] > phi[cnt] * phi[cnt]) cnt = Kooo;   /* :110-111 */
      } else {
        if (delta[Kooo] > delta[cnt]) cnt = Kooo;                      /* :114-115 */
      }
      int z = (int)(Kooo / (Nc * N * N));
      int y = (int)((Kooo - (long)Nc * N * N * z) / (Nc * N));
      int x = (int)((Kooo - (long)Nc * N * N * z - (long)Nc * N * y) / Nc);
      long Kuoo = pos3(N, mod(x + 1, N), y, z);
      long Kdoo = pos3(N, mod(x - 1, N), y, z);
      long Kouo = pos3(N, x, mod(y + 1, N), z);
      long Kodo = pos3(N, x, mod(y - 1, N), z);
      long Koou = pos3(N, x, y, z + 1);
      long Kood = pos3(N, x, y, z - 1);
      NN++;
      cGradEn += 0.5 * (sqr(phi[Kuoo] - phi[Kooo]) + sqr(phi[Kouo] - phi[Kooo])
                        + sqr(phi[Koou] - phi[Kooo])) / sqr(a);          /* :131 */
      cKinEn += 0.5 * sqr(pi[Kooo]);                                     /* :134 */
      cPotEn += (double)(p->mass_mode) * 0.5 * p->m2 * sqr(phi[Kooo])
                + (double)(p->lambda_mode) * (sqr(sqr(phi[Kooo])) / (24.0 * Nc))
                + (sqr(sqr(phi[Kooo])) * sqr(phi[Kooo]) / (720.0 * Nc * Nc * kap * kap)); /* :142 */
      cavPhi += phi[Kooo];
      if (is_island(mode, phi, pi, delta, thr, Kuoo) && b[Kuoo] == -1) { b[Kuoo] = flag; q[qt++] = (int)Kuoo; }
      if (is_island(mode, phi, pi, delta, thr, Kdoo) && b[Kdoo] == -1) { b[Kdoo] = flag; q[qt++] = (int)Kdoo; }
      if (is_island(mode, phi, pi, delta, thr, Kouo) && b[Kouo] == -1) { b[Kouo] = flag; q[qt++] = (int)Kouo; }
      if (is_island(mode, phi, pi, delta, thr, Kodo) && b[Kodo] == -1) { b[Kodo] = flag; q[qt++] = (int)Kodo; }
      if (z < N)
        if (is_island(mode, phi, pi, delta, thr, Koou) && b[Koou] == -1) { b[Koou] = flag; q[qt++] = (int)Koou; }
      if (z > 1)
        if (is_island(mode, phi, pi, delta, thr, Kood) && b[Kood] == -1) { b[Kood] = flag; q[qt++] = (int)Kood; }
    }
    cavPhi /= (double)(NN);                                               /* :240 */
    double cV = NN * a * a * a_eta;                                       /* :241 */
    cPotEn *= (a * a * a_eta); cKinEn *= (a * a * a_eta); cGradEn *= (a * a * a_eta);
    if (nisl < cap) {
      double *c = cat + 10 * nisl;
      c[0] = nisl; c[1] = NN; c[2] = cV; c[3] = (double)cnt;
      c[4] = (mode == 0) ? phi[cnt] : delta[cnt];                         /* :257,260 */
      c[5] = cavPhi; c[6] = cPotEn; c[7] = cKinEn; c[8] = cGradEn; c[9] = 0.0;
    }
    nisl++;
  }
  free(q);
  return nisl <= cap ? nisl : -1;
}

/* ---------------------------------------------------------------------------
 * FP32 form (new; no reference counterpart).  Same leapfrog and stencil in
 * binary32 with the potential written with host-rounded coefficients:
 *   force = ((c2 + c4*q) + c6*(q*q)) * phi,  q = phi*phi,
 *   c2 = Mass_Mode*M^2, c4 = Lambda_Mode/6, c6 = 1/(120 kappa^2)
 * and 1/a^2 as a multiplier.  Same pos() layout as the FP64 oracle.
 * ------------------------------------------------------------------------- */
void or_evolve_f32(int N, float inv_a2, float c2, float c4, float c6, float *phi,
                   float *pi, int nsteps, const float *dtk, const float *dtd) {
  long plane = (long)N * N;
  float *tmp = (float *)malloc(sizeof(float) * plane * (N + 2));
  for (int n = 0; n < nsteps; n++) {
    memcpy(tmp, phi, sizeof(float) * plane * (N + 2));
    for (int z = 1; z <= N; z++)
      for (int y = 0; y < N; y++)
        for (int x = 0; x < N; x++) {
          long K = pos3(N, x, y, z);
          float f = tmp[K];
          float s = tmp[pos3(N, mod(x + 1, N), y, z)] + tmp[pos3(N, mod(x - 1, N), y, z)];
          s = s + tmp[pos3(N, x, mod(y + 1, N), z)];
          s = s + tmp[pos3(N, x, mod(y - 1, N), z)];
          s = s + tmp[pos3(N, x, y, z + 1)];
          s = s + tmp[pos3(N, x, y, z - 1)];
          float lap = (s - 6.0f * f) * inv_a2;
          float q = f * f;
          float force = ((c2 + c4 * q) + c6 * (q * q)) * f;
          float p = pi[K] + (lap - force) * dtk[n];
          pi[K] = p;
          phi[K] = f + p * dtd[n];
        }
    memcpy(phi + plane * (N + 1), phi + plane, sizeof(float) * plane);
    memcpy(phi, phi + plane * N, sizeof(float) * plane);
  }
  free(tmp);
}
