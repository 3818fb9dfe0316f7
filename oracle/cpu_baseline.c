/*
 * cpu_baseline.c -- CPU baseline for the leapfrog hot path (test/bench
 * infrastructure only: bench.py's cpu_baseline leg times it; nothing in the
 * product calls it).
 *
 * It keeps the reference's loop nest and layout so that its timing stands in
 * for the reference's own evolve() (dynamics.cpp:52-237, comm.cpp:63-78):
 * AoS storage with two z halo planes, pos()/mod() index arithmetic per site
 * (util.cpp:12-37), an in-place kick over the interior followed by a drift,
 * then the z-halo copy.  The reference splits z into MPI slabs with one halo
 * exchange per step (lattice.cpp:72-78); here the slabs are POSIX threads
 * sharing one array, with barriers where the reference has MPI_Wait/Barrier.
 * Built with the reference's flags from run_PC.pbs:1.
 */
#include <math.h>
#include <pthread.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

/* externally visible globals, as the reference's parameters and LocalLattice
 * namespace members are: the compiler must assume the Pi stores may alias
 * them and reload them per site, as it does for the reference */
int g_N;
double g_a, g_m2, g_kappa;
int g_mmode, g_lmode;
double *Phi, *Pi;
double g_tau, g_dtau, g_lastF, g_lastM;
int g_nthreads;
int g_Nc = 1;  /* runtime component count, as the reference's global Nc */
static pthread_barrier_t g_bar;

static inline int mod(int i, int N) { return i >= 0 ? i % N : (N + i) % N; }
static inline double sqr(double x) { return x * x; }
static inline int pos(int x, int y, int z) {
  return g_Nc * x + g_Nc * g_N * y + g_Nc * g_N * g_N * z;
}

static void kick_slab(int z0, int z1) {
  const int N = g_N, Nc = g_Nc;
  int Kooo = pos(0, 0, z0);
  int Kuoo, Kdoo, Kouo, Kodo, Koou, Kood;
  for (int z = z0; z < z1; z++)
    for (int y = 0; y < N; y++)
      for (int x = 0; x < N; x++) {
        double q = 0.0;
        for (int a = 0; a < Nc; a++) q += sqr(Phi[Kooo + a]);
        Kuoo = pos(mod(x + 1, N), y, z); Kouo = pos(x, mod(y + 1, N), z); Koou = pos(x, y, z + 1);
        Kdoo = pos(mod(x - 1, N), y, z); Kodo = pos(x, mod(y - 1, N), z); Kood = pos(x, y, z - 1);
        for (int a = 0; a < Nc; a++) {
          double lap = (Phi[Kuoo] + Phi[Kdoo] + Phi[Kouo] + Phi[Kodo] + Phi[Koou] + Phi[Kood]
                        - 6. * Phi[Kooo]) / sqr(g_a);
          if (a == 0)
            Pi[Kooo] += (lap - (g_mmode * g_m2 + g_lmode * (q / (6. * Nc))
                                + (sqr(q / g_kappa) / (120. * Nc * Nc))) * Phi[Kooo]) *
                        (g_tau - g_lastM);
          Kooo++; Kuoo++; Kouo++; Koou++; Kdoo++; Kodo++; Kood++;
        }
      }
}

static void drift_slab(int z0, int z1) {
  for (int K = pos(0, 0, z0); K < g_Nc * g_N * g_N * z1; K++) Phi[K] += Pi[K] * (g_tau - g_lastF);
}

typedef struct { int id, nsteps; double phi0; } targ;

/* each thread first-touches its own slab, as each MPI rank allocates its own */
static void fill_slab(int z0, int z1, double phi0) {
  long plane = (long)g_N * g_N;
  for (long k = plane * z0; k < plane * z1; k++) {
    unsigned long long s = (unsigned long long)k * 0x9E3779B97F4A7C15ULL + 0x632BE59BD9B4E019ULL;
    s ^= s >> 29; s *= 0xBF58476D1CE4E5B9ULL; s ^= s >> 32;
    double u = (double)(s >> 11) * (1.0 / 9007199254740992.0);
    Phi[k] = phi0 + 0.1 * (u - 0.5);
    Pi[k] = 0.0;
  }
}

static void *worker(void *vp) {
  targ *t = (targ *)vp;
  int N = g_N, P = g_nthreads;
  int z0 = 1 + (N * t->id) / P, z1 = 1 + (N * (t->id + 1)) / P;
  long plane = (long)N * N;
  if (t->nsteps < 0) {
    fill_slab(t->id == 0 ? 0 : z0, t->id == P - 1 ? N + 2 : z1, t->phi0);
    return NULL;
  }
  for (int s = 0; s < t->nsteps; s++) {
    /* evolve() (dynamics.cpp:196-237): thread 0 keeps the time bookkeeping */
    if (t->id == 0) g_tau += 0.5 * g_dtau;
    pthread_barrier_wait(&g_bar);
    kick_slab(z0, z1);
    pthread_barrier_wait(&g_bar);
    if (t->id == 0) { g_lastM = g_tau; g_tau += 0.5 * g_dtau; }
    pthread_barrier_wait(&g_bar);
    drift_slab(z0, z1);
    pthread_barrier_wait(&g_bar);
    if (t->id == 0) {
      g_lastF = g_tau;
      memcpy(Phi + plane * (N + 1), Phi + plane, plane * sizeof(double));
      memcpy(Phi, Phi + plane * N, plane * sizeof(double));
    }
    pthread_barrier_wait(&g_bar);
  }
  return NULL;
}

static void run_steps(int nsteps, double phi0) {
  pthread_t th[256];
  targ ta[256];
  pthread_barrier_init(&g_bar, NULL, (unsigned)g_nthreads);
  for (int i = 0; i < g_nthreads; i++) {
    ta[i].id = i; ta[i].nsteps = nsteps; ta[i].phi0 = phi0;
    pthread_create(&th[i], NULL, worker, &ta[i]);
  }
  for (int i = 0; i < g_nthreads; i++) pthread_join(th[i], NULL);
  pthread_barrier_destroy(&g_bar);
}

/* Returns seconds for `nsteps` timed steps after `warmup` untimed ones, or a
 * negative value on bad arguments.  Fields start as phi0 + a deterministic
 * small perturbation so the potential terms are exercised. */
double cb_run(int N, int nthreads, int warmup, int nsteps, double a, int mmode, double m2,
              int lmode, double kappa, double dtbase, double phi0, double *checksum) {
  if (N < 4 || nthreads < 1 || nthreads > 256 || nthreads > N) return -1.0;
  g_N = N; g_a = a; g_mmode = mmode; g_m2 = m2; g_lmode = lmode; g_kappa = kappa;
  g_nthreads = nthreads;
  long total = (long)N * N * (N + 2);
  Phi = (double *)malloc(sizeof(double) * total);
  Pi = (double *)malloc(sizeof(double) * total);
  if (!Phi || !Pi) { free(Phi); free(Pi); return -2.0; }
  run_steps(-1, phi0);
  long plane = (long)N * N;
  memcpy(Phi + plane * (N + 1), Phi + plane, plane * sizeof(double));
  memcpy(Phi, Phi + plane * N, plane * sizeof(double));
  g_tau = 0.0; g_dtau = dtbase * a; g_lastF = g_tau; g_lastM = g_tau;
  run_steps(warmup, phi0);
  struct timespec t0, t1;
  clock_gettime(CLOCK_MONOTONIC, &t0);
  run_steps(nsteps, phi0);
  clock_gettime(CLOCK_MONOTONIC, &t1);
  double sec = (double)(t1.tv_sec - t0.tv_sec) + 1e-9 * (double)(t1.tv_nsec - t0.tv_nsec);
  if (checksum) {
    double c = 0;
    for (long k = plane; k < plane * (N + 1); k++) c += Phi[k];
    *checksum = c;
  }
  free(Phi); free(Pi);
  return sec;
}
