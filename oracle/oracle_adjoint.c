/*
 * oracle_adjoint.c -- CPU checker for the differentiable leapfrog (no
 * reference counterpart; test infrastructure only, like phinum_oracle.c).
 *
 * Forward model (FP64, dense x-fastest N^3 lattice, periodic in x, y, z):
 *   pi  <- pi + (Lap(phi) - f(phi)) * dtk[n],  Lap = (sum of 6 nbrs - 6 phi)/a2
 *   phi <- phi + pi * dtd[n]
 *   f(phi) = ((c2 + c4 q) + c6 q^2) phi,  q = phi^2
 * i.e. the theory-0 force of dynamics.cpp:84 with c2 = Mass_Mode*M^2,
 * c4 = Lambda_Mode/6, c6 = 1/(120 Kappa^2) as continuous couplings.
 *
 * Observable at the final step (smooth oscillon energy fraction):
 *   rho_i = pi_i^2/2 + (1/2) sum_d (phi_{i+d} - phi_i)^2 / a2 + V(phi_i),
 *   V = c2 phi^2/2 + c4 phi^4/4 + c6 phi^6/6          (the summand of Energy,
 *                                                       fieldanalysis.cpp:625)
 *   F = sum_i rho_i s(u_i) / sum_i rho_i,  u_i = (rho_i - t rhobar)/(w rhobar),
 *   s = logistic, t = 4, w = 0.1.
 *
 * Functions: or_poly_evolve (forward), or_observable (F and dF/drho),
 * or_tlm (tangent-linear directional derivative), or_adjoint (reverse pass
 * storing every step; small lattices only).  Pinned by the dot-product test
 * <TLM v, w> = <v, ADJ w>, central finite differences and the closed-form
 * derivative of F with respect to rho (tests/test_adjoint_oracle.py).
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

typedef struct {
  int N;
  double a2, c2, c4, c6;
  double t, w;  /* observable threshold and width */
} or_adj_params;

static inline long idx(int N, int x, int y, int z) { return (long)x + (long)N * (y + (long)N * z); }
static inline int wrap(int i, int N) { return i < 0 ? i + N : (i >= N ? i - N : i); }

static void poly_step(const or_adj_params *p, const double *phi, double *pi, double *phin,
                      double dtk, double dtd) {
  const int N = p->N;
  for (int z = 0; z < N; z++)
    for (int y = 0; y < N; y++)
      for (int x = 0; x < N; x++) {
        long i = idx(N, x, y, z);
        double f = phi[i];
        double s = phi[idx(N, wrap(x + 1, N), y, z)] + phi[idx(N, wrap(x - 1, N), y, z)];
        s = s + phi[idx(N, x, wrap(y + 1, N), z)];
        s = s + phi[idx(N, x, wrap(y - 1, N), z)];
        s = s + phi[idx(N, x, y, wrap(z + 1, N))];
        s = s + phi[idx(N, x, y, wrap(z - 1, N))];
        double lap = (s - 6.0 * f) / p->a2;
        double q = f * f;
        double force = ((p->c2 + p->c4 * q) + p->c6 * (q * q)) * f;
        double pn = pi[i] + (lap - force) * dtk;
        pi[i] = pn;
        phin[i] = f + pn * dtd;
      }
}

void or_poly_evolve(const or_adj_params *p, double *phi, double *pi, int nsteps, const double *dtk,
                    const double *dtd) {
  long n3 = (long)p->N * p->N * p->N;
  double *tmp = (double *)malloc(sizeof(double) * n3);
  for (int n = 0; n < nsteps; n++) {
    poly_step(p, phi, pi, tmp, dtk[n], dtd[n]);
    memcpy(phi, tmp, sizeof(double) * n3);
  }
  free(tmp);
}

static double site_rho(const or_adj_params *p, const double *phi, const double *pi, int x, int y,
                       int z) {
  const int N = p->N;
  long i = idx(N, x, y, z);
  double f = phi[i];
  double dx = phi[idx(N, wrap(x + 1, N), y, z)] - f;
  double dy = phi[idx(N, x, wrap(y + 1, N), z)] - f;
  double dz = phi[idx(N, x, y, wrap(z + 1, N))] - f;
  double q = f * f;
  double V = ((0.5 * p->c2) * q + (0.25 * p->c4) * (q * q)) + (p->c6 / 6.0) * (q * q * q);
  return 0.5 * (pi[i] * pi[i]) + 0.5 * ((dx * dx + dy * dy + dz * dz) / p->a2) + V;
}

/* F and g = dF/drho (g may be NULL); rho is written if non-NULL */
double or_observable(const or_adj_params *p, const double *phi, const double *pi, double *rho_out,
                     double *g) {
  const int N = p->N;
  long n3 = (long)N * N * N;
  double *rho = rho_out ? rho_out : (double *)malloc(sizeof(double) * n3);
  double R = 0;
  for (int z = 0; z < N; z++)
    for (int y = 0; y < N; y++)
      for (int x = 0; x < N; x++) {
        double r = site_rho(p, phi, pi, x, y, z);
        rho[idx(N, x, y, z)] = r;
        R += r;
      }
  double rbar = R / (double)n3;
  double num = 0, A = 0;
  for (long i = 0; i < n3; i++) {
    double u = (rho[i] - p->t * rbar) / (p->w * rbar);
    double s = 1.0 / (1.0 + exp(-u));
    num += rho[i] * s;
    A += rho[i] * rho[i] * s * (1.0 - s);
  }
  double F = num / R;
  if (g)
    for (long i = 0; i < n3; i++) {
      double u = (rho[i] - p->t * rbar) / (p->w * rbar);
      double s = 1.0 / (1.0 + exp(-u));
      g[i] = (s - F) / R + (rho[i] * s * (1.0 - s) - A / R) / (p->w * rbar * R);
    }
  if (!rho_out) free(rho);
  return F;
}

/* forward S steps then F */
double or_forward_F(const or_adj_params *p, const double *phi0, const double *pi0, int nsteps,
                    const double *dtk, const double *dtd) {
  long n3 = (long)p->N * p->N * p->N;
  double *phi = (double *)malloc(sizeof(double) * n3), *pi = (double *)malloc(sizeof(double) * n3);
  memcpy(phi, phi0, sizeof(double) * n3);
  memcpy(pi, pi0, sizeof(double) * n3);
  or_poly_evolve(p, phi, pi, nsteps, dtk, dtd);
  double F = or_observable(p, phi, pi, NULL, NULL);
  free(phi);
  free(pi);
  return F;
}

/* tangent-linear: directional derivative of F along (dphi0, dpi0, dc[3]) */
double or_tlm(const or_adj_params *p, const double *phi0, const double *pi0, int nsteps,
              const double *dtk, const double *dtd, const double *dphi0, const double *dpi0,
              const double *dc) {
  const int N = p->N;
  long n3 = (long)N * N * N;
  double *phi = malloc(sizeof(double) * n3), *pi = malloc(sizeof(double) * n3);
  double *dphi = malloc(sizeof(double) * n3), *dpi = malloc(sizeof(double) * n3);
  double *phin = malloc(sizeof(double) * n3), *dphin = malloc(sizeof(double) * n3);
  memcpy(phi, phi0, sizeof(double) * n3);
  memcpy(pi, pi0, sizeof(double) * n3);
  memcpy(dphi, dphi0, sizeof(double) * n3);
  memcpy(dpi, dpi0, sizeof(double) * n3);
  for (int n = 0; n < nsteps; n++) {
    for (int z = 0; z < N; z++)
      for (int y = 0; y < N; y++)
        for (int x = 0; x < N; x++) {
          long i = idx(N, x, y, z);
          long nb[6] = {idx(N, wrap(x + 1, N), y, z), idx(N, wrap(x - 1, N), y, z),
                        idx(N, x, wrap(y + 1, N), z), idx(N, x, wrap(y - 1, N), z),
                        idx(N, x, y, wrap(z + 1, N)), idx(N, x, y, wrap(z - 1, N))};
          double f = phi[i], df = dphi[i];
          double s = 0, ds = 0;
          for (int k = 0; k < 6; k++) { s += phi[nb[k]]; ds += dphi[nb[k]]; }
          double q = f * f;
          double force = ((p->c2 + p->c4 * q) + p->c6 * (q * q)) * f;
          double dforce = (p->c2 + 3.0 * p->c4 * q + 5.0 * p->c6 * q * q) * df
                          + dc[0] * f + dc[1] * q * f + dc[2] * q * q * f;
          double pn = pi[i] + ((s - 6.0 * f) / p->a2 - force) * dtk[n];
          double dpn = dpi[i] + ((ds - 6.0 * df) / p->a2 - dforce) * dtk[n];
          pi[i] = pn;
          dpi[i] = dpn;
          phin[i] = f + pn * dtd[n];
          dphin[i] = df + dpn * dtd[n];
        }
    memcpy(phi, phin, sizeof(double) * n3);
    memcpy(dphi, dphin, sizeof(double) * n3);
  }
  double *g = malloc(sizeof(double) * n3);
  or_observable(p, phi, pi, NULL, g);
  double dF = 0;
  for (int z = 0; z < N; z++)
    for (int y = 0; y < N; y++)
      for (int x = 0; x < N; x++) {
        long i = idx(N, x, y, z);
        double f = phi[i], df = dphi[i], q = f * f;
        long nb[3] = {idx(N, wrap(x + 1, N), y, z), idx(N, x, wrap(y + 1, N), z),
                      idx(N, x, y, wrap(z + 1, N))};
        double dgrad = 0;
        for (int k = 0; k < 3; k++) dgrad += (phi[nb[k]] - f) * (dphi[nb[k]] - df);
        double dV = (p->c2 + p->c4 * q + p->c6 * q * q) * f * df
                    + dc[0] * q / 2.0 + dc[1] * q * q / 4.0 + dc[2] * q * q * q / 6.0;
        dF += g[i] * (pi[i] * dpi[i] + dgrad / p->a2 + dV);
      }
  free(phi); free(pi); free(dphi); free(dpi); free(phin); free(dphin); free(g);
  return dF;
}

/* reverse pass: stores every step's phi (small lattices); returns F */
double or_adjoint(const or_adj_params *p, const double *phi0, const double *pi0, int nsteps,
                  const double *dtk, const double *dtd, double *gphi0, double *gpi0, double *gc) {
  const int N = p->N;
  long n3 = (long)N * N * N;
  double *traj = malloc(sizeof(double) * n3 * (nsteps + 1));
  double *pi = malloc(sizeof(double) * n3);
  memcpy(traj, phi0, sizeof(double) * n3);
  memcpy(pi, pi0, sizeof(double) * n3);
  for (int n = 0; n < nsteps; n++) poly_step(p, traj + n3 * n, pi, traj + n3 * (n + 1), dtk[n], dtd[n]);
  const double *phi = traj + n3 * nsteps;
  double *g = malloc(sizeof(double) * n3);
  double F = or_observable(p, phi, pi, NULL, g);
  double *bphi = malloc(sizeof(double) * n3), *bpi = malloc(sizeof(double) * n3);
  double *pb = malloc(sizeof(double) * n3);
  gc[0] = gc[1] = gc[2] = 0;
  for (int z = 0; z < N; z++)
    for (int y = 0; y < N; y++)
      for (int x = 0; x < N; x++) {
        long i = idx(N, x, y, z);
        double f = phi[i], q = f * f;
        long up[3] = {idx(N, wrap(x + 1, N), y, z), idx(N, x, wrap(y + 1, N), z), idx(N, x, y, wrap(z + 1, N))};
        long dn[3] = {idx(N, wrap(x - 1, N), y, z), idx(N, x, wrap(y - 1, N), z), idx(N, x, y, wrap(z - 1, N))};
        double v = (p->c2 + p->c4 * q + p->c6 * q * q) * f * g[i];
        for (int k = 0; k < 3; k++) v += ((f - phi[up[k]]) * g[i] + (f - phi[dn[k]]) * g[dn[k]]) / p->a2;
        bphi[i] = v;
        bpi[i] = pi[i] * g[i];
        gc[0] += g[i] * q / 2.0;
        gc[1] += g[i] * q * q / 4.0;
        gc[2] += g[i] * q * q * q / 6.0;
      }
  for (int n = nsteps - 1; n >= 0; n--) {
    const double *ph = traj + n3 * n;
    for (long i = 0; i < n3; i++) pb[i] = bpi[i] + dtd[n] * bphi[i];
    for (int z = 0; z < N; z++)
      for (int y = 0; y < N; y++)
        for (int x = 0; x < N; x++) {
          long i = idx(N, x, y, z);
          double s = pb[idx(N, wrap(x + 1, N), y, z)] + pb[idx(N, wrap(x - 1, N), y, z)]
                     + pb[idx(N, x, wrap(y + 1, N), z)] + pb[idx(N, x, wrap(y - 1, N), z)]
                     + pb[idx(N, x, y, wrap(z + 1, N))] + pb[idx(N, x, y, wrap(z - 1, N))];
          double f = ph[i], q = f * f;
          double fp = p->c2 + 3.0 * p->c4 * q + 5.0 * p->c6 * q * q;
          bphi[i] = bphi[i] + dtk[n] * ((s - 6.0 * pb[i]) / p->a2 - fp * pb[i]);
          gc[0] -= dtk[n] * pb[i] * f;
          gc[1] -= dtk[n] * pb[i] * q * f;
          gc[2] -= dtk[n] * pb[i] * q * q * f;
        }
    memcpy(bpi, pb, sizeof(double) * n3);
  }
  memcpy(gphi0, bphi, sizeof(double) * n3);
  memcpy(gpi0, bpi, sizeof(double) * n3);
  free(traj); free(pi); free(g); free(bphi); free(bpi); free(pb);
  return F;
}
