/* oracle_track.cpp -- TEST INFRASTRUCTURE ONLY.  CPU restatement of island
 * tracking over time: Full_Analysis (islands.cpp:788-835) with its BFS
 * (islands.cpp:90-279, theory 0, one process), Old_islands (:651-705),
 * print_bubble_history (:707-736), is_bubble / is_oscillon (:61-70) and the
 * volume ordering (:29-32).  C++ only because the reference orders the
 * tracked islands with std::sort, whose order among equal volumes is part of
 * the behaviour: the same library algorithm with the same comparison gives
 * the same order.
 *
 * Parity pins: bitwise against the reference's own Full_Analysis run several
 * times on evolving fields (oracle/_ref/ref_harness, tracking mode), tracked
 * list and text files; the fixture tests/golden/track_*.npz from that build.
 *
 * Fields are in the reference layout: pos(x,y,z) = x + N*y + N*N*z with the
 * z halo planes 0 and N+1 (N^2 (N+2) values); delta likewise (mode 1). */
#include <algorithm>
#include <cmath>
#include <fstream>
#include <queue>
#include <string>
#include <vector>

extern "C" {
typedef struct {
  int N;
  double a;
  int mass_mode;
  double m2;
  int lambda_mode;
  double kappa;
} or_params;

typedef struct {
  double at_time;
  long long center_pos;
  double peak_value, avg_phi, volume, pot_en, kin_en, grad_en;
} or_track_entry;

typedef struct {
  long long index;
  int status, disp, nhist;
  or_track_entry hist[5];
} or_tracked_island;
}

namespace {

inline double sqr(double x) { return x * x; }
inline int mod(int i, int N) { return i >= 0 ? i % N : (N + i) % N; }

struct Island {  // islands.cpp:13-27
  long long index = 0;
  bool disp = false;
  std::vector<long long> center;
  std::vector<double> V, at_time, PotEn, KinEn, GradEn, avPhi, Peak_value;
  int status = 0;
};

struct Tracker {
  or_params p;
  std::vector<Island> found;
  long long number_of_islands = 0;
  std::vector<int> b;
};

bool is_island(int mode, const double *phi, const double *pi, const double *delta, double thr,
               long K) {  // islands.cpp:72-87
  if (mode == 0) return phi[K] * phi[K] * 0.5 + pi[K] * pi[K] * 0.5 > thr * thr * 0.5;
  return delta[K] > thr;
}

// BFS from P (islands.cpp:90-279): labels b[] with flag, pushes a history entry
void bfs(Tracker &t, int mode, double thr, double tau, const double *phi, const double *pi,
         const double *delta, long P, long long fflag, int flag, Island *B) {
  const int N = t.p.N, Nc = 1;
  const double a = t.p.a, a_eta = t.p.a, kap = t.p.kappa;
  int NN = 0;
  double cPotEn = 0., cGradEn = 0., cKinEn = 0., cavPhi = 0.;
  long cnt = P;
  std::queue<long> q;
  t.b[P] = flag;
  q.push(P);
  while (!q.empty()) {
    const long Kooo = q.front();
    q.pop();
    if (mode == 0) {
      if (phi[Kooo] * phi[Kooo] > phi[cnt] * phi[cnt]) cnt = Kooo;
    } else {
      if (delta[Kooo] > delta[cnt]) cnt = Kooo;
    }
    const int z = (int)(Kooo / ((long)Nc * N * N));
    const int y = (int)((Kooo - (long)Nc * N * N * z) / (Nc * N));
    const int x = (int)((Kooo - (long)Nc * N * N * z - (long)Nc * N * y) / Nc);
    const long Kuoo = mod(x + 1, N) + (long)N * y + (long)N * N * z;
    const long Kdoo = mod(x - 1, N) + (long)N * y + (long)N * N * z;
    const long Kouo = x + (long)N * mod(y + 1, N) + (long)N * N * z;
    const long Kodo = x + (long)N * mod(y - 1, N) + (long)N * N * z;
    const long Koou = x + (long)N * y + (long)N * N * (z + 1);
    const long Kood = x + (long)N * y + (long)N * N * (z - 1);
    NN++;
    cGradEn += 0.5 * (sqr(phi[Kuoo] - phi[Kooo]) + sqr(phi[Kouo] - phi[Kooo]) +
                      sqr(phi[Koou] - phi[Kooo])) / sqr(a);
    cKinEn += 0.5 * sqr(pi[Kooo]);
    cPotEn += (double)(t.p.mass_mode) * 0.5 * t.p.m2 * sqr(phi[Kooo]) +
              (double)(t.p.lambda_mode) * (sqr(sqr(phi[Kooo])) / (24.0 * Nc)) +
              (sqr(sqr(phi[Kooo])) * sqr(phi[Kooo]) / (720.0 * Nc * Nc * kap * kap));
    cavPhi += phi[Kooo];
    const long nb[4] = {Kuoo, Kdoo, Kouo, Kodo};
    for (long K : nb)
      if (is_island(mode, phi, pi, delta, thr, K) && t.b[K] == -1) {
        t.b[K] = flag;
        q.push(K);
      }
    if (z < N)
      if (is_island(mode, phi, pi, delta, thr, Koou) && t.b[Koou] == -1) {
        t.b[Koou] = flag;
        q.push(Koou);
      }
    if (z > 1)
      if (is_island(mode, phi, pi, delta, thr, Kood) && t.b[Kood] == -1) {
        t.b[Kood] = flag;
        q.push(Kood);
      }
  }
  cavPhi /= (double)(NN);
  const double cV = NN * a * a * a_eta;
  cPotEn *= (a * a * a_eta);
  cKinEn *= (a * a * a_eta);
  cGradEn *= (a * a * a_eta);
  B->center.push_back(cnt);
  B->PotEn.push_back(cPotEn);
  B->KinEn.push_back(cKinEn);
  B->GradEn.push_back(cGradEn);
  B->avPhi.push_back(cavPhi);
  B->V.push_back(cV);
  B->Peak_value.push_back(mode == 0 ? phi[cnt] : delta[cnt]);
  B->at_time.push_back(tau);
  B->index = fflag;
  B->status = 0;
  if (B->center.size() > 5) {
    B->PotEn.erase(B->PotEn.begin() + 2);
    B->center.erase(B->center.begin() + 2);
    B->KinEn.erase(B->KinEn.begin() + 2);
    B->GradEn.erase(B->GradEn.begin() + 2);
    B->avPhi.erase(B->avPhi.begin() + 2);
    B->V.erase(B->V.begin() + 2);
    B->Peak_value.erase(B->Peak_value.begin() + 2);
    B->at_time.erase(B->at_time.begin() + 2);
  }
}

// point::x/y/z (fieldanalysis.cpp:19-32), id 0
long long px(const Tracker &t, long long pos) { return (pos % ((long long)t.p.N * t.p.N)) % t.p.N; }
long long py(const Tracker &t, long long pos) { return (pos % ((long long)t.p.N * t.p.N)) / t.p.N; }
long long pz(const Tracker &t, long long pos) { return -1 + pos / ((long long)t.p.N * t.p.N); }

void history_rows(std::ofstream &o, const Tracker &t, const Island &I, size_t from) {
  for (size_t j = from; j < I.V.size(); j++) {
    o << I.at_time[j] << ' ' << px(t, I.center[j]) << ' ' << py(t, I.center[j]) << ' '
      << pz(t, I.center[j]) << "        ";
    o << I.Peak_value[j] << ' ' << I.avPhi[j] << ' ' << I.V[j] << ' ';
    o << I.GradEn[j] << ' ' << I.PotEn[j] << ' ' << I.KinEn[j] << ' ' << I.index << std::endl;
  }
}

void init_row(std::ofstream &o, const Island &I, int status, double thres) {
  o << pow(I.V[0] * 3. / 4. / acos(-1.), 1. / 3.) << ' ' << I.Peak_value[0] << ' ' << I.at_time[0]
    << ' ' << I.index << ' ' << status << ' '
    << pow(I.V[0] * 3. / 4. / acos(-1.), 1. / 3.) / sqrt(log(I.Peak_value[0] * I.Peak_value[0] / thres / thres))
    << std::endl;
}

}  // namespace

extern "C" {

void *or_track_create(const or_params *p) {
  Tracker *t = new Tracker();
  t->p = *p;
  t->b.assign((size_t)p->N * p->N * (p->N + 2), -1);
  return t;
}

void or_track_destroy(void *h) { delete (Tracker *)h; }

int or_track_count(void *h) { return (int)((Tracker *)h)->found.size(); }

void or_track_get(void *h, int i, or_tracked_island *out) {
  const Island &I = ((Tracker *)h)->found[i];
  out->index = I.index;
  out->status = I.status;
  out->disp = I.disp ? 1 : 0;
  out->nhist = (int)I.V.size();
  for (int j = 0; j < out->nhist; j++) {
    or_track_entry &e = out->hist[j];
    e.at_time = I.at_time[j];
    e.center_pos = I.center[j];
    e.peak_value = I.Peak_value[j];
    e.avg_phi = I.avPhi[j];
    e.volume = I.V[j];
    e.pot_en = I.PotEn[j];
    e.kin_en = I.KinEn[j];
    e.grad_en = I.GradEn[j];
  }
}

/* One Full_Analysis call (islands.cpp:788-835).  outdir != NULL: the
 * reference's three text files are appended there. */
void or_track_update(void *h, int mode, double thr, const double *phi, const double *pi,
                     const double *delta, double tau, const char *outdir) {
  Tracker &t = *(Tracker *)h;
  const int N = t.p.N;
  std::vector<Island> &F = t.found;
  std::fill(t.b.begin(), t.b.end(), -1);
  std::sort(F.begin(), F.end(), [](Island &B1, Island &B2) { return B1.V.back() > B2.V.back(); });
  std::ofstream safe, init, bub;
  if (outdir) {
    safe.open(std::string(outdir) + "/non_critical_islands.txt", std::ios::app);
    init.open(std::string(outdir) + "/islands.txt", std::ios::app);
  }
  // Old_islands (:651-705)
  for (size_t i = 0; i < F.size();) {
    if (!is_island(mode, phi, pi, delta, thr, F[i].center.back())) {
      if (outdir) {
        if (F[i].disp) {
          history_rows(safe, t, F[i], 0);
          safe << std::endl;
        }
        init_row(init, F[i], F[i].status, thr);
      }
      F.erase(F.begin() + i);
    } else {
      i++;
    }
  }
  for (int i = 0; i < (int)F.size(); i++) {
    const long P = F[i].center.back();
    if (is_island(mode, phi, pi, delta, thr, P) && t.b[P] != -1) {  // merged
      F.erase(F.begin() + i);
      i--;
      continue;
    }
    if (is_island(mode, phi, pi, delta, thr, P) && t.b[P] == -1)
      bfs(t, mode, thr, tau, phi, pi, delta, P, F[i].index, i, &F[i]);
    const bool bubble = F[i].GradEn.back() + F[i].PotEn.back() < 0.;
    if (bubble && F[i].status == 0) {  // print_bubble_history (:707-736)
      if (outdir) {
        if (!bub.is_open()) bub.open(std::string(outdir) + "/expanding_bubbles.txt", std::ios::app);
        if (F[i].disp) init_row(init, F[i], 1, thr);
        history_rows(bub, t, F[i], 1);
        bub << std::endl;
      }
      F[i].status = 1;
    }
    if (F[i].at_time.back() - F[i].at_time.front() > 40. && F[i].status == 0) F[i].status = 2;
  }
  const long end = (long)N * N * (N + 1);
  for (long K = (long)N * N; K < end; K++)
    if (is_island(mode, phi, pi, delta, thr, K) && t.b[K] == -1) {
      Island bbb;
      bfs(t, mode, thr, tau, phi, pi, delta, K, t.number_of_islands, (int)F.size(), &bbb);
      bbb.disp = true;
      t.number_of_islands++;
      F.push_back(bbb);
    }
}

}  /* extern "C" */
