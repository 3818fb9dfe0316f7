// phinum_dropin.cpp -- reference-named host layer over the C ABI.
#include "phinum_dropin.hpp"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <iostream>
#include <map>
#include <stdexcept>
#include <sys/stat.h>

namespace phinum {

// ---- parameters.cpp defaults (parameters.cpp:10,88-100,172-195,289-301) ----
int runs = 10, runID = 0;
int Nc = 1, Mass_Mode = 1, Lambda_Mode = -1;
double M_sqr = 1., Lambda = 100., Kappa = 2.;
int renorm_tadpole = 0, max_renSteps = 2000;
double tad_precision = 1e-6;
double a_t = 0.5, a_eta = 0.5;
int N_t = 64, N_eta = 64;
double T0 = 0, TMax = 100000, TSO = 0.1, TMO = 50., dtBaseTrans = 0.1, dtBaseLong = 0.1;
int justnoise = 0;
double QS = 1., scale2 = 1., Amplitude = 0., chi = 1., Noise = 0.5;
double phi0_0 = 0., phi0_1 = 0., pi0_0 = 0., pi0_1 = 0.;
double M2_long = 0., M2_trans = 0.;
int nodes = 1, ID = 0, LASTRUN = 0;
long long seed = 1234;
int precision = PN_FP64;
int islanddetect = 0;
int island_mode = 0;
int island_every = 5;
int write_output = 1;
int save_config = -1;
int extract_config = -1;

static std::string g_err;
std::string last_error() { return g_err; }

static void check(int rc, const char *what) {
  if (rc != PN_OK) {
    g_err = std::string(what) + ": " + pn_last_error(LocalLattice::ctx);
    throw std::runtime_error(g_err);
  }
}

// Steps taken by evolve() on the host clock (tau, Last_*_update and the step
// counts advance at once, as in the reference) but not yet launched.  They
// go to the device in one pn_evolve call, so pairs of steps share a pass over
// HBM: before anything reads or writes the device fields, and every
// kMaxPending steps.  The fields are the same bits as with one launch per step.
static std::vector<double> g_pend_dtk, g_pend_dtd;
static constexpr size_t kMaxPending = 256;

static void flush_steps() {
  if (g_pend_dtk.empty()) return;
  const int n = (int)g_pend_dtk.size();
  const int rc = pn_evolve(LocalLattice::ctx, n, g_pend_dtk.data(), g_pend_dtd.data());
  g_pend_dtk.clear();
  g_pend_dtd.clear();
  check(rc, "evolve");
}

// -key value pairs as Konfig (cfile.c:55-73): a later duplicate overwrites,
// values are read with atof/atoi (cfile.c:252-278), unknown keys are ignored.
void GetParameters(int argc, char **argv) {
  std::map<std::string, std::string> kv;
  for (int i = 1; i < argc; i++)
    if (argv[i][0] == '-') {
      std::string key(argv[i] + 1);
      if (key.size() > 79) key.resize(79);
      kv[key] = (i + 1 < argc) ? argv[i + 1] : key;
    }
  auto gi = [&](const char *k, int &v) { auto it = kv.find(k); if (it != kv.end()) v = std::atoi(it->second.c_str()); };
  auto gd = [&](const char *k, double &v) { auto it = kv.find(k); if (it != kv.end()) v = std::atof(it->second.c_str()); };
  gi("runs", runs);
  gi("Nc", Nc);
  gd("M", M_sqr);
  gi("Mmode", Mass_Mode);
  gd("L", Lambda);
  gi("Lmode", Lambda_Mode);
  gd("Kappa", Kappa);
  gi("tadpole", renorm_tadpole);
  gd("tadprec", tad_precision);
  gi("renSteps", max_renSteps);
  gi("Nt", N_t);
  gd("at", a_t);
  gd("T0", T0);
  gd("TMax", TMax);
  gd("dt", dtBaseLong);
  gd("TSO", TSO);
  gd("TMO", TMO);
  gi("jn", justnoise);
  gd("Q2", scale2);
  gd("Ampl", Amplitude);
  gd("QS", QS);
  gd("chi", chi);
  gd("Noise", Noise);
  gd("phi0_0", phi0_0);
  gd("phi0_1", phi0_1);
  gd("pi0_0", pi0_0);
  gd("pi0_1", pi0_1);
  auto it = kv.find("seed");
  if (it != kv.end()) seed = std::atoll(it->second.c_str());
  gi("precision", precision);
  gi("islands", islanddetect);
  gi("islandmode", island_mode);
  gi("islandevery", island_every);
  gi("output", write_output);
  gi("saveconfig", save_config);
  gi("extractconfig", extract_config);
  if (Nc != 1) throw std::runtime_error("GetParameters: the device path evolves one component (Nc=1)");
  if (renorm_tadpole != 0)
    throw std::runtime_error("GetParameters: tadpole renormalisation is not implemented upstream either");
}

void CompleteParameters() {  // parameters.cpp:398-411
  a_eta = a_t;
  N_eta = N_t;
  dtBaseTrans = dtBaseLong;
  Noise *= Lambda;
}

// ---- lattice.cpp ----------------------------------------------------------
namespace LocalLattice {
double tau = 0, dtau = 0;
int N_t_loc = 0, N_eta_loc = 0, vol = 0;
double *Phi = nullptr, *Pi = nullptr;
double *phi = nullptr, *pi = nullptr;
double pi2, phi2, phi4, phi6, phipi, Energy, Interaction, Charge;
double *avgphi = nullptr, *avgpi = nullptr;
double avgpi2, avgphi2, avgphi4, avgphi6, avgphipi, avgEnergy, avgInteraction, avgCharge;
double maxPhi, minPhi, maxAbsPhi;
int maxpos, minpos, maxAbspos;
double GlmaxPhi, GlminPhi, GlmaxAbsPhi;
double BareM_sqr_long = 1., BareM_sqr_trans = 1.;
pn_ctx *ctx = nullptr;

void init() {  // lattice.cpp:70-104
  N_t_loc = N_t;
  N_eta_loc = N_eta / nodes;
  vol = N_t_loc * N_t_loc * N_eta_loc;
  const size_t n = (size_t)Nc * N_t * N_t * (N_eta_loc + 2);
  Phi = new double[n]();
  Pi = new double[n]();
  phi = new double[Nc]();
  pi = new double[Nc]();
  avgphi = new double[Nc]();
  avgpi = new double[Nc]();
  BareM_sqr_long = M_sqr;
  BareM_sqr_trans = M_sqr;
  pn_params p{N_t, a_t, Mass_Mode, BareM_sqr_long, Lambda_Mode, Kappa, precision, 0};
  check(pn_create(&p, &ctx), "LocalLattice::init");
}

void fin() {
  g_pend_dtk.clear();  // nothing can observe them any more
  g_pend_dtd.clear();
  delete[] Phi;
  delete[] Pi;
  delete[] phi;
  delete[] pi;
  delete[] avgphi;
  delete[] avgpi;
  Phi = Pi = phi = pi = avgphi = avgpi = nullptr;
  pn_destroy(ctx);
  ctx = nullptr;
}

void SyncToHost() {
  flush_steps();
  check(pn_get_fields(ctx, Phi, Pi, PN_LAYOUT_REF), "SyncToHost");
}
void SyncToDevice() {
  flush_steps();
  check(pn_set_fields(ctx, Phi, Pi, PN_LAYOUT_REF), "SyncToDevice");
}
}  // namespace LocalLattice

// ---- renormalization.cpp / initialconditions.cpp ---------------------------
void fix_tadpole() {  // renormalization.cpp:3-31 with renorm_tadpole == 0
  LocalLattice::BareM_sqr_long = M_sqr;
  LocalLattice::BareM_sqr_trans = M_sqr;
  M2_long = M_sqr;
  M2_trans = M_sqr;
  flush_steps();  // steps already taken used the old couplings
  check(pn_set_couplings(LocalLattice::ctx, Mass_Mode, LocalLattice::BareM_sqr_long, Lambda_Mode, Kappa),
        "fix_tadpole");
}

void SetInitialConditions() {
  pn_ic_params ic;
  ic.seed = seed + (runID > 0 ? runID - 1 : 0);
  ic.QS = QS;
  ic.scale2 = scale2;
  ic.amplitude = Amplitude;
  ic.noise = Noise;
  ic.m2_long = M2_long;
  ic.justnoise = justnoise;
  ic.phi0 = phi0_0;
  ic.pi0 = pi0_0;
  flush_steps();
  check(pn_initial_conditions(LocalLattice::ctx, &ic), "SetInitialConditions");
}

// ---- dynamics.cpp -----------------------------------------------------------
namespace numeric {
int regime = 0, total_stepcount = 0, normal_stepcount = 0, adaptive_stepcount = 0;
double Last_Field_update = 0, Last_Momenta_update = 0, LastMOoutput = 0, LastSOoutput = 0;
}  // namespace numeric

void reset() {  // dynamics.cpp:28-50
  using namespace numeric;
  regime = 0;
  total_stepcount = normal_stepcount = adaptive_stepcount = 0;
  LocalLattice::tau = T0;
  LocalLattice::dtau = dtBaseTrans * a_t;
  Last_Field_update = LocalLattice::tau;
  Last_Momenta_update = LocalLattice::tau;
}

void updateMomenta() {  // dynamics.cpp:52-175
  flush_steps();
  check(pn_kick(LocalLattice::ctx, LocalLattice::tau - numeric::Last_Momenta_update), "updateMomenta");
  numeric::Last_Momenta_update = LocalLattice::tau;
}

void updateFields() {  // dynamics.cpp:177-194
  flush_steps();
  check(pn_drift(LocalLattice::ctx, LocalLattice::tau - numeric::Last_Field_update), "updateFields");
  numeric::Last_Field_update = LocalLattice::tau;
}

void evolve() {  // dynamics.cpp:196-237: kick and drift fused, launched by flush_steps
  using LocalLattice::tau;
  tau += 0.5 * LocalLattice::dtau;
  const double dtk = tau - numeric::Last_Momenta_update;
  numeric::Last_Momenta_update = tau;
  tau += 0.5 * LocalLattice::dtau;
  const double dtd = tau - numeric::Last_Field_update;
  numeric::Last_Field_update = tau;
  g_pend_dtk.push_back(dtk);
  g_pend_dtd.push_back(dtd);
  if (g_pend_dtk.size() >= kMaxPending) flush_steps();
  numeric::normal_stepcount++;
  Communication::exchange();
  numeric::total_stepcount++;
}

namespace Communication {
// comm.cpp:63-78: the periodic z wrap is index arithmetic inside the device
// kernels, so there is nothing to exchange
void exchange() {}
}  // namespace Communication

// ---- fieldanalysis.cpp ----------------------------------------------------
static void set_averages(const pn_averages &a) {
  using namespace LocalLattice;
  phi[0] = avgphi[0] = a.phi;
  pi[0] = avgpi[0] = a.pi;
  phi2 = avgphi2 = a.phi2;
  pi2 = avgpi2 = a.pi2;
  phipi = avgphipi = a.phipi;
  phi4 = avgphi4 = a.phi4;
  phi6 = avgphi6 = a.phi6;
  Energy = avgEnergy = a.energy;
  Interaction = avgInteraction = a.interaction;
  Charge = avgCharge = a.charge;
}

void CalculateVolumeAverages() {
  pn_averages a;
  flush_steps();
  check(pn_volume_averages(LocalLattice::ctx, &a), "CalculateVolumeAverages");
  set_averages(a);
}

static void set_extrema(const pn_extrema &e) {
  using namespace LocalLattice;
  GlminPhi = minPhi = e.min;
  GlmaxPhi = maxPhi = e.max;
  GlmaxAbsPhi = maxAbsPhi = e.maxabs;
  minpos = (int)e.minpos;
  maxpos = (int)e.maxpos;
  maxAbspos = (int)e.maxabspos;
}

void CalculateExtrema() {
  pn_extrema e;
  flush_steps();
  check(pn_extrema_get(LocalLattice::ctx, &e), "CalculateExtrema");
  set_extrema(e);
}

static std::ofstream SpatialObs;

void Prepare_Output() {  // fieldanalysis.cpp:48-79
  if (!write_output || ID != 0) return;
  {
    std::ifstream fin("lastrun.txt");
    fin >> LASTRUN;
  }
  {
    std::ofstream fout("lastrun.txt");
    fout << ++LASTRUN << std::endl;
  }
  char folder[256];
  std::snprintf(folder, sizeof folder, "./output/%i", LASTRUN);
  struct stat info;
  if (stat(folder, &info) == 0 && S_ISDIR(info.st_mode))
    std::cout << "Warning: Output folder " << folder << " already exists! Files will be overwritten." << std::endl;
  mkdir("./output", 0755);
  mkdir(folder, 0755);
}

// The parameter summary of Print_Info_Console and Print_Info_File for theory
// 0 and initialize 3 (fieldanalysis.cpp:81-158,160-247).  The two differ in
// one line: the console prints the time step a_t*dtBaseTrans as "dt", the
// file the two base factors as "dtBT" and "dtBL".
static void write_info(std::ostream &out, bool file) {
  out << "Parameters" << std::endl;
  out << "Nc=" << Nc << std::endl;
  out << "m^2=" << Mass_Mode * M_sqr << std::endl;
  out << "Lambda=" << Lambda_Mode * Lambda << std::endl;
  out << "Nyu=" << (Lambda / Kappa) * (Lambda / Kappa) << std::endl;
  out << "N_s=" << N_t << std::endl;
  out << "a_s=" << a_t << std::endl;
  out << "T0=" << T0 << std::endl;
  out << "TMax=" << TMax << std::endl;
  if (file) {
    out << "dtBT=" << dtBaseTrans << std::endl;
    out << "dtBL=" << dtBaseLong << std::endl;
  } else {
    out << "dt=" << a_t * dtBaseTrans << std::endl;
  }
  out << "TSO=" << TSO << std::endl;
  out << "TMO=" << TMO << std::endl;
  out << "f0(p)=" << Amplitude << "/Lambda, for p<" << QS << std::endl;
  out << "phi0_0=" << phi0_0 << std::endl;
  out << "phi0_1=" << phi0_1 << std::endl;
  out << "pi0_0=" << pi0_0 << std::endl;
  out << "pi0_1=" << pi0_1 << std::endl;
  out << std::endl << runs;
  out << (runs == 1 ? " run with" : " runs with");
  out << (renorm_tadpole == 1 ? " " : "out ");
  out << "renormalization." << std::endl;
}

void Print_Info_Console() {  // fieldanalysis.cpp:81-158
  if (ID != 0) return;
  write_info(std::cout, false);
}

void Print_Info_File() {  // fieldanalysis.cpp:160-247: ./output/<LASTRUN>/info.txt
  if (ID != 0 || !write_output) return;
  char fname[256];
  std::snprintf(fname, sizeof fname, "./output/%i/info.txt", LASTRUN);
  std::ofstream out(fname);
  write_info(out, true);
}

void Print_AveragesExtrema() {  // fieldanalysis.cpp:748-799
  using namespace LocalLattice;
  if (numeric::total_stepcount > 0 && tau < numeric::LastSOoutput + TSO) return;
  if (numeric::total_stepcount == 0) {
    if (ID == 0 && write_output) {
      if (SpatialObs.is_open()) SpatialObs.close();
      char fname[256];
      std::snprintf(fname, sizeof fname, "./output/%i/SO-%i.txt", LASTRUN, runID);
      SpatialObs.open(fname, std::fstream::app);
    }
    numeric::LastSOoutput = tau;
  } else if (tau >= numeric::LastSOoutput + TSO) {
    numeric::LastSOoutput += TSO;
  }
  // CalculateExtrema() and CalculateVolumeAverages() (fieldanalysis.cpp:765-766)
  // from one pass over the fields
  pn_averages a;
  pn_extrema e;
  flush_steps();
  check(pn_volume_analysis(ctx, &a, &e), "Print_AveragesExtrema");
  set_extrema(e);
  set_averages(a);
  if (ID == 0 && write_output && SpatialObs.is_open()) {
    SpatialObs << numeric::LastSOoutput << ' ';
    for (int i = 0; i < Nc; i++) SpatialObs << avgphi[i] << ' ' << avgpi[i] << ' ';
    SpatialObs << avgphi2 << ' ' << avgpi2 << ' ' << avgphipi << ' ' << avgphi4 << ' ' << avgphi6 << ' ';
    SpatialObs << avgEnergy << ' ' << avgInteraction << ' ' << avgCharge << ' ';
    SpatialObs << GlminPhi << ' ' << GlmaxPhi << ' ' << GlmaxAbsPhi << std::endl;
  }
}

bool false_vacuum_decayed() {  // fieldanalysis.cpp:887-893
  const double I = LocalLattice::avgInteraction;
  return (I < 0 || I != I || I * 0. != 0.);
}

// ---- momentumspace.cpp ------------------------------------------------------
namespace MomentumSpace {
double pTmin = 0, pTmax = 0, binsizeT = 0;
int nTBins = 0;
std::vector<double> TotalCorrF, TotalCorrK, TotalCorrQ;
std::vector<int> TotalBCounts;
void init() {  // momentumspace.cpp:32-65
  pTmin = (2.0 / (a_t)) * std::sin(std::acos(-1.) / N_t);
  pTmax = std::sqrt(3.0 * (4.0 / (a_t * a_t)));
  binsizeT = pTmin;
  nTBins = int(pTmax / binsizeT + 0.5) + 2;
  TotalCorrF.assign(nTBins, 0.);
  TotalCorrK.assign(nTBins, 0.);
  TotalCorrQ.assign(nTBins, 0.);
  TotalBCounts.assign(nTBins, 0);
}
void fin() {
  TotalCorrF.clear();
  TotalCorrK.clear();
  TotalCorrQ.clear();
  TotalBCounts.clear();
}
}  // namespace MomentumSpace

void Calculate_CorrelationFunctions() {  // momentumspace.cpp:140-278
  using namespace MomentumSpace;
  pn_spectrum s;
  s.nbins = nTBins;
  s.counts = TotalBCounts.data();
  s.corr_f = TotalCorrF.data();
  s.corr_k = TotalCorrK.data();
  s.corr_q = TotalCorrQ.data();
  flush_steps();
  check(pn_spectrum_get(LocalLattice::ctx, &s), "Calculate_CorrelationFunctions");
}

void Print_fp() {  // momentumspace.cpp:280-384
  using namespace LocalLattice;
  using namespace MomentumSpace;
  if (numeric::total_stepcount > 0 && tau < numeric::LastMOoutput + TMO) return;
  if (numeric::total_stepcount == 0) numeric::LastMOoutput = tau;
  else if (tau >= numeric::LastMOoutput + TMO) numeric::LastMOoutput += TMO;
  Calculate_CorrelationFunctions();
  if (ID != 0 || !write_output) return;
  char fname[256];
  std::snprintf(fname, sizeof fname, "./output/%i/Distrib-%i-%g.txt", LASTRUN, runID,
                double(numeric::LastMOoutput));
  std::ofstream specOut(fname);
  double NormCorr = 1. / ((N_t * a_t) * (N_t * a_t) * N_eta * a_eta);
  for (int k = 0; k < nTBins; k++) {
    double norm2 = (TotalBCounts[k] == 0 ? 0 : 1 / (double)(TotalBCounts[k]));
    double pabs = binsizeT * k;
    if (norm2 != 0 && TotalCorrF[k * Nc] != 0) {
      if (k != 0) specOut << '\n';
      specOut << pabs << " ";
      specOut << TotalBCounts[k] << " ";
      double FF = TotalCorrF[k] / Nc, KK = TotalCorrK[k] / Nc, QQ = TotalCorrQ[k] / Nc;
      double omg = std::sqrt(KK / FF);
      specOut << omg << ' ';
      specOut << FF * NormCorr * norm2 << ' ' << KK * NormCorr * norm2 << ' '
              << std::sqrt(FF * KK) * NormCorr * norm2 << " ";
      specOut << FF * NormCorr * norm2 * pabs * pabs * pabs / 2. / std::acos(-1.) / std::acos(-1.) << ' ';
      if (FF * KK - QQ * QQ < 0) std::cout << "negative at " << k << std::endl;
    }
  }
}

// ---- islands.cpp -------------------------------------------------------------
namespace island_analysis {
double thres = 1.125;
int number_of_islands = 0;
std::vector<pn_tracked_island> found_islands;
std::vector<pn_island> catalog;
static pn_tracker *tracker = nullptr;
void init() {  // islands.cpp:46-51
  number_of_islands = 0;
  found_islands.clear();
  catalog.clear();
  if (tracker) pn_tracker_destroy(tracker);
  tracker = nullptr;
  if (pn_tracker_create(&tracker) != PN_OK) throw std::runtime_error("island_analysis::init: " + last_error());
}
void fin() {
  found_islands.clear();
  catalog.clear();
  if (tracker) pn_tracker_destroy(tracker);
  tracker = nullptr;
}
}  // namespace island_analysis

// islands.cpp:778-840: the tracked islands re-grow from their last peaks,
// retired ones and expanding bubbles go to the run's text files, new
// components become new islands (pn_tracker_update); `catalog` keeps the
// components of this call.
void Full_Analysis() {
  using namespace island_analysis;
  if (island_mode == 1) {  // islands.cpp:783-786
    CalculateVolumeAverages();
    thres = LocalLattice::avgEnergy * 4.;
  }
  if (!tracker) init();
  char dir[64];
  std::snprintf(dir, sizeof dir, "./output/%i", LASTRUN);
  const bool files = write_output && ID == 0;
  flush_steps();
  int rc = pn_tracker_update(tracker, LocalLattice::ctx, island_mode, thres, LocalLattice::tau,
                             files ? dir : nullptr);
  if (rc != PN_OK) throw std::runtime_error(std::string("Full_Analysis: ") + pn_tracker_last_error(tracker));
  int n = 0;
  pn_tracker_count(tracker, &n);
  found_islands.resize(n);
  for (int i = 0; i < n; i++) {
    pn_tracker_get(tracker, i, &found_islands[i]);
    number_of_islands = std::max<int>(number_of_islands, (int)found_islands[i].index + 1);
  }
  // the components of this call, as numbered by the raster scan (the
  // tracker's update already computed them)
  pn_catalog cat{0, 0, 0.0, nullptr, nullptr};
  pn_tracker_catalog(tracker, &cat);  // the count
  catalog.resize(cat.count);
  cat.capacity = cat.count;
  cat.islands = catalog.data();
  if (pn_tracker_catalog(tracker, &cat) != PN_OK)
    throw std::runtime_error(std::string("Full_Analysis: ") + pn_tracker_last_error(tracker));
}

// ---- checkpoint text (fieldanalysis.cpp:252-313) ---------------------------
// ./field-<Name>-<ID>.out: one "Phi Pi" line per interior site in pos()
// order, %.17e for FP64 fields and %.9e for FP32 ones (the reference's
// doubleprecision 0 / -1 formats); ./moment-<Name>.out: tau.
void SaveConfig(int Name) {
  using namespace LocalLattice;
  SyncToHost();
  char fname[256];
  std::snprintf(fname, sizeof fname, "./field-%i-%i.out", Name, ID);
  FILE *of = std::fopen(fname, "w");
  if (!of) throw std::runtime_error(std::string("SaveConfig: cannot write ") + fname);
  const long end = (long)Nc * N_t * N_t * (N_eta_loc + 1);
  const char *fmt = precision == 1 ? "%.9e %.9e\n" : "%.17e %.17e\n";
  for (long K = (long)Nc * N_t * N_t; K < end; K++) std::fprintf(of, fmt, Phi[K], Pi[K]);
  std::fclose(of);
  if (ID == 0) {
    std::snprintf(fname, sizeof fname, "./moment-%i.out", Name);
    std::ofstream fout(fname);
    fout << tau << std::endl;
  }
}

// Reads a checkpoint back: the fields go to the device, T0 takes the saved
// tau (the next reset() starts from it).
void ExtractConfig(int Name) {
  using namespace LocalLattice;
  char fname[256];
  std::snprintf(fname, sizeof fname, "./field-%i-%i.out", Name, ID);
  FILE *of = std::fopen(fname, "r");
  if (!of) throw std::runtime_error(std::string("ExtractConfig: cannot read ") + fname);
  const long end = (long)Nc * N_t * N_t * (N_eta_loc + 1);
  for (long K = (long)Nc * N_t * N_t; K < end; K++)
    if (std::fscanf(of, "%lf %lf\n", Phi + K, Pi + K) != 2) {
      std::fclose(of);
      throw std::runtime_error(std::string("ExtractConfig: short file ") + fname);
    }
  std::fclose(of);
  SyncToDevice();
  std::snprintf(fname, sizeof fname, "./moment-%i.out", Name);
  std::ifstream fin(fname);
  fin >> T0;
}

}  // namespace phinum
