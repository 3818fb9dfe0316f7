// phinum_dropin.hpp -- host C++ drop-in for PhiNum's lattice hot path.
//
// Keeps the reference's call surface (free functions and namespace globals
// with the same names and meaning) on top of the C ABI in phinum_gpu.h, so a
// driver written against the reference (main.cpp:21-117) keeps working with
// the lattice resident on one MI355X:
//   parameters.cpp : GetParameters, CompleteParameters and the globals
//   lattice.cpp    : LocalLattice::{init, fin, tau, dtau, Phi, Pi, avg*, ...}
//   dynamics.cpp   : reset, updateMomenta, updateFields, evolve, numeric::*
//   comm.cpp       : Communication::exchange (device-side wrap: a no-op)
//   fieldanalysis  : CalculateVolumeAverages, CalculateExtrema,
//                    Print_AveragesExtrema, false_vacuum_decayed, Prepare_Output
//   momentumspace  : Calculate_CorrelationFunctions, Print_fp, MomentumSpace::*
//   islands.cpp    : Full_Analysis, island_analysis::*
//   initialconditions.cpp / renormalization.cpp : SetInitialConditions, fix_tadpole
// Phi and Pi are host mirrors in the reference layout (pos() with the two z
// halo planes); they are filled by LocalLattice::SyncToHost() and pushed back
// by LocalLattice::SyncToDevice() after host-side edits.
#pragma once
#include <string>
#include <vector>

#include "../../../include/phinum_gpu.h"

namespace phinum {

// ---- parameters.cpp -------------------------------------------------------
extern int runs, runID;
extern int Nc, Mass_Mode, Lambda_Mode;
extern double M_sqr, Lambda, Kappa;
extern int renorm_tadpole, max_renSteps;
extern double tad_precision;
extern double a_t, a_eta;
extern int N_t, N_eta;
extern double T0, TMax, TSO, TMO, dtBaseTrans, dtBaseLong;
extern int justnoise;
extern double QS, scale2, Amplitude, chi, Noise, phi0_0, phi0_1, pi0_0, pi0_1;
extern double M2_long, M2_trans;
extern int nodes, ID, LASTRUN;
// additions: parameters the reference fixes at compile time or by the clock
extern long long seed;        // -seed   (reference: time(NULL)*(ID+1))
extern int precision;         // -precision 0 FP64, 1 FP32 fields
extern int islanddetect;      // -islands 0/1 (reference: #define islanddetect)
extern int island_mode;       // -islandmode 0: phi^2/2+pi^2/2>thr^2/2, 1: Delta>4<E>
extern int island_every;      // -islandevery n (reference: 5, main.cpp:82)
extern int write_output;      // -output 0/1: write ./output/<run>/ files
extern int save_config;       // -saveconfig n: SaveConfig(n) at the end of each run (-1: off)
extern int extract_config;    // -extractconfig n: ExtractConfig(n) after the initial conditions

void GetParameters(int argc, char **argv);
void CompleteParameters();

// ---- lattice.cpp ----------------------------------------------------------
namespace LocalLattice {
extern double tau, dtau;
extern int N_t_loc, N_eta_loc, vol;
extern double *Phi, *Pi;                 // host mirrors, pos() layout with halos
extern double *phi, *pi;                 // per-component local averages
extern double pi2, phi2, phi4, phi6, phipi, Energy, Interaction, Charge;
extern double *avgphi, *avgpi;
extern double avgpi2, avgphi2, avgphi4, avgphi6, avgphipi, avgEnergy, avgInteraction, avgCharge;
extern double maxPhi, minPhi, maxAbsPhi;
extern int maxpos, minpos, maxAbspos;
extern double GlmaxPhi, GlminPhi, GlmaxAbsPhi;
extern double BareM_sqr_long, BareM_sqr_trans;
extern pn_ctx *ctx;                      // the device lattice
void init();
void fin();
void SyncToHost();
void SyncToDevice();
}  // namespace LocalLattice

// ---- dynamics.cpp ---------------------------------------------------------
namespace numeric {
extern int regime, total_stepcount, normal_stepcount, adaptive_stepcount;
extern double Last_Field_update, Last_Momenta_update, LastMOoutput, LastSOoutput;
}  // namespace numeric
void reset();
void updateMomenta();
void updateFields();
void evolve();
namespace Communication {
void exchange();
}

// ---- analysis ---------------------------------------------------------------
void fix_tadpole();
void SetInitialConditions();
void CalculateVolumeAverages();
void CalculateExtrema();
void Prepare_Output();
void Print_Info_Console();
void Print_Info_File();
void Print_AveragesExtrema();
bool false_vacuum_decayed();
void SaveConfig(int Name);     // fieldanalysis.cpp:252-282
void ExtractConfig(int Name);  // fieldanalysis.cpp:284-313

namespace MomentumSpace {
extern double pTmin, pTmax, binsizeT;
extern int nTBins;
extern std::vector<double> TotalCorrF, TotalCorrK, TotalCorrQ;
extern std::vector<int> TotalBCounts;
void init();
void fin();
}  // namespace MomentumSpace
void Calculate_CorrelationFunctions();
void Print_fp();

namespace island_analysis {
extern double thres;
extern int number_of_islands;
extern std::vector<pn_tracked_island> found_islands;  // tracked over time
extern std::vector<pn_island> catalog;                // components of the last call
void init();
void fin();
}  // namespace island_analysis
void Full_Analysis();

// first error text of the device layer (the reference aborts through MPI;
// the drop-in throws std::runtime_error with this text)
std::string last_error();

}  // namespace phinum
