// phinum_main.cpp -- the reference driver (main.cpp:21-117) on the drop-in
// layer: same parameters (-key value), same run loop, output cadence, island
// cadence and false-vacuum break, with the lattice on one MI355X.  The MPI
// world is one process; statistics.cpp's ensemble histograms are not kept.
#include <cmath>
#include <iostream>
#include <stdexcept>

#include "phinum_dropin.hpp"

int main(int argc, char **argv) {
  using namespace phinum;
  try {
    nodes = 1;
    ID = 0;
    std::cout << "\nPhiNum lattice on one MI355X: " << nodes << " process(es).\n";
    GetParameters(argc, argv);
    CompleteParameters();
    Prepare_Output();
    Print_Info_Console();
    Print_Info_File();
    LocalLattice::init();
    MomentumSpace::init();
    for (runID = 1; runID <= runs; runID++) {
      if (runs != 1) std::cout << "Run: " << runID << std::endl;
      fix_tadpole();
      if (islanddetect) island_analysis::init();
      SetInitialConditions();
      if (extract_config >= 0) ExtractConfig(extract_config);  // resume from a checkpoint
      if (islanddetect) {  // main.cpp:69-72
        CalculateVolumeAverages();
        island_analysis::thres = 2.57583 * std::sqrt(LocalLattice::avgphi2);
        Full_Analysis();
      }
      for (reset(); LocalLattice::tau < TMax; evolve()) {  // main.cpp:77-91
        Print_AveragesExtrema();
        if (islanddetect && numeric::total_stepcount % island_every == 0) Full_Analysis();
        if (false_vacuum_decayed()) {
          std::cout << "false vacuum decayed at tau=" << LocalLattice::tau << std::endl;
          break;
        }
        Print_fp();
      }
      if (save_config >= 0) SaveConfig(save_config);
      if (islanddetect) island_analysis::fin();
    }
    MomentumSpace::fin();
    LocalLattice::fin();
    std::cout << "Done!\n";
  } catch (const std::exception &e) {
    std::cerr << "phinum_main: " << e.what() << std::endl;
    return 1;
  }
  return 0;
}
