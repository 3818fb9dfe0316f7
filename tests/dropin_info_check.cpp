// The drop-in layer's parameter handling on the host alone (no device call):
// GetParameters and CompleteParameters (parameters.cpp:308-411) from the
// command line, then Print_Info_Console to stdout and Print_Info_File to
// ./output/0/info.txt (fieldanalysis.cpp:81-247).  Built and run by
// tests/test_dropin_host.py.
#include <sys/stat.h>

#include <iostream>

#include "phinum_amd/csrc/dropin/phinum_dropin.hpp"

int main(int argc, char **argv) {
  using namespace phinum;
  try {
    GetParameters(argc, argv);
    CompleteParameters();
  } catch (const std::exception &e) {
    std::cerr << e.what() << std::endl;
    return 2;
  }
  LASTRUN = 0;
  mkdir("output", 0755);
  mkdir("output/0", 0755);
  Print_Info_Console();
  Print_Info_File();
  return 0;
}
