// ref_timing.cpp -- CPU baseline only: times the reference's own evolve()
// (dynamics.cpp:196-237, with the halo exchange of comm.cpp:63-78 between
// MPI z-slabs, lattice.cpp:70-78), compiled unmodified from the reference
// sources with the reference's flags (run_PC.pbs:1) into oracle/_ref/.  No
// output but the time; bench.py's cpu_baseline leg runs it under mpirun.
//
// Usage: mpirun -np P ref_timing <nsteps> -Nt N -at a -M m2 -Mmode s -L Lambda
//                                 -Lmode s -Kappa k -T0 t0 -dt base
// Each rank fills its slab with phi0 + a small deterministic pattern (the time
// of a step does not depend on the values) and runs one warm-up step, then
// nsteps timed steps between two barriers; rank 0 prints the seconds.
#define MPICH_SKIP_MPICXX 1
#include "islands.cpp"
#include "renormalization.cpp"

#include <cstdio>
#include <cstdlib>

int main(int argc, char **argv) {
  if (argc < 2) return 2;
  MPI_Init(&argc, &argv);
  MPI::init();
  const int nsteps = std::atoi(argv[1]);
  // the parameter parser skips argv[0]: nsteps sits where the program name would
  GetParameters(argc - 1, argv + 1);
  CompleteParameters();
  LocalLattice::init();
  fix_tadpole();
  const long N = N_t;
  unsigned long long h = 0x9E3779B97F4A7C15ULL * (unsigned long long)(ID + 1);
  for (long z = 1; z <= LocalLattice::N_eta_loc; z++)
    for (long y = 0; y < N; y++)
      for (long x = 0; x < N; x++) {
        h = h * 6364136223846793005ULL + 1442695040888963407ULL;
        const double r = (double)(h >> 11) * 0x1p-53 - 0.5;
        LocalLattice::Phi[pos(x, y, z)] = phi0_0 + 0.1 * r;
        LocalLattice::Pi[pos(x, y, z)] = 0.05 * r;
      }
  Communication::exchange();
  reset();
  evolve();
  MPI_Barrier(MPI_COMM_WORLD);
  const double t0 = MPI_Wtime();
  for (int s = 0; s < nsteps; s++) evolve();
  MPI_Barrier(MPI_COMM_WORLD);
  const double t1 = MPI_Wtime();
  if (ID == 0) std::printf("%.6f\n", t1 - t0);
  MPI_Finalize();
  return 0;
}
