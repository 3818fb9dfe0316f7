// ref_harness.cpp -- drives the reference's own hot-path sources, compiled
// unmodified from where they lie (oracle/Makefile adds them with -I), to
// produce bit-exact reference outputs for pinning the oracle.  Test
// infrastructure only; it is built only where the reference tree exists and
// its binary lands in oracle/_ref/.
//
// Usage: ref_harness <in.bin> <out.bin> -Nt N -at a -M m2 -Mmode s -L Lambda
//                    -Lmode s -Kappa k -T0 t0 -dt base
// in.bin : int32 nsteps, int32 islands, double thr (<0: 2.57583*sqrt(<phi^2>)),
//          double phi[N^3], double pi[N^3]        (x fastest, interior only)
// out.bin: double phi[N^3], pi[N^3], avg[10], ext[3], int32 extpos[3],
//          double thr, int32 nisl, double cat[nisl*10], int32 labels[N^3]
// islands 0/1: nsteps steps, then (1) one Full_Analysis.  islands R >= 2
// (tracking): Full_Analysis at the start and after each of R-1 blocks of
// nsteps steps, the reference writing its island files into ./output/0/;
// out.bin then has nisl = 0, labels -1 and, appended, int32 ntrack and per
// tracked island int64 index, int32 status, disp, nhist and nhist x 8 doubles
// (at_time, peak pos, Peak_value, avPhi, V, PotEn, KinEn, GradEn).
// With REF_SAVECONFIG=n in the environment the final state is also written by
// the reference's SaveConfig(n) (./field-n-0.out, ./moment-n.out).
// ref_harness --info - <parameters>: the reference's Print_Info_Console (to
// stdout) and Print_Info_File (./output/0/info.txt) for those parameters.
#define MPICH_SKIP_MPICXX 1
#include "islands.cpp"
#include "renormalization.cpp"
#include <sys/stat.h>

#include <cstdio>
#include <string>
#include <vector>

static void die(const char *m) { std::fprintf(stderr, "ref_harness: %s\n", m); std::exit(2); }

int main(int argc, char **argv) {
  if (argc < 3) die("usage");
  MPI_Init(&argc, &argv);
  MPI::init();
  // the parameter parser skips argv[0]; hand it argv[2..] so that the output
  // path sits where the program name would be
  GetParameters(argc - 2, argv + 2);
  CompleteParameters();
  if (std::string(argv[1]) == "--info") {
    mkdir("output", 0755);
    mkdir("output/0", 0755);  // LASTRUN is 0
    Print_Info_Console();
    Print_Info_File();
    MPI_Finalize();
    return 0;
  }
  LocalLattice::init();
  fix_tadpole();

  FILE *fi = std::fopen(argv[1], "rb");
  if (!fi) die("open in");
  int nsteps = 0, islands = 0;
  double thr = -1;
  if (std::fread(&nsteps, 4, 1, fi) != 1 || std::fread(&islands, 4, 1, fi) != 1 ||
      std::fread(&thr, 8, 1, fi) != 1) die("read header");
  const long N = N_t, n3 = N * N * N;
  std::vector<double> buf(n3);
  if ((long)std::fread(buf.data(), 8, n3, fi) != n3) die("read phi");
  for (long z = 0; z < N; z++)
    for (long y = 0; y < N; y++)
      for (long x = 0; x < N; x++) LocalLattice::Phi[pos(x, y, z + 1)] = buf[x + N * y + N * N * z];
  if ((long)std::fread(buf.data(), 8, n3, fi) != n3) die("read pi");
  for (long z = 0; z < N; z++)
    for (long y = 0; y < N; y++)
      for (long x = 0; x < N; x++) LocalLattice::Pi[pos(x, y, z + 1)] = buf[x + N * y + N * N * z];
  std::fclose(fi);
  Communication::exchange();

  reset();
  const bool tracking = islands >= 2;
  if (tracking) {
    mkdir("output", 0755);
    mkdir("output/0", 0755);  // LASTRUN is 0
    island_analysis::init();
    CalculateVolumeAverages();
    island_analysis::thres = thr >= 0 ? thr : 2.57583 * sqrt(LocalLattice::avgphi2);
    thr = island_analysis::thres;
    for (int r = 0; r < islands; r++) {
      if (r > 0)
        for (int s = 0; s < nsteps; s++) evolve();
      Full_Analysis();
    }
  } else {
    for (int s = 0; s < nsteps; s++) evolve();
  }

  CalculateVolumeAverages();
  CalculateExtrema();

  FILE *fo = std::fopen(argv[2], "wb");
  if (!fo) die("open out");
  for (int f = 0; f < 2; f++) {
    double *src = f == 0 ? LocalLattice::Phi : LocalLattice::Pi;
    for (long z = 0; z < N; z++)
      for (long y = 0; y < N; y++)
        for (long x = 0; x < N; x++) buf[x + N * y + N * N * z] = src[pos(x, y, z + 1)];
    std::fwrite(buf.data(), 8, n3, fo);
  }
  double avg[10] = {LocalLattice::avgphi[0], LocalLattice::avgpi[0], LocalLattice::avgphi2,
                    LocalLattice::avgpi2, LocalLattice::avgphipi, LocalLattice::avgphi4,
                    LocalLattice::avgphi6, LocalLattice::avgEnergy, LocalLattice::avgInteraction,
                    LocalLattice::avgCharge};
  std::fwrite(avg, 8, 10, fo);
  double ext[3] = {LocalLattice::GlminPhi, LocalLattice::GlmaxPhi, LocalLattice::GlmaxAbsPhi};
  int extpos[3] = {LocalLattice::minpos, LocalLattice::maxpos, LocalLattice::maxAbspos};
  std::fwrite(ext, 8, 3, fo);
  std::fwrite(extpos, 4, 3, fo);

  int nisl = 0;
  std::vector<double> cat;
  std::vector<int> lab(n3, -1);
  if (islands == 1) {
    island_analysis::init();
    island_analysis::thres = thr >= 0 ? thr : 2.57583 * sqrt(LocalLattice::avgphi2);
    thr = island_analysis::thres;
    Full_Analysis();
    nisl = (int)island_analysis::found_islands.size();
    for (int i = 0; i < nisl; i++) {
      island &B = island_analysis::found_islands[i];
      double c[10] = {(double)B.index, (double)llround(B.V.back() / (a_t * a_t * a_eta)), B.V.back(),
                      (double)B.center.back().pos, B.Peak_value.back(), B.avPhi.back(),
                      B.PotEn.back(), B.KinEn.back(), B.GradEn.back(), 0.0};
      cat.insert(cat.end(), c, c + 10);
    }
    for (long z = 0; z < N; z++)
      for (long y = 0; y < N; y++)
        for (long x = 0; x < N; x++) lab[x + N * y + N * N * z] = island_analysis::b[pos(x, y, z + 1)];
  }
  std::fwrite(&thr, 8, 1, fo);
  std::fwrite(&nisl, 4, 1, fo);
  if (nisl) std::fwrite(cat.data(), 8, cat.size(), fo);
  std::fwrite(lab.data(), 4, n3, fo);
  if (const char *sc = std::getenv("REF_SAVECONFIG")) SaveConfig(std::atoi(sc));
  if (tracking) {
    std::vector<island> &F = island_analysis::found_islands;
    const int ntrack = (int)F.size();
    std::fwrite(&ntrack, 4, 1, fo);
    for (int i = 0; i < ntrack; i++) {
      const long long index = F[i].index;
      const int hdr[3] = {F[i].status, (int)F[i].disp, (int)F[i].V.size()};
      std::fwrite(&index, 8, 1, fo);
      std::fwrite(hdr, 4, 3, fo);
      for (size_t j = 0; j < F[i].V.size(); j++) {
        const double e[8] = {F[i].at_time[j], (double)F[i].center[j].pos, F[i].Peak_value[j],
                             F[i].avPhi[j], F[i].V[j], F[i].PotEn[j], F[i].KinEn[j], F[i].GradEn[j]};
        std::fwrite(e, 8, 8, fo);
      }
    }
  }
  std::fclose(fo);
  MPI_Finalize();
  return 0;
}
