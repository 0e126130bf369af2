// A stand-alone executable around the serial emulation (emu.cpp), for builds with
// AddressSanitizer / UndefinedBehaviorSanitizer and PELE_CHECK_BOUNDS: a sanitised
// runtime has to come first in a process, which a library loaded into Python cannot
// arrange.
//
// usage: emu_check MECH NCELL DT RTOL ATOL LU_MODE IR_SWEEPS GMRES MAXL IN OUT
//   IN : rY [nspec][ncell], T [ncell], rE [ncell] (doubles)
//   OUT: rY, T, rE, FC [ncell] (doubles), then 16 counters (uint64) and the trip count
// The workspace starts filled with NaNs (byte 0xFF), as in the library-form tests.
#include <cstdio>
#include <vector>

#include "emu.cpp"

int main(int argc, char** argv) {
  if (argc != 12) {
    std::fprintf(stderr, "usage: emu_check MECH NCELL DT RTOL ATOL LU_MODE IR_SWEEPS GMRES MAXL IN OUT\n");
    return 2;
  }
  const char* mech = argv[1];
  const long n = std::atol(argv[2]);
  const double dt = std::atof(argv[3]), rtol = std::atof(argv[4]), atol = std::atof(argv[5]);
  const int m = mech_lookup(mech);
  if (m < 0 || n <= 0) return 2;
  const int ns = mech_nspec(m);
  std::vector<double> rY((size_t)ns * n), T(n), rE(n), FC(n, 0.0);
  FILE* f = std::fopen(argv[10], "rb");
  if (!f) return 2;
  const bool ok = std::fread(rY.data(), sizeof(double), rY.size(), f) == rY.size() &&
                  std::fread(T.data(), sizeof(double), n, f) == (size_t)n &&
                  std::fread(rE.data(), sizeof(double), n, f) == (size_t)n;
  std::fclose(f);
  if (!ok) return 2;
  emu_set_fill(0xFF);
  emu_set_lu(std::atoi(argv[6]), std::atoi(argv[7]));
  emu_set_gmres(std::atoi(argv[8]), std::atoi(argv[9]));
  unsigned long long cnt[16] = {0};
  long trips = 0;
  const int rc = emu_react(mech, n, rY.data(), T.data(), rE.data(), FC.data(), dt, rtol, atol, 5, 1, cnt, &trips);
  if (rc != 0) {
    std::fprintf(stderr, "emu_react returned %d\n", rc);
    return 1;
  }
  f = std::fopen(argv[11], "wb");
  if (!f) return 2;
  std::fwrite(rY.data(), sizeof(double), rY.size(), f);
  std::fwrite(T.data(), sizeof(double), n, f);
  std::fwrite(rE.data(), sizeof(double), n, f);
  std::fwrite(FC.data(), sizeof(double), n, f);
  std::fwrite(cnt, sizeof(cnt[0]), 16, f);
  std::fwrite(&trips, sizeof(trips), 1, f);
  std::fclose(f);
  return 0;
}
