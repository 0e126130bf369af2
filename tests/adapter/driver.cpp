// A PeleLM-shaped driver written against include/pelelm_reactor.hpp: the calls and
// argument order of Source/PeleLM.cpp:701-709 (create/init), 1497-1516 (typical
// values), 5752-5760 (react on one box, dt/time reset after) and 930 (close).
// Mode "host": lifecycle and option calls only (no device memory touched).
// Mode "gpu <ncell-per-side>": reacts one box of seeded LiDryer cells read from the
// binary file given as argv[3] (rY planes, T, rE as written by the test), writes the
// results back into argv[4].
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <string>
#include <vector>

#include "pelelm_reactor.hpp"

int main(int argc, char** argv) {
  const std::string mode = argc > 1 ? argv[1] : "host";
  std::unique_ptr<mi::ReactorBase> reactor = mi::ReactorBase::create("ReactorCvode");
  reactor->set("mi.mechanism", "LiDryer");
  reactor->set("ode.rtol", "1.0e-6");
  reactor->set("ode.atol", "1.0e-10");
  reactor->set("cvode.solve_type", "denseAJ_direct");
  reactor->set("cvode.max_order", "4");
  reactor->init(2, 1);
  const int nspec = reactor->nspec();
  if (mode == "host") {
    std::vector<double> typ(nspec + 1, 1.0);
    reactor->set_typ_vals_ode(typ);
    reactor->close();
    std::printf("host ok nspec %d\n", nspec);
    return 0;
  }
  const int n = std::atoi(argv[2]);
  const long ncell = (long)n * n * n;
  std::vector<double> h((nspec + 2) * ncell);
  {
    std::ifstream f(argv[3], std::ios::binary);
    f.read(reinterpret_cast<char*>(h.data()), h.size() * sizeof(double));
    if (!f) return 3;
  }
  // STemp-like box: comps 0..N-1 = rhoY, N = rhoH, N+1 = T; FC and mask separate
  double *s = nullptr, *fc = nullptr;
  int* mask = nullptr;
  if (hipMalloc(&s, (nspec + 2) * ncell * sizeof(double)) != hipSuccess) return 4;
  if (hipMalloc(&fc, ncell * sizeof(double)) != hipSuccess) return 4;
  if (hipMalloc(&mask, ncell * sizeof(int)) != hipSuccess) return 4;
  // input file order: rY planes, T, rE -> STemp order rY, rE, T
  std::vector<double> st((nspec + 2) * ncell);
  std::memcpy(st.data(), h.data(), nspec * ncell * sizeof(double));
  std::memcpy(st.data() + nspec * ncell, h.data() + (nspec + 1) * ncell, ncell * sizeof(double));
  std::memcpy(st.data() + (nspec + 1) * ncell, h.data() + nspec * ncell, ncell * sizeof(double));
  if (hipMemcpy(s, st.data(), st.size() * sizeof(double), hipMemcpyHostToDevice) != hipSuccess) return 4;
  std::vector<int> hm(ncell, 0);
  if (hipMemcpy(mask, hm.data(), ncell * sizeof(int), hipMemcpyHostToDevice) != hipSuccess) return 4;
  mi::Box bx{{{0, 0, 0}}, {{n - 1, n - 1, n - 1}}};
  auto arr = [&](double* p, int nc) {
    mi::Array4<double> a;
    a.p = p;
    a.begin = {{0, 0, 0}};
    a.end = {{n, n, n}};
    a.jstride = n;
    a.kstride = (long)n * n;
    a.nstride = ncell;
    a.ncomp = nc;
    return a;
  };
  mi::Array4<double> rY = arr(s, nspec), rE = arr(s + nspec * ncell, 1), T = arr(s + (nspec + 1) * ncell, 1);
  mi::Array4<double> FC = arr(fc, 1), none;
  mi::Array4<int> m;
  m.p = mask;
  m.begin = {{0, 0, 0}};
  m.end = {{n, n, n}};
  m.jstride = n;
  m.kstride = (long)n * n;
  m.nstride = ncell;
  m.ncomp = 1;
  double dt = 1e-6, time = 0.0;
  const double dt_incr = dt, time_chem = time;
  const int rc = reactor->react(bx, rY, none, T, rE, none, FC, m, dt, time, nullptr);
  (void)hipDeviceSynchronize();
  dt = dt_incr;       // the caller resets both after the call
  time = time_chem;
  std::vector<double> fch(ncell);
  if (hipMemcpy(st.data(), s, st.size() * sizeof(double), hipMemcpyDeviceToHost) != hipSuccess) return 4;
  if (hipMemcpy(fch.data(), fc, ncell * sizeof(double), hipMemcpyDeviceToHost) != hipSuccess) return 4;
  std::memcpy(h.data(), st.data(), nspec * ncell * sizeof(double));
  std::memcpy(h.data() + nspec * ncell, st.data() + (nspec + 1) * ncell, ncell * sizeof(double));
  std::memcpy(h.data() + (nspec + 1) * ncell, st.data() + nspec * ncell, ncell * sizeof(double));
  std::ofstream o(argv[4], std::ios::binary);
  o.write(reinterpret_cast<const char*>(h.data()), h.size() * sizeof(double));
  o.write(reinterpret_cast<const char*>(fch.data()), ncell * sizeof(double));
  (void)hipFree(s);
  (void)hipFree(fc);
  (void)hipFree(mask);
  reactor->close();
  std::printf("gpu rc %d\n", rc);
  return rc == 0 ? 0 : 5;
}
