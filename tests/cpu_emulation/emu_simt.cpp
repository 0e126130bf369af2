// Runs the reactor's kernels on the CPU with wave semantics (simt/hip/hip_runtime.h:
// a workgroup's threads are fibers that meet at every cross-lane operation), so the
// cooperative kernels -- LU and solves with 32 or 64 lanes per cell, and the persistent
// schedule -- run here as they are written for the GPU.
#define PELE_WAVES_PER_EU(n)
#include <cstring>
#include <vector>

#include "../../pelelm_amd/csrc/reactor_bdf.hip"

using namespace pelelm;

namespace {
int g_lu_mfma = 0;
struct Host {
  std::vector<double> ws, mat;
  std::vector<int> iws, lists;
  unsigned int counts[8] = {0};
};

// a call's arguments over flat cell-fastest arrays, with a workspace for `slots` lanes
// and work lists for `ncell` cells, every word starting as a NaN (lists as -1)
ReactArgs make_args(Host& h, int m, long ncell, long slots, double* rY, double* T, double* rE, double* FC,
                    double dt, double rtol, double atol, int lu_mode, unsigned long long* counters) {
  const int N = mech_nspec(m);
  ReactArgs a{};
  a.rY.p = rY; a.rY.ns = ncell;
  a.T.p = T; a.T.ns = ncell;
  a.rE.p = rE; a.rE.ns = ncell;
  a.FC.p = FC; a.FC.ns = ncell;
  a.nx = (int)ncell; a.ny = 1; a.ncell = ncell;
  a.dt = dt; a.rtol = rtol;
  for (int i = 0; i <= N; ++i) a.atol[i] = atol;
  a.qmax = 4; a.mxstep = 20000;
  const long nw = (slots + 63) / 64, nl = (ncell + 63) / 64 * 64;
  h.ws.assign((size_t)nw * ws_slots(m) * 64, 0.0);
  h.mat.assign((size_t)nw * 64 * mat_stride(m), 0.0);
  h.iws.assign((size_t)nw * iws_slots(m) * 64, -1);
  h.lists.assign(6 * (size_t)(nl > nw * 64 ? nl : nw * 64), -1);
  std::memset(h.ws.data(), 0xFF, h.ws.size() * sizeof(double));
  std::memset(h.mat.data(), 0xFF, h.mat.size() * sizeof(double));
  const size_t ls = h.lists.size() / 6;
  a.ws = h.ws.data(); a.iws = h.iws.data(); a.mat = h.mat.data();
  for (int i = 0; i < 2; ++i) {
    a.list[i] = h.lists.data() + (size_t)i * ls;
    a.slist[i] = h.lists.data() + (size_t)(2 + i) * ls;
  }
  a.plist = h.lists.data() + 4 * ls;
  a.nlist = h.lists.data() + 5 * ls;
  a.counts = h.counts;
  a.lu_lane = 0;
  a.lu_mfma = g_lu_mfma;
  a.gmres = 0;
  a.gmres_maxl = 5;
  a.lu_mode = lu_mode;
  a.ir_sweeps = 3;
  a.ir_tol = 1e-4;
  a.counters = counters;
  return a;
}
}  // namespace

// 1: the following calls factor on the (emulated) matrix cores (mi.lu_form = mfma)
extern "C" void simt_set_lu_mfma(int on) { g_lu_mfma = on; }
// emulated matrix-core products so far (per lane)
extern "C" long simt_mfma_count() { return simt_mfma_calls; }

// the host-driven trips with the cooperative LU and solves (compact: 0 plain, 1 queue)
extern "C" int simt_react_trips(const char* mech, long ncell, double* rY, double* T, double* rE, double* FC,
                                double dt, double rtol, double atol, int compact, int lu_mode,
                                unsigned long long* counters, long* trips) {
  const int m = mech_lookup(mech);
  if (m < 0) return -1;
  Host h;
  ReactArgs a = make_args(h, m, ncell, ncell, rY, T, rE, FC, dt, rtol, atol, lu_mode, counters);
  a.compact = compact;
  return (int)run_react(m, a, nullptr, 3 * 20000 + 100, trips, nullptr, nullptr);
}

// the persistent schedule: `waves` waves of 64 lane slots, claims whenever `chunk` lanes
// of a wave are idle; iters_out: the most iterations any wave ran
extern "C" int simt_react_persist(const char* mech, long ncell, double* rY, double* T, double* rE, double* FC,
                                  double dt, double rtol, double atol, int waves, int chunk, int lu_mode,
                                  long max_trips, unsigned long long* counters, long* iters_out) {
  const int m = mech_lookup(mech);
  if (m < 0) return -1;
  Host h;
  ReactArgs a = make_args(h, m, ncell, (long)waves * 64, rY, T, rE, FC, dt, rtol, atol, lu_mode, counters);
  PersistPlan plan;
  plan.waves = waves;
  plan.chunk = chunk;
  plan.occupancy = 1;
  const hipError_t e = run_react_persist(m, a, plan, max_trips, nullptr);
  if (iters_out) *iters_out = h.counts[2];
  return (int)e;
}
