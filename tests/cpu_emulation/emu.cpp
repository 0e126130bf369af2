// Runs the reactor kernels' device code serially on the CPU (see hip/hip_runtime.h
// in this directory) so tests can check the kernel logic against the oracle without a GPU.
#define PELE_WAVES_PER_EU(n)
#include <limits>
#include <vector>

#include "../../pelelm_amd/csrc/reactor_bdf.hip"

thread_local Idx3 threadIdx, blockIdx, blockDim, gridDim;

using namespace pelelm;

static int g_lu_mode = LU_FP64, g_ir_sweeps = 3, g_pred = 0, g_fill = -1, g_gmres = 0, g_maxl = 5;
static const int* g_order = nullptr;
static const double *g_fy = nullptr, *g_fh = nullptr;
static const int* g_mask = nullptr;
// forcing (rYsrc [nspec][ncell], rEsrc [ncell]) and mask of the following emu_react calls
extern "C" void emu_set_forcing(const double* fy, const double* fh, const int* mask) {
  g_fy = fy;
  g_fh = fh;
  g_mask = mask;
}
// slot -> cell map of the following emu_react calls (the cost ordering; null: identity)
extern "C" void emu_set_order(const int* slot_cell) { g_order = slot_cell; }
// byte pattern the workspace, lists and matrix region are filled with before a run
// (-1: left as allocated), so a read of a word the integrator never wrote shows up
extern "C" void emu_set_fill(int byte) { g_fill = byte; }
// 1: attach a predictor whose network is all zeros (its iterate is the state at tn,
// the "previous value" guess) -- the predictor's plumbing without a trained network
extern "C" void emu_set_pred(int on) { g_pred = on; }
// factor precision of the following emu_react calls (LU_FP64 = 0, LU_FP32_IR = 1)
extern "C" void emu_set_lu(int mode, int ir_sweeps) {
  g_lu_mode = mode;
  g_ir_sweeps = ir_sweeps;
}

// 1: the following emu_react calls solve the Newton systems by GMRES (cvode.solve_type =
// GMRES) with maxl Krylov vectors
extern "C" void emu_set_gmres(int on, int maxl) {
  g_gmres = on;
  g_maxl = maxl;
}

extern "C" int emu_react(const char* mech, long ncell, double* rY, double* T, double* rE, double* FC, double dt,
                         double rtol, double atol, int max_order, int compact, unsigned long long* counters,
                         long* trips) {
  const int m = mech_lookup(mech);
  if (m < 0) return -1;
  const int N = mech_nspec(m);
  ReactArgs a{};
  a.rY.p = rY; a.rY.ns = ncell;
  a.T.p = T; a.T.ns = ncell;
  a.rE.p = rE; a.rE.ns = ncell;
  a.FC.p = FC; a.FC.ns = ncell;
  a.nx = (int)ncell; a.ny = 1; a.ncell = ncell;
  a.dt = dt; a.rtol = rtol;
  for (int i = 0; i <= N; ++i) a.atol[i] = atol;
  a.qmax = max_order; a.mxstep = 20000;
  const long nw = (ncell + 63) / 64;
  double* ws = new double[(size_t)nw * ws_slots(m) * 64];
  int* iws = new int[(size_t)nw * iws_slots(m) * 64];
  int* lists = new int[6 * (size_t)nw * 64];
  double* mat = new double[(size_t)nw * 64 * mat_stride(m)];
  if (g_fill >= 0) {
    std::memset(ws, g_fill, (size_t)nw * ws_slots(m) * 64 * sizeof(double));
    std::memset(iws, g_fill, (size_t)nw * iws_slots(m) * 64 * sizeof(int));
    std::memset(lists, g_fill, 6 * (size_t)nw * 64 * sizeof(int));
    std::memset(mat, g_fill, (size_t)nw * 64 * mat_stride(m) * sizeof(double));
  }
  unsigned int counts[8] = {0};
  a.ws = ws; a.iws = iws; a.mat = mat;
  a.plist = lists + (size_t)4 * nw * 64;
  a.nlist = lists + (size_t)5 * nw * 64;
  for (int i = 0; i < 2; ++i) {
    a.list[i] = lists + (size_t)i * nw * 64;
    a.slist[i] = lists + (size_t)(2 + i) * nw * 64;
  }
  a.counts = counts;
  a.compact = compact;
  a.lu_lane = 1;   // lanes run one at a time here, so they cannot cooperate
  a.lu_mode = g_lu_mode;
  a.ir_sweeps = g_ir_sweeps;
  a.ir_tol = 1e-4;
  a.gmres = g_gmres;
  a.gmres_maxl = g_maxl;
  a.counters = counters;
  a.slot_cell = g_order;
  if (g_fy) { a.rYsrc.p = const_cast<double*>(g_fy); a.rYsrc.ns = ncell; }
  if (g_fh) { a.rEsrc.p = const_cast<double*>(g_fh); a.rEsrc.ns = ncell; }
  a.mask.p = g_mask;
  MlpDev md{};
  std::vector<unsigned short> wz(2 * 64 * 128 * 8, 0);
  std::vector<float> bz(4 * 128, 0.f), norm(256, 0.f);
  for (int i = 0; i < 64; ++i) norm[64 + i] = norm[192 + i] = 1.f;
  if (g_pred) {
    const int N1 = N + 1;
    md.n_in = N + 3;
    md.n_out = N1;
    md.kp[0] = 32; md.np[3] = 32;
    for (int l = 0; l < 4; ++l) {
      md.wt[l] = md.wk[l] = wz.data();
      md.b[l] = bz.data();
    }
    md.norm = norm.data();
    a.use_pred = 1;
  }
  const hipError_t e = run_react(m, a, nullptr, 3 * 20000 + 100, trips, nullptr, g_pred ? &md : nullptr);
  delete[] ws;
  delete[] iws;
  delete[] lists;
  delete[] mat;
  return (int)e;
}

extern "C" int emu_rates(const char* mech, long ncell, const double* rY, const double* rH, double* T, double* out) {
  const int m = mech_lookup(mech);
  if (m < 0) return -1;
  RatesArgs ra;
  ra.rhoY.p = const_cast<double*>(rY); ra.rhoY.ns = ncell;
  ra.rhoH.p = const_cast<double*>(rH); ra.rhoH.ns = ncell;
  ra.T.p = T; ra.T.ns = ncell;
  ra.rdot.p = out; ra.rdot.ns = ncell;
  ra.Tout = T;
  ra.nx = (int)ncell;
  ra.ncell = ncell;
  return (int)launch_rates(m, ra, nullptr);
}

// the cost-ordering keys (mi.order) of flat cells with scalar tolerances
extern "C" int emu_cost_keys(const char* mech, long ncell, double* rY, double* T, double* rE, double* FC,
                             double dt, double rtol, double atol, int mode, unsigned int* keys, int* vals) {
  const int m = mech_lookup(mech);
  if (m < 0) return -1;
  ReactArgs a{};
  a.rY.p = rY; a.rY.ns = ncell;
  a.T.p = T; a.T.ns = ncell;
  a.rE.p = rE; a.rE.ns = ncell;
  a.FC.p = FC; a.FC.ns = ncell;
  a.nx = (int)ncell; a.ny = 1; a.ncell = ncell;
  a.dt = dt; a.rtol = rtol;
  for (int i = 0; i <= mech_nspec(m); ++i) a.atol[i] = atol;
  return (int)launch_cost_keys(m, a, mode, keys, vals, nullptr);
}

// ReactorRK64's kernel over flat cells (with the forcing and mask of emu_set_forcing)
extern "C" int emu_rk64(const char* mech, long ncell, double* rY, double* T, double* rE, double* FC, double dt,
                        double rtol, double atol, int guess, int nmin, int nmax, unsigned long long* counters) {
  const int m = mech_lookup(mech);
  if (m < 0) return -1;
  const int N = mech_nspec(m);
  ReactArgs a{};
  a.rY.p = rY; a.rY.ns = ncell;
  a.T.p = T; a.T.ns = ncell;
  a.rE.p = rE; a.rE.ns = ncell;
  a.FC.p = FC; a.FC.ns = ncell;
  if (g_fy) { a.rYsrc.p = const_cast<double*>(g_fy); a.rYsrc.ns = ncell; }
  if (g_fh) { a.rEsrc.p = const_cast<double*>(g_fh); a.rEsrc.ns = ncell; }
  a.mask.p = g_mask;
  a.nx = (int)ncell; a.ny = 1; a.ncell = ncell;
  a.dt = dt; a.rtol = rtol;
  for (int i = 0; i <= N; ++i) a.atol[i] = atol;
  const long nw = (ncell + 63) / 64;
  std::vector<double> ws((size_t)nw * ws_slots(m) * 64, std::numeric_limits<double>::quiet_NaN());
  a.ws = ws.data();
  a.counters = counters;
  return (int)launch_rk64(m, a, guess, nmin, nmax, nullptr);
}
