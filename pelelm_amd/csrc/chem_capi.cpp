// C ABI of the MI355X reactor (include/pelelm_chem.h): reactor objects, the runtime
// options the reference decks set through ParmParse, workspace management and the
// launch of the integration kernels on the caller's stream.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <utility>
#include <chrono>
#include <map>
#include <mutex>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/pelelm_chem.h"
#include "../../include/pelelm_mlp.h"
#include "mlp_internal.h"
#include "reactor_internal.h"

using namespace pelelm;

namespace {

// one configuration of the integration among those the autotuner (mi.schedule = auto)
// chooses from
struct TuneConfig {
  int schedule = SCHED_PLAIN;   // SCHED_PLAIN, SCHED_COMPACT or SCHED_PERSIST
  int graph_trips = 0;          // trip schedules: 0 (host-driven) or K trips per graph
  int order = ORDER_RATE;       // trip schedules: cost ordering
  int chunk = 8, occ = 1;       // persistent schedule: claim chunk, waves per SIMD
};

struct Reactor {
  std::string integrator;
  bool null_reactor = false;
  // ReactorRK64: explicit sub-steps within [dt / rk_max, dt / rk_min] from dt / rk_guess
  bool rk64 = false;
  int rk_guess = 10, rk_min = 5, rk_max = 500;
  bool inited = false;
  int reactor_type = 2;
  int mech = MECH_DRM19;
  // ode.* / cvode.* keys
  double rtol = 1e-6;
  double atol = 1e-10;
  int max_order = 4;
  long max_steps = 20000;
  bool use_typ = false;
  std::vector<double> typ;
  // this reactor's keys
  int schedule = SCHED_PLAIN;     // measured fastest on every workload here
  int lu_mode = LU_FP64;
  int ir_sweeps = 3;
  double ir_tol = 1e-4;         // refinement acceptance (WRMS of the correction, error-weight units)
  long max_trips = 0;           // 0: 3 x cvode.max_num_steps + 100 trips
  int lu_lane = 0;
  int lu_mfma = 0;
  // cvode.solve_type = GMRES: the Newton systems by matrix-free scaled GMRES, Krylov
  // dimension mi.gmres_maxl (CVODE's default 5)
  int gmres = 0, gmres_maxl = 5;
  void* predictor = nullptr;    // attached network (mi_chem_attach_predictor), not owned
  int graph_trips = 0;          // 0: host-driven trips; K (even): graphs of K trips
  // mi.schedule = persistent: waves (0: as many as the device holds resident), refill
  // chunk, waves per SIMD the kernel is compiled for
  int persist_waves = 0, persist_chunk = 8, persist_occ = 1;
  int persist_cap[NMECH][2] = {};   // resident waves of the persistent kernel (0: not queried)
  hipStream_t capture_stream = nullptr;
  // mi.order: cost ordering of the cells over the workspace slots (ORDER_*), its keys,
  // values (2 x order_cells each, double-buffered) and the sort's temporary storage
  int order_mode = ORDER_RATE;   // measured faster on every workload here, results unchanged
  unsigned int* order_keys = nullptr;
  int* order_vals = nullptr;
  void* order_temp = nullptr;
  size_t order_temp_bytes = 0;
  long order_cells = 0;
  // mi.schedule = auto: the configuration chosen per box size (how it was found and what
  // it cost: final-round ms of the choice, total tuning ms, candidates, 1 when read from
  // the cache file), and the cache file (mi.tune_cache; empty: none)
  struct Tuned {
    TuneConfig cfg;
    double ms[4];
  };
  std::string tune_cache;
  std::map<long, Tuned> tuned;
  int last_schedule = SCHED_COMPACT, last_graph = 0;
  // a cell that exhausts every fallback: abort (default) makes react return MI_ECELLFAIL,
  // as the reference's integrator aborts inside react; continue returns MI_OK, the
  // failed cells carrying FC < 0 and counted in the statistics
  bool on_fail_continue = false;
  bool profile = false;
  KernelProf prof{};
  // device state
  int device = -1;
  int ncu = 0;
  double* ws = nullptr;
  int* iws = nullptr;
  double* mat = nullptr;        // per-cell matrix region (factors, Newton vectors)
  int* lists = nullptr;         // 6 lists of ws_cells ints (active x2, set-up x2, predictor, solve)
  long ws_cells = 0;            // cells the workspace and lists hold
  double* glue = nullptr;       // advance_box: the forcing in CGS, (nspec + 1) x glue_cells
  long glue_cells = 0;
  unsigned int* counts = nullptr;
  unsigned long long* counters = nullptr;
  // where the last call's counters are (the reactor's, or the no-allocation array's)
  unsigned int* stat_counts = nullptr;
  unsigned long long* stat_counters = nullptr;
  long last_trips = 0;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  hipStream_t last_stream = nullptr;
  bool have_stats = false;
  long last_ncell = 0;
  std::string err;
};

int fail(Reactor* r, int code, const std::string& msg) {
  if (r) r->err = msg;
  return code;
}

#define HIPCHK(r, call)                                                                 \
  do {                                                                                  \
    hipError_t e_ = (call);                                                             \
    if (e_ != hipSuccess) return fail((r), MI_EHIP, std::string(#call ": ") + hipGetErrorString(e_)); \
  } while (0)

// the integration workspace only: advance_box's forcing buffer is filled before the
// react call it wraps, whose first call of a larger box regrows the workspace
void free_workspace(Reactor* r) {
  if (r->ws) (void)hipFree(r->ws);
  if (r->iws) (void)hipFree(r->iws);
  if (r->lists) (void)hipFree(r->lists);
  if (r->mat) (void)hipFree(r->mat);
  r->mat = nullptr;
  r->ws = nullptr;
  r->iws = nullptr;
  r->lists = nullptr;
  r->ws_cells = 0;
}

int ensure_device_state(Reactor* r, long ncell) {
  if (r->device < 0) {
    HIPCHK(r, hipGetDevice(&r->device));
    HIPCHK(r, hipMalloc(&r->counts, 8 * sizeof(unsigned int)));
    HIPCHK(r, hipMalloc(&r->counters, NCOUNTERS * sizeof(unsigned long long)));
    if (!r->ev0) HIPCHK(r, hipEventCreate(&r->ev0));
    if (!r->ev1) HIPCHK(r, hipEventCreate(&r->ev1));
  }
  if (ncell > r->ws_cells) {
    free_workspace(r);
    const long nw = (ncell + 63) / 64;
    const size_t nd = (size_t)nw * ws_slots(r->mech) * 64;
    const size_t ni = (size_t)nw * iws_slots(r->mech) * 64;
    if (hipMalloc(&r->ws, nd * sizeof(double)) != hipSuccess)
      return fail(r, MI_ENOMEM, "workspace allocation failed (" + std::to_string(nd * 8) + " bytes)");
    if (hipMalloc(&r->iws, ni * sizeof(int)) != hipSuccess) return fail(r, MI_ENOMEM, "workspace allocation failed");
    if (hipMalloc(&r->lists, 6 * (size_t)nw * 64 * sizeof(int)) != hipSuccess)
      return fail(r, MI_ENOMEM, "list allocation failed");
    const size_t nm = (size_t)nw * 64 * mat_stride(r->mech);
    if (hipMalloc(&r->mat, nm * sizeof(double)) != hipSuccess)
      return fail(r, MI_ENOMEM, "matrix region allocation failed (" + std::to_string(nm * 8) + " bytes)");
    r->ws_cells = nw * 64;
  }
  return MI_OK;
}

void free_glue(Reactor* r) {
  if (r->glue) (void)hipFree(r->glue);
  r->glue = nullptr;
  r->glue_cells = 0;
}

void free_order(Reactor* r) {
  if (r->order_keys) (void)hipFree(r->order_keys);
  if (r->order_vals) (void)hipFree(r->order_vals);
  if (r->order_temp) (void)hipFree(r->order_temp);
  r->order_keys = nullptr;
  r->order_vals = nullptr;
  r->order_temp = nullptr;
  r->order_temp_bytes = 0;
  r->order_cells = 0;
}

int ensure_order_buffers(Reactor* r, long ncell) {
  if (ncell <= r->order_cells) return MI_OK;
  free_order(r);
  const size_t tb = cost_sort_temp_bytes(ncell);
  if (tb == 0) return fail(r, MI_EHIP, "cannot size the cost-ordering sort");
  if (hipMalloc(&r->order_keys, 2 * (size_t)ncell * sizeof(unsigned int)) != hipSuccess ||
      hipMalloc(&r->order_vals, 2 * (size_t)ncell * sizeof(int)) != hipSuccess ||
      hipMalloc(&r->order_temp, tb) != hipSuccess)
    return fail(r, MI_ENOMEM, "cost-ordering buffers: allocation failed");
  r->order_temp_bytes = tb;
  r->order_cells = ncell;
  return MI_OK;
}

bool parse_double(const char* s, double* out) {
  char* end = nullptr;
  const double v = std::strtod(s, &end);
  if (end == s || *end != 0 || !std::isfinite(v)) return false;
  *out = v;
  return true;
}
bool parse_long(const char* s, long* out) {
  char* end = nullptr;
  const long v = std::strtol(s, &end, 10);
  if (end == s || *end != 0) return false;
  *out = v;
  return true;
}

// the return code of a finished call under mi.on_fail: one read of the failure counter
// (the only synchronisation of a persistent call, and none under continue)
int failure_code(Reactor* r, hipStream_t stream) {
  if (r->on_fail_continue || r->null_reactor) return MI_OK;
  unsigned long long nfail = 0;
  HIPCHK(r, hipMemcpyAsync(&nfail, r->stat_counters + CNT_NFAIL, sizeof(nfail), hipMemcpyDeviceToHost, stream));
  HIPCHK(r, hipStreamSynchronize(stream));
  if (nfail > 0)
    return fail(r, MI_ECELLFAIL, std::to_string(nfail) + " cell(s) failed to integrate (FC < 0 marks them)");
  return MI_OK;
}

// Device memory placed by the caller of react_persist instead of the reactor's workspace
// (mi_chem_react_flat_noalloc): lane slots for at most max_waves waves and the counters.
struct PersistMem {
  double* ws;
  int* iws;
  double* mat;
  unsigned int* counts;
  unsigned long long* counters;
  long max_waves;
};

// mi.schedule = persistent: the workspace holds the kernel's lane slots (whatever the box
// size); the kernels are enqueued on `stream` with no host synchronisation.  With `fixed`
// the slots and counters are that memory and nothing is allocated.
int react_persist(Reactor* r, ReactArgs& a, double* dt, double* time, hipStream_t stream,
                  const PersistMem* fixed = nullptr, const TuneConfig* cfg = nullptr) {
  PersistPlan plan;
  plan.occupancy = cfg ? cfg->occ : r->persist_occ;
  plan.chunk = cfg ? cfg->chunk : r->persist_chunk;
  int& cap = r->persist_cap[r->mech][plan.occupancy - 1];
  if (cap <= 0) cap = persist_capacity(r->mech, plan.occupancy);
  if (cap <= 0) return fail(r, MI_EHIP, "cannot query the persistent kernel's occupancy");
  // under stream capture (a caller building a graph) nothing may synchronise: no timing
  // events and no failure read-back (failed cells still carry FC < 0)
  hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
  HIPCHK(r, hipStreamIsCapturing(stream, &cs));
  const bool capturing = cs != hipStreamCaptureStatusNone;
  long waves = r->persist_waves > 0 ? r->persist_waves : cap;
  waves = std::min(waves, (a.ncell + 63) / 64);
  if (fixed) waves = std::min(waves, fixed->max_waves);
  plan.waves = (int)std::max(1L, waves);
  if (fixed) {
    a.ws = fixed->ws;
    a.iws = fixed->iws;
    a.mat = fixed->mat;
    a.counts = fixed->counts;
    a.counters = fixed->counters;
  } else {
    int rc = ensure_device_state(r, (long)plan.waves * 64);
    if (rc) return rc;
    a.ws = r->ws;
    a.iws = r->iws;
    a.mat = r->mat;
    a.counts = r->counts;
    a.counters = r->counters;
  }
  r->stat_counts = a.counts;
  r->stat_counters = a.counters;
  a.compact = 0;
  a.use_pred = 0;
  const long max_trips = r->max_trips > 0 ? r->max_trips : 3 * r->max_steps + 100;
  r->last_schedule = SCHED_PERSIST;
  r->last_graph = 0;
  const bool timed = !capturing && r->ev0 && r->ev1;
  HIPCHK(r, launch_counters_reset(a.counters, stream));
  if (timed) HIPCHK(r, hipEventRecord(r->ev0, stream));
  HIPCHK(r, run_react_persist(r->mech, a, plan, max_trips, stream));
  if (timed) HIPCHK(r, hipEventRecord(r->ev1, stream));
  r->prof = KernelProf{};
  r->last_stream = stream;
  r->have_stats = true;
  r->last_trips = -1;   // read from the device by mi_chem_get_stats
  *time += *dt;
  return capturing ? MI_OK : failure_code(r, stream);
}

// ---------------------------------------------------------------------------
// mi.schedule = auto: the autotuner.  Candidates: the plain and compacted trip schedules,
// host-driven or in graphs of 8 trips, with cost ordering (and plain without it), and the
// persistent schedule at three claim chunks and two occupancies.  Successive halving on
// samples of the call's own cells (every 16th cell, then every 4th, then all of them;
// 3 and then 2 survivors): about three box integrations, once per box size.  The choice is
// kept per box size and, with mi.tune_cache = <file>, in a JSON file keyed by mechanism,
// GPU, tolerances and box size, so later processes skip the search.
// ---------------------------------------------------------------------------
std::vector<TuneConfig> tune_candidates(bool with_persistent) {
  std::vector<TuneConfig> c;
  for (int sched : {SCHED_PLAIN, SCHED_COMPACT})
    for (int g : {0, 8}) {
      TuneConfig t;
      t.schedule = sched;
      t.graph_trips = g;
      t.order = ORDER_RATE;
      c.push_back(t);
    }
  TuneConfig u;
  u.order = ORDER_NONE;
  c.push_back(u);
  if (!with_persistent) return c;
  for (int chunk : {1, 8, 32}) {
    TuneConfig t;
    t.schedule = SCHED_PERSIST;
    t.chunk = chunk;
    t.occ = 1;
    c.push_back(t);
  }
  TuneConfig t2;
  t2.schedule = SCHED_PERSIST;
  t2.chunk = 8;
  t2.occ = 2;
  c.push_back(t2);
  return c;
}

const char* sched_name(int s) { return s == SCHED_COMPACT ? "queue" : s == SCHED_PERSIST ? "persistent" : "plain"; }
const char* order_name(int o) { return o == ORDER_RATE ? "rate" : o == ORDER_FC ? "fc" : "none"; }

std::string tune_key(Reactor* r, long n) {
  hipDeviceProp_t prop{};
  int dev = 0;
  std::string gpu = "unknown";
  if (hipGetDevice(&dev) == hipSuccess && hipGetDeviceProperties(&prop, dev) == hipSuccess)
    gpu = std::string(prop.gcnArchName) + " " + std::to_string(prop.multiProcessorCount) + "CU";
  char buf[128];
  std::snprintf(buf, sizeof(buf), "|%.6g|%.6g|%d|%ld", r->rtol, r->atol, r->use_typ ? 1 : 0, n);
  return mech_name(r->mech) + std::string("|") + gpu + buf;
}

// the cache file's entries: key -> the object's text
std::map<std::string, std::string> cache_read(const std::string& path) {
  std::map<std::string, std::string> out;
  FILE* f = std::fopen(path.c_str(), "r");
  if (!f) return out;
  std::string txt;
  char buf[4096];
  size_t got;
  while ((got = std::fread(buf, 1, sizeof(buf), f)) > 0) txt.append(buf, got);
  std::fclose(f);
  size_t pos = 0;
  for (;;) {
    const size_t k0 = txt.find('"', pos);
    if (k0 == std::string::npos) break;
    const size_t k1 = txt.find('"', k0 + 1);
    const size_t o0 = txt.find('{', k1);
    const size_t o1 = txt.find('}', o0);
    if (k1 == std::string::npos || o0 == std::string::npos || o1 == std::string::npos) break;
    out[txt.substr(k0 + 1, k1 - k0 - 1)] = txt.substr(o0, o1 - o0 + 1);
    pos = o1 + 1;
  }
  return out;
}

std::string json_field(const std::string& obj, const char* name) {
  const std::string key = std::string("\"") + name + "\":";
  size_t p = obj.find(key);
  if (p == std::string::npos) return "";
  p += key.size();
  while (p < obj.size() && (obj[p] == ' ' || obj[p] == '"')) ++p;
  size_t e = p;
  while (e < obj.size() && obj[e] != '"' && obj[e] != ',' && obj[e] != '}') ++e;
  return obj.substr(p, e - p);
}

bool cache_lookup(const std::string& path, const std::string& key, TuneConfig* cfg, double* ms) {
  const auto entries = cache_read(path);
  const auto it = entries.find(key);
  if (it == entries.end()) return false;
  const std::string& o = it->second;
  const std::string sc = json_field(o, "schedule"), od = json_field(o, "order");
  TuneConfig c;
  if (sc == "plain") c.schedule = SCHED_PLAIN;
  else if (sc == "queue") c.schedule = SCHED_COMPACT;
  else if (sc == "persistent") c.schedule = SCHED_PERSIST;
  else return false;
  c.order = od == "rate" ? ORDER_RATE : od == "fc" ? ORDER_FC : ORDER_NONE;
  c.graph_trips = std::atoi(json_field(o, "graph_trips").c_str());
  c.chunk = std::atoi(json_field(o, "persist_chunk").c_str());
  c.occ = std::atoi(json_field(o, "persist_occupancy").c_str());
  if (c.graph_trips < 0 || (c.graph_trips & 1) || c.chunk < 1 || c.chunk > 64 || (c.occ != 1 && c.occ != 2))
    return false;
  *cfg = c;
  *ms = std::atof(json_field(o, "ms").c_str());
  return true;
}

void cache_store(const std::string& path, const std::string& key, const TuneConfig& c, double ms) {
  auto entries = cache_read(path);
  char buf[256];
  std::snprintf(buf, sizeof(buf),
                "{\"schedule\": \"%s\", \"graph_trips\": %d, \"order\": \"%s\", \"persist_chunk\": %d, "
                "\"persist_occupancy\": %d, \"ms\": %.4f}",
                sched_name(c.schedule), c.graph_trips, order_name(c.order), c.chunk, c.occ, ms);
  entries[key] = buf;
  const std::string tmp = path + ".tmp";
  FILE* f = std::fopen(tmp.c_str(), "w");
  if (!f) return;
  std::fprintf(f, "{\n");
  size_t i = 0;
  for (const auto& e : entries)
    std::fprintf(f, " \"%s\": %s%s\n", e.first.c_str(), e.second.c_str(), ++i < entries.size() ? "," : "");
  std::fprintf(f, "}\n");
  std::fclose(f);
  (void)std::rename(tmp.c_str(), path.c_str());
}

// one trial integration of the flat cells of `a` (the reactor's workspace) in
// configuration c, timed with the reactor's events
int trial(Reactor* r, ReactArgs a, const TuneConfig& c, hipStream_t stream, long max_trips, float* ms) {
  HIPCHK(r, launch_counters_reset(r->counters, stream));
  HIPCHK(r, hipEventRecord(r->ev0, stream));
  if (c.schedule == SCHED_PERSIST) {
    int& cap = r->persist_cap[r->mech][c.occ - 1];
    if (cap <= 0) cap = persist_capacity(r->mech, c.occ);
    if (cap <= 0) return fail(r, MI_EHIP, "cannot query the persistent kernel's occupancy");
    PersistPlan plan;
    plan.occupancy = c.occ;
    plan.chunk = c.chunk;
    plan.waves = (int)std::max(1L, std::min((long)cap, (a.ncell + 63) / 64));
    a.compact = 0;
    HIPCHK(r, run_react_persist(r->mech, a, plan, max_trips, stream));
  } else {
    a.compact = c.schedule == SCHED_COMPACT ? 1 : 0;
    a.slot_cell = nullptr;
    if (c.order != ORDER_NONE) {
      const int rc = ensure_order_buffers(r, a.ncell);
      if (rc) return rc;
      HIPCHK(r, launch_cost_keys(r->mech, a, c.order, r->order_keys, r->order_vals, stream));
      const int* sorted = nullptr;
      HIPCHK(r, cost_sort(r->order_temp, r->order_temp_bytes, r->order_keys, r->order_vals, a.ncell, stream,
                          &sorted));
      a.slot_cell = sorted;
    }
    TripPlan plan;
    plan.graph_trips = c.graph_trips;
    plan.capture_stream = r->capture_stream;
    long trips = 0;
    HIPCHK(r, run_react(r->mech, a, stream, max_trips, &trips, nullptr, nullptr, plan));
  }
  HIPCHK(r, hipEventRecord(r->ev1, stream));
  HIPCHK(r, hipEventSynchronize(r->ev1));
  HIPCHK(r, hipEventElapsedTime(ms, r->ev0, r->ev1));
  return MI_OK;
}

// the search; a0 carries the reactor's workspace (sized for the call's cells)
int autotune(Reactor* r, const ReactArgs& a0, hipStream_t stream, long max_trips) {
  const int N = mech_nspec(r->mech);
  const long n = a0.ncell;
  Reactor::Tuned t{};
  const std::string key = tune_key(r, n);
  if (!r->tune_cache.empty() && cache_lookup(r->tune_cache, key, &t.cfg, &t.ms[0])) {
    t.ms[1] = 0.0;
    t.ms[2] = 0.0;
    t.ms[3] = 1.0;
    r->tuned[n] = t;
    return MI_OK;
  }
  const auto t_start = std::chrono::steady_clock::now();
  double *rY = nullptr, *T = nullptr, *rE = nullptr, *FC = nullptr, *cY = nullptr, *cT = nullptr, *cE = nullptr,
         *fY = nullptr, *fE = nullptr;
  struct Free {
    double** p[9];
    ~Free() {
      for (auto q : p)
        if (*q) (void)hipFree(*q);
    }
  } fr{{&rY, &T, &rE, &FC, &cY, &cT, &cE, &fY, &fE}};
  for (double** p : {&rY, &cY, &fY}) HIPCHK(r, hipMalloc(p, sizeof(double) * N * n));
  for (double** p : {&T, &rE, &FC, &cT, &cE, &fE}) HIPCHK(r, hipMalloc(p, sizeof(double) * n));
  std::vector<TuneConfig> cand = tune_candidates(!r->gmres && mech_nspec(r->mech) + 1 <= COOP_MAX_N1);
  const long strides[3] = {16, 4, 1};
  const size_t keep[3] = {3, 2, 1};
  std::vector<float> last_ms;
  for (int round = 0; round < 3; ++round) {
    const long m = (n + strides[round] - 1) / strides[round];
    HIPCHK(r, launch_gather_state(N, a0, strides[round], m, cY, cT, cE, fY, fE, stream));
    ReactArgs a = a0;
    auto flat = [m](double* p) {
      ArrView v{};
      v.p = p;
      v.js = m;
      v.ks = m;
      v.ns = m;
      return v;
    };
    a.rY = flat(rY);
    a.T = flat(T);
    a.rE = flat(rE);
    a.FC = flat(FC);
    a.rYsrc = flat(fY);
    a.rEsrc = flat(fE);
    a.mask = IntView{};
    a.nx = (int)m;
    a.ny = 1;
    a.ncell = m;
    a.use_pred = 0;
    std::vector<std::pair<float, size_t>> res;
    for (size_t i = 0; i < cand.size(); ++i) {
      HIPCHK(r, hipMemcpyAsync(rY, cY, sizeof(double) * N * m, hipMemcpyDeviceToDevice, stream));
      HIPCHK(r, hipMemcpyAsync(T, cT, sizeof(double) * m, hipMemcpyDeviceToDevice, stream));
      HIPCHK(r, hipMemcpyAsync(rE, cE, sizeof(double) * m, hipMemcpyDeviceToDevice, stream));
      float ms = 0.f;
      const int rc = trial(r, a, cand[i], stream, max_trips, &ms);
      if (rc) return rc;
      res.push_back({ms, i});
    }
    std::stable_sort(res.begin(), res.end());
    std::vector<TuneConfig> next;
    last_ms.clear();
    for (size_t i = 0; i < std::min(keep[round], res.size()); ++i) {
      next.push_back(cand[res[i].second]);
      last_ms.push_back(res[i].first);
    }
    if (round == 0) t.ms[2] = (double)cand.size();
    cand = next;
  }
  t.cfg = cand[0];
  t.ms[0] = last_ms[0];
  t.ms[1] = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t_start).count();
  t.ms[3] = 0.0;
  r->tuned[n] = t;
  if (!r->tune_cache.empty()) cache_store(r->tune_cache, key, t.cfg, t.ms[0]);
  return MI_OK;
}

int react_common(Reactor* r, ReactArgs& a, double* dt, double* time, hipStream_t stream,
                 const PersistMem* fixed = nullptr) {
  if (!r->inited) return fail(r, MI_ESTATE, "react called before init");
  if (!dt || !(*dt > 0.0)) return fail(r, MI_EINVAL, "dt must be positive");
  if (!a.rY.p || !a.T.p || !a.rE.p || !a.FC.p) return fail(r, MI_EINVAL, "rY, T, rE and FC are required");
  const int N = mech_nspec(r->mech);
  const int N1 = N + 1;
  if (N1 > NMAX1) return fail(r, MI_EINVAL, "mechanism too large");
  a.dt = *dt;
  a.rtol = r->rtol;
  for (int i = 0; i < N1; ++i) a.atol[i] = r->use_typ && (int)r->typ.size() == N1 ? r->atol * r->typ[i] : r->atol;
  a.qmax = r->max_order;
  a.mxstep = r->max_steps;
  a.lu_mode = r->lu_mode;
  a.ir_sweeps = r->ir_sweeps;
  a.ir_tol = r->ir_tol;
  a.lu_lane = r->lu_lane;
  a.lu_mfma = r->lu_mfma;
  const bool wide = N1 > COOP_MAX_N1;
  if (wide && !r->rk64 && !r->null_reactor) {
    // one matrix row per lane of a wave does not reach: one wave per cell with two rows
    // per lane, FP64 factors, the rows in registers (mi.lu_form = mfma: staged in LDS;
    // lane: one lane per cell)
    a.lu_mode = LU_FP64;
    if (fixed || r->schedule == SCHED_PERSIST)
      return fail(r, MI_EINVAL, "mechanisms above 63 species run on the trip schedules (plain, queue, auto), not "
                                "the persistent kernel");
    if (r->predictor) return fail(r, MI_EINVAL, "the learned predictor takes at most 61 species");
  }
  a.gmres = (r->gmres && !r->rk64 && !r->null_reactor) ? 1 : 0;
  a.gmres_maxl = r->gmres_maxl;
  if (a.gmres) {
    if (r->gmres_maxl > N1 - 1)
      return fail(r, MI_EINVAL, "mi.gmres_maxl must not exceed the number of species (the Krylov vectors share "
                                "the Jacobian's workspace)");
    if (fixed || r->schedule == SCHED_PERSIST)
      return fail(r, MI_EINVAL, "cvode.solve_type = GMRES runs on the trip schedules (plain, queue, auto), not "
                                "the persistent kernel");
    if (r->predictor) return fail(r, MI_EINVAL, "the learned predictor needs the direct solve");
    a.lu_mode = LU_FP64;   // no factorisation
  }
  a.compact = (r->schedule == SCHED_COMPACT) ? 1 : 0;
  int graph_trips = r->graph_trips;
  a.cell_status = nullptr;
  r->last_ncell = a.ncell;
  if (a.ncell == 0) {
    *time += *dt;
    return MI_OK;
  }
  if (r->rk64) {   // explicit sub-steps, one lane per cell, one kernel
    if (fixed)
      return fail(r, MI_EINVAL, "the no-allocation entry point runs the BDF reactor (ReactorRK64 keeps a workspace "
                                "slot per cell)");
    int rc = ensure_device_state(r, a.ncell);
    if (rc) return rc;
    a.ws = r->ws;
    a.iws = r->iws;
    a.counts = r->counts;
    a.counters = r->counters;
    r->stat_counts = r->counts;
    r->stat_counters = r->counters;
    r->last_schedule = SCHED_PLAIN;
    r->last_graph = 0;
    r->last_trips = 0;
    r->prof = KernelProf{};
    HIPCHK(r, launch_counters_reset(r->counters, stream));
    HIPCHK(r, hipEventRecord(r->ev0, stream));
    HIPCHK(r, launch_rk64(r->mech, a, r->rk_guess, r->rk_min, r->rk_max, stream));
    HIPCHK(r, hipEventRecord(r->ev1, stream));
    r->last_stream = stream;
    r->have_stats = true;
    *time += *dt;
    return failure_code(r, stream);
  }
  if (fixed) {   // the no-allocation entry point: the persistent schedule in fixed memory
    if (r->null_reactor) {
      HIPCHK(r, launch_null(N, a, stream));
      r->have_stats = false;
      *time += *dt;
      return MI_OK;
    }
    return react_persist(r, a, dt, time, stream, fixed);
  }
  if (r->schedule == SCHED_PERSIST && !r->null_reactor) return react_persist(r, a, dt, time, stream);
  int rc = ensure_device_state(r, r->null_reactor ? 0 : a.ncell);
  if (rc) return rc;
  a.ws = r->ws;
  a.iws = r->iws;
  for (int i = 0; i < 2; ++i) {
    a.list[i] = r->lists + (size_t)i * r->ws_cells;
    a.slist[i] = r->lists + (size_t)(2 + i) * r->ws_cells;
  }
  a.plist = r->lists + (size_t)4 * r->ws_cells;
  a.nlist = r->lists + (size_t)5 * r->ws_cells;
  a.mat = r->mat;
  a.use_pred = (r->predictor && !r->null_reactor) ? 1 : 0;
  a.counts = r->counts;
  a.counters = r->counters;
  r->stat_counts = r->counts;
  r->stat_counters = r->counters;
  // every step takes at least one trip; 3 per step bounds retries generously
  const long max_trips = r->max_trips > 0 ? r->max_trips : 3 * r->max_steps + 100;
  if (graph_trips > 0 || r->schedule == SCHED_AUTO) {
    if (!r->capture_stream) HIPCHK(r, hipStreamCreateWithFlags(&r->capture_stream, hipStreamNonBlocking));
  }
  int order_mode = r->order_mode;
  if (r->schedule == SCHED_AUTO && !r->null_reactor) {
    auto it = r->tuned.find(a.ncell);
    if (it == r->tuned.end()) {
      rc = autotune(r, a, stream, max_trips);
      if (rc) return rc;
      it = r->tuned.find(a.ncell);
    }
    const TuneConfig cfg = it->second.cfg;
    if (cfg.schedule == SCHED_PERSIST) return react_persist(r, a, dt, time, stream, nullptr, &cfg);
    a.compact = (cfg.schedule == SCHED_COMPACT) ? 1 : 0;
    graph_trips = cfg.graph_trips;
    order_mode = cfg.order;
  }
  if (a.use_pred && !a.compact) return fail(r, MI_EINVAL, "the learned predictor needs mi.schedule = queue");
  r->last_schedule = a.compact ? SCHED_COMPACT : SCHED_PLAIN;
  r->last_graph = (r->profile || a.use_pred) ? 0 : graph_trips;
  HIPCHK(r, launch_counters_reset(r->counters, stream));
  HIPCHK(r, hipEventRecord(r->ev0, stream));
  if (r->null_reactor) {
    HIPCHK(r, launch_null(N, a, stream));
  } else {
    if (order_mode != ORDER_NONE) {
      // workspace slot s integrates cell slot_cell[s], the cells in ascending cost, so a
      // wave's lanes finish together (the keys and the sort are part of the call)
      rc = ensure_order_buffers(r, a.ncell);
      if (rc) return rc;
      HIPCHK(r, launch_cost_keys(r->mech, a, order_mode, r->order_keys, r->order_vals, stream));
      const int* sorted = nullptr;
      HIPCHK(r, cost_sort(r->order_temp, r->order_temp_bytes, r->order_keys, r->order_vals, a.ncell, stream,
                          &sorted));
      a.slot_cell = sorted;
    }
    r->prof = KernelProf{};
    TripPlan plan;
    plan.graph_trips = graph_trips;
    plan.capture_stream = r->capture_stream;
    HIPCHK(r, run_react(r->mech, a, stream, max_trips, &r->last_trips, r->profile ? &r->prof : nullptr,
                        mlp_device(r->predictor), plan));
  }
  HIPCHK(r, hipEventRecord(r->ev1, stream));
  r->last_stream = stream;
  r->have_stats = true;
  *time += *dt;
  return failure_code(r, stream);
}

ArrView flat_view(double* p, long ncell) {
  ArrView v{};
  v.p = p;
  v.js = 0;
  v.ks = 0;
  v.ns = ncell;
  return v;
}

ArrView box_view(const mi_real4& x, const int lo[3]) {
  ArrView v{};
  v.p = x.p;
  v.js = x.jstride;
  v.ks = x.kstride;
  v.ns = x.nstride;
  v.ox = lo[0] - x.begin[0];
  v.oy = lo[1] - x.begin[1];
  v.oz = lo[2] - x.begin[2];
  return v;
}

IntView box_iview(const mi_int4& x, const int lo[3]) {
  IntView v{};
  v.p = x.p;
  v.js = x.jstride;
  v.ks = x.kstride;
  v.ox = lo[0] - x.begin[0];
  v.oy = lo[1] - x.begin[1];
  v.oz = lo[2] - x.begin[2];
  return v;
}

template <class V>
bool box_covers(const V& x, const int lo[3], const int hi[3], int ncomp) {
  if (!x.p) return true;
  for (int d = 0; d < 3; ++d)
    if (lo[d] < x.begin[d] || hi[d] >= x.end[d]) return false;
  return x.ncomp >= ncomp;
}

// ---------------------------------------------------------------------------
// Device memory of mi_chem_react_flat_noalloc: the persistent schedule's lane slots
// (workspace tiles and matrix regions) and the call's queue and statistics counters,
// carved from one file-scope array.  Every GPU the library runs on holds its own copy
// until the process ends.  Host code never uses the array's name as an address: the
// address of the current device's copy comes from hipGetSymbolAddress, once per device.
// ---------------------------------------------------------------------------
constexpr size_t NOALLOC_BYTES = size_t(256) << 20;
static __device__ __align__(256) unsigned char scratch[NOALLOC_BYTES];

constexpr int NOALLOC_MAX_DEVICES = 64;
std::mutex g_noalloc_mutex;
unsigned char* g_noalloc_ptr[NOALLOC_MAX_DEVICES] = {};

// the scratch array's address on the current device
int noalloc_scratch(Reactor* r, unsigned char** out) {
  int dev = 0;
  HIPCHK(r, hipGetDevice(&dev));
  if (dev < 0 || dev >= NOALLOC_MAX_DEVICES) return fail(r, MI_EHIP, "device index beyond the scratch table");
  std::lock_guard<std::mutex> lock(g_noalloc_mutex);
  if (!g_noalloc_ptr[dev]) {
    void* p = nullptr;
    HIPCHK(r, hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    g_noalloc_ptr[dev] = static_cast<unsigned char*>(p);
  }
  *out = g_noalloc_ptr[dev];
  return MI_OK;
}

size_t align256(size_t x) { return (x + 255) & ~size_t(255); }

// the scratch array's layout for `waves` waves of lane slots of mechanism `mech`; returns
// the bytes used
size_t noalloc_layout(int mech, long waves, unsigned char* base, PersistMem* m) {
  size_t off = 0;
  auto take = [&](size_t bytes) {
    unsigned char* p = base ? base + off : nullptr;
    off = align256(off + bytes);
    return p;
  };
  unsigned char* counts = take(8 * sizeof(unsigned int));
  unsigned char* counters = take(NCOUNTERS * sizeof(unsigned long long));
  unsigned char* ws = take((size_t)waves * ws_slots(mech) * 64 * sizeof(double));
  unsigned char* iws = take((size_t)waves * iws_slots(mech) * 64 * sizeof(int));
  unsigned char* mat = take((size_t)waves * 64 * mat_stride(mech) * sizeof(double));
  if (m) {
    m->counts = reinterpret_cast<unsigned int*>(counts);
    m->counters = reinterpret_cast<unsigned long long*>(counters);
    m->ws = reinterpret_cast<double*>(ws);
    m->iws = reinterpret_cast<int*>(iws);
    m->mat = reinterpret_cast<double*>(mat);
    m->max_waves = waves;
  }
  return off;
}

// the most waves of lane slots of `mech` the scratch array holds
long noalloc_max_waves(int mech) {
  const size_t fixed = noalloc_layout(mech, 0, nullptr, nullptr);
  const size_t per_wave = noalloc_layout(mech, 1, nullptr, nullptr) - fixed + 3 * 256;
  return NOALLOC_BYTES > fixed ? (long)((NOALLOC_BYTES - fixed) / per_wave) : 0;
}

}  // namespace

extern "C" {

int mi_chem_create(const char* integrator, void** reactor) {
  if (!reactor || !integrator) return MI_EINVAL;
  *reactor = nullptr;
  const std::string name(integrator);
  auto* r = new Reactor();
  r->integrator = name;
  if (name == "ReactorNull") {
    r->null_reactor = true;
  } else if (name == "ReactorRK64") {
    r->rk64 = true;
  } else if (name == "ReactorCvode" || name == "ReactorBDF") {
    r->null_reactor = false;
  } else {
    delete r;
    return MI_EINVAL;
  }
  *reactor = r;
  return MI_OK;
}

int mi_chem_init(void* reactor, int reactor_type, int ncells) {
  auto* r = static_cast<Reactor*>(reactor);
  if (!r) return MI_EINVAL;
  if (reactor_type != 2) return fail(r, MI_EINVAL, "only reactor_type 2 (h-reactor, constant pressure) is supported");
  if (ncells < 0) return fail(r, MI_EINVAL, "ncells must be >= 0");
  r->reactor_type = reactor_type;
  // set-up for the no-allocation entry point, when a device is present: the call's
  // timing events, the scratch array's address on this device and the persistent
  // kernel's resident waves (without them that entry point takes the address and the
  // occupancy on its first call and records no timing)
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) == hipSuccess && ndev > 0) {
    if (!r->ev0) HIPCHK(r, hipEventCreate(&r->ev0));
    if (!r->ev1) HIPCHK(r, hipEventCreate(&r->ev1));
    unsigned char* base = nullptr;
    const int rc = noalloc_scratch(r, &base);
    if (rc) return rc;
    if (!r->null_reactor) {
      int& cap = r->persist_cap[r->mech][r->persist_occ - 1];
      if (cap <= 0) cap = persist_capacity(r->mech, r->persist_occ);
    }
  } else {
    (void)hipGetLastError();
  }
  r->inited = true;
  return MI_OK;
}

int mi_chem_set(void* reactor, const char* key, const char* value) {
  auto* r = static_cast<Reactor*>(reactor);
  if (!r) return MI_EINVAL;
  if (!key || !value) return fail(r, MI_EINVAL, "null key or value");
  const std::string k(key), v(value);
  double d;
  long l;
  if (k == "ode.rtol") {
    if (!parse_double(value, &d) || d <= 0.0) return fail(r, MI_EINVAL, "ode.rtol must be a positive number");
    r->rtol = d;
  } else if (k == "ode.atol") {
    if (!parse_double(value, &d) || d <= 0.0) return fail(r, MI_EINVAL, "ode.atol must be a positive number");
    r->atol = d;
  } else if (k == "cvode.max_order") {
    if (!parse_long(value, &l) || l < 1 || l > 5) return fail(r, MI_EINVAL, "cvode.max_order must be 1..5");
    r->max_order = (int)l;
  } else if (k == "cvode.max_num_steps") {
    if (!parse_long(value, &l) || l < 1) return fail(r, MI_EINVAL, "cvode.max_num_steps must be >= 1");
    r->max_steps = l;
  } else if (k == "cvode.solve_type") {
    if (v == "denseAJ_direct" || v == "dense_direct") r->gmres = 0;
    else if (v == "GMRES") r->gmres = 1;
    else return fail(r, MI_EINVAL, "cvode.solve_type must be denseAJ_direct, dense_direct or GMRES");
  } else if (k == "mi.gmres_maxl") {
    if (!parse_long(value, &l) || l < 1 || l > 64) return fail(r, MI_EINVAL, "mi.gmres_maxl must be 1..64");
    r->gmres_maxl = (int)l;
  } else if (k == "peleLM.use_typ_vals_chem") {
    if (!parse_long(value, &l) || (l != 0 && l != 1)) return fail(r, MI_EINVAL, "use_typ_vals_chem must be 0 or 1");
    r->use_typ = (l == 1);
  } else if (k == "mi.mechanism") {
    const int m = mech_lookup(value);
    if (m < 0) return fail(r, MI_EINVAL, "unknown mechanism " + v);
    if (m != r->mech) {
      r->mech = m;
      free_workspace(r);
      free_glue(r);   // sized for the previous mechanism's species
      r->typ.clear();
    }
  } else if (k == "mi.schedule") {
    if (v == "plain") r->schedule = SCHED_PLAIN;
    else if (v == "queue") r->schedule = SCHED_COMPACT;
    else if (v == "auto") r->schedule = SCHED_AUTO;
    else if (v == "persistent") r->schedule = SCHED_PERSIST;
    else return fail(r, MI_EINVAL, "mi.schedule must be plain, queue, auto or persistent");
  } else if (k == "ode.rk64_nsubsteps_guess" || k == "ode.rk64_nsubsteps_min" || k == "ode.rk64_nsubsteps_max") {
    if (!parse_long(value, &l) || l < 1 || l > (1L << 24)) return fail(r, MI_EINVAL, k + " must be 1..2^24");
    if (k == "ode.rk64_nsubsteps_guess") r->rk_guess = (int)l;
    else if (k == "ode.rk64_nsubsteps_min") r->rk_min = (int)l;
    else r->rk_max = (int)l;
    if (r->rk_min > r->rk_max) return fail(r, MI_EINVAL, "ode.rk64_nsubsteps_min exceeds ode.rk64_nsubsteps_max");
  } else if (k == "mi.tune_cache") {
    r->tune_cache = v;
  } else if (k == "mi.order") {
    if (v == "none") r->order_mode = ORDER_NONE;
    else if (v == "rate") r->order_mode = ORDER_RATE;
    else if (v == "fc") r->order_mode = ORDER_FC;
    else return fail(r, MI_EINVAL, "mi.order must be none, rate or fc");
  } else if (k == "mi.persist_waves") {
    if (!parse_long(value, &l) || l < 0 || l > (1L << 20)) return fail(r, MI_EINVAL, "mi.persist_waves must be 0..2^20");
    r->persist_waves = (int)l;
  } else if (k == "mi.persist_chunk") {
    if (!parse_long(value, &l) || l < 1 || l > 64) return fail(r, MI_EINVAL, "mi.persist_chunk must be 1..64");
    r->persist_chunk = (int)l;
  } else if (k == "mi.persist_occupancy") {
    if (!parse_long(value, &l) || (l != 1 && l != 2)) return fail(r, MI_EINVAL, "mi.persist_occupancy must be 1 or 2");
    r->persist_occ = (int)l;
  } else if (k == "mi.graph_trips") {
    if (!parse_long(value, &l) || l < 0 || l > 64 || (l & 1)) return fail(r, MI_EINVAL, "mi.graph_trips must be 0 or even up to 64");
    r->graph_trips = (int)l;
  } else if (k == "mi.lu") {
    if (v == "fp64") r->lu_mode = LU_FP64;
    else if (v == "fp32_ir") r->lu_mode = LU_FP32_IR;
    else return fail(r, MI_EINVAL, "mi.lu must be fp64 or fp32_ir");
  } else if (k == "mi.lu_form") {
    if (v == "coop") { r->lu_lane = 0; r->lu_mfma = 0; }
    else if (v == "lane") { r->lu_lane = 1; r->lu_mfma = 0; }
    else if (v == "mfma") { r->lu_lane = 0; r->lu_mfma = 1; }
    else return fail(r, MI_EINVAL, "mi.lu_form must be coop, lane or mfma");
  } else if (k == "mi.ir_sweeps") {
    if (!parse_long(value, &l) || l < 0 || l > 10) return fail(r, MI_EINVAL, "mi.ir_sweeps must be 0..10");
    r->ir_sweeps = (int)l;
  } else if (k == "mi.ir_tol") {
    if (!parse_double(value, &d) || d < 0.0) return fail(r, MI_EINVAL, "mi.ir_tol must be a number >= 0");
    r->ir_tol = d;
  } else if (k == "mi.max_trips") {
    if (!parse_long(value, &l) || l < 0) return fail(r, MI_EINVAL, "mi.max_trips must be >= 0 (0: automatic)");
    r->max_trips = l;
  } else if (k == "mi.profile") {
    if (!parse_long(value, &l) || (l != 0 && l != 1)) return fail(r, MI_EINVAL, "mi.profile must be 0 or 1");
    r->profile = (l == 1);
  } else if (k == "mi.on_fail") {
    if (v == "continue") r->on_fail_continue = true;
    else if (v == "abort") r->on_fail_continue = false;
    else return fail(r, MI_EINVAL, "mi.on_fail must be continue or abort");
  } else {
    return fail(r, MI_EINVAL, "unknown key " + k);
  }
  return MI_OK;
}

int mi_chem_set_typ_vals(void* reactor, const double* typ, int n) {
  auto* r = static_cast<Reactor*>(reactor);
  if (!r) return MI_EINVAL;
  if (!typ || n != mech_nspec(r->mech) + 1) return fail(r, MI_EINVAL, "set_typ_vals expects nspec + 1 values");
  for (int i = 0; i < n; ++i)
    if (!(typ[i] > 0.0)) return fail(r, MI_EINVAL, "typical values must be positive");
  r->typ.assign(typ, typ + n);
  return MI_OK;
}

int mi_chem_react_flat(void* reactor, double* rY, const double* rYsrc, double* T, double* rE, const double* rEsrc,
                       double* FC, const int* mask, int ncells, double* dt, double* time, void* hip_stream) {
  auto* r = static_cast<Reactor*>(reactor);
  if (!r) return MI_EINVAL;
  if (ncells < 0) return fail(r, MI_EINVAL, "ncells must be >= 0");
  if (!time) return fail(r, MI_EINVAL, "time is required");
  ReactArgs a{};
  a.rY = flat_view(rY, ncells);
  a.rYsrc = flat_view(const_cast<double*>(rYsrc), ncells);
  a.T = flat_view(T, ncells);
  a.rE = flat_view(rE, ncells);
  a.rEsrc = flat_view(const_cast<double*>(rEsrc), ncells);
  a.FC = flat_view(FC, ncells);
  a.mask.p = mask;
  a.nx = std::max(ncells, 1);
  a.ny = 1;
  a.ncell = ncells;
  return react_common(r, a, dt, time, static_cast<hipStream_t>(hip_stream));
}

int mi_chem_react_flat_noalloc(void* reactor, double* rY, const double* rYsrc, double* T, double* rE,
                               const double* rEsrc, double* FC, const int* mask, int ncells, double* dt,
                               double* time, void* hip_stream) {
  auto* r = static_cast<Reactor*>(reactor);
  if (!r) return MI_EINVAL;
  if (ncells < 0) return fail(r, MI_EINVAL, "ncells must be >= 0");
  if (!time) return fail(r, MI_EINVAL, "time is required");
  if (!r->inited) return fail(r, MI_ESTATE, "react called before init");
  const long maxw = noalloc_max_waves(r->mech);
  if (maxw < 1)
    return fail(r, MI_ENOMEM, "one wave of lane slots of this mechanism needs more than the " +
                                  std::to_string(NOALLOC_BYTES >> 20) + " MiB no-allocation array");
  unsigned char* base = nullptr;
  const int rc = noalloc_scratch(r, &base);
  if (rc) return rc;
  PersistMem m{};
  noalloc_layout(r->mech, maxw, base, &m);
  ReactArgs a{};
  a.rY = flat_view(rY, ncells);
  a.rYsrc = flat_view(const_cast<double*>(rYsrc), ncells);
  a.T = flat_view(T, ncells);
  a.rE = flat_view(rE, ncells);
  a.rEsrc = flat_view(const_cast<double*>(rEsrc), ncells);
  a.FC = flat_view(FC, ncells);
  a.mask.p = mask;
  a.nx = std::max(ncells, 1);
  a.ny = 1;
  a.ncell = ncells;
  return react_common(r, a, dt, time, static_cast<hipStream_t>(hip_stream), &m);
}

int mi_chem_react_box(void* reactor, const int lo[3], const int hi[3], mi_real4 rY, mi_real4 rYsrc, mi_real4 T,
                      mi_real4 rE, mi_real4 rEsrc, mi_real4 FC, mi_int4 mask, double* dt, double* time,
                      void* hip_stream) {
  auto* r = static_cast<Reactor*>(reactor);
  if (!r) return MI_EINVAL;
  if (!lo || !hi || !time) return fail(r, MI_EINVAL, "lo, hi and time are required");
  for (int d = 0; d < 3; ++d)
    if (hi[d] < lo[d]) return fail(r, MI_EINVAL, "empty box");
  const int N = mech_nspec(r->mech);
  if (!box_covers(rY, lo, hi, N) || !box_covers(rYsrc, lo, hi, N) || !box_covers(T, lo, hi, 1) ||
      !box_covers(rE, lo, hi, 1) || !box_covers(rEsrc, lo, hi, 1) || !box_covers(FC, lo, hi, 1) ||
      !box_covers(mask, lo, hi, 1))
    return fail(r, MI_EINVAL, "an array view does not cover the box or has too few components");
  ReactArgs a{};
  a.rY = box_view(rY, lo);
  a.rYsrc = box_view(rYsrc, lo);
  a.T = box_view(T, lo);
  a.rE = box_view(rE, lo);
  a.rEsrc = box_view(rEsrc, lo);
  a.FC = box_view(FC, lo);
  a.mask = box_iview(mask, lo);
  a.nx = hi[0] - lo[0] + 1;
  a.ny = hi[1] - lo[1] + 1;
  a.ncell = (long)a.nx * a.ny * (hi[2] - lo[2] + 1);
  return react_common(r, a, dt, time, static_cast<hipStream_t>(hip_stream));
}

int mi_chem_reaction_rates_flat(void* reactor, const double* rY, const double* rH, const double* T,
                                const int* mask, double* rYdot, double* T_out, int ncells, void* hip_stream) {
  auto* r = static_cast<Reactor*>(reactor);
  if (!r) return MI_EINVAL;
  if (ncells < 0 || !rY || !rH || !T || !rYdot) return fail(r, MI_EINVAL, "bad arguments");
  if (ncells == 0) return MI_OK;
  RatesArgs ra;
  ra.rhoY = flat_view(const_cast<double*>(rY), ncells);
  ra.rhoH = flat_view(const_cast<double*>(rH), ncells);
  ra.T = flat_view(const_cast<double*>(T), ncells);
  ra.rdot = flat_view(rYdot, ncells);
  ra.mask.p = mask;
  ra.Tout = T_out;
  ra.nx = ncells;
  ra.ncell = ncells;
  HIPCHK(r, launch_rates(r->mech, ra, static_cast<hipStream_t>(hip_stream)));
  return MI_OK;
}

int mi_chem_reaction_rates_box(void* reactor, const int lo[3], const int hi[3], mi_real4 rhoY, mi_real4 rhoH,
                               mi_real4 T, mi_int4 mask, mi_real4 rhoYdot, void* hip_stream) {
  auto* r = static_cast<Reactor*>(reactor);
  if (!r) return MI_EINVAL;
  if (!lo || !hi) return fail(r, MI_EINVAL, "lo and hi are required");
  for (int d = 0; d < 3; ++d)
    if (hi[d] < lo[d]) return fail(r, MI_EINVAL, "empty box");
  const int N = mech_nspec(r->mech);
  if (!rhoY.p || !rhoH.p || !T.p || !rhoYdot.p) return fail(r, MI_EINVAL, "rhoY, rhoH, T and rhoYdot are required");
  if (!box_covers(rhoY, lo, hi, N) || !box_covers(rhoH, lo, hi, 1) || !box_covers(T, lo, hi, 1) ||
      !box_covers(rhoYdot, lo, hi, N) || !box_covers(mask, lo, hi, 1))
    return fail(r, MI_EINVAL, "an array view does not cover the box or has too few components");
  RatesArgs ra;
  ra.rhoY = box_view(rhoY, lo);
  ra.rhoH = box_view(rhoH, lo);
  ra.T = box_view(T, lo);
  ra.rdot = box_view(rhoYdot, lo);
  ra.mask = box_iview(mask, lo);
  ra.nx = hi[0] - lo[0] + 1;
  ra.ny = hi[1] - lo[1] + 1;
  ra.ncell = (long)ra.nx * ra.ny * (hi[2] - lo[2] + 1);
  // PeleLM's MKS state: h = rhoH / rho x 1e4 (J/kg -> erg/g), rho x 1e-3 (kg/m^3 -> g/cm^3),
  // rates x 1e3 (g/cm^3/s -> kg/m^3/s), Source/PeleLM_K.H:87-97
  ra.hscale = 1.0e4;
  ra.rscale = 1.0e-3;
  ra.oscale = 1.0e3;
  HIPCHK(r, launch_rates(r->mech, ra, static_cast<hipStream_t>(hip_stream)));
  return MI_OK;
}

int mi_chem_advance_box(void* reactor, const int lo[3], const int hi[3], mi_real4 s_old, mi_real4 s_new,
                        mi_real4 force, mi_real4 rhoYdot, mi_real4 FC, mi_int4 mask, double dt, void* hip_stream) {
  auto* r = static_cast<Reactor*>(reactor);
  if (!r) return MI_EINVAL;
  if (!r->inited) return fail(r, MI_ESTATE, "advance_box called before init");
  if (!lo || !hi) return fail(r, MI_EINVAL, "lo and hi are required");
  if (!(dt > 0.0)) return fail(r, MI_EINVAL, "dt must be positive");
  for (int d = 0; d < 3; ++d)
    if (hi[d] < lo[d]) return fail(r, MI_EINVAL, "empty box");
  const int N = mech_nspec(r->mech);
  if (!s_old.p || !s_new.p || !force.p || !rhoYdot.p || !FC.p)
    return fail(r, MI_EINVAL, "s_old, s_new, force, rhoYdot and FC are required");
  if (!box_covers(s_old, lo, hi, N + 2) || !box_covers(s_new, lo, hi, N + 2) || !box_covers(force, lo, hi, N + 1) ||
      !box_covers(rhoYdot, lo, hi, N) || !box_covers(FC, lo, hi, 1))
    return fail(r, MI_EINVAL, "an array view does not cover the box or has too few components");
  {  // s_old is read after the reaction (I_R): it must not overlap s_new's storage
    const double* a0 = s_old.p;
    const double* a1 = s_old.p + (long)(s_old.ncomp) * s_old.nstride;
    const double* b0 = s_new.p;
    const double* b1 = s_new.p + (long)(s_new.ncomp) * s_new.nstride;
    if (a0 < b1 && b0 < a1) return fail(r, MI_EINVAL, "s_old and s_new must not share memory");
  }
  const int nx = hi[0] - lo[0] + 1, ny = hi[1] - lo[1] + 1;
  const long ncell = (long)nx * ny * (hi[2] - lo[2] + 1);
  hipStream_t st = static_cast<hipStream_t>(hip_stream);
  if (ncell > r->glue_cells) {
    if (r->glue) HIPCHK(r, hipFree(r->glue));
    r->glue = nullptr;
    r->glue_cells = 0;
    if (hipMalloc(&r->glue, sizeof(double) * (N + 1) * (size_t)ncell) != hipSuccess)
      return fail(r, MI_ENOMEM, "glue buffer allocation failed");
    r->glue_cells = ncell;
  }
  const ArrView vo = box_view(s_old, lo), vn = box_view(s_new, lo), vf = box_view(force, lo);
  HIPCHK(r, launch_glue_in(vo, vn, vf, r->glue, N, nx, ny, ncell, st));
  // the forcing's CGS copy as a flat box-shaped view (cell fastest)
  mi_real4 fy{};
  fy.p = r->glue;
  for (int d = 0; d < 3; ++d) {
    fy.begin[d] = lo[d];
    fy.end[d] = hi[d] + 1;
  }
  fy.jstride = nx;
  fy.kstride = (long)nx * ny;
  fy.nstride = ncell;
  fy.ncomp = N + 1;
  mi_real4 fh = fy;
  fh.p = r->glue + (long)N * ncell;
  fh.ncomp = 1;
  auto comp = [](mi_real4 v, int n) {
    v.p += (long)n * v.nstride;
    v.ncomp -= n;
    return v;
  };
  double dt_io = dt, time = 0.0;
  const int rc = mi_chem_react_box(reactor, lo, hi, s_new, fy, comp(s_new, N + 1), comp(s_new, N), fh, FC, mask,
                                   &dt_io, &time, hip_stream);
  if (rc != MI_OK && rc != MI_ECELLFAIL) return rc;
  const ArrView vd = box_view(rhoYdot, lo);
  HIPCHK(r, launch_glue_out(vo, vn, vf, vd, 1.0 / dt, N, nx, ny, ncell, st));
  return rc;
}

int mi_chem_rhoh_to_temp_box(void* reactor, const int lo[3], const int hi[3], mi_real4 rho, mi_real4 rhoY,
                             mi_real4 rhoH, mi_real4 T, void* hip_stream) {
  auto* r = static_cast<Reactor*>(reactor);
  if (!r) return MI_EINVAL;
  if (!lo || !hi) return fail(r, MI_EINVAL, "lo and hi are required");
  for (int d = 0; d < 3; ++d)
    if (hi[d] < lo[d]) return fail(r, MI_EINVAL, "empty box");
  const int N = mech_nspec(r->mech);
  if (!rho.p || !rhoY.p || !rhoH.p || !T.p) return fail(r, MI_EINVAL, "rho, rhoY, rhoH and T are required");
  if (!box_covers(rho, lo, hi, 1) || !box_covers(rhoY, lo, hi, N) || !box_covers(rhoH, lo, hi, 1) ||
      !box_covers(T, lo, hi, 1))
    return fail(r, MI_EINVAL, "an array view does not cover the box or has too few components");
  const int nx = hi[0] - lo[0] + 1, ny = hi[1] - lo[1] + 1;
  const long ncell = (long)nx * ny * (hi[2] - lo[2] + 1);
  HIPCHK(r, launch_rhoh_to_temp(r->mech, box_view(rho, lo), box_view(rhoY, lo), box_view(rhoH, lo),
                                box_view(T, lo), nx, ny, ncell, static_cast<hipStream_t>(hip_stream)));
  return MI_OK;
}

int mi_chem_get_stats(void* reactor, mi_chem_stats* out) {
  auto* r = static_cast<Reactor*>(reactor);
  if (!r || !out) return MI_EINVAL;
  std::memset(out, 0, sizeof(*out));
  if (!r->have_stats) return MI_OK;
  if (r->ev1) HIPCHK(r, hipEventSynchronize(r->ev1));
  else HIPCHK(r, hipStreamSynchronize(r->last_stream));
  unsigned long long c[NCOUNTERS];
  HIPCHK(r, hipMemcpy(c, r->stat_counters, sizeof(c), hipMemcpyDeviceToHost));
  out->ncells = (long)c[CNT_CELLS];
  out->nsteps = (long)c[CNT_NST];
  out->nfe = (long)c[CNT_NFE];
  out->nje = (long)c[CNT_NJE];
  out->nsetups = (long)c[CNT_NSETUPS];
  out->nni = (long)c[CNT_NNI];
  out->ncfn = (long)c[CNT_NCFN];
  out->netf = (long)c[CNT_NETF];
  out->nfail = (long)c[CNT_NFAIL];
  out->max_steps_cell = (long)c[CNT_MAXSTEPS];
  out->nlu_refine = (long)c[CNT_IR];
  out->nlu_fallback = (long)c[CNT_LUFALLBACK];
  out->nhy2t_iters = (long)c[CNT_HY2T];
  out->npred_used = (long)c[CNT_PRED];
  out->npred_fallback = (long)c[CNT_PREDFB];
  out->nli = (long)c[CNT_NLI];
  float ms = 0.f;
  if (r->ev0 && r->ev1) HIPCHK(r, hipEventElapsedTime(&ms, r->ev0, r->ev1));
  out->ms_kernel = ms;
  out->ntrips = r->last_trips;
  if (r->last_schedule == SCHED_PERSIST) {   // the most iterations (trips) of a wave
    unsigned int it = 0;
    HIPCHK(r, hipMemcpy(&it, r->stat_counts + 2, sizeof(it), hipMemcpyDeviceToHost));
    out->ntrips = it;
  }
  for (int kk = 0; kk < MI_NKERNEL_KINDS && kk < NKK; ++kk) {
    out->kernel_ms[kk] = r->prof.ms[kk];
    out->kernel_launches[kk] = r->prof.launches[kk];
    out->kernel_cells[kk] = r->prof.cells[kk];
  }
  out->schedule_used = r->last_schedule;
  out->graph_trips_used = r->last_graph;
  auto it = r->tuned.find(r->last_ncell);
  if (it != r->tuned.end())
    for (int v = 0; v < 4; ++v) out->tune_ms[v] = it->second.ms[v];
  return MI_OK;
}

int mi_chem_attach_predictor(void* reactor, void* mlp) {
  auto* r = static_cast<Reactor*>(reactor);
  if (!r) return MI_EINVAL;
  if (!mlp) {
    r->predictor = nullptr;
    return MI_OK;
  }
  const MlpDev* d = mlp_device(mlp);
  const int N = mech_nspec(r->mech);
  if (d->n_in != N + 3 || d->n_out != N + 1)
    return fail(r, MI_EINVAL, "network shape must be (nspec + 3) inputs and (nspec + 1) outputs");
  if (r->schedule != SCHED_COMPACT) return fail(r, MI_EINVAL, "the learned predictor needs mi.schedule = queue");
  r->predictor = mlp;
  return MI_OK;
}

int mi_chem_nspec(void* reactor) {
  auto* r = static_cast<Reactor*>(reactor);
  if (!r) return MI_EINVAL;
  return mech_nspec(r->mech);
}

const char* mi_chem_last_error(void* reactor) {
  auto* r = static_cast<Reactor*>(reactor);
  if (!r) return "null reactor";
  return r->err.c_str();
}

int mi_chem_close(void* reactor) {
  auto* r = static_cast<Reactor*>(reactor);
  if (!r) return MI_EINVAL;
  free_workspace(r);
  free_order(r);
  free_glue(r);
  if (r->counts) (void)hipFree(r->counts);
  if (r->counters) (void)hipFree(r->counters);
  if (r->ev0) (void)hipEventDestroy(r->ev0);
  if (r->ev1) (void)hipEventDestroy(r->ev1);
  if (r->capture_stream) (void)hipStreamDestroy(r->capture_stream);
  delete r;
  return MI_OK;
}

}  // extern "C"
