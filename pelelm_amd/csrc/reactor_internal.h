// Internal interface between the C ABI (chem_capi.cpp) and the HIP kernels
// (reactor_bdf.hip).  Plain structs only: they are passed by value as kernel
// arguments.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

namespace pelelm {

constexpr int NMAX1 = 96;   // largest ODE size (species + 1) a kernel argument block carries

// One strided device array over a box: element (cell, n) with cell -> (i, j, k),
// i fastest, lives at p[(i + ox) + (j + oy) * js + (k + oz) * ks + n * ns].
struct ArrView {
  double* p;
  long js, ks, ns;
  int ox, oy, oz;
};
struct IntView {
  const int* p;
  long js, ks;
  int ox, oy, oz;
};

// SCHED_AUTO: the first call of each box size times the plain and compacted schedules,
// each host-driven and graph-captured, on a copy of its cells and keeps the fastest
// SCHED_PERSIST: one persistent kernel per call (run_react_persist)
enum Schedule : int { SCHED_PLAIN = 0, SCHED_COMPACT = 1, SCHED_AUTO = 2, SCHED_PERSIST = 3 };
// LU_FP64: FP64 factors of I - gamma J; LU_FP32_IR: FP32 factors with FP64-residual
// iterative refinement of every solve (a cell whose refinement fails falls back to FP64)
enum LuMode : int { LU_FP64 = 0, LU_FP32_IR = 1 };

// Per-call options shared by every schedule.
struct ReactArgs {
  ArrView rY, rYsrc, T, rE, rEsrc, FC;   // rYsrc.p / rEsrc.p may be null (zero forcing)
  IntView mask;                          // mask.p may be null (all active)
  int nx, ny;                            // box extents in i and j (cells are i-fastest)
  long ncell;
  double dt;
  double rtol;
  double atol[NMAX1];
  int qmax;
  long mxstep;
  int lu_mode;
  int ir_sweeps;
  double ir_tol;                         // a refinement sweep whose correction has WRMS norm (in
                                         // units of the error weights) <= ir_tol ends the refinement
  int lu_mfma;                           // 1: the factorisations on the matrix cores (one cell
                                         // per wave, blocked; the solves stay cooperative)
  int lu_lane;                           // 1: one lane factors a cell (the form that runs when
                                         // the kernels execute serially on a CPU); 0: 32 lanes
  int gmres;                             // 1: cvode.solve_type = GMRES (no matrix: the Newton
                                         // systems by scaled GMRES with difference-quotient J v)
  int gmres_maxl;                        // its Krylov dimension (1 .. N1 - 1)
  // per-cell workspace, tiled by 64 cells: element (slot s) of cell c at
  // ws[((c / 64) * nslot + s) * 64 + c % 64]
  double* ws;
  int* iws;
  // per-cell matrix region, cell-major: cell c's words at mat[c * mat_stride ...]
  double* mat;
  // trip pipeline: active-cell lists (current / next), set-up lists and their counts
  int compact;                           // 1: finished cells leave the lists
  int* list[2];
  int* slist[2];
  int* plist;                            // cells whose next attempt awaits the learned prediction
  int* nlist;                            // cells due a Newton solve this trip
  int use_pred;                          // 1: the learned predictor supplies each Newton attempt's
                                         // initial iterate (compacted schedule only)
  unsigned int* counts;                  // 8 counters, see run_react
  // optional cost ordering (mi.order): workspace slot s integrates the caller's cell
  // slot_cell[s]; null = slot s is cell s
  const int* slot_cell;
  // outputs
  unsigned long long* counters;          // per-call aggregate counters (see CounterIndex)
  int* cell_status;                      // optional per-cell status (nullable)
};

enum CounterIndex : int {
  CNT_CELLS = 0, CNT_NST, CNT_NFE, CNT_NJE, CNT_NSETUPS, CNT_NNI, CNT_NCFN, CNT_NETF, CNT_NFAIL,
  CNT_MAXSTEPS, CNT_IR, CNT_LUFALLBACK, CNT_HY2T, CNT_PRED, CNT_PREDFB, CNT_NLI, NCOUNTERS = 16
};

// Mechanism ids (order of the dispatch table in reactor_bdf.hip).
enum MechId : int {
  MECH_LIDRYER = 0, MECH_DRM19 = 1, MECH_DODECANE53 = 2, MECH_DODECANE56 = 3, MECH_HEPTANE88 = 4, NMECH = 5
};
// the largest ODE size whose linear algebra runs cooperatively (one matrix row per lane
// of a wave); larger mechanisms factor and solve one cell per lane (mi.lu_form = lane)
constexpr int COOP_MAX_N1 = 64;

int mech_nspec(int mech);
const char* mech_name(int mech);
int mech_lookup(const char* name);
long ws_slots(int mech);        // doubles per lane
long iws_slots(int mech);       // ints per lane
long mat_stride(int mech);      // doubles per cell in the matrix region

// Integrate every cell described by `a`: trips of bounded kernels on `stream`, driven
// from the host until no cell is active or max_trips trips have run (cells still
// active then are reported failed).  Synchronises the stream once per trip.
// Optional per-kernel timing: when `prof` is non-null every launch of the trip
// kernels is bracketed by events and the per-kind device time (ms), launch count and
// cells processed are accumulated into prof[kind].
enum KernelKind : int { KK_EVAL = 0, KK_SETUP, KK_CONSUME, KK_LU, KK_PRED, KK_SOLVE, NKK };
struct KernelProf {
  double ms[NKK];
  long launches[NKK];
  long cells[NKK];
};
struct MlpDev;
// How the trips are issued: graph_trips = 0 -> the host launches every kernel of every
// trip and reads the list counts after each; graph_trips = K (even) -> K trips are
// captured once per call into a HIP graph (on capture_stream, which must not be the
// null stream) and the graph is launched on `stream` until no cell is active, reading
// the counts once per K trips.  Per-kernel timing and the learned predictor need the
// host-driven form.
struct TripPlan {
  int graph_trips = 0;
  hipStream_t capture_stream = nullptr;
};
// pred: the learned predictor's device view, required when a.use_pred = 1
hipError_t run_react(int mech, const ReactArgs& a, hipStream_t stream, long max_trips, long* trips_out,
                     KernelProf* prof, const MlpDev* pred = nullptr, TripPlan plan = TripPlan{});

// The persistent schedule: one kernel whose waves keep cells in their lanes.  Lane
// slot s (wave s / 64, lane s % 64) owns workspace tile s and matrix region s, so the
// workspace holds waves x 64 cells however large the box; idle lanes claim the next
// cells (in `order` when given, through a queue of the cells with mask >= 0; else in cell
// order, masked cells settled as they are claimed) from a device counter whenever
// `chunk` lanes of a wave are idle, and every
// cell is carried through all its trips inside the kernel.  Nothing is read back by the
// host: the kernel ends when the queue is drained, each cell being bounded by max_trips.
struct PersistPlan {
  int waves = 0;          // workgroups of one wave (the slots are waves x 64)
  int chunk = 1;          // idle lanes of a wave that trigger a claim (1..64)
  int occupancy = 1;      // waves per SIMD the kernel is compiled for (1 or 2)
  const int* order = nullptr;   // optional: cells in the order they are handed out
};
// the persistent kernels' maximum resident waves on this device for an occupancy
int persist_capacity(int mech, int occupancy);
// a.ws / a.mat hold plan.waves x 64 slots; with plan.order, a.list[0] holds ncell
// entries (the queue); counts[0..3] are the queue length, its head, the largest
// iteration count of a wave and a spare
hipError_t run_react_persist(int mech, const ReactArgs& a, const PersistPlan& plan, long max_trips,
                             hipStream_t stream);
// cost-ordering keys of every caller cell (ORDER_FC: the |FC| the caller's array holds;
// ORDER_RATE: the fastest species' change over dt in tolerance units at the input state)
// and the identity as values, for the sort that gives ReactArgs::slot_cell
enum OrderMode : int { ORDER_NONE = 0, ORDER_RATE = 1, ORDER_FC = 2 };
hipError_t launch_cost_keys(int mech, const ReactArgs& a, int mode, unsigned int* keys, int* vals,
                            hipStream_t stream);
// stable ascending radix sort of (keys, vals) pairs with hipCUB (cost_order.hip); the
// buffers hold 2 n entries each (double buffering); *sorted_vals receives the half that
// holds the result
size_t cost_sort_temp_bytes(long n);
hipError_t cost_sort(void* temp, size_t temp_bytes, unsigned int* keys, int* vals, long n, hipStream_t stream,
                     const int** sorted_vals);
hipError_t launch_counters_reset(unsigned long long* counters, hipStream_t stream);
// reactionRateRhoY over a box or a flat array of cells (rates_kernel)
struct RatesArgs {
  ArrView rhoY, rhoH, T, rdot;
  IntView mask;            // mask.p may be null (every cell active)
  double* Tout = nullptr;  // optional: the re-solved T of each cell, flat
  int nx = 1, ny = 1;
  long ncell = 0;
  double hscale = 1.0, rscale = 1.0, oscale = 1.0;
};
hipError_t launch_rates(int mech, const RatesArgs& ra, hipStream_t stream);
hipError_t launch_null(int nspec, const ReactArgs& a, hipStream_t stream);
// ReactorRK64 over every cell of a call (explicit sub-steps; the workspace's FY and
// Nordsieck slots hold the forcing and the stage derivatives)
hipError_t launch_rk64(int mech, const ReactArgs& a, int nsub_guess, int nsub_min, int nsub_max,
                       hipStream_t stream);
// copy a call's rY [nspec][ncell], T and rE [ncell] into flat device arrays
hipError_t launch_rhoh_to_temp(int mech, const ArrView& rho, const ArrView& rhoY, const ArrView& rhoH,
                               const ArrView& T, int nx, int ny, long ncell, hipStream_t stream);

// One mechanism's entry table.  Each mechanism's kernels compile in a translation unit of
// their own (reactor_bdf.hip built with PELE_TU_MECH); the dispatch functions above go
// through these tables.
struct MechTable {
  int nspec;
  const char* name;
  long ws_slots, iws_slots, mat_stride;
  hipError_t (*run_react)(const ReactArgs&, hipStream_t, long, long*, KernelProf*, const MlpDev*, TripPlan);
  hipError_t (*run_react_persist)(const ReactArgs&, const PersistPlan&, long, hipStream_t);
  int (*persist_capacity)(int);
  hipError_t (*launch_rates)(const RatesArgs&, hipStream_t);
  hipError_t (*launch_rk64)(const ReactArgs&, int, int, int, hipStream_t);
  hipError_t (*launch_cost_keys)(const ReactArgs&, int, unsigned int*, int*, hipStream_t);
  hipError_t (*launch_rhoh_to_temp)(const ArrView&, const ArrView&, const ArrView&, const ArrView&, int, int, long,
                                    hipStream_t);
};
const MechTable* mech_table(int mech);
const MechTable* mech_table_LiDryer();
const MechTable* mech_table_drm19();
const MechTable* mech_table_dodecane53();
const MechTable* mech_table_dodecane56();
const MechTable* mech_table_heptane88();
hipError_t launch_glue_in(const ArrView& sold, const ArrView& snew, const ArrView& force, double* fcgs, int nspec,
                          int nx, int ny, long ncell, hipStream_t stream);
hipError_t launch_glue_out(const ArrView& sold, const ArrView& snew, const ArrView& force, const ArrView& rdot,
                           double dt_inv, int nspec, int nx, int ny, long ncell, hipStream_t stream);
// every stride-th cell of a call (m of them) as flat arrays, forcing included (zeros
// where the call has none)
hipError_t launch_gather_state(int nspec, const ReactArgs& a, long stride, long m, double* rY, double* T,
                               double* rE, double* fY, double* fE, hipStream_t stream);

}  // namespace pelelm
