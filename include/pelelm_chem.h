/*
 * pelelm_chem -- MI355X (gfx950) per-cell stiff chemistry reactor, C ABI.
 *
 * Drop-in for the call PeleLM makes through pele::physics::reactions::ReactorBase
 * (Source/PeleLM.H:16,672-674): one reactor object per process, created from the
 * `peleLM.chem_integrator` name, initialised as an h-reactor, given typical values for
 * the absolute tolerances, asked to react boxes, closed at teardown.  Every entry point
 * takes plain pointers and sizes; device arrays are HIP device pointers and all work is
 * enqueued on the caller's HIP stream.  The host-driven schedules (plain, queue, auto)
 * read the work-list counts back once per trip, so react returns with the integration
 * finished (the reference's caller synchronises right after the call anyway,
 * Source/PeleLM.cpp:5773-5775); the persistent schedule enqueues one kernel and
 * synchronises only to read the failure count under mi.on_fail = abort (never under
 * stream capture).
 *
 * Units are CGS in and out, as the reference converts before and after the call
 * (Source/PeleLM.cpp:5737-5746, 5764-5771).
 *
 * Return codes: 0 success, < 0 error (mi_chem_last_error describes it).
 */
#ifndef PELELM_CHEM_H
#define PELELM_CHEM_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define MI_OK 0
#define MI_EINVAL -1      /* bad argument / unknown key or value            */
#define MI_ESTATE -2      /* call order (react before init, ...)            */
#define MI_EHIP -3        /* HIP runtime error                              */
#define MI_ECELLFAIL -4   /* at least one cell failed to integrate          */
#define MI_ENOMEM -5

#define MI_NKERNEL_KINDS 6

/* amrex::Array4<Real>-compatible view: element (i,j,k,n) lives at
   p[(i-begin[0]) + (j-begin[1])*jstride + (k-begin[2])*kstride + n*nstride]. */
typedef struct {
  double* p;
  int begin[3];
  int end[3];           /* exclusive, as amrex::Array4::end */
  long jstride, kstride, nstride;
  int ncomp;
} mi_real4;

typedef struct {
  int* p;
  int begin[3];
  int end[3];
  long jstride, kstride, nstride;
  int ncomp;
} mi_int4;

/* Per-call statistics (summed over cells of the last react call). */
typedef struct {
  long ncells;          /* cells integrated (mask >= 0)                         */
  long nsteps, nfe, nje, nsetups, nni, ncfn, netf;
  long nfail;           /* cells that did not reach dt                          */
  long max_steps_cell;  /* largest step count of a single cell                  */
  long nlu_refine;      /* iterative-refinement sweeps (mixed-precision LU)     */
  long nlu_fallback;    /* FP32 factorisations replaced by FP64 ones            */
  long npred_used;      /* Newton predictions taken from the learned predictor  */
  long npred_fallback;  /* steps where the learned prediction was rejected      */
  double ms_kernel;     /* device time of the integration kernel(s), ms         */
  long ntrips;          /* right-hand-side trips the call took                  */
  long nhy2t_iters;     /* Newton iterations of the T-from-enthalpy solves      */
  /* with mi.profile = 1: device time (ms), launches and cells processed per kernel
     kind [0] right-hand side, [1] Jacobian set-up, [2] Newton/step logic, [3] LU,
     [4] learned predictor, [5] Newton linear solves */
  double kernel_ms[MI_NKERNEL_KINDS];
  long kernel_launches[MI_NKERNEL_KINDS];
  long kernel_cells[MI_NKERNEL_KINDS];
  /* how the trips ran: schedule 0 plain / 1 queue / 3 persistent, graph_trips 0
     host-driven or K trips per captured graph; with mi.schedule = auto, for this box size:
     the chosen configuration's final-round time (ms), the whole search's time (ms), the
     candidates searched, and 1 when the choice came from mi.tune_cache (all 0 when not
     tuned) */
  long schedule_used;
  long graph_trips_used;
  double tune_ms[4];
  long nli;             /* GMRES iterations (cvode.solve_type = GMRES); each evaluates the
                           right-hand side once for its J v product, not counted in nfe */
} mi_chem_stats;

/* ReactorBase::create(name) -- Source/PeleLM.cpp:706.
   Accepted names: "ReactorCvode" / "ReactorBDF" (this reactor; dense analytic-Jacobian
   BDF on the GPU), "ReactorRK64" (explicit six-stage Runge-Kutta sub-steps for cells that
   are not stiff, the reactor Tools/CMake/BuildPeleLMExe.cmake:124-125 builds: fourth-order
   solution with a fifth-order error estimate in the rtol/atol norm, sub-steps within
   [dt / ode.rk64_nsubsteps_max, dt / ode.rk64_nsubsteps_min] from
   dt / ode.rk64_nsubsteps_guess (defaults 500, 5, 10); a cell that fails the error test at
   the smallest sub-step fails) and "ReactorNull" (pass-through: rY += dt rYsrc,
   rE += dt rEsrc). */
int mi_chem_create(const char* integrator, void** reactor);

/* ReactorBase::init(reactor_type, ncells) -- Source/PeleLM.cpp:707-709.
   reactor_type 2 = h-reactor (constant pressure, energy = rho h); 1 = e-reactor is
   rejected.  ncells is a hint for workspace sizing. */
int mi_chem_init(void* reactor, int reactor_type, int ncells);

/* Runtime options, the ParmParse keys the reference decks use
   (Exec/RegTests/FlameSheet/inputs.3d-regt:84-90) plus this reactor's own:
     ode.rtol, ode.atol, cvode.max_order, cvode.max_num_steps (default 20000),
     cvode.solve_type (denseAJ_direct: LU of I - gamma J with the analytic Jacobian, default
     | dense_direct: the same | GMRES: matrix-free scaled GMRES, J v by a difference quotient
     of the right-hand side, no set-up; the trip schedules only, mi.gmres_maxl Krylov
     vectors, default 5 as CVODE's, at most the number of species), peleLM.use_typ_vals_chem (0/1),
     mi.mechanism (drm19 | LiDryer | dodecane53 | dodecane56 | heptane88; above 63 species the
     factorisations take two waves per cell with one matrix row per thread and the solves
     one wave per cell with two rows per lane, in FP64, and the persistent schedule and the
     learned predictor are not available), mi.schedule (plain: fixed cell-to-lane mapping
     for the evaluation and step kernels, default | queue: finished cells leave the work
     lists | auto: the first call of each box size searches the plain and queue schedules
     (host-driven or in graphs of 8 trips, with cost ordering, plain also without) and the
     persistent one (claim chunks 1, 8, 32; 1 or 2 waves per SIMD) by successive halving on
     every 16th, every 4th and then every cell of a copy of the call's cells, and keeps the
     fastest; mi.tune_cache = <file> keeps the choice in a JSON file keyed by mechanism,
     GPU, tolerances and box size for later processes | persistent: one kernel per
     call, waves keep their cells through all trips and claim new ones from a device
     counter), mi.persist_waves (0: as many as the device holds resident), mi.persist_chunk
     (8: idle lanes of a wave that trigger a claim, 1..64), mi.persist_occupancy (1 | 2
     waves per SIMD), mi.order (rate: cells sorted by their initial rates' change over dt
     in tolerance units, so a wave's cells finish together, default | fc: sorted by the
     |FC| the FC array holds on entry, the previous step's counts when the caller passes
     them | none; plain, queue and auto schedules; results do not depend on it),
     mi.graph_trips (0: the host
     issues every trip, default | K even: K trips per captured HIP graph), mi.lu (fp64 |
     fp32_ir: FP32 factors with FP64-residual refinement), mi.ir_sweeps (3), mi.lu_form
     (coop: 32 lanes per cell, a whole wave above 32 unknowns, default | lane: one lane per
     cell | mfma: one wave per cell, blocked LU in LDS with the trailing updates on 16x16x4
     FP64 or FP32 matrix-core tiles; the trip schedules only, the persistent kernel factors
     cooperatively; above 64 unknowns it selects the LU staged in LDS), mi.ir_tol (1e-4: a refinement sweep whose correction has WRMS norm at most this,
     in error-weight units, ends the refinement; a solve that does not get there in
     mi.ir_sweeps sweeps makes its cell refactor in FP64), mi.max_trips (0: 3 x
     cvode.max_num_steps + 100; cells still integrating after this many trips fail),
     mi.on_fail (abort: react returns MI_ECELLFAIL when a cell failed, default | continue:
     react returns MI_OK, failed cells carry FC < 0), mi.profile (0 | 1: per-kernel event
     timing). */
int mi_chem_set(void* reactor, const char* key, const char* value);

/* ReactorBase::set_typ_vals_ode -- Source/PeleLM.cpp:1497-1516: n = nspec + 1 typical
   values; with peleLM.use_typ_vals_chem = 1 the absolute tolerance of component i
   becomes ode.atol * typ[i]. */
int mi_chem_set_typ_vals(void* reactor, const double* typ, int n);

/* ReactorBase::react(box, rY, rYsrc, T, rE, rEsrc, FC, mask, dt, time, stream) --
   Source/PeleLM.cpp:5752-5758.  rY holds nspec components, rYsrc nspec, T/rE/rEsrc/FC
   one each; cells of [lo, hi] (inclusive) with mask >= 0 are advanced from time to
   time + dt in place.  On return *time += dt.  FC receives the number of RHS
   evaluations per cell; a failed cell keeps its input state and gets FC = -(its RHS
   count).  Returns MI_ECELLFAIL when a cell failed under mi.on_fail = abort (the
   default), MI_OK under mi.on_fail = continue. */
int mi_chem_react_box(void* reactor, const int lo[3], const int hi[3], mi_real4 rY, mi_real4 rYsrc,
                      mi_real4 T, mi_real4 rE, mi_real4 rEsrc, mi_real4 FC, mi_int4 mask, double* dt,
                      double* time, void* hip_stream);

/* Flat structure-of-arrays form of the same call: rY[k*ncells + c] (cell fastest),
   rYsrc likewise (nullable = zero forcing), T/rE/FC[c], rEsrc[c] (nullable), mask[c]
   (nullable = all cells active). */
int mi_chem_react_flat(void* reactor, double* rY, const double* rYsrc, double* T, double* rE,
                       const double* rEsrc, double* FC, const int* mask, int ncells, double* dt, double* time,
                       void* hip_stream);

/* mi_chem_react_flat that allocates nothing.  No counterpart in the reference: it is for
   callers that cannot allocate device memory during the call (code under HIP graph
   capture, tight serving loops, harnesses that forbid hipMalloc).  The cells integrate in
   the persistent schedule (one kernel; lanes claim cells from a device counter), whose
   lane slots, queue counters and statistics counters live in a file-scope device array of
   256 MiB per GPU, held until the process ends: at most as many waves of 64 lane slots as
   fit there (drm19 about 250, the 53-species mechanism about 50), however many cells the
   call has.  A mechanism whose one wave of slots would not fit is rejected with
   MI_ENOMEM.  The reactor's settings (tolerances, mechanism, mi.lu, mi.persist_chunk,
   mi.on_fail) apply; mi.schedule and mi.order do not.  Every call on a GPU shares that
   GPU's array: calls that overlap on one GPU race, and mi_chem_get_stats reports the last
   call on the array.  mi_chem_init must have run on the device (it records the array's
   address and the kernel's occupancy). */
int mi_chem_react_flat_noalloc(void* reactor, double* rY, const double* rYsrc, double* T, double* rE,
                               const double* rEsrc, double* FC, const int* mask, int ncells, double* dt,
                               double* time, void* hip_stream);

/* Net production rates rho*wdot (CGS, g/cm^3/s) of flat cells from (rhoY, rhoH, T) --
   reactionRateRhoY (Source/PeleLM_K.H:57-98) in CGS units: T is re-solved from
   h = rhoH/rho by HY2T starting from T, which is read only; the re-solved temperatures
   are written to T_out when it is not NULL.  SoA flat layout (rY[k*ncells + c]);
   mask < 0 cells (nullable mask) get zero rates. */
int mi_chem_reaction_rates_flat(void* reactor, const double* rY, const double* rH, const double* T,
                                const int* mask, double* rYdot, double* T_out, int ncells, void* hip_stream);

/* PeleLM::compute_instantaneous_reaction_rates on one box (Source/PeleLM.cpp:2013-2073;
   the per-cell reactionRateRhoY of Source/PeleLM_K.H:57-98, called at
   Source/PeleLM.cpp:2046-2055).  MKS in and out, as PeleLM's state holds them: rho = the
   sum of rhoY (nspec components), Y = rhoY / rho, h = rhoH / rho x 1e4, T re-solved by
   HY2T from the T it holds (T is read only), and rhoYdot_k = W_k wdot_k(rho x 1e-3, T, Y)
   x 1e3 (kg/m^3/s).  Cells with mask < 0 (EB-covered; mask.p NULL = all active) get zero
   rates.  Asynchronous on hip_stream. */
int mi_chem_reaction_rates_box(void* reactor, const int lo[3], const int hi[3], mi_real4 rhoY, mi_real4 rhoH,
                               mi_real4 T, mi_int4 mask, mi_real4 rhoYdot, void* hip_stream);

/* PeleLM::advance_chemistry on one box -- its per-box part, Source/PeleLM.cpp:5626-5844
   without the parallel copies: rhoY x 1e-3 and rhoH x 10 (MKS -> CGS, with the forcing
   converted the same way, Source/PeleLM.cpp:5736-5746), react over dt from time 0
   (Source/PeleLM.cpp:5752-5758), rhoY x 1e3 and rhoH x 0.1 back (Source/PeleLM.cpp:
   5764-5771), and I_R = -(rhoY_old - rhoY_new) / dt - F_Y (Source/PeleLM.cpp:5798-5802).
   s_old / s_new: rhoY (nspec components), rhoH, T at components 0 .. nspec+1 of each view
   (the state from first_spec on); s_old is read only and must not share memory with
   s_new, whose rhoY, rhoH and T are written.  force: F_Y (nspec), F_h (MKS, read only);
   rhoYdot: nspec components written; FC: one component written (RHS evaluations).
   Asynchronous on hip_stream; the forcing's CGS copy lives in a reactor-owned buffer
   (grown on a larger box, kept between calls). */
int mi_chem_advance_box(void* reactor, const int lo[3], const int hi[3], mi_real4 s_old, mi_real4 s_new,
                        mi_real4 force, mi_real4 rhoYdot, mi_real4 FC, mi_int4 mask, double dt,
                        void* hip_stream);

/* PeleLM::RhoH_to_Temp on one box (getTfromHY, Source/PeleLM_K.H:511-530, called at
   Source/PeleLM.cpp:5176): Y = rhoY / rho, h = rhoH / rho x 1e4 (MKS to erg/g), and T
   re-solved in place by HY2T starting from the T it holds.  rho: 1 component, rhoY:
   nspec, rhoH: 1, T: 1 (read and written). */
int mi_chem_rhoh_to_temp_box(void* reactor, const int lo[3], const int hi[3], mi_real4 rho, mi_real4 rhoY,
                             mi_real4 rhoH, mi_real4 T, void* hip_stream);

/* Statistics of the last react call (synchronises the stream of that call); returns
   MI_OK whether or not cells failed (nfail counts them). */
int mi_chem_get_stats(void* reactor, mi_chem_stats* out);

/* Number of species of the selected mechanism, or < 0. */
int mi_chem_nspec(void* reactor);

const char* mi_chem_last_error(void* reactor);

/* ReactorBase::close() -- Source/PeleLM.cpp:930.  Frees the reactor and its workspace. */
int mi_chem_close(void* reactor);

#ifdef __cplusplus
}
#endif
#endif
