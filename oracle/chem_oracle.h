/*
 * CPU oracle for the per-cell stiff chemistry advance -- TEST INFRASTRUCTURE ONLY.
 *
 * A plain FP64 restatement of the algorithm behind PeleLM's
 * ReactorBase::react(box, rY, rYsrc, T, rE, rEsrc, FC, mask, dt, time)
 * (call site Source/PeleLM.cpp:5752-5758), used by tests/, by the smoke check and
 * by the CPU-baseline leg of bench.py -- never by the product path.
 *
 *   - species thermo (NASA-7), net production rates, third-body / Lindemann / Troe
 *     fall-off and equilibrium-constant reverse rates: Docs/source/manual/Model.rst:227-296;
 *   - the constant-pressure reactor right-hand side with T re-solved from
 *     (rho h)(t) = (rho h)_0 + F_h t on every call: Docs/source/manual/Model.rst:683-698,
 *     per-cell N+1 layout Source/PeleLM_K.H:14-52, HY2T use Source/PeleLM_K.H:89,511-530;
 *   - a variable-order, fixed-leading-coefficient BDF in Nordsieck form with Newton
 *     (maxcor 3), dense partial-pivot LU of I - gamma J, WRMS error control and
 *     CVODE's step/order heuristics (the solver the reference decks select with
 *     cvode.solve_type = denseAJ_direct, cvode.max_order = 4,
 *     Exec/RegTests/FlameSheet/inputs.3d-regt:84-90), or scaled GMRES with difference-
 *     quotient J v products (cvode.solve_type = GMRES, the GPU decks,
 *     Exec/RegTests/FlameSheet/inputs.3d-regt_GPU:89).  SUNDIALS is not vendored in the
 *     reference (Submodules/sundials is empty), so this restates CVODE's published
 *     algorithm (Hindmarsh et al., ACM TOMS 31(3), 2005; CVODE user guide, Mathematical
 *     Considerations).
 *
 * Parity pins (tests/): thermo continuity and heats of formation, PMF far-field
 * enthalpy balance and ideal-gas density, mass/element conservation of wdot,
 * detailed balance at HP equilibrium, complex-step derivatives for the analytic
 * Jacobian, mpmath golden vectors for wdot, closed-form linear ODEs for the BDF,
 * and SciPy Radau fixtures at rtol 1e-13 for reactor advances.
 *
 * Layout: every multi-cell array is structure-of-arrays, component-major with the
 * cell index fastest (x[k*ncell + c]), the flattened form of an amrex::Array4 box.
 */
#ifndef PELELM_CHEM_ORACLE_H
#define PELELM_CHEM_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct orc_mech orc_mech;

orc_mech* orc_load(const char* tbl_path);
void orc_free(orc_mech* m);
int orc_nspec(const orc_mech* m);
int orc_nreac(const orc_mech* m);
int orc_nelem(const orc_mech* m);
void orc_molwt(const orc_mech* m, double* W);
/* element composition, row-major [nspec][nelem] */
void orc_elem_comp(const orc_mech* m, int* ecomp);
double orc_gas_constant(const orc_mech* m);

/* per-species cp/R, h/RT, s/R at T */
void orc_thermo(const orc_mech* m, double T, double* cp_R, double* h_RT, double* s_R);
/* net molar production rates [mol/cm^3/s] from concentrations C [mol/cm^3] and T [K];
   qnet (nullable) receives the net rate of progress of every reaction */
void orc_wdot(const orc_mech* m, const double* C, double T, double* wdot, double* qnet);
/* forward and reverse rates of progress of every reaction [mol/cm^3/s] */
void orc_rop(const orc_mech* m, const double* C, double T, double* qf, double* qr);
/* T such that sum_k Y_k h_k(T) = h [erg/g]; Newton from Tguess.  *iters (nullable) */
double orc_hy2t(const orc_mech* m, double h, const double* Y, double Tguess, int* iters);

/* Reactor right-hand side of y = (rhoY_0..rhoY_{N-1}, T) at time t (CGS):
   T is re-solved from (rhoH0 + fH t)/rho with y[N] as initial guess. */
void orc_rhs(const orc_mech* m, double t, const double* y, const double* fY, double rhoH0, double fH,
             double* ydot);
/* Same right-hand side with T = y[N] taken as given (no HY2T). */
void orc_rhs_fixedT(const orc_mech* m, const double* y, const double* fY, double fH, double* ydot);
/* Analytic Jacobian d(orc_rhs_fixedT)/dy, row-major (N+1)x(N+1). */
void orc_jac(const orc_mech* m, const double* y, const double* fY, double fH, double* J);
/* Complex-step derivative of orc_rhs_fixedT (pins orc_jac), row-major. */
void orc_jac_cs(const orc_mech* m, const double* y, const double* fY, double fH, double* J);

typedef struct {
  double rtol;
  double atol;               /* scalar absolute tolerance ...                       */
  const double* atol_vec;    /* ... or per-component (N+1) values, nullable          */
  int max_order;             /* 1..5 */
  long max_steps;
  int solve_type;            /* 0 dense direct LU, 1 GMRES (matrix-free, spgmr.hpp) */
  int gmres_maxl;            /* Krylov dimension of the GMRES solves, 0: CVODE's default 5 */
} orc_opts;

typedef struct {
  long nst, nfe, nje, nsetups, nni, ncfn, netf;
  int status;                /* 0 ok, <0 failure code */
  int qlast;
  double hlast;
  long nli;                  /* GMRES iterations (solve_type 1) */
} orc_cellstats;

/* Integrate every cell with mask >= 0 from 0 to dt (h-reactor, constant p).
   Arrays are SoA component-major.  On return rY, T, rE hold the state at dt and
   FC the number of RHS evaluations; mask < 0 cells are left untouched (FC = 0).
   Returns the number of failed cells (0 on success). */
int orc_react(const orc_mech* m, int ncell, double* rY, const double* rYsrc, double* T, double* rE,
              const double* rEsrc, double* FC, const int* mask, double dt, const orc_opts* opts,
              int nthreads, orc_cellstats* stats);

/* The same advance (scalar tolerances, no mask) with the three-stage Radau IIA method of
   order 5 (radau5.cpp), an integrator independent of the BDF.  nsteps (nullable) gets
   each cell's accepted steps (-1 on failure).  Returns the number of failed cells. */
int orc_radau_react(const orc_mech* m, int ncell, double* rY, const double* rYsrc, double* T, double* rE,
                    const double* rEsrc, double dt, double rtol, double atol, int nthreads, long* nsteps);

/* The explicit reactor (ReactorRK64, rk64.cpp): six-stage Cash-Karp pair, fourth-order
   solution propagated, sub-steps in [dt/nsub_max, dt/nsub_min] starting at dt/nsub_guess;
   a cell that fails the error test at the smallest sub-step fails (FC = -evaluations).
   Returns the number of failed cells. */
int orc_rk64_react(const orc_mech* m, int ncell, double* rY, const double* rYsrc, double* T, double* rE,
                   const double* rEsrc, double* FC, const int* mask, double dt, double rtol, double atol,
                   int nsub_guess, int nsub_min, int nsub_max, int nthreads);
/* The same integrator on y' = A y; st[0..2] = steps, evaluations, rejections. */
int orc_rk64_linear(int n, const double* A, double* y, double tout, double rtol, double atol, int nsub_guess,
                    int nsub_min, int nsub_max, long* st);
/* c[6], A[6][5], b4[6], b5[6] of the pair. */
void orc_rk64_tableau(double* out30);

/* Radau IIA-5 coefficients as radau5.cpp derives them: c[3], A[3][3], e[3], g0. */
void orc_radau_coefficients(double* out16);

/* The GMRES core (spgmr.hpp) on a dense A: out[0] residual estimate, out[1] iterations. */
int orc_spgmr_dense(int n, const double* A, const double* b, const double* w, int maxl, double delta, double* x,
                    double* out);

/* Integrate the linear test system y' = A y (n x n, row-major) with the same BDF
   machinery (pins the integrator against exp(A t)).  Returns status. */
int orc_bdf_linear(int n, const double* A, double* y, double tout, const orc_opts* opts, orc_cellstats* st);

#ifdef __cplusplus
}
#endif
#endif
