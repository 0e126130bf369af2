// MI355X per-cell stiff chemistry advance: variable-order BDF with Newton and dense
// LU of I - gamma J, one cell per lane.
//
// The integrator is the algorithm the reference selects with
// cvode.solve_type = denseAJ_direct, cvode.max_order = 4 (Exec/RegTests/FlameSheet/
// inputs.3d-regt:84-90) behind ReactorBase::react (Source/PeleLM.cpp:5752-5758):
// Nordsieck history, fixed-leading-coefficient BDF coefficients, Newton with at most
// three corrections, Jacobian/iteration-matrix reuse, WRMS error test and CVODE's
// step/order selection; the right-hand side re-solves T from the enthalpy on every
// call (Source/PeleLM_K.H:89) and advances (rho h) linearly with its forcing.
//
// GPU structure.  Each lane runs its cell as a state machine whose every trip makes
// exactly one right-hand-side evaluation at a single call site, so lanes that are in
// different phases (initial-step search, Newton iteration k, order-1 restart) still
// share that evaluation instead of serialising several inlined copies of it.  The
// per-lane history (Nordsieck array, saved Jacobian, LU factors, weights) lives in a
// wave-tiled workspace -- element (slot, lane) of wave w at ws[(w*nslot+slot)*64+lane]
// -- so every access of one slot by a wave is one contiguous 512-byte line.
// Trips of bounded kernels (evaluation, set-up, LU, solve, step logic) issued by the
// host, with a fixed cell-to-lane mapping (plain; the cells optionally sorted by an
// estimate of their cost so that a wave's cells finish together) or compacted work lists
// (queue); or one persistent kernel whose waves keep their cells through every trip and
// claim new cells from a device counter.
#include <hip/hip_runtime.h>

#include <cfloat>
#include <cmath>
#include <cstddef>
#include <utility>
#ifdef PELE_CHECK_BOUNDS
#include <cstdio>
#include <cstdlib>
#endif

#include "reactor_internal.h"
#include "mlp_internal.h"

// minimum resident waves per SIMD the heavy trip kernels are compiled for (the
// register budget follows: 512 / n VGPRs per lane)
#ifndef PELE_CONSUME_WAVES
#define PELE_CONSUME_WAVES 3
#endif
// k_setup: two waves per SIMD up to 32 unknowns, one above (the 53-species Jacobian
// measured 7 % faster with the larger register file)
#ifndef PELE_SETUP_WAVES
#define PELE_SETUP_WAVES(N1) ((N1) <= 32 ? 2 : 1)
#endif
// k_eval: one wave per SIMD (256 VGPRs plus the accumulation registers, which take the
// spills) except for the smallest mechanisms, whose right-hand side fits in 128
// registers; measured on drm19 (spill-free at one wave: 76 -> 68 ms on 128^3) and the
// 53-species mechanism
#ifndef PELE_EVAL_WAVES
#define PELE_EVAL_WAVES(N1) ((N1) <= 16 ? 4 : 1)
#endif
#ifndef PELE_WAVES_PER_EU
#define PELE_WAVES_PER_EU(n) __attribute__((amdgpu_waves_per_eu(n)))
#endif

// Build modes.  By default this file is one translation unit holding every mechanism and
// the shared kernels.  The library builds it once per mechanism (PELE_TU_MECH set, with
// PELE_TU_MECH_INC / _TYPE / _FN naming the generated code, its type and its table
// function: the unit holds that mechanism's kernels and its entry table) and once with
// PELE_TU_COMMON (the shared kernels and the dispatch over the tables), so the
// mechanisms compile in parallel.
// A single unit can leave out the two largest mechanisms (PELE_BASE_MECHS_ONLY: their
// tables are then absent and mi.mechanism rejects them; PELE_WITH_HEPTANE88 keeps the
// 88-species one), which keeps test builds short.
#if defined(PELE_BASE_MECHS_ONLY)
#define PELE_HAVE_DODECANE56 0
#if defined(PELE_WITH_HEPTANE88)
#define PELE_HAVE_HEPTANE88 1
#else
#define PELE_HAVE_HEPTANE88 0
#endif
#else
#define PELE_HAVE_DODECANE56 1
#define PELE_HAVE_HEPTANE88 1
#endif
#if defined(PELE_TU_MECH)
#include PELE_TU_MECH_INC
#elif !defined(PELE_TU_COMMON)
#include "mech/LiDryer.inc"
#include "mech/drm19.inc"
#include "mech/dodecane53.inc"
#if PELE_HAVE_DODECANE56
#include "mech/dodecane56.inc"
#endif
#if PELE_HAVE_HEPTANE88
#include "mech/heptane88.inc"
#endif
#endif

namespace pelelm {
namespace {

constexpr double UROUND = DBL_EPSILON;
constexpr double ADDON = 1e-6, BIAS1 = 6.0, BIAS2 = 6.0, BIAS3 = 10.0;
constexpr double THRESH = 1.5, ETAMX1 = 1e4, ETAMX2 = 10.0, ETAMX3 = 10.0, ETAMXF = 0.2;
constexpr double ETAMIN = 0.1, ETACF = 0.25, CRDOWN = 0.3, RDIV = 2.0, DGMAX = 0.3, DGMAX_LS = 0.2;
constexpr double HLB_FACTOR = 100.0, HUB_FACTOR = 0.1, H_BIAS = 0.5, NLSCOEF = 0.1;
constexpr int SMALL_NST = 10, MXNCF = 10, MXNEF = 7, MXNEF1 = 3, SMALL_NEF = 2, LONG_WAIT = 10;
constexpr int MAXCOR = 3, MSBP = 20, MSBJ = 51, HIN_ITERS = 4;
constexpr int LMAX = 6;
constexpr int HY2T_MAXIT = 50;
// Newton convergence threshold factor for attempts that start from a learned iterate
constexpr double PRED_NLS_TIGHTEN = 0.01;
constexpr double HY2T_RTOL = 1e-10;

enum { FIRST_CALL, PREV_CONV_FAIL, PREV_ERR_FAIL };
enum { NO_FAILURES, FAIL_BAD_J, FAIL_OTHER };
enum Phase : int { PH_NEED_CELL, PH_INIT_F0, PH_HIN, PH_NEWTON, PH_RESTART_F };
enum Status : int { ST_OK = 0, ST_TOO_MUCH_WORK = -1, ST_TOO_MUCH_ACC = -2, ST_ERR_FAIL = -3,
                    ST_CONV_FAIL = -4, ST_TOO_CLOSE = -27, ST_NONFINITE = -50 };

// workspace slot map (doubles per lane)
template <int N1>
struct Slots {
  static constexpr int ZN = 0;                       // LMAX * N1
  static constexpr int ACOR = ZN + LMAX * N1;
  static constexpr int EWT = ACOR + N1;
  static constexpr int JS = EWT + N1;                 // Jacobian accumulation (copied to the matrix region)
  static constexpr int FY = JS + N1 * N1;             // species forcing (N1-1)
  static constexpr int TMP = FY + N1;                 // N1 scratch vector
  static constexpr int ATOL = TMP + N1;               // N1
  static constexpr int Y = ATOL + N1;                 // evaluation point of this trip
  static constexpr int F = Y + N1;                    // right-hand side at Y
  static constexpr int TS = F + N1;                   // temperature re-solved at Y
  static constexpr int SING = TS + 1;                 // 0, or 1 + first zero-pivot column of the last LU
  static constexpr int LANE = SING + 1;               // integrator state words
  static constexpr int LANE_WORDS = 96;
  static constexpr int TOTAL = LANE + LANE_WORDS;
  static constexpr int ITOTAL = 1;                    // int slots (none in use)
};

// Per-cell matrix region, cell-major (a cell's words are contiguous, so the 32 lanes
// that factor or solve one cell read and write whole lines): the Jacobian, the factors
// of M = I - gamma J (row i holds row PERM[i] of M; FP64, or FP32 packed), the Newton
// right-hand side B and solution X, refinement information and PERM.
template <int N1>
struct MatSlots {
  static constexpr int JS = 0;                        // Jacobian copy, row-major (FP64)
  static constexpr int LU = N1 * N1;
  static constexpr int B = LU + N1 * N1;
  static constexpr int X = B + N1;
  static constexpr int IRS = X + N1;                  // refinement sweeps of the last solve
  static constexpr int IRF = IRS + 1;                 // 1: refinement did not converge
  static constexpr int PERM = IRF + 1;                // N1 ints packed two per word
  static constexpr int R = PERM + (N1 + 1) / 2;       // refinement residual (one-lane solves)
  static constexpr int D = R + N1;                    // refinement correction (one-lane solves)
  static constexpr int USED = D + N1;
  static constexpr int STRIDE = (USED + 15) / 16 * 16; // whole 128-byte lines per cell
};

// PELE_CHECK_BOUNDS (host builds of this file only, never a device build): every slot
// index of a cell's workspace and matrix region, and every FP32 factor index, is
// checked against that cell's extent, and a violation aborts with the index.
#ifdef PELE_CHECK_BOUNDS
inline void pele_check(bool ok, const char* what, long i, long lim) {
  if (!ok) {
    std::fprintf(stderr, "bounds: %s index %ld outside [0, %ld)\n", what, i, lim);
    std::abort();
  }
}
#define PELE_CHECK(ok, what, i, lim) pele_check((ok), (what), (long)(i), (long)(lim))
#define PELE_LIM(x) , (x)
#define PELE_LIM_FIELD int lim = 1 << 30;
#else
#define PELE_CHECK(ok, what, i, lim) ((void)0)
#define PELE_LIM(x)
#define PELE_LIM_FIELD
#endif

struct WS {
  double* p;
  PELE_LIM_FIELD
  __device__ __forceinline__ double& operator[](int s) const {
    PELE_CHECK(s >= 0 && s < lim, "workspace slot", s, lim);
    return p[(long)s * 64];
  }
};
struct IWS {
  int* p;
  PELE_LIM_FIELD
  __device__ __forceinline__ int& operator[](int s) const {
    PELE_CHECK(s >= 0 && s < lim, "int workspace slot", s, lim);
    return p[(long)s * 64];
  }
};
// one cell's matrix region
struct CM {
  double* p;
  PELE_LIM_FIELD
  __device__ __forceinline__ double& operator[](int s) const {
    PELE_CHECK(s >= 0 && s < lim, "matrix-region word", s, lim);
    return p[s];
  }
  __device__ __forceinline__ int& perm(int base, int i) const {
    PELE_CHECK(base >= 0 && i >= 0 && 2L * base + i < 2L * lim, "packed int", 2L * base + i, 2L * lim);
    return reinterpret_cast<int*>(p + base)[i];
  }
};

__device__ __forceinline__ long cell_off(const ArrView& v, int i, int j, int k) {
  return (long)(i + v.ox) + (long)(j + v.oy) * v.js + (long)(k + v.oz) * v.ks;
}
__device__ __forceinline__ long cell_off(const IntView& v, int i, int j, int k) {
  return (long)(i + v.ox) + (long)(j + v.oy) * v.js + (long)(k + v.oz) * v.ks;
}

// ---------------------------------------------------------------------------
// right-hand side: T from (rho h)(t) by Newton, then rates and the energy equation
// ---------------------------------------------------------------------------
template <class M>
__device__ __forceinline__ double hy2t(const double h, const double* Y, double T, int* iters = nullptr) {
  double cpR[M::NS], hRT[M::NS];
  int it = 0;
  for (; it < HY2T_MAXIT; ++it) {
    M::thermo_cph(T, cpR, hRT);
    double hT = 0.0, cp = 0.0;
#pragma unroll
    for (int k = 0; k < M::NS; ++k) {
      hT += Y[k] * hRT[k] * M::imw(k);
      cp += Y[k] * cpR[k] * M::imw(k);
    }
    hT *= M::RU * T;
    cp *= M::RU;
    const double dT = (h - hT) / cp;
    T += dT;
    if (fabs(dT) <= HY2T_RTOL * T) { ++it; break; }
  }
  if (iters) *iters = it;
  return T;
}

template <class M>
__device__ __forceinline__ void rhs(const double* __restrict__ y, const double t, const WS& ws, const double rE0,
                                    const double fH, double* __restrict__ f, double& Tsol, int* hy2t_iters = nullptr) {
  constexpr int N = M::NS;
  using S = Slots<M::N1>;
  double rho = 0.0;
#pragma unroll
  for (int k = 0; k < N; ++k) rho += y[k];
  const double irho = 1.0 / rho;
  double Y[N];
#pragma unroll
  for (int k = 0; k < N; ++k) Y[k] = y[k] * irho;
  const double T = hy2t<M>((rE0 + fH * t) * irho, Y, y[N], hy2t_iters);
  Tsol = T;
  const double logT = log(T), invT = 1.0 / T;
  double C[N], g[N], wd[N];
#pragma unroll
  for (int k = 0; k < N; ++k) C[k] = y[k] * M::imw(k);
  M::gibbs(T, logT, g);
  M::wdot(C, T, logT, invT, g, wd);
  double cpR[N], hRT[N];
  M::thermo_cph(T, cpR, hRT);
  // energy closure from rho h = sum_k rhoY_k h_k(T): rho cp dT/dt = F_h - sum_k h_k d(rhoY_k)/dt
  // with d(rhoY_k)/dt = W_k wdot_k + F_Y,k (the species forcing carries its enthalpy)
  double rhocp = 0.0, hw = 0.0;
#pragma unroll
  for (int k = 0; k < N; ++k) {
    const double fy = ws[S::FY + k];
    f[k] = wd[k] * M::mw(k) + fy;
    rhocp += y[k] * cpR[k] * M::imw(k);
    hw += (wd[k] + fy * M::imw(k)) * hRT[k];
  }
  f[N] = (fH - hw * M::RU * T) / (rhocp * M::RU);
}

// analytic Jacobian of the right-hand side at (y, T) with T held fixed, into ws[JS].
// Every entry is written once (no accumulation in memory); the T row's sum over a
// column is formed as that column's entries are produced.
template <class M>
__device__ __noinline__ void jacobian(const double* __restrict__ y, const double T, const double fT, const WS& ws) {
  constexpr int N = M::NS, N1 = M::N1;
  using S = Slots<N1>;
  const double logT = log(T), invT = 1.0 / T;
  double C[N], g[N], hRT[N], cpR[N], wd[N], dwdT[N];
#pragma unroll
  for (int k = 0; k < N; ++k) C[k] = y[k] * M::imw(k);
  M::gibbs(T, logT, g);
  M::thermo_cph(T, cpR, hRT);
  double rhocp = 0.0;
#pragma unroll
  for (int k = 0; k < N; ++k) rhocp += y[k] * cpR[k] * M::imw(k);
  rhocp *= M::RU;
  // T row: fT = (F_h - R T sum_i hRT_i (wdot_i + F_Y,i / W_i)) / rhocp (fT is passed in);
  // d fT / d rhoY_k = (-R T sum_i hRT_i/W_i J_ik - fT cp_k) / rhocp, the forcing being constant
  double s = 0.0;
  M::jacobian(C, T, logT, invT, g, hRT, wd, dwdT,
              [&](int i, int k, double v) {
                ws[S::JS + i * N1 + k] = v;
                s += hRT[i] * M::imw(i) * v;
              },
              [&](int k) {
                ws[S::JS + N * N1 + k] = (-M::RU * T * s - fT * M::RU * cpR[k] * M::imw(k)) / rhocp;
                s = 0.0;
              });
  double dcp[N];
  M::thermo_dcp(T, dcp);
  double drhocp = 0.0, sTT = 0.0;
#pragma unroll
  for (int k = 0; k < N; ++k) {
    drhocp += y[k] * dcp[k] * M::imw(k);
    sTT += cpR[k] * (wd[k] + ws[S::FY + k] * M::imw(k)) + hRT[k] * T * dwdT[k];
    ws[S::JS + k * N1 + N] = M::mw(k) * dwdT[k];
  }
  drhocp *= M::RU;
  ws[S::JS + N * N1 + N] = (-M::RU * sTT - fT * drhocp) / rhocp;
}

// ---------------------------------------------------------------------------
// dense LU of M = I - gamma J (partial pivoting, whole-row interchanges) and solve
// ---------------------------------------------------------------------------
// Factors of I - gamma J in the matrix region: FP64 one per word, or FP32 contiguous
// (element e at float index e of the region).
template <int N1, class Real>
struct LUAcc;
template <int N1>
struct LUAcc<N1, double> {
  static __device__ __forceinline__ double& at(const CM& m, int e) { return m[MatSlots<N1>::LU + e]; }
};
template <int N1>
struct LUAcc<N1, float> {
  static __device__ __forceinline__ float& at(const CM& m, int e) {
    PELE_CHECK(e >= 0 && e < N1 * N1, "FP32 factor", e, N1 * N1);
    return reinterpret_cast<float*>(&m[MatSlots<N1>::LU])[e];
  }
};

// one lane factors its cell's M = I - gamma J (partial pivoting, arithmetic in Real);
// returns 0, or 1 + the first column without a nonzero pivot
template <int N1, class Real>
__device__ __noinline__ int lu_factor(const CM& mat, const double gamma) {
  using MS = MatSlots<N1>;
  using A = LUAcc<N1, Real>;
#pragma unroll 1
  for (int i = 0; i < N1; ++i) {
    mat.perm(MS::PERM, i) = i;
#pragma unroll
    for (int j = 0; j < N1; ++j) A::at(mat, i * N1 + j) = (Real)((i == j ? 1.0 : 0.0) - gamma * mat[MS::JS + i * N1 + j]);
  }
  int sing = 0;
#pragma unroll 1
  for (int k = 0; k < N1; ++k) {
    int p = k;
    Real amax = fabs(A::at(mat, k * N1 + k));
#pragma unroll 1
    for (int i = k + 1; i < N1; ++i) {
      const Real a = fabs(A::at(mat, i * N1 + k));
      if (a > amax) { amax = a; p = i; }
    }
    if (amax == (Real)0) { if (!sing) sing = k + 1; continue; }
    if (p != k) {
      const int t = mat.perm(MS::PERM, k);
      mat.perm(MS::PERM, k) = mat.perm(MS::PERM, p);
      mat.perm(MS::PERM, p) = t;
#pragma unroll 1
      for (int j = 0; j < N1; ++j) {
        const Real t2 = A::at(mat, k * N1 + j);
        A::at(mat, k * N1 + j) = A::at(mat, p * N1 + j);
        A::at(mat, p * N1 + j) = t2;
      }
    }
    const Real inv = (Real)1 / A::at(mat, k * N1 + k);
#pragma unroll 1
    for (int i = k + 1; i < N1; ++i) {
      const Real lik = A::at(mat, i * N1 + k) * inv;
      A::at(mat, i * N1 + k) = lik;
      if (lik != (Real)0) {
#pragma unroll 1
        for (int j = k + 1; j < N1; ++j) A::at(mat, i * N1 + j) -= lik * A::at(mat, k * N1 + j);
      }
    }
  }
  return sing;
}

// one lane solves M x = b with the factors of the matrix region, FP64 arithmetic on
// factors stored in Real.  P M = L U, row i of the factors holding row PERM[i] of M.
// b and x are vectors of the cell's matrix region (word offsets src, dst; distinct),
// so the solve holds no N1-long array in registers.
template <int N1, class Real = double>
__device__ __forceinline__ void lu_solve(const CM& mat, const int src, const int dst) {
  using MS = MatSlots<N1>;
  using A = LUAcc<N1, Real>;
#pragma unroll 1
  for (int i = 0; i < N1; ++i) mat[dst + i] = mat[src + mat.perm(MS::PERM, i)];
#pragma unroll 1
  for (int k = 0; k < N1; ++k) {
    const double xk = mat[dst + k];
#pragma unroll 1
    for (int i = k + 1; i < N1; ++i) mat[dst + i] -= (double)A::at(mat, i * N1 + k) * xk;
  }
#pragma unroll 1
  for (int k = N1 - 1; k >= 0; --k) {
    double s = mat[dst + k];
#pragma unroll 1
    for (int j = k + 1; j < N1; ++j) s -= (double)A::at(mat, k * N1 + j) * mat[dst + j];
    mat[dst + k] = s / (double)A::at(mat, k * N1 + k);
  }
}

template <class M>
struct Lane {
  static constexpr int N = M::NS, N1 = M::N1;
  using S = Slots<N1>;
  // -- persistent state: the words stored in the LANE slots between kernels (8-byte
  // members first, then 4-byte ones, then flags, so that no word is padding) --
  // history / step control
  double tn, h, hprime, hscale, eta, etamax, gamma, gammap, gamrat, rl1, crate, delp, saved_tq5;
  double tq[6];
  // the BDF coefficients l share their words with the initial step search, which
  // ends before the first attempt computes them (set_coeffs writes l before any read)
  union {
    double l[LMAX];
    struct {
      double hg, hub, hlb, hnew;
      int count1;
    };
  };
  double tau[LMAX + 1];
  // current attempt
  double saved_t;
  // evaluation request for the next trip: y = zn0 + a1 * zn1 + a2 * acor at t = teval
  double a1, a2, teval;
  // cell constants: (rho h)_0 and its forcing
  double rE0, fH;
  int q, qprime, L, qwait;
  int nst, nfe, nje, nsetups, nni, ncfn, netf, nstlp, nstlj, nstloc;
  int ncf, nef, nflag, m, convfail;
  int phase, status;
  // setup_pending = 1 when this trip's evaluation started an iteration-matrix set-up
  // whose factorisation is to be consumed
  int setup_pending;
  // linear algebra: lu64 = 1 once refinement of FP32 factors has failed for this cell
  // (its later factorisations are FP64)
  int lu64;
  int nir, nlufb;
  // learned predictor: pred_state 0 none, 1 requested for this attempt (y_n saved in
  // TMP), 2 applied (acor holds prediction - zn0), 3 declined
  int pred_state;
  int npred, npredfb;
  bool jcur, callSetup;
  // -- set by load_lane from the call and the cell's addresses, not stored --
  WS ws;
  IWS iws;
  CM mat;
  const double* xs;   // this trip's Newton solution (matrix region, or its LDS copy)
  int qmax;
  double rtol;
  long mxstep;
  double tout;
  int lu_mode, ir_sweeps, use_pred;
  int gmres;   // the Newton systems by GMRES: no set-up, no matrix
  // the converged correction's norm and the error test's measure, used in the trip that
  // computes them
  double acnrm, dsm;

  __device__ __forceinline__ double& zn(int j, int i) const { return ws[S::ZN + j * N1 + i]; }

  __device__ void ewtset() {
    for (int i = 0; i < N1; ++i) ws[S::EWT + i] = 1.0 / (rtol * fabs(zn(0, i)) + ws[S::ATOL + i]);
  }
  // ewtset() and then the WRMS norm of zn0 with the new weights, in one pass
  __device__ double ewtset_wrms_zn0() {
    double s = 0.0;
#pragma unroll 2
    for (int i = 0; i < N1; ++i) {
      const double z = zn(0, i);
      const double e = 1.0 / (rtol * fabs(z) + ws[S::ATOL + i]);
      ws[S::EWT + i] = e;
      const double p = z * e;
      s += p * p;
    }
    return sqrt(s / N1);
  }
  __device__ double wrms_ws(int slot) const {
    double s = 0.0;
#pragma unroll
    for (int i = 0; i < N1; ++i) {
      const double p = ws[slot + i] * ws[S::EWT + i];
      s += p * p;
    }
    return sqrt(s / N1);
  }
  __device__ double wrms_reg(const double* v) const {
    double s = 0.0;
#pragma unroll
    for (int i = 0; i < N1; ++i) {
      const double p = v[i] * ws[S::EWT + i];
      s += p * p;
    }
    return sqrt(s / N1);
  }

  // The Nordsieck updates below take each component's column zn(0..q, i) into
  // registers once, apply the same operations in the same order as the column-at-a-
  // time loops (so the results are identical), and store it once.
  //
  // start of an attempt: optionally rescale (zn(j) *= eta^j, h = hscale eta) and save
  // y_n = zn0 in TMP, then predict (tn += h; zn(j-1) += zn(j) for k = 1..q, j = q..k),
  // and clear acor
  __device__ void rescale_predict(const bool do_rescale, const bool save_y) {
    double fac[LMAX];
    double factor = eta;
    fac[0] = 1.0;
#pragma unroll
    for (int j = 1; j < LMAX; ++j) {
      fac[j] = factor;
      factor *= eta;
    }
    if (do_rescale) {
      h = hscale * eta;
      hscale = h;
    }
    tn += h;
#pragma unroll 2
    for (int i = 0; i < N1; ++i) {
      double z[LMAX];
#pragma unroll
      for (int j = 0; j < LMAX; ++j) z[j] = (j <= q) ? zn(j, i) : 0.0;
      if (save_y) ws[S::TMP + i] = z[0];
      if (do_rescale) {
#pragma unroll
        for (int j = 1; j < LMAX; ++j)
          if (j <= q) z[j] *= fac[j];
      }
#pragma unroll
      for (int k = 1; k < LMAX; ++k)
#pragma unroll
        for (int j = LMAX - 1; j >= k; --j)
          if (k <= q && j <= q) z[j - 1] += z[j];
#pragma unroll
      for (int j = 0; j < LMAX; ++j)
        if (j < q || (j == q && do_rescale)) zn(j, i) = z[j];
      ws[S::ACOR + i] = 0.0;
    }
  }
  // undo the prediction: zn(j-1) -= zn(j) for k = 1..q, j = q..k
  __device__ void restore() {
    tn = saved_t;
#pragma unroll 2
    for (int i = 0; i < N1; ++i) {
      double z[LMAX];
#pragma unroll
      for (int j = 0; j < LMAX; ++j) z[j] = (j <= q) ? zn(j, i) : 0.0;
#pragma unroll
      for (int k = 1; k < LMAX; ++k)
#pragma unroll
        for (int j = LMAX - 1; j >= k; --j)
          if (k <= q && j <= q) z[j - 1] -= z[j];
#pragma unroll
      for (int j = 0; j < LMAX - 1; ++j)
        if (j < q) zn(j, i) = z[j];
    }
  }
  __device__ void increase_bdf() {
    for (int i = 0; i <= qmax; ++i) l[i] = 0.0;
    l[2] = 1.0;
    double alpha1 = 1.0, prod = 1.0, xiold = 1.0, alpha0 = -1.0, hsum = hscale;
    if (q > 1) {
      for (int j = 1; j < q; ++j) {
        hsum += tau[j + 1];
        const double xi = hsum / hscale;
        prod *= xi;
        alpha0 -= 1.0 / (j + 1);
        alpha1 += 1.0 / xi;
        for (int i = j + 2; i >= 2; --i) l[i] = l[i] * xiold + l[i - 1];
        xiold = xi;
      }
    }
    const double A1 = (-alpha0 - alpha1) / prod;
#pragma unroll
    for (int i = 0; i < N1; ++i) zn(L, i) = A1 * zn(qmax, i);
    for (int j = 2; j <= q; ++j)
#pragma unroll
      for (int i = 0; i < N1; ++i) zn(j, i) += l[j] * zn(L, i);
  }
  __device__ void decrease_bdf() {
    for (int i = 0; i <= qmax; ++i) l[i] = 0.0;
    l[2] = 1.0;
    double hsum = 0.0;
    for (int j = 1; j <= q - 2; ++j) {
      hsum += tau[j];
      const double xi = hsum / hscale;
      for (int i = j + 2; i >= 2; --i) l[i] = l[i] * xi + l[i - 1];
    }
    for (int j = 2; j < q; ++j)
#pragma unroll
      for (int i = 0; i < N1; ++i) zn(j, i) -= l[j] * zn(q, i);
  }
  __device__ void adjust_order(int dq) {
    if (q == 2 && dq != 1) return;
    if (dq == 1) increase_bdf(); else decrease_bdf();
  }
  __device__ void set_coeffs() {
    l[0] = l[1] = 1.0;
    double xi_inv = 1.0, xistar_inv = 1.0;
    for (int i = 2; i <= q; ++i) l[i] = 0.0;
    double alpha0 = -1.0, alpha0_hat = -1.0, hsum = h;
    if (q > 1) {
      for (int j = 2; j < q; ++j) {
        hsum += tau[j - 1];
        xi_inv = h / hsum;
        alpha0 -= 1.0 / j;
        for (int i = j; i >= 1; --i) l[i] += l[i - 1] * xi_inv;
      }
      alpha0 -= 1.0 / q;
      xistar_inv = -l[1] - alpha0;
      hsum += tau[q - 1];
      xi_inv = h / hsum;
      alpha0_hat = -l[1] - xi_inv;
      for (int i = q; i >= 1; --i) l[i] += l[i - 1] * xistar_inv;
    }
    const double A1 = 1.0 - alpha0_hat + alpha0;
    const double A2 = 1.0 + q * A1;
    tq[2] = fabs(A1 / (alpha0 * A2));
    tq[5] = fabs(A2 * xistar_inv / (l[q] * xi_inv));
    if (qwait == 1) {
      if (q > 1) {
        const double C = xistar_inv / l[q];
        const double A3 = alpha0 + 1.0 / q;
        const double A4 = alpha0_hat + xi_inv;
        const double Cpinv = (1.0 - A4 + A3) / A3;
        tq[1] = fabs(C * Cpinv);
      } else {
        tq[1] = 1.0;
      }
      hsum += tau[q];
      xi_inv = h / hsum;
      const double A5 = alpha0 - 1.0 / (q + 1);
      const double A6 = alpha0_hat - xi_inv;
      const double Cppinv = (1.0 - A6 + A5) / A2;
      tq[3] = fabs(Cppinv / (xi_inv * (q + 2) * A5));
    }
    tq[4] = NLSCOEF / tq[2];
    rl1 = 1.0 / l[1];
    gamma = h * rl1;
    if (nst == 0) gammap = gamma;
    gamrat = (nst > 0) ? gamma / gammap : 1.0;
  }

  // start one attempt of the current step: predict, coefficients, Newton set-up
  // (do_rescale: the step size changes by eta first)
  __device__ void attempt(const bool do_rescale) {
    // with the predictor, the state at tn (before the Nordsieck prediction) is the
    // network's input
    pred_state = use_pred ? 1 : 0;
    rescale_predict(do_rescale, use_pred != 0);
    set_coeffs();
    convfail = (nflag == FIRST_CALL || nflag == PREV_ERR_FAIL) ? NO_FAILURES : FAIL_OTHER;
    callSetup = (nflag == PREV_CONV_FAIL) || (nflag == PREV_ERR_FAIL) || (nst == 0) ||
                (nst >= nstlp + MSBP) || (fabs(gamrat - 1.0) > DGMAX);
    if (gmres) {   // no set-up function: nothing to refresh, the rate estimate restarts
      callSetup = false;
      crate = 1.0;
    }
    m = 0;
    a1 = 0.0; a2 = 0.0; teval = tn;
    phase = PH_NEWTON;
  }
  __device__ void finish(int st) {
    status = st;
    phase = PH_NEED_CELL;
  }
  __device__ void loop_top() {
    const double zn0_norm = (nst > 0) ? ewtset_wrms_zn0() : wrms_ws(S::ZN);
    if (nstloc >= mxstep) { finish(ST_TOO_MUCH_WORK); return; }
    const double tolsf = UROUND * zn0_norm;
    if (tolsf > 1.0) { finish(ST_TOO_MUCH_ACC); return; }
    saved_t = tn;
    ncf = nef = 0;
    nflag = FIRST_CALL;
    const bool new_h = nst > 0 && hprime != h;
    if (new_h && qprime != q) {
      adjust_order(qprime - q);
      q = qprime;
      L = q + 1;
      qwait = L;
    }
    attempt(new_h);
  }
  __device__ void begin_integration() {
    hscale = h;
    hprime = h;
#pragma unroll
    for (int i = 0; i < N1; ++i) zn(1, i) *= h;
    nstloc = 0;
    loop_top();
  }
  __device__ void hin_done() {
    double h0 = H_BIAS * hnew;
    if (h0 < hlb) h0 = hlb;
    if (h0 > hub) h0 = hub;
    h = h0;
    begin_integration();
  }

  // the first right-hand side, f(0, y0), has been evaluated into zn1
  __device__ void after_f0() {
    const double tdist = tout;
    const double tround = UROUND * tout;
    if (tdist < 2.0 * tround) { finish(ST_TOO_CLOSE); return; }
    hlb = HLB_FACTOR * tround;
    double hub_inv = 0.0;
#pragma unroll
    for (int i = 0; i < N1; ++i) {
      const double t1 = HUB_FACTOR * fabs(zn(0, i)) + ws[S::ATOL + i];
      hub_inv = fmax(hub_inv, fabs(zn(1, i)) / t1);
    }
    hub = HUB_FACTOR * tdist;
    if (hub * hub_inv > 1.0) hub = 1.0 / hub_inv;
    hg = sqrt(hlb * hub);
    if (hub < hlb) {
      h = hg;
      begin_integration();
      return;
    }
    count1 = 1;
    phase = PH_HIN;
    a1 = hg; a2 = 0.0; teval = hg;
  }
  __device__ void after_hin(const double* f) {
    double s = 0.0;
#pragma unroll
    for (int i = 0; i < N1; ++i) {
      const double d = (f[i] - zn(1, i)) / hg * ws[S::EWT + i];
      s += d * d;
    }
    const double yddnrm = sqrt(s / N1);
    hnew = (yddnrm * hub * hub > 2.0) ? sqrt(2.0 / yddnrm) : sqrt(hg * hub);
    if (count1 == HIN_ITERS) { hin_done(); return; }
    const double hrat = hnew / hg;
    if (hrat > 0.5 && hrat < 2.0) { hin_done(); return; }
    if (count1 > 1 && hrat > 2.0) { hnew = hg; hin_done(); return; }
    hg = hnew;
    ++count1;
    a1 = hg; a2 = 0.0; teval = hg;
  }

  __device__ void conv_failure() {
    ++ncfn;
    restore();
    ++ncf;
    etamax = 1.0;
    if (ncf == MXNCF) { finish(ST_CONV_FAIL); return; }
    eta = ETACF;
    nflag = PREV_CONV_FAIL;
    attempt(true);
  }
  __device__ void nls_fail() {
    if (pred_state == 2) {
      // the learned initial iterate did not converge: retry the same step from the
      // Nordsieck prediction (the matrix set up in this attempt stays valid)
      ++npredfb;
      pred_state = 3;
      callSetup = false;
      m = 0;
#pragma unroll
      for (int i = 0; i < N1; ++i) ws[S::ACOR + i] = 0.0;
      a1 = 0.0; a2 = 0.0; teval = tn;
      return;
    }
    if (!jcur && !gmres) {
      callSetup = true;
      convfail = FAIL_BAD_J;
      m = 0;
#pragma unroll
      for (int i = 0; i < N1; ++i) ws[S::ACOR + i] = 0.0;
      a1 = 0.0; a2 = 0.0; teval = tn;
      return;
    }
    conv_failure();
  }

  __device__ void set_eta() {
    if (eta < THRESH) {
      eta = 1.0;
      hprime = h;
    } else {
      eta = fmin(eta, etamax);
      hprime = h * eta;
    }
  }
  __device__ void prepare_next_step() {
    if (etamax == 1.0) {
      qwait = max(qwait, 2);
      qprime = q;
      hprime = h;
      eta = 1.0;
      return;
    }
    const double etaq = 1.0 / (pow(BIAS2 * dsm, 1.0 / L) + ADDON);
    if (qwait != 0) {
      eta = etaq;
      qprime = q;
      set_eta();
      return;
    }
    qwait = 2;
    double etaqm1 = 0.0, etaqp1 = 0.0;
    if (q > 1) {
      const double ddn = wrms_ws(S::ZN + q * N1) * tq[1];
      etaqm1 = 1.0 / (pow(BIAS1 * ddn, 1.0 / q) + ADDON);
    }
    if (q != qmax && saved_tq5 != 0.0) {
      const double cquot = (tq[5] / saved_tq5) * pow(h / tau[2], (double)L);
      double s = 0.0;
#pragma unroll
      for (int i = 0; i < N1; ++i) {
        const double d = (ws[S::ACOR + i] - cquot * zn(qmax, i)) * ws[S::EWT + i];
        s += d * d;
      }
      const double dup = sqrt(s / N1) * tq[3];
      etaqp1 = 1.0 / (pow(BIAS3 * dup, 1.0 / (L + 1)) + ADDON);
    }
    const double etam = fmax(etaqm1, fmax(etaq, etaqp1));
    if (etam < THRESH) {
      eta = 1.0;
      qprime = q;
    } else if (etam == etaq) {
      eta = etaq;
      qprime = q;
    } else if (etam == etaqm1) {
      eta = etaqm1;
      qprime = q - 1;
    } else {
      eta = etaqp1;
      qprime = q + 1;
#pragma unroll
      for (int i = 0; i < N1; ++i) zn(qmax, i) = ws[S::ACOR + i];
    }
    set_eta();
  }

  // Newton converged: error test, then complete the step or retry
  __device__ bool after_converged() {   // returns true when tout has been reached
    dsm = acnrm * tq[2];
    if (!(dsm <= 1.0)) {   // a NaN norm fails the test, as in CVODE
      ++nef;
      ++netf;
      nflag = PREV_ERR_FAIL;
      restore();
      if (nef == MXNEF) { finish(ST_ERR_FAIL); return false; }
      etamax = 1.0;
      if (nef <= MXNEF1) {
        eta = 1.0 / (pow(BIAS2 * dsm, 1.0 / L) + ADDON);
        eta = fmax(ETAMIN, eta);
        if (nef >= SMALL_NEF) eta = fmin(eta, ETAMXF);
        attempt(true);
        return false;
      }
      if (q > 1) {
        eta = ETAMIN;
        adjust_order(-1);
        L = q;
        --q;
        qwait = L;
        attempt(true);
        return false;
      }
      eta = ETAMIN;
      h *= eta;
      hscale = h;
      qwait = LONG_WAIT;
      phase = PH_RESTART_F;
      a1 = 0.0; a2 = 0.0; teval = tn;
      return false;
    }
    // complete the step
    ++nst;
    for (int i = q; i >= 2; --i) tau[i] = tau[i - 1];
    if (q == 1 && nst > 1) tau[2] = tau[1];
    tau[1] = h;
#pragma unroll 2
    for (int i = 0; i < N1; ++i) {
      const double ac = ws[S::ACOR + i];
#pragma unroll
      for (int j = 0; j < LMAX; ++j)
        if (j <= q) zn(j, i) += l[j] * ac;
    }
    --qwait;
    if (qwait == 1 && q != qmax) {
#pragma unroll
      for (int i = 0; i < N1; ++i) zn(qmax, i) = ws[S::ACOR + i];
      saved_tq5 = tq[5];
    }
    prepare_next_step();
    etamax = (nst <= SMALL_NST) ? ETAMX2 : ETAMX3;
    ++nstloc;
    if ((tn - tout) * h >= 0.0) return true;
    loop_top();
    return false;
  }

  __device__ bool needs_setup() const { return phase == PH_NEWTON && m == 0 && callSetup; }

  // iteration-matrix set-up at the first evaluation of a Newton attempt: the Jacobian
  // when stale (or after a convergence failure) and the set-up bookkeeping; the LU of
  // I - gamma J follows in its own kernel
  __device__ void setup(const double* y, const double* f, const double Tsol) {
    const double dgamma = fabs(gamma / gammap - 1.0);
    const bool jbad = (nst == 0) || (nst >= nstlj + MSBJ) ||
                      (convfail == FAIL_BAD_J && dgamma < DGMAX_LS) || (convfail == FAIL_OTHER);
    if (jbad) {
      ++nje;
      nstlj = nst;
      jcur = true;
      jacobian<M>(y, Tsol, f[N], ws);
    } else {
      jcur = false;
    }
    ++nsetups;
    gammap = gamma;          // gamrat = 1 and crate = 1 once the factorisation is in place
    nstlp = nst;
    setup_pending = 1;
  }
  // after the factorisation: false when it was singular (the attempt is abandoned)
  __device__ bool after_setup(int sing) {
    setup_pending = 0;
    gamrat = 1.0;
    crate = 1.0;
    if (sing) {
      conv_failure();
      return false;
    }
    return true;
  }

  // one Newton iteration: x solves M x = gamma f(tn, zn0 + acor) - rl1 zn1 - acor (the
  // right-hand side is formed by the evaluation kernel, the solve by the solve kernel,
  // with FP64-residual refinement of FP32 factors); returns true when tout is reached
  __device__ bool iterate() {
    using MS = MatSlots<N1>;
    ++nni;
    double x[N1];
#pragma unroll
    for (int i = 0; i < N1; ++i) x[i] = xs[i];
    if (gmres) {
      // the linear solve failed (no residual reduction, or a reduction without
      // convergence after the first Newton iteration): the Newton iteration fails
      if (mat[MS::IRF] != 0.0) {
        nls_fail();
        return false;
      }
    } else if (lu_mode == LU_FP32_IR && !lu64) {
      nir += (long)mat[MS::IRS];
      if (mat[MS::IRF] != 0.0) {
        // the FP32 factors cannot deliver this solve: refactor in FP64 and restart the
        // Newton iteration from the prediction (the Jacobian is kept unless stale)
        ++nlufb;
        lu64 = 1;
        callSetup = true;
        convfail = NO_FAILURES;
        m = 0;
#pragma unroll
        for (int i = 0; i < N1; ++i) ws[S::ACOR + i] = 0.0;
        a1 = 0.0; a2 = 0.0; teval = tn;
        return false;
      }
    }
    if (!gmres && gamrat != 1.0) {   // matrix-free solves use the current gamma
      const double sc = 2.0 / (1.0 + gamrat);
#pragma unroll
      for (int i = 0; i < N1; ++i) x[i] *= sc;
    }
#pragma unroll
    for (int i = 0; i < N1; ++i) ws[S::ACOR + i] += x[i];
    const double del = wrms_reg(x);
    if (m > 0) crate = fmax(CRDOWN * crate, del / delp);
    // A learned iterate starts farther from the BDF solution than the Nordsieck
    // prediction, so the iteration stops closer to the convergence threshold: its test
    // is PRED_NLS_TIGHTEN times stricter, and its first correction (no convergence-rate
    // history) is never taken as converged
    const double dcon = del * fmin(1.0, crate) / (pred_state == 2 ? PRED_NLS_TIGHTEN * tq[4] : tq[4]);
    if (dcon <= 1.0 && !(m == 0 && pred_state == 2)) {
      // with a learned initial iterate acor is not the first correction alone
      acnrm = (m == 0 && pred_state != 2) ? del : wrms_ws(S::ACOR);
      jcur = false;
      return after_converged();
    }
    if (m >= 1 && del > RDIV * delp) {
      nls_fail();
      return false;
    }
    delp = del;
    ++m;
    if (m >= MAXCOR) {
      nls_fail();
      return false;
    }
    a1 = 0.0; a2 = 1.0; teval = tn;
    return false;
  }

  __device__ void restart_f(const double* f) {
#pragma unroll
    for (int i = 0; i < N1; ++i) zn(1, i) = h * f[i];
    attempt(false);
  }

  __device__ void init_cell(const double* y0) {
    tn = 0.0;
    q = 1; L = 2; qwait = L; qprime = 1;
    etamax = ETAMX1;
    eta = 1.0;
    nst = nfe = nje = nsetups = nni = ncfn = netf = 0;
    nstlp = nstlj = 0;
    nstloc = 0;
    saved_tq5 = 0.0;
    crate = 1.0;
    delp = 0.0;
    gamma = gammap = 0.0;
    gamrat = 1.0;
    jcur = false;
    h = hprime = hscale = 0.0;
    for (int i = 0; i < 6; ++i) tq[i] = 0.0;
    for (int i = 0; i < LMAX; ++i) l[i] = 0.0;
    for (int i = 0; i <= LMAX; ++i) tau[i] = 0.0;
    for (int j = 0; j < LMAX; ++j)
#pragma unroll
      for (int i = 0; i < N1; ++i) zn(j, i) = (j == 0) ? y0[i] : 0.0;
    // the first evaluations (f0 and the initial-step search) form zn0 + a1 zn1 + a2 acor
    // with a2 = 0: acor must hold numbers, since 0 x NaN is NaN in whatever the
    // workspace held before
#pragma unroll
    for (int i = 0; i < N1; ++i) ws[S::ACOR + i] = 0.0;
    ewtset();
    status = ST_OK;
    phase = PH_INIT_F0;
    setup_pending = 0;
    lu64 = 0;
    nir = 0;
    nlufb = 0;
    pred_state = 0;
    npred = 0;
    npredfb = 0;
    a1 = 0.0; a2 = 0.0; teval = 0.0;
  }
};

template <class M>
__device__ __forceinline__ WS cell_ws(double* ws, long c) {
  return WS{ws + (c >> 6) * (long)Slots<M::N1>::TOTAL * 64 + (c & 63) PELE_LIM(Slots<M::N1>::TOTAL)};
}
template <class M>
__device__ __forceinline__ IWS cell_iws(int* iws, long c) {
  return IWS{iws + (c >> 6) * (long)Slots<M::N1>::ITOTAL * 64 + (c & 63) PELE_LIM(Slots<M::N1>::ITOTAL)};
}

template <class M>
__device__ __forceinline__ CM cell_mat(double* mat, long c) {
  return CM{mat + c * (long)MatSlots<M::N1>::STRIDE PELE_LIM(MatSlots<M::N1>::STRIDE)};
}

// The lane state persists between the kernels of a trip as words of the workspace (the
// Lane<M> object copied word by word).  The copy lives in scratch memory while a kernel
// runs: on the step kernel that measured 1.4x faster than holding the whole state in
// registers, which costs occupancy.
// words of Lane<M> kept between kernels: the persistent members, which precede `ws`
template <class M>
__host__ __device__ constexpr int lane_words() {
  static_assert(offsetof(Lane<M>, ws) % 8 == 0, "the persistent lane state is whole words");
  static_assert(offsetof(Lane<M>, ws) / 8 <= Slots<M::N1>::LANE_WORDS, "lane state does not fit its slots");
  return (int)(offsetof(Lane<M>, ws) / 8);
}

template <class M>
__device__ __forceinline__ void load_lane(Lane<M>& st, const ReactArgs& a, long c) {
  using S = Slots<M::N1>;
  constexpr int NW = lane_words<M>();
  const WS w = cell_ws<M>(a.ws, c);
  double tmp[NW];
#pragma unroll
  for (int i = 0; i < NW; ++i) tmp[i] = w[S::LANE + i];
  __builtin_memcpy(&st, tmp, NW * 8);
  st.ws = w;
  st.iws = cell_iws<M>(a.iws, c);
  st.mat = cell_mat<M>(a.mat, c);
  st.xs = &st.mat[MatSlots<M::N1>::X];
  st.qmax = a.qmax;
  st.lu_mode = a.lu_mode;
  st.ir_sweeps = a.ir_sweeps;
  st.use_pred = a.use_pred;
  st.gmres = a.gmres;
  st.rtol = a.rtol;
  st.mxstep = a.mxstep;
  st.tout = a.dt;
}
template <class M>
__device__ __forceinline__ void store_lane(const Lane<M>& st, const WS& w) {
  using S = Slots<M::N1>;
  constexpr int NW = lane_words<M>();
  double tmp[NW];
  __builtin_memcpy(tmp, &st, NW * 8);
#pragma unroll
  for (int i = 0; i < NW; ++i) w[S::LANE + i] = tmp[i];
}

// a field of the lane state (at byte offset `off` of Lane<M>), read straight from its
// workspace word without loading the whole state
template <class M, class T>
__device__ __forceinline__ T lane_field(const WS& w, size_t off) {
  static_assert(sizeof(T) <= 8, "one word");
  using S = Slots<M::N1>;
  const double word = w[S::LANE + (int)(off / 8)];
  T v;
  __builtin_memcpy(&v, reinterpret_cast<const char*>(&word) + off % 8, sizeof(T));
  return v;
}
#define LANE_FIELD(M, w, f) lane_field<M, decltype(Lane<M>::f)>((w), offsetof(Lane<M>, f))

// Copies between 64 lanes' rows in LDS (sm[l * LEN + i], lane l's cell cells[l], or -1)
// and the cell-major region: consecutive lanes of a store or load handle consecutive
// words of one cell's row, so each instruction touches a few whole lines instead of
// one line per lane.  Called by all 64 threads of a one-wave block.
template <int LEN>
__device__ __forceinline__ void rows_to_cells(const double* sm, const long* cells, double* mat, long stride, int off) {
  for (int e = threadIdx.x; e < 64 * LEN; e += 64) {
    const int l = e / LEN, i = e - l * LEN;
    const long c = cells[l];
    if (c >= 0) mat[c * stride + off + i] = sm[l * LEN + i];
  }
}
template <int LEN>
__device__ __forceinline__ void cells_to_rows(double* sm, const long* cells, const double* mat, long stride, int off) {
  for (int e = threadIdx.x; e < 64 * LEN; e += 64) {
    const int l = e / LEN, i = e - l * LEN;
    const long c = cells[l];
    sm[l * LEN + i] = (c >= 0) ? mat[c * stride + off + i] : 0.0;
  }
}

// ---------------------------------------------------------------------------
// the integration kernel
// ---------------------------------------------------------------------------
__device__ __forceinline__ void wave_add(unsigned long long* dst, unsigned long long v) {
  // wave-level sum, one atomic per wave
  for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
  if (__lane_id() == 0 && v) atomicAdd(dst, v);
}
__device__ __forceinline__ void wave_max(unsigned long long* dst, unsigned long long v) {
  for (int off = 32; off > 0; off >>= 1) {
    const unsigned long long o = __shfl_down(v, off, 64);
    v = o > v ? o : v;
  }
  if (__lane_id() == 0 && v) atomicMax(dst, v);
}

// ---------------------------------------------------------------------------
// the integration as trips of bounded kernels
//
// A trip evaluates one right-hand side for every active cell (k_eval), sets up the
// iteration matrix of the cells whose Newton attempt starts now (k_setup, over its
// own compacted list), and advances every cell's CVODE state machine by that
// evaluation (k_consume).  With compaction, k_consume appends the cells that are
// still integrating to the next trip's list, so finished cells leave the pipeline
// and no lane waits for a stiffer neighbour.
// ---------------------------------------------------------------------------
__device__ __forceinline__ void wave_append(bool pred, int value, int* list, unsigned int* count) {
  const unsigned long long b = __ballot(pred);
  if (b == 0ull) return;
  const int lane = __lane_id();
  const int leader = __ffsll((long long)b) - 1;
  unsigned int base = 0;
  if (lane == leader) base = atomicAdd(count, (unsigned int)__popcll(b));
  base = __shfl(base, leader, 64);
  if (pred) list[base + __popcll(b & ((1ull << lane) - 1ull))] = value;
}

__device__ __forceinline__ void cell_ijk(const ReactArgs& a, long c, int& i, int& j, int& k) {
  i = (int)(c % a.nx);
  j = (int)((c / a.nx) % a.ny);
  k = (int)(c / ((long)a.nx * a.ny));
}
// the caller's cell that workspace slot s integrates (cost ordering), and its (i, j, k)
__device__ __forceinline__ long slot_cell(const ReactArgs& a, long s) { return a.slot_cell ? a.slot_cell[s] : s; }
__device__ __forceinline__ void slot_ijk(const ReactArgs& a, long s, int& i, int& j, int& k) {
  cell_ijk(a, slot_cell(a, s), i, j, k);
}

template <class M>
__global__ void __launch_bounds__(64) k_init(const ReactArgs a) {
  constexpr int N = M::NS, N1 = M::N1;
  using S = Slots<N1>;
  const long c = (long)blockIdx.x * 64 + threadIdx.x;
  bool active = false;
  if (c < a.ncell) {
    int ci, cj, ck;
    slot_ijk(a, c, ci, cj, ck);
    const int mk = a.mask.p ? a.mask.p[cell_off(a.mask, ci, cj, ck)] : 1;
    const WS w = cell_ws<M>(a.ws, c);
    Lane<M> st;
    st.ws = w;
    st.iws = cell_iws<M>(a.iws, c);
    st.mat = cell_mat<M>(a.mat, c);
    st.qmax = a.qmax;
    st.lu_mode = a.lu_mode;
    st.ir_sweeps = a.ir_sweeps;
    st.use_pred = a.use_pred;
    st.gmres = a.gmres;
    st.rtol = a.rtol;
    st.mxstep = a.mxstep;
    st.tout = a.dt;
    if (mk < 0) {
      a.FC.p[cell_off(a.FC, ci, cj, ck)] = 0.0;
      st.phase = PH_NEED_CELL;
      st.status = ST_OK;
      st.setup_pending = 0;
      st.lu64 = 0;
      st.pred_state = 0;
      if (a.cell_status) a.cell_status[slot_cell(a, c)] = 0;
    } else {
#pragma unroll
      for (int i = 0; i < N1; ++i) w[S::ATOL + i] = a.atol[i];
      double y0[N1];
#pragma unroll
      for (int k = 0; k < N; ++k) {
        y0[k] = a.rY.p[cell_off(a.rY, ci, cj, ck) + k * a.rY.ns];
        w[S::FY + k] = a.rYsrc.p ? a.rYsrc.p[cell_off(a.rYsrc, ci, cj, ck) + k * a.rYsrc.ns] : 0.0;
      }
      y0[N] = a.T.p[cell_off(a.T, ci, cj, ck)];
      st.rE0 = a.rE.p[cell_off(a.rE, ci, cj, ck)];
      st.fH = a.rEsrc.p ? a.rEsrc.p[cell_off(a.rEsrc, ci, cj, ck)] : 0.0;
      st.init_cell(y0);
      active = true;
    }
    store_lane(st, w);
  }
  wave_append(active, (int)c, a.list[0], &a.counts[0]);
}

// evaluation point y = zn0 + a1 zn1 + a2 acor at teval, and f(teval, y)
template <class M>
__global__ void __launch_bounds__(64) PELE_WAVES_PER_EU(PELE_EVAL_WAVES(M::N1)) k_eval(const ReactArgs a, const int* __restrict__ list,
                                                                  const unsigned int* __restrict__ count) {
  constexpr int N1 = M::N1;
  using S = Slots<N1>;
  using MS = MatSlots<N1>;
  __shared__ double sB[64 * N1];
  __shared__ long scell[64];
  const long idx = (long)blockIdx.x * 64 + threadIdx.x;
  int iters = 0;
  bool newton = false;
  long c = -1;
  if (idx < (long)*count) {
    c = list[idx];
    const WS w = cell_ws<M>(a.ws, c);
    // only the fields the evaluation needs are read, not the whole lane state
    const int phase = LANE_FIELD(M, w, phase);
    if (phase != PH_NEED_CELL) {
      const double a1 = LANE_FIELD(M, w, a1), a2 = LANE_FIELD(M, w, a2);
      double y[N1], f[N1], Tsol;
#pragma unroll
      for (int i = 0; i < N1; ++i) y[i] = w[S::ZN + i] + a1 * w[S::ZN + N1 + i] + a2 * w[S::ACOR + i];
      rhs<M>(y, LANE_FIELD(M, w, teval), w, LANE_FIELD(M, w, rE0), LANE_FIELD(M, w, fH), f, Tsol, &iters);
#pragma unroll
      for (int i = 0; i < N1; ++i) {
        w[S::Y + i] = y[i];
        w[S::F + i] = f[i];
      }
      w[S::TS] = Tsol;
      if (phase == PH_NEWTON) {
        // the Newton system's right-hand side gamma f - rl1 zn1 - acor
        newton = true;
        const double gamma = LANE_FIELD(M, w, gamma), rl1 = LANE_FIELD(M, w, rl1);
#pragma unroll
        for (int i = 0; i < N1; ++i) f[i] = gamma * f[i] - rl1 * w[S::ZN + N1 + i] - w[S::ACOR + i];
        if (a.lu_lane) {
          const CM mat = cell_mat<M>(a.mat, c);
#pragma unroll
          for (int i = 0; i < N1; ++i) mat[MS::B + i] = f[i];
        } else {
#pragma unroll
          for (int i = 0; i < N1; ++i) sB[threadIdx.x * N1 + i] = f[i];
        }
      }
    }
  }
  if (!a.lu_lane) {
    scell[threadIdx.x] = newton ? c : -1;
    __syncthreads();
    rows_to_cells<N1>(sB, scell, a.mat, MS::STRIDE, MS::B);
  }
  wave_append(newton, (int)c, a.nlist, &a.counts[6]);
  if (a.counters) wave_add(&a.counters[CNT_HY2T], (unsigned long long)iters);
}

// Jacobian (when stale) and LU of I - gamma J for the cells starting a Newton attempt
template <class M>
__global__ void __launch_bounds__(64) PELE_WAVES_PER_EU(PELE_SETUP_WAVES(M::N1)) k_setup(const ReactArgs a, const int* __restrict__ list,
                                              const unsigned int* __restrict__ count) {
  constexpr int N1 = M::N1;
  using S = Slots<N1>;
  using MS = MatSlots<N1>;
  __shared__ double sJ[64 * N1];
  __shared__ long scell[64];
  const long idx = (long)blockIdx.x * 64 + threadIdx.x;
  long c = -1;
  bool fresh = false;   // this lane evaluated a new Jacobian
  if (idx < (long)*count) {
    c = list[idx];
    const WS w = cell_ws<M>(a.ws, c);
    Lane<M> st;
    load_lane(st, a, c);
    if (st.needs_setup() && !st.setup_pending) {
      double y[N1], f[N1];
#pragma unroll
      for (int i = 0; i < N1; ++i) {
        y[i] = w[S::Y + i];
        f[i] = w[S::F + i];
      }
      st.setup(y, f, w[S::TS]);
      fresh = st.jcur;
      store_lane(st, w);
    }
  }
  // the Jacobian (accumulated in the wave-tiled slots) goes to the matrix region, where
  // the factorisation and refinement kernels read it row by row
  if (a.lu_lane) {
    if (fresh) {
      const WS w = cell_ws<M>(a.ws, c);
      const CM mat = cell_mat<M>(a.mat, c);
#pragma unroll 1
      for (int e = 0; e < N1 * N1; ++e) mat[MS::JS + e] = w[S::JS + e];
    }
  } else {
    scell[threadIdx.x] = fresh ? c : -1;
    const WS w = cell_ws<M>(a.ws, fresh ? c : 0);
#pragma unroll 1
    for (int r = 0; r < N1; ++r) {
      __syncthreads();
      if (fresh) {
#pragma unroll
        for (int j = 0; j < N1; ++j) sJ[threadIdx.x * N1 + j] = w[S::JS + r * N1 + j];
      }
      __syncthreads();
      rows_to_cells<N1>(sJ, scell, a.mat, MS::STRIDE, MS::JS + r * N1);
    }
  }
}

// Learned initial iterate for the cells whose attempt requested one (pred_state 1):
// features (T, log p, log10 Y_k, log h) of the state y_n saved in TMP, normalised;
// the network (three GELU layers of 128 on the matrix cores) gives (dT, dY_k) over h;
// acor = (T + dT, rho (Y + dY)) - zn0, so the first evaluation is at the learned guess
// while the error test keeps measuring against the Nordsieck prediction zn0.
// 64 cells per workgroup of four waves.
template <class M, int KIN, int NOUT>
__global__ void __launch_bounds__(MLP_THREADS) k_predict(const ReactArgs a, const MlpDev w,
                                                         const int* __restrict__ list,
                                                         const unsigned int* __restrict__ count) {
  constexpr int N = M::NS, N1 = M::N1;
  static_assert(N + 3 <= KIN && N1 <= NOUT, "network shape does not fit the mechanism");
  using S = Slots<N1>;
  constexpr int H_WORD = (int)(offsetof(Lane<M>, h) / 8);
  __shared__ unsigned short h0[MLP_ROWS * MLP_LDW], h1[MLP_ROWS * MLP_LDW];
  __shared__ float out[MLP_ROWS * NOUT];
  const long base = (long)blockIdx.x * MLP_ROWS;
  const long n = (long)*count;
  const int t = threadIdx.x;
  const float* in_mean = w.norm;
  const float* in_std = w.norm + 64;
  const float* out_mean = w.norm + 128;
  const float* out_std = w.norm + 192;
  if (t < MLP_ROWS) {
    const long idx = base + t;
    unsigned short* row = &h0[t * MLP_LDW];
    for (int k = 0; k < KIN; ++k) row[k] = 0;
    if (idx < n) {
      const WS wc = cell_ws<M>(a.ws, list[idx]);
      double rho = 0.0, sw = 0.0;
      for (int k = 0; k < N; ++k) {
        rho += wc[S::TMP + k];
        sw += wc[S::TMP + k] * M::imw(k);
      }
      const double T = wc[S::TMP + N];
      const double h = wc[S::LANE + H_WORD];
      const double irho = 1.0 / rho;
      double x = T;
      row[0] = f2bf((float)((x - in_mean[0]) / in_std[0]));
      x = log(M::RU * T * sw);                     // p = rho R T sum(Y_k / W_k)
      row[1] = f2bf((float)((x - in_mean[1]) / in_std[1]));
      for (int k = 0; k < N; ++k) {
        x = log10(fmax(wc[S::TMP + k] * irho, 1e-20));
        row[2 + k] = f2bf((float)((x - in_mean[2 + k]) / in_std[2 + k]));
      }
      x = log(h);
      row[N + 2] = f2bf((float)((x - in_mean[N + 2]) / in_std[N + 2]));
    }
  }
  __syncthreads();
  unsigned short* hin = h0;
  unsigned short* hout = h1;
#pragma unroll 1
  for (int l = 0; l < 3; ++l) {
    auto emit = [&](int r, int c, float z) { hout[r * MLP_LDW + c] = f2bf(mlp_gelu(z)); };
    if (l == 0)
      mlp_dense<KIN, MLP_HIDDEN>(hin, w.wt[0], w.b[0], emit);
    else
      mlp_dense<MLP_HIDDEN, MLP_HIDDEN>(hin, w.wt[l], w.b[l], emit);
    __syncthreads();
    unsigned short* tmp = hin;
    hin = hout;
    hout = tmp;
  }
  mlp_dense<MLP_HIDDEN, NOUT>(hin, w.wt[3], w.b[3], [&](int r, int c, float z) { out[r * NOUT + c] = z; });
  __syncthreads();
  if (t < MLP_ROWS && base + t < n) {
    const long c = list[base + t];
    const WS wc = cell_ws<M>(a.ws, c);
    Lane<M> st;
    load_lane(st, a, c);
    if (st.pred_state == 1) {
      double rho = 0.0;
      for (int k = 0; k < N; ++k) rho += wc[S::TMP + k];
      double g[N1];
      bool finite = true;
      for (int i = 0; i < N1; ++i) {
        const double d = (double)out_mean[i] + (double)out_std[i] * (double)out[t * NOUT + i];
        // output 0 is dT, outputs 1..N are dY_k
        g[i] = (i == 0) ? wc[S::TMP + N] + d : 0.0;
        if (i > 0) g[i] = wc[S::TMP + i - 1] + rho * d;
        finite = finite && isfinite(g[i]);
      }
      if (finite) {
        // state order is (rho Y_0 .. rho Y_{N-1}, T)
        for (int k = 0; k < N; ++k) wc[S::ACOR + k] = g[k + 1] - st.zn(0, k);
        wc[S::ACOR + N] = g[0] - st.zn(0, N);
        st.a2 = 1.0;
        st.pred_state = 2;
        ++st.npred;
      } else {
        st.pred_state = 3;
      }
      store_lane(st, wc);
    }
  }
}

// LU of M = I - gamma J for every cell whose set-up is pending, one lane per cell
// (used where lanes cannot cooperate, e.g. when the kernels run serially on a CPU)
template <class M>
__global__ void __launch_bounds__(64) k_lu_lane(const ReactArgs a, const int* __restrict__ list,
                                                const unsigned int* __restrict__ count) {
  using S = Slots<M::N1>;
  const long idx = (long)blockIdx.x * 64 + threadIdx.x;
  if (idx >= (long)*count) return;
  const long c = list[idx];
  const WS w = cell_ws<M>(a.ws, c);
  Lane<M> st;
  load_lane(st, a, c);
  if (!st.setup_pending) return;
  const bool fp32 = (st.lu_mode == LU_FP32_IR) && !st.lu64;
  w[S::SING] = (double)(fp32 ? lu_factor<M::N1, float>(st.mat, st.gamma)
                             : lu_factor<M::N1, double>(st.mat, st.gamma));
}


// f(std::integral_constant<int, k>) for k = 0 .. N-1, each call with k a compile-time
// constant (so arrays indexed by k stay in registers however long the sequence)
template <class F, int... K>
__device__ __forceinline__ void static_for_impl(F&& f, std::integer_sequence<int, K...>) {
  (f(std::integral_constant<int, K>{}), ...);
}
template <int N, class F>
__device__ __forceinline__ void static_for(F&& f) {
  static_for_impl(f, std::make_integer_sequence<int, N>{});
}

// x as held by lane `src` of the wave (src uniform, possibly known only at run time),
// read with v_readlane: a scalar broadcast, no trip through the LDS crossbar
__device__ __forceinline__ double lane_read(const double x, const int src) {
  const long long b = __double_as_longlong(x);
  const unsigned lo = (unsigned)__builtin_amdgcn_readlane((int)b, src);
  const unsigned hi = (unsigned)__builtin_amdgcn_readlane((int)(b >> 32), src);
  return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
}

// The largest 64-bit key over each group of G lanes (32: a half-wave, 64: the wave), in
// every lane of the group: butterfly maxima within each 16-lane row through DPP
// (quad_perm [1,0,3,2] and [2,3,0,1], then the half-row and row mirrors), then the four
// rows' maxima read with v_readlane.  It stays in the vector and scalar units, where a
// __shfl_xor ladder sends each of its levels through the LDS crossbar.  Every lane of
// the wave must take part.
template <int G>
__device__ __forceinline__ unsigned long long group_max_u64(unsigned long long v) {
  static_assert(G == 32 || G == 64, "a half-wave or the whole wave");
  auto step = [&](auto ctrl) {
    constexpr int C = decltype(ctrl)::value;
    const unsigned lo = (unsigned)__builtin_amdgcn_update_dpp(0, (int)(unsigned)v, C, 0xF, 0xF, false);
    const unsigned hi = (unsigned)__builtin_amdgcn_update_dpp(0, (int)(unsigned)(v >> 32), C, 0xF, 0xF, false);
    const unsigned long long o = ((unsigned long long)hi << 32) | lo;
    v = o > v ? o : v;
  };
  step(std::integral_constant<int, 0xB1>{});    // quad_perm [1,0,3,2]: lane ^ 1
  step(std::integral_constant<int, 0x4E>{});    // quad_perm [2,3,0,1]: lane ^ 2
  step(std::integral_constant<int, 0x141>{});   // row_half_mirror: the other quad of the half-row
  step(std::integral_constant<int, 0x140>{});   // row_mirror: the other half of the row
  auto row_max = [&](int lane) {
    const unsigned lo = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)v, lane);
    const unsigned hi = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)(v >> 32), lane);
    return ((unsigned long long)hi << 32) | lo;
  };
  const unsigned long long r0 = row_max(0), r1 = row_max(16), r2 = row_max(32), r3 = row_max(48);
  const unsigned long long a = r0 > r1 ? r0 : r1, b = r2 > r3 ? r2 : r3;
  if (G == 64) return a > b ? a : b;
  return (threadIdx.x & 32) ? b : a;
}

// x as held by lane `src` (a compile-time constant) of this lane's group of G lanes:
// v_readlane for a whole-wave group (measured faster than a permute); two groups per
// wave would need both lanes read and a select, which measured no faster than the
// permute kept for them
template <int G>
__device__ __forceinline__ double group_bcast(const double x, const int src) {
  if (G == 64) return lane_read(x, src);
  return __shfl(x, src, G);
}

// Lanes per cell of the cooperative LU and solves: a 32-lane half-wave (two cells per
// wave) up to 32 unknowns, the whole wave (one cell) up to 64.
template <int N1>
struct Coop {
  static_assert(N1 <= 64, "the cooperative kernels hold one matrix row per lane of a wave");
  static constexpr int G = N1 <= 32 ? 32 : 64;
  static constexpr int CPW = 64 / G;
};

// Cooperative LU of M = I - gamma J in Real arithmetic: G lanes per cell (row r of M
// in lane r), 64 / G cells per wave.  Column k's pivot is the largest |m_rk| over the rows
// not yet pivoted (a G-lane max-reduction); its row is broadcast through LDS and every
// other remaining row is eliminated in registers.  Row r ends as row `step` of the
// factors, where `step` is the column at which it was chosen (PERM[step] = r), which
// is partial pivoting with the interchanges kept as a permutation.  A launch handles
// the pending cells whose factor precision is Real (FP32 under LU_FP32_IR unless the
// cell has fallen back to FP64).
// One cell's factorisation by its group of G lanes (lane r of the group holds row r):
// `work` (uniform over the group) says whether the group factors at all, `mat` is the
// cell's matrix region; every lane of the wave must call it (it synchronises the block).
// Returns 0, or 1 + the first column without a nonzero pivot (uniform over the group).
template <int N1, class Real>
__device__ __forceinline__ int lu_coop_core(const CM& mat, const bool work, const double gamma) {
  constexpr int G = Coop<N1>::G, CPW = Coop<N1>::CPW;
  using A = LUAcc<N1, Real>;
  constexpr bool IS32 = sizeof(Real) == 4;
  __shared__ Real prow[CPW][N1];
  const int lane = threadIdx.x;
  const int half = lane / G, r = lane % G;
  const bool row = work && r < N1;
  Real m[N1];
  if (row) {
#pragma unroll
    for (int j = 0; j < N1; ++j) m[j] = (Real)((r == j ? 1.0 : 0.0) - gamma * mat[MatSlots<N1>::JS + r * N1 + j]);
  } else {
#pragma unroll
    for (int j = 0; j < N1; ++j) m[j] = (Real)0;
  }
  bool done = !row;
  int my_step = 0, sing = 0;
  static_for<N1>([&](auto kc) {
    constexpr int k = decltype(kc)::value;
    // Pivot: the largest |m_rk| over the remaining rows, ties to the lowest row.  The
    // bits of a non-negative double order as unsigned integers, so the search is one
    // 64-bit max over the cell's lanes and a ballot for the lowest lane holding it;
    // remaining rows carry bits + 1 so that a zero entry still beats a finished row, and
    // a NaN row counts as infinite (it still takes a definite place in the order).
    unsigned long long key = 0ull;
    if (!done) {
      double av = fabs((double)m[k]);
      if (av != av) av = HUGE_VAL;
      key = (unsigned long long)__double_as_longlong(av) + 1ull;
    }
    const unsigned long long kmax = group_max_u64<G>(key);
    const unsigned long long ball = __ballot(key == kmax);
    const unsigned long long mine = (G == 64) ? ball : ((ball >> (half * 32)) & 0xFFFFFFFFull);
    const int who = __ffsll((long long)mine) - 1;
    const bool nonzero = kmax > 1ull;   // the pivot magnitude is > 0 (equal in all G lanes)
    if (!nonzero && sing == 0) sing = k + 1;
    if constexpr (G == 64 && !IS32) {
      // one cell per wave, FP64: the pivot row's entries are read from lane `who`
      // (uniform) with v_readlane; that lane's row is not updated in this step.  (The
      // FP32 factorisation measured faster with the LDS broadcast below.)
      const bool upd = !done && r != who && nonzero;
      const Real pk = lane_read(m[k], who);
      Real l = (Real)0;
      if (upd) {
        l = m[k] * ((Real)1 / pk);
        m[k] = l;
      }
#pragma unroll
      for (int j = k + 1; j < N1; ++j) {
        const Real pj = lane_read(m[j], who);
        if (upd && l != (Real)0) m[j] -= l * pj;
      }
    } else {
      if (row && r == who) {   // the pivot row's columns k.. (those left of k are multipliers)
#pragma unroll
        for (int j = k; j < N1; ++j) prow[half][j] = m[j];
      }
      __syncthreads();
      if (!done && r != who && nonzero) {
        const Real inv = (Real)1 / prow[half][k];
        const Real l = m[k] * inv;
        m[k] = l;
        if (l != (Real)0) {
#pragma unroll
          for (int j = k + 1; j < N1; ++j) m[j] -= l * prow[half][j];
        }
      }
      __syncthreads();
    }
    if (r == who && !done) {
      done = true;
      my_step = k;
    }
  });
  if (row) {
#pragma unroll
    for (int j = 0; j < N1; ++j) A::at(mat, my_step * N1 + j) = m[j];
    mat.perm(MatSlots<N1>::PERM, my_step) = r;
  }
  __syncthreads();   // prow is reused by the next group of this block
  return sing;
}

// Cooperative LU of M = I - gamma J in Real arithmetic: G lanes per cell (row r of M
// in lane r), 64 / G cells per wave.  Column k's pivot is the largest |m_rk| over the rows
// not yet pivoted (a G-lane max-reduction); its row is broadcast through LDS and every
// other remaining row is eliminated in registers.  Row r ends as row `step` of the
// factors, where `step` is the column at which it was chosen (PERM[step] = r), which
// is partial pivoting with the interchanges kept as a permutation.  A launch handles
// the pending cells whose factor precision is Real (FP32 under LU_FP32_IR unless the
// cell has fallen back to FP64).
template <class M, class Real>
__device__ __forceinline__ void lu_coop_pair(const ReactArgs& a, const int* __restrict__ list, const long n,
                                             const long blk) {
  constexpr int N1 = M::N1;
  constexpr int G = Coop<N1>::G, CPW = Coop<N1>::CPW;
  using S = Slots<N1>;
  constexpr bool IS32 = sizeof(Real) == 4;
  const int half = threadIdx.x / G, r = threadIdx.x % G;
  const long idx = blk * CPW + half;
  long c = -1;
  bool work = false;
  double gamma = 0.0;
  if (idx < n) {
    c = list[idx];
    const WS w = cell_ws<M>(a.ws, c);
    const bool pend = LANE_FIELD(M, w, setup_pending) != 0;
    const bool fp32_cell = (a.lu_mode == LU_FP32_IR) && LANE_FIELD(M, w, lu64) == 0;
    work = pend && (fp32_cell == IS32);
    gamma = LANE_FIELD(M, w, gamma);
  }
  const int sing = lu_coop_core<N1, Real>(cell_mat<M>(a.mat, c < 0 ? 0 : c), work, gamma);
  if (work && r == 0) cell_ws<M>(a.ws, c)[S::SING] = (double)sing;
}

// Workgroups are dealt to the 8 XCDs round-robin: each XCD gets contiguous ranges of
// list entries, so neighbouring cells (whose tiles share lines) meet in one L2.  The
// grid is a multiple of 8 (so the mapping is a permutation) and strides over the list,
// so a fixed grid serves any count (as a captured graph needs).
__device__ __forceinline__ long xcd_block() {
  const unsigned per_xcd = gridDim.x / 8;
  return (long)(blockIdx.x % 8) * per_xcd + blockIdx.x / 8;
}

// STRIDE = false: one wave's cells per block, the grid covering the list (host-driven
// trips); STRIDE = true: a fixed grid striding over any count (captured graphs)
template <class M, class Real, bool STRIDE>
__global__ void __launch_bounds__(64) k_lu_coop(const ReactArgs a, const int* __restrict__ list,
                                                const unsigned int* __restrict__ count) {
  const long n = (long)*count;
  if (STRIDE) {
    for (long blk = xcd_block(); blk * Coop<M::N1>::CPW < n; blk += gridDim.x) lu_coop_pair<M, Real>(a, list, n, blk);
  } else {
    lu_coop_pair<M, Real>(a, list, n, xcd_block());
  }
}

// ---------------------------------------------------------------------------
// LU of M = I - gamma J on the matrix cores (mi.lu_form = mfma): one cell per wave, M
// staged in LDS (NP x NP, NP the unknowns rounded up to 16, 32 or 64), blocked right-
// looking elimination with panels of four columns.  Within a panel each column's pivot
// is the largest |m_rk| over the rows not yet pivoted (ties to the lowest row, a NaN as
// infinite -- the cooperative factorisation's choice), multipliers m_rk (1 / pivot) and
// the panel's other columns are updated on the vector ALUs, row by lane; the panel's
// pivot rows then take the updates of the earlier pivots of the panel in the trailing
// columns (a forward substitution, column by lane, in the unblocked order); and the
// trailing matrix of the rows not yet pivoted takes the rank-4 update L21 U12 as 16 x 16
// tiles on v_mfma_f64_16x16x4_f64 (FP64 factors) or v_mfma_f32_16x16x4_f32 (FP32 factors).
// Rows are not moved: a row ends as row `step` of the factors, the step at which it was
// chosen (PERM[step] = row), which is the layout the cooperative solves read.
//
// The 16 x 16 x 4 operand layout (one block, 64 lanes; measured by a GPU unit test): lane l
// holds A[l % 16][l / 16] and B[l / 16][l % 16]; its result register v holds
// D[l / 16 + 4 v][l % 16] for the FP64 product and D[4 (l / 16) + v][l % 16] for the FP32 one.
// ---------------------------------------------------------------------------
typedef double pele_double4 __attribute__((ext_vector_type(4)));
typedef float pele_float4 __attribute__((ext_vector_type(4)));

template <class Real>
struct MatCore;
template <>
struct MatCore<double> {
  using V4 = pele_double4;
  static __device__ __forceinline__ V4 mac(double a, double b, V4 c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
  }
  // the result row register v of lanes 16 g .. 16 g + 15 holds
  static __device__ __forceinline__ int drow(int g, int v) { return g + 4 * v; }
};
template <>
struct MatCore<float> {
  using V4 = pele_float4;
  static __device__ __forceinline__ V4 mac(float a, float b, V4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
  }
  static __device__ __forceinline__ int drow(int g, int v) { return 4 * g + v; }
};

template <int N1>
struct MfmaDim {
  static constexpr int NP = N1 <= 16 ? 16 : N1 <= 32 ? 32 : 64;
  static constexpr int LD = NP + 1;   // LDS row pitch (an odd word count spreads the banks)
  static_assert(N1 <= 64, "one wave factors at most 64 unknowns");
};

// factors the cell of `mat` when `work` (all 64 lanes call it); sm: NP x LD Reals of LDS,
// sdone / sstep: 64 ints each; returns 0, or 1 + the first column without a nonzero pivot
template <int N1, class Real>
__device__ __forceinline__ int lu_mfma_core(const CM& mat, const bool work, const double gamma, Real* sm,
                                            int* sdone, int* sstep) {
  constexpr int NP = MfmaDim<N1>::NP, LD = MfmaDim<N1>::LD, NT = NP / 16;
  using A = LUAcc<N1, Real>;
  using MS = MatSlots<N1>;
  using V4 = typename MatCore<Real>::V4;
  const int l = threadIdx.x;
  if (!work) return 0;   // uniform over the block (one cell per block)
  // M = I - gamma J from the Jacobian copy, whole lines per load; padding zero
#pragma unroll 1
  for (int e = l; e < NP * NP; e += 64) {
    const int i = e / NP, j = e - (e / NP) * NP;
    Real v = (Real)0;
    if (i < N1 && j < N1) v = (Real)((i == j ? 1.0 : 0.0) - gamma * mat[MS::JS + i * N1 + j]);
    sm[i * LD + j] = v;
  }
  bool done = l >= N1;
  sdone[l] = done ? 1 : 0;
  int sing = 0;
  __syncthreads();
#pragma unroll 1
  for (int k0 = 0; k0 < N1; k0 += 4) {
    const int kb = (N1 - k0 < 4) ? N1 - k0 : 4;
    int prow0 = 0, prow1 = 0, prow2 = 0, prow3 = 0;
#pragma unroll
    for (int kk = 0; kk < 4; ++kk) {
      if (kk < kb) {
        const int k = k0 + kk;
        unsigned long long key = 0ull;
        if (!done) {
          double av = fabs((double)sm[l * LD + k]);
          if (av != av) av = HUGE_VAL;
          key = (unsigned long long)__double_as_longlong(av) + 1ull;
        }
        const unsigned long long kmax = group_max_u64<64>(key);
        const int who = __ffsll((long long)__ballot(key == kmax)) - 1;
        const bool nonzero = kmax > 1ull;
        if (!nonzero && sing == 0) sing = k + 1;
        if (kk == 0) prow0 = who;
        if (kk == 1) prow1 = who;
        if (kk == 2) prow2 = who;
        if (kk == 3) prow3 = who;
        if (!done && l != who && nonzero) {
          const Real li = sm[l * LD + k] * ((Real)1 / sm[who * LD + k]);
          sm[l * LD + k] = li;
          if (li != (Real)0) {
#pragma unroll 1
            for (int c = k + 1; c < k0 + kb; ++c) sm[l * LD + c] -= li * sm[who * LD + c];
          }
        }
        if (l == who && !done) {
          done = true;
          sdone[l] = 1;
          sstep[k] = l;
        }
        __syncthreads();
      }
    }
    if (k0 + kb >= N1) break;
    // the panel's pivot rows in the trailing columns: the updates of the panel's earlier
    // pivots, in the order the unblocked elimination applies them (column c in lane c)
    const int prow[4] = {prow0, prow1, prow2, prow3};
    {
      const int c = l;
      if (c >= k0 + kb && c < N1) {
#pragma unroll
        for (int kp = 0; kp < 3; ++kp)
#pragma unroll
          for (int kk = kp + 1; kk < 4; ++kk)
            if (kk < kb) {
              const Real lf = sm[prow[kk] * LD + k0 + kp];
              if (lf != (Real)0) sm[prow[kk] * LD + c] -= lf * sm[prow[kp] * LD + c];
            }
      }
    }
    __syncthreads();
    // trailing rows not yet pivoted: M_rc -= sum_kk L_r,k0+kk U_kk,c on the matrix cores
    {
      const int kk = l >> 4, lo = l & 15;
      const int tj0 = (k0 + kb) / 16;
      const int prk = prow[kk];
#pragma unroll 1
      for (int ti = 0; ti < NT; ++ti) {
        const int arow = ti * 16 + lo;
        const Real av = (kk < kb && arow < N1 && sdone[arow] == 0) ? -sm[arow * LD + k0 + kk] : (Real)0;
#pragma unroll 1
        for (int tj = tj0; tj < NT; ++tj) {
          const int col = tj * 16 + lo;
          const bool colok = col >= k0 + kb && col < N1;
          const Real bv = (kk < kb && colok) ? sm[prk * LD + col] : (Real)0;
          V4 cv;
#pragma unroll
          for (int v = 0; v < 4; ++v) cv[v] = sm[(ti * 16 + MatCore<Real>::drow(kk, v)) * LD + col];
          cv = MatCore<Real>::mac(av, bv, cv);
#pragma unroll
          for (int v = 0; v < 4; ++v) {
            const int row = ti * 16 + MatCore<Real>::drow(kk, v);
            if (colok && row < N1 && sdone[row] == 0) sm[row * LD + col] = cv[v];
          }
        }
      }
    }
    __syncthreads();
  }
  // rows to their factor rows: factor row s is matrix row sstep[s]
#pragma unroll 1
  for (int e = l; e < N1 * N1; e += 64) {
    const int s = e / N1, j = e - (e / N1) * N1;
    A::at(mat, e) = sm[sstep[s] * LD + j];
  }
  if (l < N1) mat.perm(MS::PERM, l) = sstep[l];
  __syncthreads();   // the LDS is reused by the next cell of this block
  return sing;
}

template <class M, class Real>
__device__ __forceinline__ void lu_mfma_one(const ReactArgs& a, const int* __restrict__ list, const long n,
                                            const long idx) {
  constexpr int N1 = M::N1;
  constexpr int NP = MfmaDim<N1>::NP, LD = MfmaDim<N1>::LD;
  using S = Slots<N1>;
  constexpr bool IS32 = sizeof(Real) == 4;
  __shared__ Real sm[NP * LD];
  __shared__ int sdone[64], sstep[64];
  long c = -1;
  bool work = false;
  double gamma = 0.0;
  if (idx < n) {
    c = list[idx];
    const WS w = cell_ws<M>(a.ws, c);
    const bool pend = LANE_FIELD(M, w, setup_pending) != 0;
    const bool fp32_cell = (a.lu_mode == LU_FP32_IR) && LANE_FIELD(M, w, lu64) == 0;
    work = pend && (fp32_cell == IS32);
    gamma = LANE_FIELD(M, w, gamma);
  }
  const int sing = lu_mfma_core<N1, Real>(cell_mat<M>(a.mat, c < 0 ? 0 : c), work, gamma, sm, sdone, sstep);
  if (work && threadIdx.x == 0) cell_ws<M>(a.ws, c)[S::SING] = (double)sing;
}

template <class M, class Real, bool STRIDE>
__global__ void __launch_bounds__(64) k_lu_mfma(const ReactArgs a, const int* __restrict__ list,
                                                const unsigned int* __restrict__ count) {
  const long n = (long)*count;
  if (STRIDE) {
    for (long blk = xcd_block(); blk < n; blk += gridDim.x) lu_mfma_one<M, Real>(a, list, n, blk);
  } else {
    lu_mfma_one<M, Real>(a, list, n, xcd_block());
  }
}

// true when the cell's Newton iteration proceeds this trip (not a singular set-up)
template <class M>
__device__ __forceinline__ bool solve_due(const WS& w) {
  using S = Slots<M::N1>;
  return !(LANE_FIELD(M, w, setup_pending) != 0 && w[S::SING] != 0.0);
}

// Newton solves, one lane per cell (the form that runs when the kernels execute
// serially on a CPU): x = M^-1 B with FP64 factors, or with FP32 factors and up to
// ir_sweeps refinement sweeps r = B - M x (M = I - gammap J from the FP64 Jacobian),
// x += M^-1 r, stopping once the correction's WRMS norm is at most a.ir_tol.
template <class M>
__global__ void __launch_bounds__(64) k_solve_lane(const ReactArgs a, const int* __restrict__ list,
                                                   const unsigned int* __restrict__ count) {
  constexpr int N1 = M::N1;
  using S = Slots<N1>;
  using MS = MatSlots<N1>;
  const long idx = (long)blockIdx.x * 64 + threadIdx.x;
  if (idx >= (long)*count) return;
  const long c = list[idx];
  const WS w = cell_ws<M>(a.ws, c);
  if (!solve_due<M>(w)) return;
  const CM mat = cell_mat<M>(a.mat, c);
  const bool fp32 = (a.lu_mode == LU_FP32_IR) && LANE_FIELD(M, w, lu64) == 0;
  if (!fp32) {
    lu_solve<N1, double>(mat, MS::B, MS::X);
    return;
  }
  lu_solve<N1, float>(mat, MS::B, MS::X);
  const double gp = LANE_FIELD(M, w, gammap);
  int sweeps = 0;
  bool ok = (a.ir_sweeps == 0);
#pragma unroll 1
  for (; sweeps < a.ir_sweeps && !ok;) {
    ++sweeps;
#pragma unroll 1
    for (int i = 0; i < N1; ++i) {
      double v = mat[MS::B + i] - mat[MS::X + i];
#pragma unroll 1
      for (int j = 0; j < N1; ++j) v += gp * mat[MS::JS + i * N1 + j] * mat[MS::X + j];
      mat[MS::R + i] = v;
    }
    lu_solve<N1, float>(mat, MS::R, MS::D);
    double nrm = 0.0;
#pragma unroll 1
    for (int i = 0; i < N1; ++i) {
      const double d = mat[MS::D + i];
      mat[MS::X + i] += d;
      const double q = d * w[S::EWT + i];
      nrm += q * q;
    }
    ok = (sqrt(nrm / N1) <= a.ir_tol);   // false for a NaN norm
  }
  mat[MS::IRS] = (double)sweeps;
  mat[MS::IRF] = ok ? 0.0 : 1.0;
}

// The same solves with G lanes per cell (row r of the factors in lane r), 64 / G cells
// per wave: forward and back substitution broadcast each finished component with a
// G-lane shuffle; a refinement residual gathers x the same way.
//
// One cell's solve by its group: `work` (uniform over the group), the cell's matrix
// region and, for FP32 factors, its error weights (ewt of row r in lane r) and gammap.
template <int N1, class Real>
__device__ __forceinline__ void solve_coop_core(const CM& mat, const bool work, const double ewt, const double gp,
                                                const int ir_sweeps, const double ir_tol) {
  constexpr int G = Coop<N1>::G;
  using MS = MatSlots<N1>;
  using A = LUAcc<N1, Real>;
  constexpr bool IS32 = sizeof(Real) == 4;
  const int r = threadIdx.x % G;
  const bool row = work && r < N1;
  Real lrow[N1];
  int pr = 0;
  if (row) {
#pragma unroll
    for (int j = 0; j < N1; ++j) lrow[j] = A::at(mat, r * N1 + j);
    pr = mat.perm(MS::PERM, r);
  } else {
#pragma unroll
    for (int j = 0; j < N1; ++j) lrow[j] = (Real)0;
  }
  // x = U^-1 L^-1 (P v) where lane r holds v[PERM[r]] on entry
  auto subst = [&](double xv) {
#pragma unroll
    for (int k = 0; k < N1 - 1; ++k) {
      const double xk = group_bcast<G>(xv, k);
      if (r > k) xv -= (double)lrow[k] * xk;
    }
#pragma unroll
    for (int k = N1 - 1; k >= 0; --k) {
      if (r == k) xv /= (double)lrow[k];
      const double xk = group_bcast<G>(xv, k);
      if (r < k) xv -= (double)lrow[k] * xk;
    }
    return row ? xv : 0.0;
  };
  double x = subst(row ? mat[MS::B + pr] : 0.0);
  if (IS32) {
    // refinement of the FP32 solve against the FP64 matrix (uniform per cell's lanes)
    const double bi = row ? mat[MS::B + r] : 0.0;
    int sweeps = 0;
    bool ok = !work || (ir_sweeps == 0);
    // the loop runs while any group of the wave still refines (its cross-lane operations
    // stay wave-uniform); a group that has finished computes a sweep it does not apply
#pragma unroll 1
    for (;;) {
      const bool more = !ok && sweeps < ir_sweeps;   // uniform over the group
      if (__ballot(more) == 0ull) break;
      double res = bi - x;
#pragma unroll
      for (int j = 0; j < N1; ++j) {
        const double xj = group_bcast<G>(x, j);
        if (row) res += gp * mat[MS::JS + r * N1 + j] * xj;
      }
      const double d = subst(__shfl(res, pr, G));
      double q = row ? d * ewt : 0.0;
      q *= q;
#pragma unroll
      for (int off = G / 2; off > 0; off >>= 1) q += __shfl_xor(q, off, G);
      if (more) {
        ++sweeps;
        x += d;
        ok = (sqrt(q / N1) <= ir_tol);   // false for a NaN norm
      }
    }
    if (work && r == 0) {
      mat[MS::IRS] = (double)sweeps;
      mat[MS::IRF] = ok ? 0.0 : 1.0;
    }
  }
  if (row) mat[MS::X + r] = x;
}

// the solves of a list's cells whose factor precision is Real
template <class M, class Real>
__device__ __forceinline__ void solve_coop_pair(const ReactArgs& a, const int* __restrict__ list, const long n,
                                                const long blk) {
  constexpr int N1 = M::N1;
  constexpr int G = Coop<N1>::G, CPW = Coop<N1>::CPW;
  using S = Slots<N1>;
  constexpr bool IS32 = sizeof(Real) == 4;
  const int half = threadIdx.x / G, r = threadIdx.x % G;
  const long idx = blk * CPW + half;
  bool work = false;
  long c = 0;
  if (idx < n) {
    c = list[idx];
    const WS w = cell_ws<M>(a.ws, c);
    const bool fp32_cell = (a.lu_mode == LU_FP32_IR) && LANE_FIELD(M, w, lu64) == 0;
    work = solve_due<M>(w) && (fp32_cell == IS32);
  }
  double ewt = 0.0, gp = 0.0;
  if (IS32 && work) {
    const WS w = cell_ws<M>(a.ws, c);
    gp = LANE_FIELD(M, w, gammap);
    if (r < N1) ewt = w[S::EWT + r];
  }
  solve_coop_core<N1, Real>(cell_mat<M>(a.mat, c), work, ewt, gp, a.ir_sweeps, a.ir_tol);
}

template <class M, class Real, bool STRIDE>
__global__ void __launch_bounds__(64) k_solve_coop(const ReactArgs a, const int* __restrict__ list,
                                                   const unsigned int* __restrict__ count) {
  const long n = (long)*count;
  if (STRIDE) {
    for (long blk = xcd_block(); blk * Coop<M::N1>::CPW < n; blk += gridDim.x) solve_coop_pair<M, Real>(a, list, n, blk);
  } else {
    solve_coop_pair<M, Real>(a, list, n, xcd_block());
  }
}

// ---------------------------------------------------------------------------
// cvode.solve_type = GMRES: the Newton system M x = b, M = I - gamma J, by scaled GMRES
// with no preconditioner and no matrix (CVODE's iterative linear-solver interface with
// SUNLinSol_SPGMR, as the reference's GPU decks select it).  One lane per cell.  A
// right-hand side already below eplifac tq[4] in the WRMS norm gives x = b on the first
// Newton iteration and x = 0 later.  Otherwise the Krylov space of S M S^-1 on S b,
// S = diag(error weights), grows to at most maxl vectors (modified Gram-Schmidt with one
// reorthogonalisation pass when the new vector's norm drops below 1e-3 of its norm before
// orthogonalisation; Givens rotations reduce the Hessenberg matrix as it grows) until
// the residual estimate |prod s_i| beta reaches eplifac tq[4] sqrt(n).  Each product
// M v = v - gamma (f(tn, y + sig v) - f(tn, y)) / sig, sig = 1 / ||v||_WRMS, evaluates
// the right-hand side (T by HY2T) at the trip's iterate y, whose f the evaluation kernel
// left in the workspace.  A residual that fell but did not converge is accepted on the
// first Newton iteration only.  Returns 1 when the Newton iteration must fail.
// Storage: the Krylov vectors in the workspace's Jacobian slots (unused without a
// matrix); the Hessenberg matrix, rotations and small solution in LDS up to 5 Krylov
// vectors, beyond in the matrix region and TMP slots.
// ---------------------------------------------------------------------------
constexpr double GMRES_EPLIFAC = 0.05, GMRES_GS_FACTOR = 1000.0;
// Krylov dimensions up to this keep the Hessenberg matrix, rotations and small solution
// in LDS (46 words per lane at 5); larger ones keep them in the cell's matrix region
constexpr int GMRES_LDS_MAXL = 5;
constexpr int GMRES_LDS_WORDS = (GMRES_LDS_MAXL + 1) * GMRES_LDS_MAXL + 2 * GMRES_LDS_MAXL + GMRES_LDS_MAXL + 1;

// the small GMRES arrays of one lane: H (maxl + 1 rows of maxl), the rotations (c, s
// pairs) and the small solution y.  In LDS, word k of lane l at sm[k * 64 + l] (the 64
// lanes of an access hit 64 banks); in the matrix region, H in the factor block (row
// stride N1), the rotations in R and D, y in the TMP slots of the workspace.
struct GmresLds {
  double* sm;
  int lane;
  __device__ __forceinline__ double& H(int i, int j) const { return sm[(i * GMRES_LDS_MAXL + j) * 64 + lane]; }
  __device__ __forceinline__ double& G(int k) const {
    return sm[((GMRES_LDS_MAXL + 1) * GMRES_LDS_MAXL + k) * 64 + lane];
  }
  __device__ __forceinline__ double& Y(int k) const {
    return sm[((GMRES_LDS_MAXL + 1) * GMRES_LDS_MAXL + 2 * GMRES_LDS_MAXL + k) * 64 + lane];
  }
};
template <int N1>
struct GmresMat {
  CM mat;
  WS w;
  __device__ __forceinline__ double& H(int i, int j) const { return mat[MatSlots<N1>::LU + i * N1 + j]; }
  __device__ __forceinline__ double& G(int k) const { return mat[MatSlots<N1>::R + k]; }
  __device__ __forceinline__ double& Y(int k) const { return w[Slots<N1>::TMP + k]; }
};

template <class M, class HS>
__device__ __forceinline__ int gmres_cell(const WS& w, const CM& mat, const HS& hs, const int maxl,
                                          const double gamma, const double tq4, const int mnewt, const double t,
                                          const double rE0, const double fH, int& nli) {
  constexpr int N1 = M::N1;
  using S = Slots<N1>;
  using MS = MatSlots<N1>;
  nli = 0;
  const double deltar = GMRES_EPLIFAC * tq4;
  double s = 0.0;
#pragma unroll
  for (int i = 0; i < N1; ++i) {
    const double p = mat[MS::B + i] * w[S::EWT + i];
    s += p * p;
  }
  if (sqrt(s / N1) <= deltar) {
#pragma unroll
    for (int i = 0; i < N1; ++i) mat[MS::X + i] = (mnewt > 0) ? 0.0 : mat[MS::B + i];
    return 0;
  }
  const double delta = deltar * sqrt((double)N1);
  double beta2 = 0.0;
#pragma unroll
  for (int i = 0; i < N1; ++i) {
    const double v = w[S::EWT + i] * mat[MS::B + i];
    w[S::JS + i] = v;
    beta2 += v * v;
  }
  const double beta = sqrt(beta2);
  if (beta <= delta) {
#pragma unroll
    for (int i = 0; i < N1; ++i) mat[MS::X + i] = 0.0;
    return 0;
  }
  const double rb = 1.0 / beta;
#pragma unroll
  for (int i = 0; i < N1; ++i) w[S::JS + i] *= rb;
  double rho = beta, rotp = 1.0;
  int krydim = 0;
  bool converged = false;
#pragma unroll 1
  for (int l = 0; l < maxl; ++l) {
    ++nli;
    krydim = l + 1;
    const int vl = S::JS + l * N1, vk = vl + N1;
    // V[l+1] = S M S^-1 V[l]
    double vt[N1], work[N1], fw[N1], Tw;
    double sv = 0.0;
#pragma unroll
    for (int i = 0; i < N1; ++i) {
      vt[i] = w[vl + i] / w[S::EWT + i];
      const double p = vt[i] * w[S::EWT + i];
      sv += p * p;
    }
    const double sig = 1.0 / sqrt(sv / N1);
#pragma unroll
    for (int i = 0; i < N1; ++i) work[i] = sig * vt[i] + w[S::Y + i];
    rhs<M>(work, t, w, rE0, fH, fw, Tw);
    const double siginv = 1.0 / sig;
    // the new vector stays in registers through the orthogonalisation; each basis
    // vector is read once per pass
    double z[N1];
    double vn2 = 0.0;
#pragma unroll
    for (int i = 0; i < N1; ++i) {
      const double jv = siginv * fw[i] - siginv * w[S::F + i];
      z[i] = w[S::EWT + i] * (vt[i] - gamma * jv);
      vn2 += z[i] * z[i];
    }
    // modified Gram-Schmidt against V[0..l]
    const double vk_norm = sqrt(vn2);
#pragma unroll 1
    for (int j = 0; j <= l; ++j) {
      const int vj = S::JS + j * N1;
      double v[N1];
      double h = 0.0;
#pragma unroll
      for (int i = 0; i < N1; ++i) {
        v[i] = w[vj + i];
        h += v[i] * z[i];
      }
      hs.H(j, l) = h;
#pragma unroll
      for (int i = 0; i < N1; ++i) z[i] = z[i] - h * v[i];
    }
    double nrm2 = 0.0;
#pragma unroll
    for (int i = 0; i < N1; ++i) nrm2 += z[i] * z[i];
    double nrm = sqrt(nrm2);
    const double big = GMRES_GS_FACTOR * vk_norm;
    if (big + nrm == big) {
      double newp = 0.0;
#pragma unroll 1
      for (int j = 0; j <= l; ++j) {
        const int vj = S::JS + j * N1;
        double v[N1];
        double tdot = 0.0;
#pragma unroll
        for (int i = 0; i < N1; ++i) {
          v[i] = w[vj + i];
          tdot += v[i] * z[i];
        }
        newp += tdot * tdot;
        hs.H(j, l) += tdot;
#pragma unroll
        for (int i = 0; i < N1; ++i) z[i] = z[i] - tdot * v[i];
      }
      if (newp != 0.0) nrm = sqrt(nrm * nrm - newp);
    }
    hs.H(l + 1, l) = nrm;
    // the earlier rotations on column l, then the rotation that zeroes H[l+1][l]
#pragma unroll 1
    for (int i = 0; i < l; ++i) {
      const double c = hs.G(2 * i), sn = hs.G(2 * i + 1);
      const double t1 = hs.H(i, l), t2 = hs.H(i + 1, l);
      hs.H(i, l) = c * t1 - sn * t2;
      hs.H(i + 1, l) = sn * t1 + c * t2;
    }
    const double t1 = hs.H(l, l), t2 = nrm;
    double c, sn;
    if (t2 == 0.0) {
      c = 1.0;
      sn = 0.0;
    } else if (fabs(t2) >= fabs(t1)) {
      const double t3 = t1 / t2;
      sn = -1.0 / sqrt(1.0 + t3 * t3);
      c = -sn * t3;
    } else {
      const double t3 = t2 / t1;
      c = 1.0 / sqrt(1.0 + t3 * t3);
      sn = -c * t3;
    }
    hs.G(2 * l) = c;
    hs.G(2 * l + 1) = sn;
    const double hll = c * t1 - sn * t2;
    hs.H(l, l) = hll;
    if (hll == 0.0) return 1;   // breakdown of the Givens-reduced Hessenberg matrix
    rotp *= sn;
    rho = fabs(rotp * beta);
    if (rho <= delta) {
      converged = true;
      break;
    }
    const double rinv = 1.0 / nrm;
#pragma unroll
    for (int i = 0; i < N1; ++i) w[vk + i] = z[i] * rinv;
  }
  if (!converged && !(rho < beta)) return 1;
  if (!converged && mnewt > 0) return 1;
  // y = R^-1 Q^T (beta e1), x = S^-1 V y
  hs.Y(0) = beta;
#pragma unroll 1
  for (int i = 1; i <= krydim; ++i) hs.Y(i) = 0.0;
#pragma unroll 1
  for (int k = 0; k < krydim; ++k) {
    const double c = hs.G(2 * k), sn = hs.G(2 * k + 1);
    const double a1 = hs.Y(k), a2 = hs.Y(k + 1);
    hs.Y(k) = c * a1 - sn * a2;
    hs.Y(k + 1) = sn * a1 + c * a2;
  }
#pragma unroll 1
  for (int k = krydim - 1; k >= 0; --k) {
    const double hkk = hs.H(k, k);
    if (hkk == 0.0) return 1;
    const double yk = hs.Y(k) / hkk;
    hs.Y(k) = yk;
#pragma unroll 1
    for (int i = 0; i < k; ++i) hs.Y(i) -= yk * hs.H(i, k);
  }
#pragma unroll
  for (int i = 0; i < N1; ++i) {
    double x = hs.Y(0) * w[S::JS + i];
#pragma unroll 1
    for (int k = 1; k < krydim; ++k) x += hs.Y(k) * w[S::JS + k * N1 + i];
    mat[MS::X + i] = x / w[S::EWT + i];
  }
  return 0;
}

// the GMRES solves of this trip's Newton systems (cells in the solve list); the flag
// (MatSlots IRF) tells the step kernel whether the linear solve failed, IRS holds its
// Krylov iterations
template <class M>
__global__ void __launch_bounds__(64) PELE_WAVES_PER_EU(PELE_EVAL_WAVES(M::N1)) k_solve_gmres(
    const ReactArgs a, const int* __restrict__ list, const unsigned int* __restrict__ count) {
  using MS = MatSlots<M::N1>;
  __shared__ double sg[GMRES_LDS_WORDS * 64];
  const long idx = (long)blockIdx.x * 64 + threadIdx.x;
  unsigned long long nli = 0;
  if (idx < (long)*count) {
    const long c = list[idx];
    const WS w = cell_ws<M>(a.ws, c);
    const CM mat = cell_mat<M>(a.mat, c);
    const double tq4 = lane_field<M, double>(w, offsetof(Lane<M>, tq) + 4 * sizeof(double));
    const double gamma = LANE_FIELD(M, w, gamma), teval = LANE_FIELD(M, w, teval);
    const double rE0 = LANE_FIELD(M, w, rE0), fH = LANE_FIELD(M, w, fH);
    const int mnewt = LANE_FIELD(M, w, m);
    int it = 0, fail;
    if (a.gmres_maxl <= GMRES_LDS_MAXL)
      fail = gmres_cell<M>(w, mat, GmresLds{sg, (int)threadIdx.x}, a.gmres_maxl, gamma, tq4, mnewt, teval, rE0, fH,
                           it);
    else
      fail = gmres_cell<M>(w, mat, GmresMat<M::N1>{mat, w}, a.gmres_maxl, gamma, tq4, mnewt, teval, rE0, fH, it);
    mat[MS::IRF] = (double)fail;
    mat[MS::IRS] = (double)it;
    nli = (unsigned long long)it;
  }
  if (a.counters) wave_add(&a.counters[CNT_NLI], nli);
}

// advance every active cell's state machine by this trip's evaluation
template <class M>
__global__ void __launch_bounds__(64) PELE_WAVES_PER_EU(PELE_CONSUME_WAVES) k_consume(const ReactArgs a, const int* __restrict__ list,
                                                const unsigned int* __restrict__ count, int* next,
                                                unsigned int* next_count, int* snext, unsigned int* snext_count) {
  constexpr int N = M::NS, N1 = M::N1;
  using S = Slots<N1>;
  const long idx = (long)blockIdx.x * 64 + threadIdx.x;
  const long n = (long)*count;
  __shared__ double sX[64 * N1];
  __shared__ long scell[64];
  bool keep = false, want_setup = false, want_pred = false;
  unsigned long long t_cells = 0, t_nst = 0, t_nfe = 0, t_nje = 0, t_nsetups = 0, t_nni = 0, t_ncfn = 0,
                     t_netf = 0, t_fail = 0, t_maxst = 0, t_ir = 0, t_lufb = 0, t_pred = 0, t_predfb = 0;
  long c = -1;
  if (idx < n) {
    c = list[idx];
  }
  const WS w0 = cell_ws<M>(a.ws, c < 0 ? 0 : c);
  const int phase0 = (c >= 0) ? LANE_FIELD(M, w0, phase) : (int)PH_NEED_CELL;
  if (!a.lu_lane) {
    // this trip's Newton solutions, read from the matrix region as whole lines
    scell[threadIdx.x] = (phase0 == PH_NEWTON) ? c : -1;
    __syncthreads();
    cells_to_rows<N1>(sX, scell, a.mat, MatSlots<N1>::STRIDE, MatSlots<N1>::X);
    __syncthreads();
  }
  if (phase0 != PH_NEED_CELL) {
    const WS w = w0;
    Lane<M> st;
    load_lane(st, a, c);
    if (!a.lu_lane) st.xs = &sX[threadIdx.x * N1];
    {
      ++st.nfe;
      // the right-hand side at this trip's point is read only by the phases that use it
      // (a Newton iteration consumes the solve's result instead)
      double f[N1];
      auto load_f = [&]() {
#pragma unroll
        for (int i = 0; i < N1; ++i) f[i] = w[S::F + i];
      };
      bool reached = false;
      bool proceed = true;
      if (st.setup_pending) proceed = st.after_setup((int)w[S::SING]);
      if (proceed) {
        switch (st.phase) {
          case PH_INIT_F0:
#pragma unroll
            for (int i = 0; i < N1; ++i) st.zn(1, i) = w[S::F + i];
            st.after_f0();
            break;
          case PH_HIN:
            load_f();
            st.after_hin(f);
            break;
          case PH_NEWTON:
            reached = st.iterate();
            break;
          case PH_RESTART_F:
            load_f();
            st.restart_f(f);
            break;
          default:
            break;
        }
      }
      int ci, cj, ck;
      slot_ijk(a, c, ci, cj, ck);
      if (reached) {
        // CV_NORMAL output: interpolate the Nordsieck array back to tout
        const double sfac = (st.tout - st.tn) / st.h;
        double yo[N1];
#pragma unroll
        for (int i = 0; i < N1; ++i) yo[i] = st.zn(st.q, i);
        for (int j = st.q - 1; j >= 0; --j)
#pragma unroll
          for (int i = 0; i < N1; ++i) yo[i] = st.zn(j, i) + sfac * yo[i];
        bool finite = true;
#pragma unroll
        for (int i = 0; i < N1; ++i) finite = finite && isfinite(yo[i]);
        st.finish(finite ? ST_OK : ST_NONFINITE);
        if (finite) {
#pragma unroll
          for (int k = 0; k < N; ++k) a.rY.p[cell_off(a.rY, ci, cj, ck) + k * a.rY.ns] = yo[k];
          a.T.p[cell_off(a.T, ci, cj, ck)] = yo[N];
          a.rE.p[cell_off(a.rE, ci, cj, ck)] = st.rE0 + st.fH * st.tout;
        }
      }
      if (st.phase == PH_NEED_CELL) {
        const bool ok = (st.status == ST_OK);
        a.FC.p[cell_off(a.FC, ci, cj, ck)] = ok ? (double)st.nfe : -(double)st.nfe;
        if (a.cell_status) a.cell_status[slot_cell(a, c)] = st.status;
        t_cells = 1;
        t_nst = st.nst; t_nfe = st.nfe; t_nje = st.nje; t_nsetups = st.nsetups; t_nni = st.nni;
        t_ncfn = st.ncfn; t_netf = st.netf; t_fail = ok ? 0 : 1; t_maxst = st.nst;
        t_ir = st.nir; t_lufb = st.nlufb; t_pred = st.npred; t_predfb = st.npredfb;
      } else {
        keep = true;
        want_setup = st.needs_setup();
        want_pred = (st.pred_state == 1);
      }
      store_lane(st, w);
    }
  }
  if (a.compact) {
    wave_append(keep, (int)c, next, next_count);
    wave_append(want_setup, (int)c, snext, snext_count);
    if (a.use_pred) wave_append(want_pred, (int)c, a.plist, &a.counts[7]);
  } else {
    // fixed cell-to-lane mapping: count the cells still integrating; the set-up list is
    // compacted as in the queue schedule
    const unsigned long long b = __ballot(keep);
    if (__lane_id() == 0 && b) atomicAdd(next_count, (unsigned int)__popcll(b));
    wave_append(want_setup, (int)c, snext, snext_count);
  }
  // the statistics change only when a cell of this wave finished in this trip (a
  // wave-uniform test: most waves skip the fourteen reductions)
  if (a.counters && __ballot(t_cells != 0) != 0ull) {
    wave_add(&a.counters[CNT_CELLS], t_cells);
    wave_add(&a.counters[CNT_NST], t_nst);
    wave_add(&a.counters[CNT_NFE], t_nfe);
    wave_add(&a.counters[CNT_NJE], t_nje);
    wave_add(&a.counters[CNT_NSETUPS], t_nsetups);
    wave_add(&a.counters[CNT_NNI], t_nni);
    wave_add(&a.counters[CNT_NCFN], t_ncfn);
    wave_add(&a.counters[CNT_NETF], t_netf);
    wave_add(&a.counters[CNT_NFAIL], t_fail);
    wave_max(&a.counters[CNT_MAXSTEPS], t_maxst);
    wave_add(&a.counters[CNT_IR], t_ir);
    wave_add(&a.counters[CNT_LUFALLBACK], t_lufb);
    wave_add(&a.counters[CNT_PRED], t_pred);
    wave_add(&a.counters[CNT_PREDFB], t_predfb);
  }
}

// cells still integrating after the trip budget: report them as failed
template <class M>
__global__ void __launch_bounds__(64) k_abandon(const ReactArgs a, const int* __restrict__ list,
                                                const unsigned int* __restrict__ count) {
  const long idx = (long)blockIdx.x * 64 + threadIdx.x;
  bool failed = false;
  if (idx < (long)*count) {
    const long c = list[idx];
    const WS w = cell_ws<M>(a.ws, c);
    Lane<M> st;
    load_lane(st, a, c);
    if (st.phase != PH_NEED_CELL) {
      int ci, cj, ck;
      slot_ijk(a, c, ci, cj, ck);
      st.finish(ST_TOO_MUCH_WORK);
      a.FC.p[cell_off(a.FC, ci, cj, ck)] = -(double)st.nfe;
      if (a.cell_status) a.cell_status[slot_cell(a, c)] = ST_TOO_MUCH_WORK;
      store_lane(st, w);
      failed = true;
    }
  }
  if (a.counters) {
    wave_add(&a.counters[CNT_NFAIL], failed ? 1ull : 0ull);
    wave_add(&a.counters[CNT_CELLS], failed ? 1ull : 0ull);
  }
}

// ---------------------------------------------------------------------------
// the persistent schedule
//
// One kernel per call.  Each wave owns 64 lane slots (workspace tile and matrix region
// per slot, so the accesses of one slot by the wave stay whole lines however the cells
// are spread); a lane keeps its cell's integrator state (Lane<M>) in private memory for
// the cell's whole integration, and every iteration of the wave's loop is one trip of
// every cell it holds: evaluation, set-up, cooperative LU and solves over the wave's
// own cells, step logic.  Idle lanes claim the next cells of the queue with one atomic
// per claim; the wave ends when the queue is drained and its lanes are idle.  Every
// cross-lane operation below is reached by all 64 lanes (wave-uniform control flow).
// ---------------------------------------------------------------------------

// the cells with mask >= 0 into the queue (in cell order, or in `order`); masked cells
// get FC = 0
template <class M>
__global__ void __launch_bounds__(64) k_queue(const ReactArgs a, const int* __restrict__ order, int* queue,
                                              unsigned int* qcount) {
  const long idx = (long)blockIdx.x * 64 + threadIdx.x;
  bool active = false;
  long c = -1;
  if (idx < a.ncell) {
    c = order ? order[idx] : idx;
    int ci, cj, ck;
    cell_ijk(a, c, ci, cj, ck);
    const int mk = a.mask.p ? a.mask.p[cell_off(a.mask, ci, cj, ck)] : 1;
    if (mk < 0) {
      a.FC.p[cell_off(a.FC, ci, cj, ck)] = 0.0;
      if (a.cell_status) a.cell_status[c] = 0;
    } else {
      active = true;
    }
  }
  wave_append(active, (int)c, queue, qcount);
}

// The wave's cells that `sel` marks, as a list of lanes in LDS; returns their count
// (wave-uniform).  Called by all 64 lanes.
__device__ __forceinline__ int wave_lanes(const bool sel, int* lanes) {
  const unsigned long long b = __ballot(sel);
  const int lane = threadIdx.x;
  if (sel) lanes[__popcll(b & ((1ull << lane) - 1ull))] = lane;
  __syncthreads();
  return __popcll(b);
}

template <class M, int OCC>
__global__ void __launch_bounds__(64) PELE_WAVES_PER_EU(OCC) k_persist(const ReactArgs a, const int* __restrict__ queue,
                                                                      const unsigned int* __restrict__ qcount,
                                                                      unsigned int* head, unsigned int* iters_max,
                                                                      const int chunk, const long max_trips) {
  constexpr int N = M::NS, N1 = M::N1;
  constexpr int G = Coop<N1>::G, CPW = Coop<N1>::CPW;
  using S = Slots<N1>;
  using MS = MatSlots<N1>;
  __shared__ double stage[64 * N1];   // rows between lanes and matrix regions
  __shared__ long sslot[64];
  __shared__ int slanes[64];
  __shared__ double sgam[64];         // gamma, then gammap, of each lane's cell
  __shared__ int ssing[64];
  const int lane = threadIdx.x;
  const int half = lane / G, r = lane % G;
  const long wave0 = (long)blockIdx.x * 64;
  const long slot = wave0 + lane;
  const WS w = cell_ws<M>(a.ws, slot);
  const CM mat = cell_mat<M>(a.mat, slot);
  // with a queue, its entries are the cells (mask >= 0 already); without one the cells
  // are claimed in cell order and masked cells are settled as they are claimed
  const unsigned int n = queue ? *qcount : (unsigned int)a.ncell;
  Lane<M> st;
  st.ws = w;
  st.iws = cell_iws<M>(a.iws, slot);
  st.mat = mat;
  st.xs = &stage[lane * N1];
  st.qmax = a.qmax;
  st.lu_mode = a.lu_mode;
  st.ir_sweeps = a.ir_sweeps;
  st.use_pred = 0;
  st.gmres = 0;   // the persistent kernel solves with the factorisation (GMRES runs in trips)
  st.rtol = a.rtol;
  st.mxstep = a.mxstep;
  st.tout = a.dt;
  st.phase = PH_NEED_CELL;
  st.setup_pending = 0;
  st.lu64 = 0;
  st.pred_state = 0;
  long cell = -1, trips = 0;
  int ci = 0, cj = 0, ck = 0;
  bool exhausted = false;
  // statistics of the cells this lane finishes, added to the call's counters at the end
  unsigned long long t_cells = 0, t_nst = 0, t_nfe = 0, t_nje = 0, t_nsetups = 0, t_nni = 0, t_ncfn = 0,
                     t_netf = 0, t_fail = 0, t_maxst = 0, t_ir = 0, t_lufb = 0, t_hy2t = 0;
  unsigned int iters = 0;
  auto lane_done = [&](const bool ok) {
    a.FC.p[cell_off(a.FC, ci, cj, ck)] = ok ? (double)st.nfe : -(double)st.nfe;
    if (a.cell_status) a.cell_status[cell] = st.status;
    ++t_cells;
    t_nst += st.nst; t_nfe += st.nfe; t_nje += st.nje; t_nsetups += st.nsetups; t_nni += st.nni;
    t_ncfn += st.ncfn; t_netf += st.netf; t_fail += ok ? 0 : 1;
    t_maxst = t_maxst > (unsigned long long)st.nst ? t_maxst : (unsigned long long)st.nst;
    t_ir += st.nir; t_lufb += st.nlufb;
    cell = -1;
  };
#pragma unroll 1
  for (;; ++iters) {
    // -- claim cells for idle lanes -------------------------------------------------
    {
      const bool idle = cell < 0;
      const unsigned long long bidle = __ballot(idle);
      const int nidle = __popcll(bidle);
      if (!exhausted && nidle > 0 && (nidle >= chunk || nidle == 64)) {
        const int leader = __ffsll((long long)bidle) - 1;
        unsigned int base = 0;
        if (lane == leader) base = atomicAdd(head, (unsigned int)nidle);
        base = __shfl(base, leader, 64);
        if (base + (unsigned int)nidle >= n) exhausted = true;
        if (idle) {
          const unsigned int idx = base + (unsigned int)__popcll(bidle & ((1ull << lane) - 1ull));
          if (idx < n) {
            cell = queue ? queue[idx] : (long)idx;
            cell_ijk(a, cell, ci, cj, ck);
            if (!queue && a.mask.p && a.mask.p[cell_off(a.mask, ci, cj, ck)] < 0) {
              a.FC.p[cell_off(a.FC, ci, cj, ck)] = 0.0;
              if (a.cell_status) a.cell_status[cell] = 0;
              cell = -1;
            }
          }
          if (cell >= 0) {
#pragma unroll
            for (int i = 0; i < N1; ++i) w[S::ATOL + i] = a.atol[i];
            double y0[N1];
#pragma unroll
            for (int k = 0; k < N; ++k) {
              y0[k] = a.rY.p[cell_off(a.rY, ci, cj, ck) + k * a.rY.ns];
              w[S::FY + k] = a.rYsrc.p ? a.rYsrc.p[cell_off(a.rYsrc, ci, cj, ck) + k * a.rYsrc.ns] : 0.0;
            }
            y0[N] = a.T.p[cell_off(a.T, ci, cj, ck)];
            st.rE0 = a.rE.p[cell_off(a.rE, ci, cj, ck)];
            st.fH = a.rEsrc.p ? a.rEsrc.p[cell_off(a.rEsrc, ci, cj, ck)] : 0.0;
            st.init_cell(y0);
            trips = 0;
          }
        }
      }
      if (__ballot(cell >= 0) == 0ull) break;   // queue drained, every lane idle
    }
    // -- trip cap: a cell that has not finished after max_trips trips fails -----------
    if (cell >= 0 && trips >= max_trips) {
      st.finish(ST_TOO_MUCH_WORK);
      lane_done(false);
    }
    const bool act = cell >= 0;
    // -- evaluation: y = zn0 + a1 zn1 + a2 acor, f(teval, y); the Newton right-hand side
    //    gamma f - rl1 zn1 - acor goes to the matrix region through LDS ---------------
    double y[N1], f[N1], Tsol = 0.0;
    bool newton = false;
    if (act) {
      int it = 0;
#pragma unroll
      for (int i = 0; i < N1; ++i) y[i] = w[S::ZN + i] + st.a1 * w[S::ZN + N1 + i] + st.a2 * w[S::ACOR + i];
      rhs<M>(y, st.teval, w, st.rE0, st.fH, f, Tsol, &it);
      t_hy2t += (unsigned long long)it;
      if (st.phase == PH_NEWTON) {
        newton = true;
#pragma unroll
        for (int i = 0; i < N1; ++i) stage[lane * N1 + i] = st.gamma * f[i] - st.rl1 * w[S::ZN + N1 + i] - w[S::ACOR + i];
      } else {
#pragma unroll
        for (int i = 0; i < N1; ++i) w[S::F + i] = f[i];
      }
    }
    sslot[lane] = newton ? slot : -1;
    __syncthreads();
    rows_to_cells<N1>(stage, sslot, a.mat, MS::STRIDE, MS::B);
    __syncthreads();
    // -- set-up: the Jacobian (when stale) into the tile, then row by row to the matrix
    //    region ---------------------------------------------------------------------
    bool fresh = false;
    if (act && st.needs_setup() && !st.setup_pending) {
      st.setup(y, f, Tsol);
      fresh = st.jcur;
    }
    if (__ballot(fresh) != 0ull) {
      sslot[lane] = fresh ? slot : -1;
#pragma unroll 1
      for (int rr = 0; rr < N1; ++rr) {
        __syncthreads();
        if (fresh) {
#pragma unroll
          for (int j = 0; j < N1; ++j) stage[lane * N1 + j] = w[S::JS + rr * N1 + j];
        }
        __syncthreads();
        rows_to_cells<N1>(stage, sslot, a.mat, MS::STRIDE, MS::JS + rr * N1);
      }
      __syncthreads();
    }
    // -- factorisations of this wave's pending cells, G lanes per cell -----------------
    const bool pend = act && st.setup_pending;
    const bool lu32 = pend && a.lu_mode == LU_FP32_IR && !st.lu64;
    sgam[lane] = st.gamma;
    ssing[lane] = 0;
    {
      const int cnt = wave_lanes(lu32, slanes);
#pragma unroll 1
      for (int b = 0; b < cnt; b += CPW) {
        const bool work = b + half < cnt;
        const int owner = work ? slanes[b + half] : 0;
        const int sg = lu_coop_core<N1, float>(cell_mat<M>(a.mat, wave0 + owner), work, sgam[owner]);
        if (work && r == 0) ssing[owner] = sg;
      }
      __syncthreads();
    }
    {
      const int cnt = wave_lanes(pend && !lu32, slanes);
#pragma unroll 1
      for (int b = 0; b < cnt; b += CPW) {
        const bool work = b + half < cnt;
        const int owner = work ? slanes[b + half] : 0;
        const int sg = lu_coop_core<N1, double>(cell_mat<M>(a.mat, wave0 + owner), work, sgam[owner]);
        if (work && r == 0) ssing[owner] = sg;
      }
      __syncthreads();
    }
    const int sing = ssing[lane];
    // -- Newton solves of this wave's cells that iterate this trip ----------------------
    const bool due = newton && !(pend && sing != 0);
    const bool s32 = due && a.lu_mode == LU_FP32_IR && !st.lu64;
    sgam[lane] = st.gammap;
    {
      const int cnt = wave_lanes(s32, slanes);
#pragma unroll 1
      for (int b = 0; b < cnt; b += CPW) {
        const bool work = b + half < cnt;
        const int owner = work ? slanes[b + half] : 0;
        const double ewt = (work && r < N1) ? cell_ws<M>(a.ws, wave0 + owner)[S::EWT + r] : 0.0;
        solve_coop_core<N1, float>(cell_mat<M>(a.mat, wave0 + owner), work, ewt, sgam[owner], a.ir_sweeps, a.ir_tol);
      }
      __syncthreads();
    }
    {
      const int cnt = wave_lanes(due && !s32, slanes);
#pragma unroll 1
      for (int b = 0; b < cnt; b += CPW) {
        const bool work = b + half < cnt;
        const int owner = work ? slanes[b + half] : 0;
        solve_coop_core<N1, double>(cell_mat<M>(a.mat, wave0 + owner), work, 0.0, 0.0, 0, 0.0);
      }
      __syncthreads();
    }
    // -- step logic: the solutions come back through LDS -------------------------------
    sslot[lane] = newton ? slot : -1;
    __syncthreads();
    cells_to_rows<N1>(stage, sslot, a.mat, MS::STRIDE, MS::X);
    __syncthreads();
    if (act) {
      ++st.nfe;
      ++trips;
      bool reached = false, proceed = true;
      if (st.setup_pending) proceed = st.after_setup(sing);
      if (proceed) {
        switch (st.phase) {
          case PH_INIT_F0:
#pragma unroll
            for (int i = 0; i < N1; ++i) st.zn(1, i) = w[S::F + i];
            st.after_f0();
            break;
          case PH_HIN:
#pragma unroll
            for (int i = 0; i < N1; ++i) f[i] = w[S::F + i];
            st.after_hin(f);
            break;
          case PH_NEWTON:
            reached = st.iterate();
            break;
          case PH_RESTART_F:
#pragma unroll
            for (int i = 0; i < N1; ++i) f[i] = w[S::F + i];
            st.restart_f(f);
            break;
          default:
            break;
        }
      }
      if (reached) {
        // CV_NORMAL output: interpolate the Nordsieck array back to tout
        const double sfac = (st.tout - st.tn) / st.h;
        double yo[N1];
#pragma unroll
        for (int i = 0; i < N1; ++i) yo[i] = st.zn(st.q, i);
        for (int j = st.q - 1; j >= 0; --j)
#pragma unroll
          for (int i = 0; i < N1; ++i) yo[i] = st.zn(j, i) + sfac * yo[i];
        bool finite = true;
#pragma unroll
        for (int i = 0; i < N1; ++i) finite = finite && isfinite(yo[i]);
        st.finish(finite ? ST_OK : ST_NONFINITE);
        if (finite) {
#pragma unroll
          for (int k = 0; k < N; ++k) a.rY.p[cell_off(a.rY, ci, cj, ck) + k * a.rY.ns] = yo[k];
          a.T.p[cell_off(a.T, ci, cj, ck)] = yo[N];
          a.rE.p[cell_off(a.rE, ci, cj, ck)] = st.rE0 + st.fH * st.tout;
        }
      }
      if (st.phase == PH_NEED_CELL) lane_done(st.status == ST_OK);
    }
    __syncthreads();   // stage and sslot are rewritten by the next trip
  }
  if (a.counters) {
    wave_add(&a.counters[CNT_CELLS], t_cells);
    wave_add(&a.counters[CNT_NST], t_nst);
    wave_add(&a.counters[CNT_NFE], t_nfe);
    wave_add(&a.counters[CNT_NJE], t_nje);
    wave_add(&a.counters[CNT_NSETUPS], t_nsetups);
    wave_add(&a.counters[CNT_NNI], t_nni);
    wave_add(&a.counters[CNT_NCFN], t_ncfn);
    wave_add(&a.counters[CNT_NETF], t_netf);
    wave_add(&a.counters[CNT_NFAIL], t_fail);
    wave_max(&a.counters[CNT_MAXSTEPS], t_maxst);
    wave_add(&a.counters[CNT_IR], t_ir);
    wave_add(&a.counters[CNT_LUFALLBACK], t_lufb);
    wave_add(&a.counters[CNT_HY2T], t_hy2t);
  }
  if (lane == 0) atomicMax(iters_max, iters);
}

// list bookkeeping (small kernels rather than memsets, so the trip can be captured)
__global__ void k_identity(int* list, long n) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) list[i] = (int)i;
}
__global__ void k_zero_counts(unsigned int* counts) {
  if (threadIdx.x < 8) counts[threadIdx.x] = 0u;
}
__global__ void k_zero_two(unsigned int* counts, int i0, int i1) {
  if (threadIdx.x == 0) {
    counts[i0] = 0u;
    counts[i1] = 0u;
  }
}
__global__ void k_plain_counts(unsigned int* counts, unsigned int n) {
  if (threadIdx.x == 0) {
    counts[4] = n;
    counts[5] = counts[0];
  }
}

// net production rates rho*wdot (g/cm^3/s) with T re-solved from h
// reactionRateRhoY per cell (Source/PeleLM_K.H:57-98): rho = sum_k rhoY_k, Y = rhoY / rho,
// h = rhoH / rho x hscale, T = HY2T(h, Y) started from the T the cell holds (T itself is
// read only; the re-solved value goes to Tout when given), and
// rhoYdot_k = W_k wdot_k(rho x rscale, T, Y) x oscale.  The box form passes PeleLM's MKS
// factors (hscale 1e4, rscale 1e-3, oscale 1e3), the flat CGS form 1, 1, 1.  Cells with
// mask < 0 (EB-covered) get zero rates.
template <class M>
__global__ void __launch_bounds__(64) PELE_WAVES_PER_EU(4) rates_kernel(const ArrView rhoY, const ArrView rhoH,
                                                                        const ArrView T, const IntView mask,
                                                                        const ArrView rdot, double* __restrict__ Tout,
                                                                        const int nx, const int ny, const long ncell,
                                                                        const double hscale, const double rscale,
                                                                        const double oscale) {
  constexpr int N = M::NS;
  const long c = (long)blockIdx.x * 64 + threadIdx.x;
  if (c >= ncell) return;
  const int i = (int)(c % nx), j = (int)((c / nx) % ny), k = (int)(c / ((long)nx * ny));
  const long od = cell_off(rdot, i, j, k);
  if (mask.p && mask.p[cell_off(mask, i, j, k)] < 0) {
#pragma unroll
    for (int n = 0; n < N; ++n) rdot.p[od + n * rdot.ns] = 0.0;
    return;
  }
  const long oy = cell_off(rhoY, i, j, k);
  double Y[N], rho = 0.0;
#pragma unroll
  for (int n = 0; n < N; ++n) { Y[n] = rhoY.p[oy + n * rhoY.ns]; rho += Y[n]; }
  const double irho = 1.0 / rho;
#pragma unroll
  for (int n = 0; n < N; ++n) Y[n] *= irho;
  const double Tc = hy2t<M>(rhoH.p[cell_off(rhoH, i, j, k)] * irho * hscale, Y, T.p[cell_off(T, i, j, k)]);
  if (Tout) Tout[c] = Tc;
  const double rho_c = rho * rscale, logT = log(Tc), invT = 1.0 / Tc;
  double C[N], g[N], wd[N];
#pragma unroll
  for (int n = 0; n < N; ++n) C[n] = rho_c * Y[n] * M::imw(n);
  M::gibbs(Tc, logT, g);
  M::wdot(C, Tc, logT, invT, g, wd);
#pragma unroll
  for (int n = 0; n < N; ++n) rdot.p[od + n * rdot.ns] = wd[n] * M::mw(n) * oscale;
}

// Cost-ordering keys (mi.order), one per caller cell c: the float bits of a cost estimate
// (non-negative floats order like their bit patterns) and c itself as the value to sort.
//   ORDER_FC:   |FC| as the caller's array holds it -- the cell's RHS count of the previous
//               step, which PeleLM keeps between steps as FuncCount (0: unknown; the
//               ordering changes only which lanes run which cells, so any content is safe)
//   ORDER_RATE: max_k |W_k wdot_k + F_Y,k| dt / (rtol rhoY_k + atol_k) at the input state,
//               T from HY2T(rhoH / rho, Y): how many tolerance units the fastest-moving
//               species changes over dt
// Cells with mask < 0 get key 0.
template <class M>
__global__ void __launch_bounds__(64) PELE_WAVES_PER_EU(4) k_cost_key(const ReactArgs a, const int mode,
                                                                      unsigned int* __restrict__ keys,
                                                                      int* __restrict__ vals) {
  constexpr int N = M::NS;
  const long c = (long)blockIdx.x * 64 + threadIdx.x;
  if (c >= a.ncell) return;
  int i, j, k;
  cell_ijk(a, c, i, j, k);
  float key = 0.0f;
  const bool masked = a.mask.p && a.mask.p[cell_off(a.mask, i, j, k)] < 0;
  if (!masked && mode == ORDER_FC) {
    const double fc = fabs(a.FC.p[cell_off(a.FC, i, j, k)]);   // any value orders: NaN and inf sort last
    key = isfinite(fc) ? (float)fmin(fc, 3.0e38) : 3.0e38f;
  } else if (!masked && mode == ORDER_RATE) {
    const long oy = cell_off(a.rY, i, j, k);
    double y[N], Y[N], rho = 0.0;
#pragma unroll
    for (int n = 0; n < N; ++n) { y[n] = a.rY.p[oy + n * a.rY.ns]; rho += y[n]; }
    const double irho = 1.0 / rho;
#pragma unroll
    for (int n = 0; n < N; ++n) Y[n] = y[n] * irho;
    const double T = hy2t<M>(a.rE.p[cell_off(a.rE, i, j, k)] * irho, Y, a.T.p[cell_off(a.T, i, j, k)]);
    const double logT = log(T), invT = 1.0 / T;
    double C[N], g[N], wd[N];
#pragma unroll
    for (int n = 0; n < N; ++n) C[n] = y[n] * M::imw(n);
    M::gibbs(T, logT, g);
    M::wdot(C, T, logT, invT, g, wd);
    double kmax = 0.0;
    const long of = a.rYsrc.p ? cell_off(a.rYsrc, i, j, k) : 0;
#pragma unroll
    for (int n = 0; n < N; ++n) {
      const double fy = wd[n] * M::mw(n) + (a.rYsrc.p ? a.rYsrc.p[of + n * a.rYsrc.ns] : 0.0);
      kmax = fmax(kmax, fabs(fy) * a.dt / (a.rtol * fabs(y[n]) + a.atol[n]));
    }
    key = isfinite(kmax) ? (float)fmin(kmax, 3.0e38) : 3.0e38f;
  }
  keys[c] = __float_as_uint(key);
  vals[c] = (int)c;
}

__global__ void counters_reset_kernel(unsigned long long* counters) {
  const int t = threadIdx.x;
  if (t < NCOUNTERS) counters[t] = 0ull;
}

// ReactorNull: rY += dt rYsrc, rE += dt rEsrc, FC = 0
__global__ void null_kernel(const ReactArgs a, int nspec) {
  const long c = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= a.ncell) return;
  const int i = (int)(c % a.nx), j = (int)((c / a.nx) % a.ny), k = (int)(c / ((long)a.nx * a.ny));
  if (a.mask.p && a.mask.p[cell_off(a.mask, i, j, k)] < 0) { a.FC.p[cell_off(a.FC, i, j, k)] = 0.0; return; }
  for (int n = 0; n < nspec; ++n) {
    const double s = a.rYsrc.p ? a.rYsrc.p[cell_off(a.rYsrc, i, j, k) + n * a.rYsrc.ns] : 0.0;
    a.rY.p[cell_off(a.rY, i, j, k) + n * a.rY.ns] += a.dt * s;
  }
  const double se = a.rEsrc.p ? a.rEsrc.p[cell_off(a.rEsrc, i, j, k)] : 0.0;
  a.rE.p[cell_off(a.rE, i, j, k)] += a.dt * se;
  a.FC.p[cell_off(a.FC, i, j, k)] = 0.0;
}

// ---------------------------------------------------------------------------
// ReactorRK64: explicit Runge-Kutta sub-steps for cells whose chemistry is not stiff
// (PelePhysics' six-stage, fourth-order explicit reactor, selected by chem_integrator =
// "ReactorRK64", Tools/CMake/BuildPeleLMExe.cmake:124-125).  One lane per cell: the
// six-stage Cash-Karp pair (Cash & Karp 1990), fourth-order solution propagated, the
// difference from the fifth-order one as the local error in the WRMS norm of
// rtol max(|y|, |y_new|) + atol; accepted steps grow by min(5, 0.9 err^-1/5), rejected
// ones (a non-finite error included) shrink by max(0.2, 0.9 err^-1/4), sub-steps within
// [dt / nsub_max, dt / nsub_min] from dt / nsub_guess; failing the error test at the
// smallest sub-step fails the cell (FC = -evaluations).  The stage derivatives live in
// the cell's Nordsieck slots of the workspace (6 n words), the forcing in its FY slot.
// ---------------------------------------------------------------------------
struct RK64Opts {
  int nsub_guess, nsub_min, nsub_max;
};

__device__ constexpr double RK_C[6] = {0.0, 1.0 / 5.0, 3.0 / 10.0, 3.0 / 5.0, 1.0, 7.0 / 8.0};
__device__ constexpr double RK_A[6][5] = {
    {0, 0, 0, 0, 0},
    {1.0 / 5.0, 0, 0, 0, 0},
    {3.0 / 40.0, 9.0 / 40.0, 0, 0, 0},
    {3.0 / 10.0, -9.0 / 10.0, 6.0 / 5.0, 0, 0},
    {-11.0 / 54.0, 5.0 / 2.0, -70.0 / 27.0, 35.0 / 27.0, 0},
    {1631.0 / 55296.0, 175.0 / 512.0, 575.0 / 13824.0, 44275.0 / 110592.0, 253.0 / 4096.0}};
__device__ constexpr double RK_B4[6] = {2825.0 / 27648.0, 0.0, 18575.0 / 48384.0, 13525.0 / 55296.0,
                                        277.0 / 14336.0, 1.0 / 4.0};
__device__ constexpr double RK_B5[6] = {37.0 / 378.0, 0.0, 250.0 / 621.0, 125.0 / 594.0, 0.0, 512.0 / 1771.0};

template <class M>
__global__ void __launch_bounds__(64) PELE_WAVES_PER_EU(PELE_EVAL_WAVES(M::N1)) k_rk64(const ReactArgs a,
                                                                                       const RK64Opts o) {
  constexpr int N = M::NS, N1 = M::N1;
  using S = Slots<N1>;
  const long c = (long)blockIdx.x * 64 + threadIdx.x;
  unsigned long long t_cells = 0, t_nst = 0, t_nfe = 0, t_rej = 0, t_fail = 0, t_hy2t = 0;
  if (c < a.ncell) {
    int ci, cj, ck;
    cell_ijk(a, c, ci, cj, ck);
    const bool masked = a.mask.p && a.mask.p[cell_off(a.mask, ci, cj, ck)] < 0;
    if (masked) {
      a.FC.p[cell_off(a.FC, ci, cj, ck)] = 0.0;
    } else {
      const WS w = cell_ws<M>(a.ws, c);
      double y[N1];
#pragma unroll
      for (int k = 0; k < N; ++k) {
        y[k] = a.rY.p[cell_off(a.rY, ci, cj, ck) + k * a.rY.ns];
        w[S::FY + k] = a.rYsrc.p ? a.rYsrc.p[cell_off(a.rYsrc, ci, cj, ck) + k * a.rYsrc.ns] : 0.0;
      }
      y[N] = a.T.p[cell_off(a.T, ci, cj, ck)];
      const double rE0 = a.rE.p[cell_off(a.rE, ci, cj, ck)];
      const double fH = a.rEsrc.p ? a.rEsrc.p[cell_off(a.rEsrc, ci, cj, ck)] : 0.0;
      const double dt = a.dt, hmin = dt / o.nsub_max, hmax = dt / o.nsub_min;
      double h = dt / o.nsub_guess, t = 0.0;
      long nst = 0, nfe = 0, nrej = 0;
      const long max_attempts = 4L * o.nsub_max + 100;
      bool ok = true;
#pragma unroll 1
      while (t < dt) {
        if (nst + nrej >= max_attempts) { ok = false; break; }
        bool last = false;
        if (t + h >= dt) { h = dt - t; last = true; }
        // one call site of the right-hand side for the six stages
#pragma unroll 1
        for (int s = 0; s < 6; ++s) {
          double yt[N1], f[N1], Tsol;
#pragma unroll
          for (int i = 0; i < N1; ++i) yt[i] = 0.0;
#pragma unroll 1
          for (int j = 0; j < s; ++j) {
            const double asj = RK_A[s][j];
#pragma unroll
            for (int i = 0; i < N1; ++i) yt[i] += asj * w[S::ZN + j * N1 + i];
          }
#pragma unroll
          for (int i = 0; i < N1; ++i) yt[i] = y[i] + h * yt[i];
          int it = 0;
          rhs<M>(yt, t + RK_C[s] * h, w, rE0, fH, f, Tsol, &it);
          t_hy2t += (unsigned long long)it;
          ++nfe;
#pragma unroll
          for (int i = 0; i < N1; ++i) w[S::ZN + s * N1 + i] = f[i];
        }
        double en = 0.0, y4[N1];
#pragma unroll
        for (int i = 0; i < N1; ++i) {
          double a4 = 0.0, e = 0.0;
#pragma unroll
          for (int s = 0; s < 6; ++s) {
            const double k = w[S::ZN + s * N1 + i];
            a4 += RK_B4[s] * k;
            e += (RK_B5[s] - RK_B4[s]) * k;
          }
          y4[i] = y[i] + h * a4;
          const double sc = a.rtol * fmax(fabs(y[i]), fabs(y4[i])) + a.atol[i];
          const double r = h * e / sc;
          en += r * r;
        }
        en = sqrt(en / N1);
        if (!isfinite(en)) en = HUGE_VAL;
        if (en <= 1.0) {
#pragma unroll
          for (int i = 0; i < N1; ++i) y[i] = y4[i];
          t = last ? dt : t + h;
          ++nst;
          h = fmin(hmax, h * (en > 0.0 ? fmin(5.0, 0.9 * pow(en, -0.2)) : 5.0));
        } else {
          ++nrej;
          if (h <= hmin * (1.0 + 1e-12)) { ok = false; break; }
          h = fmax(hmin, h * fmax(0.2, 0.9 * pow(en, -0.25)));
        }
      }
      if (ok) {
#pragma unroll
        for (int k = 0; k < N; ++k) a.rY.p[cell_off(a.rY, ci, cj, ck) + k * a.rY.ns] = y[k];
        a.T.p[cell_off(a.T, ci, cj, ck)] = y[N];
        a.rE.p[cell_off(a.rE, ci, cj, ck)] = rE0 + fH * dt;
      }
      a.FC.p[cell_off(a.FC, ci, cj, ck)] = ok ? (double)nfe : -(double)nfe;
      t_cells = 1;
      t_nst = nst;
      t_nfe = nfe;
      t_rej = nrej;
      t_fail = ok ? 0 : 1;
    }
  }
  if (a.counters) {
    wave_add(&a.counters[CNT_CELLS], t_cells);
    wave_add(&a.counters[CNT_NST], t_nst);
    wave_add(&a.counters[CNT_NFE], t_nfe);
    wave_add(&a.counters[CNT_NETF], t_rej);
    wave_add(&a.counters[CNT_NFAIL], t_fail);
    wave_max(&a.counters[CNT_MAXSTEPS], t_nst);
    wave_add(&a.counters[CNT_HY2T], t_hy2t);
  }
}

#define PELE_TRY(x)                     \
  do {                                  \
    const hipError_t e_ = (x);          \
    if (e_ != hipSuccess) return e_;    \
  } while (0)

// Blocks for the cooperative kernels (CPW cells per block), a multiple of 8: with
// `exact` (the count n is known) enough for every cell, else a fixed 8192 blocks that
// stride over whatever the list holds.
// ---------------------------------------------------------------------------
// Mechanisms above 64 unknowns (one matrix row per lane of a wave does not reach): one
// cell per wave, lane l owning rows l and l + 64.
//
// LU: M = I - gamma J staged in LDS (N1 x LD words, LD odd; 63.4 KB at 89 unknowns), the
// column loop at run time, each step's pivot the largest |m_rk| over the rows not yet
// pivoted (ties to the lowest row, a NaN as infinite: one 64-bit max over the wave and
// a ballot, rows below 64 first), multipliers m_rk (1 / pivot) and the row updates from
// the pivot row, read from LDS by every lane at once.  Rows stay in place; factor row s is
// the row chosen at step s (PERM[s]).  The factors are stored by columns (element (s, j)
// at LU + j N1 + s), so the solve's lanes, one factor row each, read consecutive words of
// a column.  FP64 only.
// ---------------------------------------------------------------------------
template <int N1>
struct Wide {
  static constexpr int LD = N1 | 1;
  static_assert(N1 <= 128, "two rows per lane");
  static_assert(N1 * LD * 8 + N1 * 4 <= 65536, "the matrix must fit one workgroup's LDS");
};

template <class M>
__device__ __forceinline__ void lu_wide_one(const ReactArgs& a, const int* __restrict__ list, const long n,
                                            const long idx) {
  constexpr int N1 = M::N1, LD = Wide<N1>::LD;
  using S = Slots<N1>;
  using MS = MatSlots<N1>;
  __shared__ double sm[N1 * LD];
  __shared__ int sstep[N1];
  if (idx >= n) return;   // uniform over the block (one cell per block)
  const long c = list[idx];
  const WS w = cell_ws<M>(a.ws, c);
  if (LANE_FIELD(M, w, setup_pending) == 0) return;
  const double gamma = LANE_FIELD(M, w, gamma);
  const CM mat = cell_mat<M>(a.mat, c);
  const int l = threadIdx.x;
#pragma unroll 1
  for (int e = l; e < N1 * N1; e += 64) {
    const int i = e / N1, j = e - (e / N1) * N1;
    sm[i * LD + j] = (i == j ? 1.0 : 0.0) - gamma * mat[MS::JS + e];
  }
  const int r0 = l, r1 = l + 64;
  bool done0 = r0 >= N1, done1 = r1 >= N1;
  int sing = 0;
  __syncthreads();
#pragma unroll 1
  for (int k = 0; k < N1; ++k) {
    auto key_of = [&](int r, bool done) -> unsigned long long {
      if (done) return 0ull;
      double av = fabs(sm[r * LD + k]);
      if (av != av) av = HUGE_VAL;
      return (unsigned long long)__double_as_longlong(av) + 1ull;
    };
    const unsigned long long k0 = key_of(r0, done0), k1 = key_of(r1, done1);
    const unsigned long long kmax = group_max_u64<64>(k0 > k1 ? k0 : k1);
    const unsigned long long b0 = __ballot(k0 == kmax);
    const unsigned long long b1 = __ballot(k1 == kmax);
    const int prow = b0 ? __ffsll((long long)b0) - 1 : 64 + __ffsll((long long)b1) - 1;
    const bool nonzero = kmax > 1ull;
    if (!nonzero && sing == 0) sing = k + 1;
    if (nonzero) {
      // both rows of the lane in one pass over the pivot row, eight columns at a time
      // (the pivot row's words loaded before the rows' stores, so the loads overlap)
      const double inv = 1.0 / sm[prow * LD + k];
      double l0 = 0.0, l1 = 0.0;
      if (!done0 && r0 != prow) {
        l0 = sm[r0 * LD + k] * inv;
        sm[r0 * LD + k] = l0;
      }
      if (!done1 && r1 != prow) {
        l1 = sm[r1 * LD + k] * inv;
        sm[r1 * LD + k] = l1;
      }
      const bool u0 = l0 != 0.0, u1 = l1 != 0.0;   // a zero multiplier leaves its row as is
      const double* pr = &sm[prow * LD];
      double* a0 = &sm[(u0 ? r0 : 0) * LD];
      double* a1 = &sm[(u1 ? r1 : 0) * LD];
      int j = k + 1;
#pragma unroll 1
      for (; j + 8 <= N1; j += 8) {
        double p[8];
#pragma unroll
        for (int t = 0; t < 8; ++t) p[t] = pr[j + t];
        if (u0) {
          double v[8];
#pragma unroll
          for (int t = 0; t < 8; ++t) v[t] = a0[j + t];
#pragma unroll
          for (int t = 0; t < 8; ++t) a0[j + t] = v[t] - l0 * p[t];
        }
        if (u1) {
          double v[8];
#pragma unroll
          for (int t = 0; t < 8; ++t) v[t] = a1[j + t];
#pragma unroll
          for (int t = 0; t < 8; ++t) a1[j + t] = v[t] - l1 * p[t];
        }
      }
#pragma unroll 1
      for (; j < N1; ++j) {
        const double pj = pr[j];
        if (u0) a0[j] -= l0 * pj;
        if (u1) a1[j] -= l1 * pj;
      }
    }
    if (r0 == prow) done0 = true;
    if (r1 == prow) done1 = true;
    if (l == (prow & 63)) sstep[k] = prow;
    __syncthreads();
  }
#pragma unroll 1
  for (int e = l; e < N1 * N1; e += 64) {
    const int j = e / N1, st = e - (e / N1) * N1;   // by columns: column j, factor row st
    mat[MS::LU + e] = sm[sstep[st] * LD + j];
  }
#pragma unroll 1
  for (int st = l; st < N1; st += 64) mat.perm(MS::PERM, st) = sstep[st];
  if (l == 0) w[S::SING] = (double)sing;
  __syncthreads();   // the LDS is reused by the next cell of this block
}

// The same factorisation on two waves per cell, one row per thread (thread t holds row
// t in registers, the column loop unrolled at compile time): each step's pivot from the
// two waves' 64-bit maxima exchanged through LDS (ties to the lower row), the pivot row
// broadcast through LDS, both double-buffered by step parity so one barrier pair per
// column suffices.  Factor row s is written to column positions LU + j N1 + s.
template <class M>
__device__ __forceinline__ void lu_wide2_one(const ReactArgs& a, const int* __restrict__ list, const long n,
                                             const long idx) {
  constexpr int N1 = M::N1;
  using S = Slots<N1>;
  using MS = MatSlots<N1>;
  static_assert(N1 <= 128, "one row per thread of two waves");
  __shared__ double prow_buf[2][N1];
  __shared__ unsigned long long wkey[2][2];
  __shared__ int wwho[2][2];
  if (idx >= n) return;   // uniform over the block (one cell per block)
  const long c = list[idx];
  const WS w = cell_ws<M>(a.ws, c);
  if (LANE_FIELD(M, w, setup_pending) == 0) return;
  const double gamma = LANE_FIELD(M, w, gamma);
  const CM mat = cell_mat<M>(a.mat, c);
  const int t = threadIdx.x, wave = t >> 6;
  const bool row = t < N1;
  double m[N1];
#pragma unroll
  for (int j = 0; j < N1; ++j) m[j] = row ? (t == j ? 1.0 : 0.0) - gamma * mat[MS::JS + t * N1 + j] : 0.0;
  bool done = !row;
  int my_step = 0, sing = 0;
  static_for<N1>([&](auto kc) {
    constexpr int k = decltype(kc)::value;
    constexpr int par = k & 1;
    unsigned long long key = 0ull;
    if (!done) {
      double av = fabs(m[k]);
      if (av != av) av = HUGE_VAL;
      key = (unsigned long long)__double_as_longlong(av) + 1ull;
    }
    const unsigned long long kmax = group_max_u64<64>(key);
    const unsigned long long ball = __ballot(key == kmax);
    if ((t & 63) == 0) {
      wkey[par][wave] = kmax;
      wwho[par][wave] = wave * 64 + __ffsll((long long)ball) - 1;
    }
    __syncthreads();
    const unsigned long long k0 = wkey[par][0], k1 = wkey[par][1];
    const int prow = (k0 >= k1) ? wwho[par][0] : wwho[par][1];   // rows of wave 0 are the lower ones
    const unsigned long long gmax = k0 >= k1 ? k0 : k1;
    const bool nonzero = gmax > 1ull;
    if (!nonzero && sing == 0) sing = k + 1;
    if (t == prow) {
#pragma unroll
      for (int j = k; j < N1; ++j) prow_buf[par][j] = m[j];
    }
    __syncthreads();
    if (!done && t != prow && nonzero) {
      const double l = m[k] * (1.0 / prow_buf[par][k]);
      m[k] = l;
      if (l != 0.0) {
#pragma unroll
        for (int j = k + 1; j < N1; ++j) m[j] -= l * prow_buf[par][j];
      }
    }
    if (t == prow && !done) {
      done = true;
      my_step = k;
    }
  });
  if (row) {
#pragma unroll
    for (int j = 0; j < N1; ++j) mat[MS::LU + j * N1 + my_step] = m[j];   // by columns
    mat.perm(MS::PERM, my_step) = t;
  }
  if (t == 0) w[S::SING] = (double)sing;
  __syncthreads();   // the LDS is reused by the next cell of this block
}

template <class M, bool STRIDE>
__global__ void __launch_bounds__(128) k_lu_wide2(const ReactArgs a, const int* __restrict__ list,
                                                  const unsigned int* __restrict__ count) {
  const long n = (long)*count;
  if (STRIDE) {
    for (long blk = xcd_block(); blk < n; blk += gridDim.x) lu_wide2_one<M>(a, list, n, blk);
  } else {
    lu_wide2_one<M>(a, list, n, xcd_block());
  }
}


template <class M, bool STRIDE>
__global__ void __launch_bounds__(64) k_lu_wide(const ReactArgs a, const int* __restrict__ list,
                                                const unsigned int* __restrict__ count) {
  const long n = (long)*count;
  if (STRIDE) {
    for (long blk = xcd_block(); blk < n; blk += gridDim.x) lu_wide_one<M>(a, list, n, blk);
  } else {
    lu_wide_one<M>(a, list, n, xcd_block());
  }
}

// The Newton solves for these cells: x = M^-1 B with the factors above, lane l holding
// components l and l + 64; forward and back substitution column by column, each finished
// component broadcast with v_readlane (each column of the factors read from the matrix
// region as 64 consecutive words: one load covers a few whole lines, not a line per lane).
template <class M>
__device__ __forceinline__ void solve_wide_one(const ReactArgs& a, const int* __restrict__ list, const long n,
                                               const long idx) {
  constexpr int N1 = M::N1;
  using MS = MatSlots<N1>;
  if (idx >= n) return;   // uniform over the block
  const long c = list[idx];
  const WS w = cell_ws<M>(a.ws, c);
  if (!solve_due<M>(w)) return;
  const CM mat = cell_mat<M>(a.mat, c);
  const int l = threadIdx.x;
  const int r0 = l, r1 = l + 64;
  const bool has0 = r0 < N1, has1 = r1 < N1;
  double y0 = has0 ? mat[MS::B + mat.perm(MS::PERM, r0)] : 0.0;
  double y1 = has1 ? mat[MS::B + mat.perm(MS::PERM, r1)] : 0.0;
  // eight columns of the lane's rows are loaded ahead of their eight dependent steps; the
  // factors are stored by columns, so a load reads 64 consecutive words
  const double* f0 = &mat[MS::LU + (has0 ? r0 : 0)];
  const double* f1 = &mat[MS::LU + (has1 ? r1 : 0)];
#pragma unroll 1
  for (int k0 = 0; k0 < N1 - 1; k0 += 8) {
    double c0[8], c1[8];
#pragma unroll
    for (int t = 0; t < 8; ++t) {
      const int k = k0 + t;
      c0[t] = (has0 && r0 > k && k < N1 - 1) ? f0[k * N1] : 0.0;
      c1[t] = (has1 && r1 > k && k < N1 - 1) ? f1[k * N1] : 0.0;
    }
#pragma unroll
    for (int t = 0; t < 8; ++t) {
      const int k = k0 + t;
      if (k < N1 - 1) {
        const double yk = lane_read(k < 64 ? y0 : y1, k & 63);
        if (c0[t] != 0.0) y0 -= c0[t] * yk;
        if (c1[t] != 0.0) y1 -= c1[t] * yk;
      }
    }
  }
#pragma unroll 1
  for (int k0 = N1 - 1; k0 >= 0; k0 -= 8) {
    double c0[8], c1[8];
#pragma unroll
    for (int t = 0; t < 8; ++t) {
      const int k = k0 - t;   // the diagonal for the owner, U entries for the rows above
      c0[t] = (has0 && r0 <= k && k >= 0) ? f0[k * N1] : 0.0;
      c1[t] = (has1 && r1 <= k && k >= 0) ? f1[k * N1] : 0.0;
    }
#pragma unroll
    for (int t = 0; t < 8; ++t) {
      const int k = k0 - t;
      if (k >= 0) {
        // the owner divides by the pivot; the quotient goes to every lane
        if (r0 == k) y0 = y0 / c0[t];
        if (r1 == k) y1 = y1 / c1[t];
        const double xk = lane_read(k < 64 ? y0 : y1, k & 63);
        if (r0 < k && c0[t] != 0.0) y0 -= c0[t] * xk;
        if (r1 < k && c1[t] != 0.0) y1 -= c1[t] * xk;
      }
    }
  }
  if (has0) mat[MS::X + r0] = y0;
  if (has1) mat[MS::X + r1] = y1;
}

template <class M>
__global__ void __launch_bounds__(64) k_solve_wide(const ReactArgs a, const int* __restrict__ list,
                                                   const unsigned int* __restrict__ count) {
  const long n = (long)*count;
  for (long blk = xcd_block(); blk < n; blk += gridDim.x) solve_wide_one<M>(a, list, n, blk);
}

template <int N1>
inline unsigned coop_grid(long n, bool exact) {
  constexpr int CPW = Coop<N1>::CPW;
  if (!exact) return 8192;
  const long want = ((n + CPW - 1) / CPW + 7) / 8 * 8;
  return (unsigned)(want > 0 ? want : 8);
}

// exact = true: `n` is the list's length; false: an upper bound (graph capture)
template <class M>
hipError_t launch_lu(const ReactArgs& a, const int* list, const unsigned int* count, long n, hipStream_t stream,
                     bool exact = true) {
  if (a.lu_lane) {
    hipLaunchKernelGGL(k_lu_lane<M>, dim3((unsigned)((n + 63) / 64)), dim3(64), 0, stream, a, list, count);
    return hipGetLastError();
  }
  if constexpr (M::N1 > COOP_MAX_N1) {   // one cell per block; a multiple of 8 blocks (XCD mapping)
    const dim3 g(exact ? (unsigned)(n > 0 ? ((n + 7) / 8) * 8 : 8) : 8192u);
    if (a.lu_mfma) {   // the staged-in-LDS form (mi.lu_form = mfma selects it above 64 unknowns)
      if (exact)
        hipLaunchKernelGGL((k_lu_wide<M, false>), g, dim3(64), 0, stream, a, list, count);
      else
        hipLaunchKernelGGL((k_lu_wide<M, true>), g, dim3(64), 0, stream, a, list, count);
    } else {
      if (exact)
        hipLaunchKernelGGL((k_lu_wide2<M, false>), g, dim3(128), 0, stream, a, list, count);
      else
        hipLaunchKernelGGL((k_lu_wide2<M, true>), g, dim3(128), 0, stream, a, list, count);
    }
  } else {
  if (a.lu_mfma) {   // one cell per block; a multiple of 8 blocks (XCD mapping)
    const dim3 g(exact ? (unsigned)(n > 0 ? ((n + 7) / 8) * 8 : 8) : 8192u);
    if (a.lu_mode == LU_FP32_IR) {
      if (exact)
        hipLaunchKernelGGL((k_lu_mfma<M, float, false>), g, dim3(64), 0, stream, a, list, count);
      else
        hipLaunchKernelGGL((k_lu_mfma<M, float, true>), g, dim3(64), 0, stream, a, list, count);
      const hipError_t e = hipGetLastError();
      if (e != hipSuccess) return e;
    }
    if (exact)
      hipLaunchKernelGGL((k_lu_mfma<M, double, false>), g, dim3(64), 0, stream, a, list, count);
    else
      hipLaunchKernelGGL((k_lu_mfma<M, double, true>), g, dim3(64), 0, stream, a, list, count);
    return hipGetLastError();
  }
  const dim3 grid(coop_grid<M::N1>(n, exact));
  if (a.lu_mode == LU_FP32_IR) {
    if (exact)
      hipLaunchKernelGGL((k_lu_coop<M, float, false>), grid, dim3(64), 0, stream, a, list, count);
    else
      hipLaunchKernelGGL((k_lu_coop<M, float, true>), grid, dim3(64), 0, stream, a, list, count);
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
  }
  // FP64 factorisations: every pending cell, or under LU_FP32_IR the cells that fell back
  if (exact)
    hipLaunchKernelGGL((k_lu_coop<M, double, false>), grid, dim3(64), 0, stream, a, list, count);
  else
    hipLaunchKernelGGL((k_lu_coop<M, double, true>), grid, dim3(64), 0, stream, a, list, count);
  }
  return hipGetLastError();
}

// n bounds the solve list's length (the list is built by this trip's evaluation kernel),
// so the grid strides
template <class M>
hipError_t launch_solve(const ReactArgs& a, long n, hipStream_t stream) {
  if (a.gmres) {
    hipLaunchKernelGGL(k_solve_gmres<M>, dim3((unsigned)((n + 63) / 64)), dim3(64), 0, stream, a, a.nlist,
                       &a.counts[6]);
    return hipGetLastError();
  }
  if (a.lu_lane) {
    hipLaunchKernelGGL(k_solve_lane<M>, dim3((unsigned)((n + 63) / 64)), dim3(64), 0, stream, a, a.nlist,
                       &a.counts[6]);
    return hipGetLastError();
  }
  if constexpr (M::N1 > COOP_MAX_N1) {   // one cell per block over the solve list
    const long nb = n < 8192 ? n : 8192;
    hipLaunchKernelGGL(k_solve_wide<M>, dim3((unsigned)(nb > 0 ? ((nb + 7) / 8) * 8 : 8)), dim3(64), 0, stream, a,
                       a.nlist, &a.counts[6]);
  } else {
    const dim3 grid(coop_grid<M::N1>(n, true));
    if (a.lu_mode == LU_FP32_IR) {
      hipLaunchKernelGGL((k_solve_coop<M, float, true>), grid, dim3(64), 0, stream, a, a.nlist, &a.counts[6]);
      const hipError_t e = hipGetLastError();
      if (e != hipSuccess) return e;
    }
    hipLaunchKernelGGL((k_solve_coop<M, double, true>), grid, dim3(64), 0, stream, a, a.nlist, &a.counts[6]);
  }
  return hipGetLastError();
}

// counts[0..1]: active-cell lists of the current / next trip, counts[2..3]: their set-up
// lists; without compaction counts[4] is the (fixed) list length and counts[5] the
// number of cells still integrating; counts[6]: cells due a Newton solve this trip;
// counts[7]: cells awaiting a learned prediction.
// Per-kernel timing: an event pair per launch, recorded asynchronously and read back
// once the call has finished, so profiling adds no synchronisation inside the trips.
struct Timer {
  KernelProf* prof;
  hipStream_t stream;
  static constexpr int MAXEV = 4096;
  hipEvent_t ev[2 * MAXEV];
  int kind[MAXEV];
  long cells[MAXEV];
  int n = 0, created = 0;
  hipError_t begin() {
    if (!prof || n >= MAXEV) return hipSuccess;
    while (created < 2 * (n + 1)) PELE_TRY(hipEventCreate(&ev[created++]));
    return hipEventRecord(ev[2 * n], stream);
  }
  hipError_t end(int k, long c) {
    if (!prof || n >= MAXEV) return hipSuccess;
    PELE_TRY(hipEventRecord(ev[2 * n + 1], stream));
    kind[n] = k;
    cells[n] = c;
    ++n;
    return hipSuccess;
  }
  hipError_t collect() {
    if (!prof || n == 0) return hipSuccess;
    PELE_TRY(hipEventSynchronize(ev[2 * n - 1]));
    for (int i = 0; i < n; ++i) {
      float ms = 0.f;
      PELE_TRY(hipEventElapsedTime(&ms, ev[2 * i], ev[2 * i + 1]));
      prof->ms[kind[i]] += ms;
      prof->launches[kind[i]] += 1;
      prof->cells[kind[i]] += cells[i];
    }
    return hipSuccess;
  }
  ~Timer() {
    for (int i = 0; i < created; ++i) (void)hipEventDestroy(ev[i]);
  }
};

template <class M>
hipError_t launch_predict(const ReactArgs& a, const MlpDev& w, unsigned np, hipStream_t stream) {
  if constexpr (M::NS + 3 > 64 || M::N1 > 64) {   // the network's layers hold at most 64 inputs and outputs
    return hipErrorNotSupported;
  } else {
    constexpr int KIN = (M::NS + 3 <= 32) ? 32 : 64;
    constexpr int NOUT = (M::N1 <= 32) ? 32 : 64;
    if (w.n_in != M::NS + 3 || w.n_out != M::N1 || w.kp[0] != KIN || w.np[3] != NOUT) return hipErrorInvalidValue;
    hipLaunchKernelGGL((k_predict<M, KIN, NOUT>), dim3((np + MLP_ROWS - 1) / MLP_ROWS), dim3(MLP_THREADS), 0, stream,
                       a, w, a.plist, &a.counts[7]);
    return hipGetLastError();
  }
}

// One trip with grids sized for every cell (kernels skip beyond their list counts), as
// captured into a graph: evaluation, set-up, factorisation, solves, step logic.
template <class M>
hipError_t enqueue_fixed_trip(const ReactArgs& a, int cur, hipStream_t s) {
  const long n = a.ncell;
  const unsigned nb = (unsigned)((n + 63) / 64);
  const int nxt = 1 - cur;
  if (a.compact)
    hipLaunchKernelGGL(k_zero_two, dim3(1), dim3(64), 0, s, a.counts, nxt, 2 + nxt);
  else
    hipLaunchKernelGGL(k_zero_two, dim3(1), dim3(64), 0, s, a.counts, 5, 2 + nxt);
  PELE_TRY(hipGetLastError());
  hipLaunchKernelGGL(k_zero_two, dim3(1), dim3(64), 0, s, a.counts, 6, 6);
  PELE_TRY(hipGetLastError());
  const int* act = a.compact ? a.list[cur] : a.list[1];
  const unsigned* act_n = a.compact ? &a.counts[cur] : &a.counts[4];
  hipLaunchKernelGGL(k_eval<M>, dim3(nb), dim3(64), 0, s, a, act, act_n);
  PELE_TRY(hipGetLastError());
  hipLaunchKernelGGL(k_setup<M>, dim3(nb), dim3(64), 0, s, a, a.slist[cur], &a.counts[2 + cur]);
  PELE_TRY(hipGetLastError());
  PELE_TRY(launch_lu<M>(a, a.slist[cur], &a.counts[2 + cur], n, s, false));
  // the cooperative solve strides over its list from a fixed grid; the one-lane solve
  // takes one cell per lane, so its grid covers every cell the list can hold
  PELE_TRY(launch_solve<M>(a, (a.lu_lane || a.gmres) ? n : 16384, s));
  if (a.compact)
    hipLaunchKernelGGL(k_consume<M>, dim3(nb), dim3(64), 0, s, a, act, act_n, a.list[nxt], &a.counts[nxt],
                       a.slist[nxt], &a.counts[2 + nxt]);
  else
    hipLaunchKernelGGL(k_consume<M>, dim3(nb), dim3(64), 0, s, a, act, act_n, a.list[0], &a.counts[5],
                       a.slist[nxt], &a.counts[2 + nxt]);
  return hipGetLastError();
}

// trips issued as a captured graph of plan.graph_trips trips, relaunched until done
template <class M>
hipError_t run_trips_graph(const ReactArgs& a, hipStream_t stream, long max_trips, long* trips_out,
                           const TripPlan& plan) {
  const int K = plan.graph_trips;
  if (K <= 0 || (K & 1) || !plan.capture_stream) return hipErrorInvalidValue;
  hipGraph_t g = nullptr;
  hipGraphExec_t ge = nullptr;
  PELE_TRY(hipStreamBeginCapture(plan.capture_stream, hipStreamCaptureModeThreadLocal));
  hipError_t e = hipSuccess;
  for (int t = 0; t < K && e == hipSuccess; ++t) e = enqueue_fixed_trip<M>(a, t & 1, plan.capture_stream);
  const hipError_t ec = hipStreamEndCapture(plan.capture_stream, &g);
  if (e != hipSuccess || ec != hipSuccess) {
    if (g) (void)hipGraphDestroy(g);
    return e != hipSuccess ? e : ec;
  }
  e = hipGraphInstantiate(&ge, g, nullptr, nullptr, 0);
  long trip = 0;
  unsigned int h_counts[8];
  while (e == hipSuccess && trip < max_trips) {
    e = hipMemcpyAsync(h_counts, a.counts, sizeof(h_counts), hipMemcpyDeviceToHost, stream);
    if (e == hipSuccess) e = hipStreamSynchronize(stream);
    if (e != hipSuccess) break;
    if ((a.compact ? h_counts[0] : h_counts[5]) == 0) break;
    e = hipGraphLaunch(ge, stream);
    trip += K;
  }
  if (ge) (void)hipGraphExecDestroy(ge);
  (void)hipGraphDestroy(g);
  PELE_TRY(e);
  if (trip >= max_trips) {
    const unsigned nb = (unsigned)((a.ncell + 63) / 64);
    if (a.compact)
      hipLaunchKernelGGL(k_abandon<M>, dim3(nb), dim3(64), 0, stream, a, a.list[0], &a.counts[0]);
    else
      hipLaunchKernelGGL(k_abandon<M>, dim3(nb), dim3(64), 0, stream, a, a.list[1], &a.counts[4]);
    PELE_TRY(hipGetLastError());
  }
  if (trips_out) *trips_out = trip;
  return hipSuccess;
}

template <class M>
hipError_t run_react_t(ReactArgs a, hipStream_t stream, long max_trips, long* trips_out, KernelProf* prof,
                       const MlpDev* pred, const TripPlan& plan) {
  const long n = a.ncell;
  if (a.use_pred && (!pred || !a.compact || !a.plist)) return hipErrorInvalidValue;
  const bool graph = plan.graph_trips > 0 && !prof && !a.use_pred;
  Timer* tmp = new Timer();
  tmp->prof = prof;
  tmp->stream = stream;
  Timer& tm = *tmp;
  struct Del {
    Timer* t;
    ~Del() { delete t; }
  } del_{tmp};
  const unsigned nb = (unsigned)((n + 63) / 64);
  hipLaunchKernelGGL(k_zero_counts, dim3(1), dim3(64), 0, stream, a.counts);
  PELE_TRY(hipGetLastError());
  hipLaunchKernelGGL(k_init<M>, dim3(nb), dim3(64), 0, stream, a);
  PELE_TRY(hipGetLastError());
  if (!a.compact) {
    hipLaunchKernelGGL(k_identity, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, stream, a.list[1], n);
    PELE_TRY(hipGetLastError());
    hipLaunchKernelGGL(k_plain_counts, dim3(1), dim3(64), 0, stream, a.counts, (unsigned int)n);
    PELE_TRY(hipGetLastError());
  }
  if (graph) return run_trips_graph<M>(a, stream, max_trips, trips_out, plan);
  int cur = 0;
  long trip = 0;
  unsigned int h_counts[8];
  for (; trip < max_trips; ++trip) {
    PELE_TRY(hipMemcpyAsync(h_counts, a.counts, sizeof(h_counts), hipMemcpyDeviceToHost, stream));
    PELE_TRY(hipStreamSynchronize(stream));
    if (a.compact) {
      const unsigned int active = h_counts[cur];
      const unsigned int nsetup = h_counts[2 + cur];
      if (active == 0) break;
      const int nxt = 1 - cur;
      hipLaunchKernelGGL(k_zero_two, dim3(1), dim3(64), 0, stream, a.counts, nxt, 2 + nxt);
      PELE_TRY(hipGetLastError());
      hipLaunchKernelGGL(k_zero_two, dim3(1), dim3(64), 0, stream, a.counts, 6, 6);
      PELE_TRY(hipGetLastError());
      const unsigned ga = (active + 63) / 64;
      if (a.use_pred && h_counts[7] > 0) {
        const unsigned np = h_counts[7];
        PELE_TRY(tm.begin());
        PELE_TRY(launch_predict<M>(a, *pred, np, stream));
        PELE_TRY(tm.end(KK_PRED, np));
        hipLaunchKernelGGL(k_zero_two, dim3(1), dim3(64), 0, stream, a.counts, 7, 7);
        PELE_TRY(hipGetLastError());
      }
      PELE_TRY(tm.begin());
      hipLaunchKernelGGL(k_eval<M>, dim3(ga), dim3(64), 0, stream, a, a.list[cur], &a.counts[cur]);
      PELE_TRY(hipGetLastError());
      PELE_TRY(tm.end(KK_EVAL, active));
      if (nsetup > 0) {
        PELE_TRY(tm.begin());
        hipLaunchKernelGGL(k_setup<M>, dim3((nsetup + 63) / 64), dim3(64), 0, stream, a, a.slist[cur],
                           &a.counts[2 + cur]);
        PELE_TRY(hipGetLastError());
        PELE_TRY(tm.end(KK_SETUP, nsetup));
        PELE_TRY(tm.begin());
        PELE_TRY(launch_lu<M>(a, a.slist[cur], &a.counts[2 + cur], nsetup, stream));
        PELE_TRY(tm.end(KK_LU, nsetup));
      }
      PELE_TRY(tm.begin());
      PELE_TRY(launch_solve<M>(a, active, stream));
      PELE_TRY(tm.end(KK_SOLVE, active));
      PELE_TRY(tm.begin());
      hipLaunchKernelGGL(k_consume<M>, dim3(ga), dim3(64), 0, stream, a, a.list[cur], &a.counts[cur], a.list[nxt],
                         &a.counts[nxt], a.slist[nxt], &a.counts[2 + nxt]);
      PELE_TRY(hipGetLastError());
      PELE_TRY(tm.end(KK_CONSUME, active));
      cur = nxt;
    } else {
      if (h_counts[5] == 0) break;
      const unsigned int nsetup = h_counts[2 + cur];
      const int nxt = 1 - cur;
      hipLaunchKernelGGL(k_zero_two, dim3(1), dim3(64), 0, stream, a.counts, 5, 2 + nxt);
      PELE_TRY(hipGetLastError());
      hipLaunchKernelGGL(k_zero_two, dim3(1), dim3(64), 0, stream, a.counts, 6, 6);
      PELE_TRY(hipGetLastError());
      PELE_TRY(tm.begin());
      hipLaunchKernelGGL(k_eval<M>, dim3(nb), dim3(64), 0, stream, a, a.list[1], &a.counts[4]);
      PELE_TRY(hipGetLastError());
      PELE_TRY(tm.end(KK_EVAL, h_counts[5]));
      if (nsetup > 0) {
        PELE_TRY(tm.begin());
        hipLaunchKernelGGL(k_setup<M>, dim3((nsetup + 63) / 64), dim3(64), 0, stream, a, a.slist[cur],
                           &a.counts[2 + cur]);
        PELE_TRY(hipGetLastError());
        PELE_TRY(tm.end(KK_SETUP, nsetup));
        PELE_TRY(tm.begin());
        PELE_TRY(launch_lu<M>(a, a.slist[cur], &a.counts[2 + cur], nsetup, stream));
        PELE_TRY(tm.end(KK_LU, nsetup));
      }
      PELE_TRY(tm.begin());
      PELE_TRY(launch_solve<M>(a, h_counts[5], stream));
      PELE_TRY(tm.end(KK_SOLVE, h_counts[5]));
      PELE_TRY(tm.begin());
      hipLaunchKernelGGL(k_consume<M>, dim3(nb), dim3(64), 0, stream, a, a.list[1], &a.counts[4], a.list[0],
                         &a.counts[5], a.slist[nxt], &a.counts[2 + nxt]);
      PELE_TRY(hipGetLastError());
      PELE_TRY(tm.end(KK_CONSUME, h_counts[5]));
      cur = nxt;
    }
  }
  if (trip == max_trips) {
    if (a.compact)
      hipLaunchKernelGGL(k_abandon<M>, dim3(nb), dim3(64), 0, stream, a, a.list[cur], &a.counts[cur]);
    else
      hipLaunchKernelGGL(k_abandon<M>, dim3(nb), dim3(64), 0, stream, a, a.list[1], &a.counts[4]);
    PELE_TRY(hipGetLastError());
  }
  if (trips_out) *trips_out = trip;
  PELE_TRY(tm.collect());
  return hipSuccess;
}

template <class M, int OCC>
hipError_t launch_persist_occ(const ReactArgs& a, const int* queue, const PersistPlan& plan, long max_trips,
                              hipStream_t s) {
  if constexpr (M::N1 > COOP_MAX_N1) {   // the persistent kernel's linear algebra is cooperative
    return hipErrorNotSupported;
  } else {
    hipLaunchKernelGGL((k_persist<M, OCC>), dim3((unsigned)plan.waves), dim3(64), 0, s, a, queue, &a.counts[0],
                       &a.counts[1], &a.counts[2], plan.chunk, max_trips);
    return hipGetLastError();
  }
}

// counts[0]: queue length, [1]: queue head, [2]: most iterations of a wave
template <class M>
hipError_t run_react_persist_t(const ReactArgs& a, const PersistPlan& plan, long max_trips, hipStream_t s) {
  if (plan.waves <= 0 || plan.chunk < 1 || plan.chunk > 64 || (plan.occupancy != 1 && plan.occupancy != 2) ||
      max_trips < 1)
    return hipErrorInvalidValue;
  hipLaunchKernelGGL(k_zero_counts, dim3(1), dim3(64), 0, s, a.counts);
  PELE_TRY(hipGetLastError());
  const int* queue = nullptr;
  if (plan.order) {   // the queue: the ordered cells with mask >= 0
    hipLaunchKernelGGL(k_queue<M>, dim3((unsigned)((a.ncell + 63) / 64)), dim3(64), 0, s, a, plan.order, a.list[0],
                       &a.counts[0]);
    PELE_TRY(hipGetLastError());
    queue = a.list[0];
  }
  return plan.occupancy == 2 ? launch_persist_occ<M, 2>(a, queue, plan, max_trips, s)
                             : launch_persist_occ<M, 1>(a, queue, plan, max_trips, s);
}

template <class M>
int persist_capacity_t(int occupancy) {
  if constexpr (M::N1 > COOP_MAX_N1) {   // no persistent kernel for this mechanism
    (void)occupancy;
    return 0;
  } else {
    int dev = 0, ncu = 0, per_cu = 0;
    if (hipGetDevice(&dev) != hipSuccess) return 0;
    if (hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess) return 0;
    const hipError_t e = occupancy == 2
                             ? hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_persist<M, 2>, 64, 0)
                             : hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_persist<M, 1>, 64, 0);
    if (e != hipSuccess) return 0;
    return ncu * per_cu;
  }
}

// PeleLM::RhoH_to_Temp per cell (getTfromHY, Source/PeleLM_K.H:511-530): Y = rhoY / rho
// and h = rhoH / rho x 1e4 (MKS to erg/g), T re-solved by HY2T from the incoming T
template <class M>
__global__ void __launch_bounds__(64) rhoh_to_temp_kernel(const ArrView rho, const ArrView rhoY, const ArrView rhoH,
                                                          const ArrView T, const int nx, const int ny, const long ncell) {
  constexpr int N = M::NS;
  const long c = (long)blockIdx.x * 64 + threadIdx.x;
  if (c >= ncell) return;
  const int i = (int)(c % nx), j = (int)((c / nx) % ny), k = (int)(c / ((long)nx * ny));
  const double rhoinv = 1.0 / rho.p[cell_off(rho, i, j, k)];
  double y[N];
  const long oy = cell_off(rhoY, i, j, k);
#pragma unroll
  for (int n = 0; n < N; ++n) y[n] = rhoY.p[oy + n * rhoY.ns] * rhoinv;
  const double h_cgs = rhoH.p[cell_off(rhoH, i, j, k)] * rhoinv * 1.0e4;
  const long ot = cell_off(T, i, j, k);
  T.p[ot] = hy2t<M>(h_cgs, y, T.p[ot]);
}

template <class M>
hipError_t launch_rhoh_to_temp_t(const ArrView& rho, const ArrView& rhoY, const ArrView& rhoH, const ArrView& T,
                                 int nx, int ny, long ncell, hipStream_t stream) {
  hipLaunchKernelGGL(rhoh_to_temp_kernel<M>, dim3((unsigned)((ncell + 63) / 64)), dim3(64), 0, stream, rho, rhoY,
                     rhoH, T, nx, ny, ncell);
  return hipGetLastError();
}

template <class M>
hipError_t launch_rates_t(const RatesArgs& ra, hipStream_t stream) {
  hipLaunchKernelGGL(rates_kernel<M>, dim3((unsigned)((ra.ncell + 63) / 64)), dim3(64), 0, stream, ra.rhoY, ra.rhoH,
                     ra.T, ra.mask, ra.rdot, ra.Tout, ra.nx, ra.ny, ra.ncell, ra.hscale, ra.rscale, ra.oscale);
  return hipGetLastError();
}

template <class M>
hipError_t launch_cost_keys_t(const ReactArgs& a, int mode, unsigned int* keys, int* vals, hipStream_t stream) {
  hipLaunchKernelGGL(k_cost_key<M>, dim3((unsigned)((a.ncell + 63) / 64)), dim3(64), 0, stream, a, mode, keys, vals);
  return hipGetLastError();
}

template <class M>
hipError_t launch_rk64_t(const ReactArgs& a, const RK64Opts& o, hipStream_t stream) {
  hipLaunchKernelGGL(k_rk64<M>, dim3((unsigned)((a.ncell + 63) / 64)), dim3(64), 0, stream, a, o);
  return hipGetLastError();
}

}  // namespace

// one mechanism's entry table: its sizes and its kernels' launchers
#define PELE_DEFINE_MECH_TABLE(T, FN)                                                                            \
  const MechTable* FN() {                                                                                      \
    static const MechTable t = {                                                                               \
        T::NS, T::NAME, Slots<T::N1>::TOTAL, Slots<T::N1>::ITOTAL, MatSlots<T::N1>::STRIDE,                    \
        [](const ReactArgs& a, hipStream_t s, long mt, long* to, KernelProf* p, const MlpDev* pr, TripPlan pl) { \
          return run_react_t<T>(a, s, mt, to, p, pr, pl);                                                      \
        },                                                                                                     \
        [](const ReactArgs& a, const PersistPlan& pl, long mt, hipStream_t s) {                                \
          return run_react_persist_t<T>(a, pl, mt, s);                                                         \
        },                                                                                                     \
        [](int occ) { return persist_capacity_t<T>(occ); },                                                    \
        [](const RatesArgs& ra, hipStream_t s) { return launch_rates_t<T>(ra, s); },                           \
        [](const ReactArgs& a, int g, int mn, int mx, hipStream_t s) {                                         \
          return launch_rk64_t<T>(a, RK64Opts{g, mn, mx}, s);                                                  \
        },                                                                                                     \
        [](const ReactArgs& a, int mode, unsigned int* keys, int* vals, hipStream_t s) {                       \
          return launch_cost_keys_t<T>(a, mode, keys, vals, s);                                                \
        },                                                                                                     \
        [](const ArrView& rho, const ArrView& rhoY, const ArrView& rhoH, const ArrView& Tv, int nx, int ny,    \
           long ncell, hipStream_t s) { return launch_rhoh_to_temp_t<T>(rho, rhoY, rhoH, Tv, nx, ny, ncell, s); }}; \
    return &t;                                                                                                 \
  }

#if defined(PELE_TU_MECH)
PELE_DEFINE_MECH_TABLE(PELE_TU_MECH_TYPE, PELE_TU_MECH_FN)
#else
#if !defined(PELE_TU_COMMON)
PELE_DEFINE_MECH_TABLE(mech_LiDryer, mech_table_LiDryer)
PELE_DEFINE_MECH_TABLE(mech_drm19, mech_table_drm19)
PELE_DEFINE_MECH_TABLE(mech_dodecane53, mech_table_dodecane53)
#if PELE_HAVE_DODECANE56
PELE_DEFINE_MECH_TABLE(mech_dodecane56, mech_table_dodecane56)
#endif
#if PELE_HAVE_HEPTANE88
PELE_DEFINE_MECH_TABLE(mech_heptane88, mech_table_heptane88)
#endif
#endif

const MechTable* mech_table(int mech) {
  switch (mech) {
    case MECH_LIDRYER: return mech_table_LiDryer();
    case MECH_DRM19: return mech_table_drm19();
    case MECH_DODECANE53: return mech_table_dodecane53();
#if PELE_HAVE_DODECANE56 || defined(PELE_TU_COMMON)
    case MECH_DODECANE56: return mech_table_dodecane56();
#endif
#if PELE_HAVE_HEPTANE88 || defined(PELE_TU_COMMON)
    case MECH_HEPTANE88: return mech_table_heptane88();
#endif
    default: return nullptr;
  }
}
int mech_nspec(int mech) {
  const MechTable* t = mech_table(mech);
  return t ? t->nspec : -1;
}
const char* mech_name(int mech) {
  const MechTable* t = mech_table(mech);
  return t ? t->name : "";
}
int mech_lookup(const char* name) {
  for (int m = 0; m < NMECH; ++m) {
    const char* a = mech_name(m);
    const char* b = name;
    while (*a && *b && *a == *b) { ++a; ++b; }
    if (*a == 0 && *b == 0) return m;
  }
  return -1;
}
long ws_slots(int mech) {
  const MechTable* t = mech_table(mech);
  return t ? t->ws_slots : 0;
}
long iws_slots(int mech) {
  const MechTable* t = mech_table(mech);
  return t ? t->iws_slots : 0;
}
long mat_stride(int mech) {
  const MechTable* t = mech_table(mech);
  return t ? t->mat_stride : 0;
}

hipError_t run_react(int mech, const ReactArgs& a, hipStream_t stream, long max_trips, long* trips_out,
                     KernelProf* prof, const MlpDev* pred, TripPlan plan) {
  const MechTable* t = mech_table(mech);
  return t ? t->run_react(a, stream, max_trips, trips_out, prof, pred, plan) : hipErrorInvalidValue;
}
hipError_t run_react_persist(int mech, const ReactArgs& a, const PersistPlan& plan, long max_trips,
                             hipStream_t stream) {
  const MechTable* t = mech_table(mech);
  return t ? t->run_react_persist(a, plan, max_trips, stream) : hipErrorInvalidValue;
}
int persist_capacity(int mech, int occupancy) {
  const MechTable* t = mech_table(mech);
  return t ? t->persist_capacity(occupancy) : 0;
}
hipError_t launch_rates(int mech, const RatesArgs& ra, hipStream_t stream) {
  const MechTable* t = mech_table(mech);
  return t ? t->launch_rates(ra, stream) : hipErrorInvalidValue;
}
hipError_t launch_rk64(int mech, const ReactArgs& a, int nsub_guess, int nsub_min, int nsub_max,
                       hipStream_t stream) {
  const MechTable* t = mech_table(mech);
  return t ? t->launch_rk64(a, nsub_guess, nsub_min, nsub_max, stream) : hipErrorInvalidValue;
}
hipError_t launch_cost_keys(int mech, const ReactArgs& a, int mode, unsigned int* keys, int* vals,
                            hipStream_t stream) {
  const MechTable* t = mech_table(mech);
  return t ? t->launch_cost_keys(a, mode, keys, vals, stream) : hipErrorInvalidValue;
}
hipError_t launch_rhoh_to_temp(int mech, const ArrView& rho, const ArrView& rhoY, const ArrView& rhoH,
                               const ArrView& T, int nx, int ny, long ncell, hipStream_t stream) {
  const MechTable* t = mech_table(mech);
  return t ? t->launch_rhoh_to_temp(rho, rhoY, rhoH, T, nx, ny, ncell, stream) : hipErrorInvalidValue;
}

// PeleLM::advance_chemistry's per-box glue (Source/PeleLM.cpp:5736-5746, 5764-5771,
// 5798-5802).  In: the new state's rhoY (x 1e-3) and rhoH (x 10) from the old state, T
// copied; the forcing converted the same way into a flat cell-fastest buffer.
__global__ void glue_in_kernel(const ArrView sold, const ArrView snew, const ArrView force, double* __restrict__ fcgs,
                               const int nspec, const int nx, const int ny, const long ncell) {
  const long c = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= ncell) return;
  const int i = (int)(c % nx), j = (int)((c / nx) % ny), k = (int)(c / ((long)nx * ny));
  const long oo = cell_off(sold, i, j, k), on = cell_off(snew, i, j, k), of = cell_off(force, i, j, k);
  for (int n = 0; n < nspec; ++n) {
    snew.p[on + n * snew.ns] = sold.p[oo + n * sold.ns] * 1.0e-3;
    fcgs[n * ncell + c] = force.p[of + n * force.ns] * 1.0e-3;
  }
  snew.p[on + nspec * snew.ns] = sold.p[oo + nspec * sold.ns] * 10.0;
  fcgs[nspec * ncell + c] = force.p[of + nspec * force.ns] * 10.0;
  snew.p[on + (nspec + 1) * snew.ns] = sold.p[oo + (nspec + 1) * sold.ns];
}

// Out: rhoY (x 1e3) and rhoH (x 0.1) back to MKS, and I_R = -(rhoY_old - rhoY_new) / dt - F_Y
__global__ void glue_out_kernel(const ArrView sold, const ArrView snew, const ArrView force, const ArrView rdot,
                                const double dt_inv, const int nspec, const int nx, const int ny, const long ncell) {
  // every product rounded on its own: I_R is formed from the stored new state (the
  // reference converts, stores, then forms I_R in a separate pass), so the x 1e3 must
  // not fuse into the difference
#pragma clang fp contract(off)
  const long c = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= ncell) return;
  const int i = (int)(c % nx), j = (int)((c / nx) % ny), k = (int)(c / ((long)nx * ny));
  const long oo = cell_off(sold, i, j, k), on = cell_off(snew, i, j, k), of = cell_off(force, i, j, k);
  const long od = cell_off(rdot, i, j, k);
  for (int n = 0; n < nspec; ++n) {
    const double yn = snew.p[on + n * snew.ns] * 1.0e3;
    snew.p[on + n * snew.ns] = yn;
    rdot.p[od + n * rdot.ns] = -(sold.p[oo + n * sold.ns] - yn) * dt_inv - force.p[of + n * force.ns];
  }
  snew.p[on + nspec * snew.ns] *= 0.1;
}

hipError_t launch_glue_in(const ArrView& sold, const ArrView& snew, const ArrView& force, double* fcgs, int nspec,
                          int nx, int ny, long ncell, hipStream_t stream) {
  hipLaunchKernelGGL(glue_in_kernel, dim3((unsigned)((ncell + 255) / 256)), dim3(256), 0, stream, sold, snew, force,
                     fcgs, nspec, nx, ny, ncell);
  return hipGetLastError();
}

hipError_t launch_glue_out(const ArrView& sold, const ArrView& snew, const ArrView& force, const ArrView& rdot,
                           double dt_inv, int nspec, int nx, int ny, long ncell, hipStream_t stream) {
  hipLaunchKernelGGL(glue_out_kernel, dim3((unsigned)((ncell + 255) / 256)), dim3(256), 0, stream, sold, snew, force,
                     rdot, dt_inv, nspec, nx, ny, ncell);
  return hipGetLastError();
}

hipError_t launch_counters_reset(unsigned long long* counters, hipStream_t stream) {
  hipLaunchKernelGGL(counters_reset_kernel, dim3(1), dim3(64), 0, stream, counters);
  return hipGetLastError();
}

// flat structure-of-arrays copy of every stride-th cell of a call (m of them): rY, T, rE
// and the forcing (zero where the call has none), for the autotuner's trial runs
__global__ void gather_state_kernel(const ReactArgs a, int nspec, long stride, long m, double* __restrict__ rY,
                                    double* __restrict__ T, double* __restrict__ rE, double* __restrict__ fY,
                                    double* __restrict__ fE) {
  const long o = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (o >= m) return;
  const long c = o * stride;
  int i, j, k;
  cell_ijk(a, c, i, j, k);
  for (int s = 0; s < nspec; ++s) {
    rY[s * m + o] = a.rY.p[cell_off(a.rY, i, j, k) + s * a.rY.ns];
    fY[s * m + o] = a.rYsrc.p ? a.rYsrc.p[cell_off(a.rYsrc, i, j, k) + s * a.rYsrc.ns] : 0.0;
  }
  T[o] = a.T.p[cell_off(a.T, i, j, k)];
  rE[o] = a.rE.p[cell_off(a.rE, i, j, k)];
  fE[o] = a.rEsrc.p ? a.rEsrc.p[cell_off(a.rEsrc, i, j, k)] : 0.0;
}

hipError_t launch_gather_state(int nspec, const ReactArgs& a, long stride, long m, double* rY, double* T,
                               double* rE, double* fY, double* fE, hipStream_t stream) {
  hipLaunchKernelGGL(gather_state_kernel, dim3((unsigned)((m + 255) / 256)), dim3(256), 0, stream, a, nspec, stride,
                     m, rY, T, rE, fY, fE);
  return hipGetLastError();
}

hipError_t launch_null(int nspec, const ReactArgs& a, hipStream_t stream) {
  const long nb = (a.ncell + 255) / 256;
  hipLaunchKernelGGL(null_kernel, dim3((unsigned)nb), dim3(256), 0, stream, a, nspec);
  return hipGetLastError();
}
#endif   // !PELE_TU_MECH

}  // namespace pelelm
