// CPU oracle: variable-order BDF (CVODE algorithm) and the per-cell reactor driver.
// TEST INFRASTRUCTURE ONLY (see chem_oracle.h).
//
// Restates CVODE's BDF as the reference decks configure it
// (Exec/RegTests/FlameSheet/inputs.3d-regt:84-90: denseAJ_direct, max_order 4,
// rtol 1e-6): Nordsieck history, fixed-leading-coefficient variable-step BDF
// coefficients, Newton with at most 3 corrections and the CRDOWN/RDIV convergence
// test, Jacobian reuse (re-setup every 20 steps or when gamma moves by > 30 %,
// Jacobian re-evaluation every 51 steps or after a convergence failure), dense
// partial-pivot LU of M = I - gamma J, WRMS local error test with BIAS1/2/3 order
// selection, the initial step estimate from a y'' norm, and CV_NORMAL output
// (step past tout, interpolate the Nordsieck array back).  With solve_type = 1 the
// Newton systems are solved by scaled GMRES instead (spgmr.hpp; the reference's GPU decks,
// Exec/RegTests/FlameSheet/inputs.3d-regt_GPU:89): CVODE's matrix-free iterative
// linear-solver interface, J v by a difference quotient of the right-hand side.
#include <algorithm>
#include <cfloat>
#include <cmath>
#include <cstring>

#include "oracle_internal.hpp"
#include "spgmr.hpp"

#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

constexpr double UROUND = DBL_EPSILON;
constexpr double ADDON = 1e-6, BIAS1 = 6.0, BIAS2 = 6.0, BIAS3 = 10.0, ONEPSM = 1.000001;
constexpr double THRESH = 1.5, ETAMX1 = 1e4, ETAMX2 = 10.0, ETAMX3 = 10.0, ETAMXF = 0.2;
constexpr double ETAMIN = 0.1, ETACF = 0.25, CRDOWN = 0.3, RDIV = 2.0, DGMAX = 0.3, DGMAX_LS = 0.2;
constexpr double HLB_FACTOR = 100.0, HUB_FACTOR = 0.1, H_BIAS = 0.5, NLSCOEF = 0.1;
constexpr int SMALL_NST = 10, MXNCF = 10, MXNEF = 7, MXNEF1 = 3, SMALL_NEF = 2, LONG_WAIT = 10;
constexpr int MAXCOR = 3, MSBP = 20, MSBJ = 51, HIN_ITERS = 4;
constexpr int LMAX = 6;   // qmax + 1 with qmax <= 5

enum { FIRST_CALL, PREV_CONV_FAIL, PREV_ERR_FAIL };
enum { NO_FAILURES, FAIL_BAD_J, FAIL_OTHER };
enum { DO_ERROR_TEST = 2, PREDICT_AGAIN = 3, TRY_AGAIN = 5 };
enum { ST_OK = 0, ST_TOO_MUCH_WORK = -1, ST_TOO_MUCH_ACC = -2, ST_ERR_FAIL = -3, ST_CONV_FAIL = -4,
       ST_TOO_CLOSE = -27, ST_NONFINITE = -50 };

struct Problem {
  virtual void f(double t, const double* y, double* yd) = 0;
  virtual void jac(double t, const double* y, const double* fy, double* J) = 0;   // row-major
  virtual ~Problem() = default;
};

struct Bdf {
  Problem* P;
  int n;
  double rtol;
  double atol[ORC_NMAX];
  int qmax;
  long mxstep;

  double zn[LMAX][ORC_NMAX];
  double ewt[ORC_NMAX], acor[ORC_NMAX], y[ORC_NMAX], tempv[ORC_NMAX], ftemp[ORC_NMAX];
  double Jsave[ORC_NMAX * ORC_NMAX], Mlu[ORC_NMAX * ORC_NMAX];
  int piv[ORC_NMAX];
  double tq[6], l[LMAX], tau[LMAX + 1];
  double tn = 0, h = 0, hprime = 0, hscale = 0, eta = 1, etamax = ETAMX1;
  double gamma = 0, gammap = 0, gamrat = 1, rl1 = 1, crate = 1, delp = 0, acnrm = 0, saved_tq5 = 0;
  double etaq = 0, etaqm1 = 0, etaqp1 = 0;
  int q = 1, qprime = 1, L = 2, qwait = 2;
  long nst = 0, nfe = 0, nje = 0, nsetups = 0, nni = 0, ncfn = 0, netf = 0, nstlp = 0, nstlj = 0;
  bool jcur = false;
  // solve_type 1: GMRES (no set-up, no matrix); nli Krylov iterations, nfe_ls their
  // right-hand-side evaluations (not counted in nfe, as CVODE counts them apart)
  int gmres = 0, maxl = 5;
  long nli = 0, nfe_ls = 0;

  void F(double t, const double* yy, double* yd) { ++nfe; P->f(t, yy, yd); }

  // CVODE's iterative linear solve of M x = b, M = I - gamma J (cvLsSolve with
  // SUNLinSol_SPGMR): a right-hand side already below eplifac tq[4] in the WRMS norm
  // gives x = b on the first Newton iteration and x = 0 on later ones; otherwise GMRES
  // in the error-weight scaling to 2-norm tolerance eplifac tq[4] sqrt(n), with
  // J v = (f(tn, y + sig v) - f(tn, y)) / sig, sig = 1 / ||v||_WRMS, at the current
  // iterate y (ftemp = f(tn, y)).  A reduced but unconverged residual is accepted on the
  // first Newton iteration only.  Returns 0, or 1 (recoverable: the Newton iteration
  // fails).  b is overwritten by x.
  int gmres_solve(double* b, int mnewt) {
    constexpr double EPLIFAC = 0.05;
    const double deltar = EPLIFAC * tq[4];
    if (wrms(b) <= deltar) {
      if (mnewt > 0)
        for (int i = 0; i < n; ++i) b[i] = 0.0;
      return 0;
    }
    const double delta = deltar * std::sqrt((double)n);
    double work[ORC_NMAX], fw[ORC_NMAX], x[ORC_NMAX];
    auto atimes = [&](const double* v, double* z) {
      const double sig = 1.0 / wrms(v);
      for (int i = 0; i < n; ++i) work[i] = sig * v[i] + y[i];
      P->f(tn, work, fw);
      ++nfe_ls;
      const double siginv = 1.0 / sig;
      for (int i = 0; i < n; ++i) {
        const double jv = siginv * fw[i] - siginv * ftemp[i];
        z[i] = v[i] - gamma * jv;
      }
    };
    double rho = 0.0;
    int it = 0;
    const int r = spgmr(n, maxl, b, ewt, delta, atimes, x, &rho, &it);
    nli += it;
    if (r == SPGMR_SUCCESS || (r == SPGMR_RES_REDUCED && mnewt == 0)) {
      for (int i = 0; i < n; ++i) b[i] = x[i];
      return 0;
    }
    return 1;
  }

  void ewtset(const double* yy) {
    for (int i = 0; i < n; ++i) ewt[i] = 1.0 / (rtol * std::fabs(yy[i]) + atol[i]);
  }
  double wrms(const double* v) const {
    double s = 0.0;
    for (int i = 0; i < n; ++i) { const double p = v[i] * ewt[i]; s += p * p; }
    return std::sqrt(s / n);
  }

  // ---- dense LU with partial pivoting (first maximal |a| in the column) ----
  int lu_factor() {
    for (int k = 0; k < n; ++k) {
      int p = k;
      double amax = std::fabs(Mlu[k * n + k]);
      for (int i = k + 1; i < n; ++i) {
        const double a = std::fabs(Mlu[i * n + k]);
        if (a > amax) { amax = a; p = i; }
      }
      piv[k] = p;
      if (amax == 0.0) return k + 1;
      if (p != k)
        for (int j = 0; j < n; ++j) std::swap(Mlu[k * n + j], Mlu[p * n + j]);
      const double inv = 1.0 / Mlu[k * n + k];
      for (int i = k + 1; i < n; ++i) {
        const double lik = Mlu[i * n + k] * inv;
        Mlu[i * n + k] = lik;
        if (lik != 0.0)
          for (int j = k + 1; j < n; ++j) Mlu[i * n + j] -= lik * Mlu[k * n + j];
      }
    }
    return 0;
  }
  // rows were interchanged whole during the factorisation, so the permutation is
  // applied to b before the unit-lower solve
  void lu_solve(double* b) const {
    for (int k = 0; k < n; ++k) {
      const int p = piv[k];
      if (p != k) std::swap(b[k], b[p]);
    }
    for (int k = 0; k < n; ++k)
      for (int i = k + 1; i < n; ++i) b[i] -= Mlu[i * n + k] * b[k];
    for (int k = n - 1; k >= 0; --k) {
      double s = b[k];
      for (int j = k + 1; j < n; ++j) s -= Mlu[k * n + j] * b[j];
      b[k] = s / Mlu[k * n + k];
    }
  }

  // ---- linear solver setup: M = I - gamma J, J re-evaluated when stale ----
  int lsetup(int convfail) {
    const double dgamma = std::fabs(gamma / gammap - 1.0);
    const bool jbad = (nst == 0) || (nst >= nstlj + MSBJ) ||
                      (convfail == FAIL_BAD_J && dgamma < DGMAX_LS) || (convfail == FAIL_OTHER);
    if (jbad) {
      ++nje;
      nstlj = nst;
      jcur = true;
      P->jac(tn, y, ftemp, Jsave);
    } else {
      jcur = false;
    }
    for (int i = 0; i < n * n; ++i) Mlu[i] = -gamma * Jsave[i];
    for (int i = 0; i < n; ++i) Mlu[i * n + i] += 1.0;
    return lu_factor();
  }

  // ---- initial step (cvHin / cvUpperBoundH0 / cvYddNorm) ----
  double upper_bound_h0(double tdist) {
    double hub_inv = 0.0;
    for (int i = 0; i < n; ++i) {
      const double t1 = HUB_FACTOR * std::fabs(zn[0][i]) + atol[i];
      hub_inv = std::max(hub_inv, std::fabs(zn[1][i]) / t1);
    }
    double hub = HUB_FACTOR * tdist;
    if (hub * hub_inv > 1.0) hub = 1.0 / hub_inv;
    return hub;
  }
  double ydd_norm(double hg) {
    for (int i = 0; i < n; ++i) y[i] = zn[0][i] + hg * zn[1][i];
    F(tn + hg, y, tempv);
    for (int i = 0; i < n; ++i) tempv[i] = (tempv[i] - zn[1][i]) / hg;
    return wrms(tempv);
  }
  int hin(double tout) {
    const double tdiff = tout - tn;
    const double sign = tdiff > 0 ? 1.0 : -1.0;
    const double tdist = std::fabs(tdiff);
    const double tround = UROUND * std::max(std::fabs(tn), std::fabs(tout));
    if (tdist < 2.0 * tround) return ST_TOO_CLOSE;
    const double hlb = HLB_FACTOR * tround;
    const double hub = upper_bound_h0(tdist);
    double hg = std::sqrt(hlb * hub);
    if (hub < hlb) { h = sign * hg; return ST_OK; }
    double hnew = hg;
    for (int count1 = 1; count1 <= HIN_ITERS; ++count1) {
      const double yddnrm = ydd_norm(hg * sign);
      hnew = (yddnrm * hub * hub > 2.0) ? std::sqrt(2.0 / yddnrm) : std::sqrt(hg * hub);
      if (count1 == HIN_ITERS) break;
      const double hrat = hnew / hg;
      if (hrat > 0.5 && hrat < 2.0) break;
      if (count1 > 1 && hrat > 2.0) { hnew = hg; break; }
      hg = hnew;
    }
    double h0 = H_BIAS * hnew;
    if (h0 < hlb) h0 = hlb;
    if (h0 > hub) h0 = hub;
    h = sign * h0;
    return ST_OK;
  }

  // ---- Nordsieck array manipulation ----
  void rescale() {
    double factor = eta;
    for (int j = 1; j <= q; ++j) {
      for (int i = 0; i < n; ++i) zn[j][i] *= factor;
      factor *= eta;
    }
    h = hscale * eta;
    hscale = h;
  }
  void predict() {
    tn += h;
    for (int k = 1; k <= q; ++k)
      for (int j = q; j >= k; --j)
        for (int i = 0; i < n; ++i) zn[j - 1][i] += zn[j][i];
  }
  void restore(double saved_t) {
    tn = saved_t;
    for (int k = 1; k <= q; ++k)
      for (int j = q; j >= k; --j)
        for (int i = 0; i < n; ++i) zn[j - 1][i] -= zn[j][i];
  }
  void increase_bdf() {
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
    for (int i = 0; i < n; ++i) zn[L][i] = A1 * zn[qmax][i];
    for (int j = 2; j <= q; ++j)
      for (int i = 0; i < n; ++i) zn[j][i] += l[j] * zn[L][i];
  }
  void decrease_bdf() {
    for (int i = 0; i <= qmax; ++i) l[i] = 0.0;
    l[2] = 1.0;
    double hsum = 0.0;
    for (int j = 1; j <= q - 2; ++j) {
      hsum += tau[j];
      const double xi = hsum / hscale;
      for (int i = j + 2; i >= 2; --i) l[i] = l[i] * xi + l[i - 1];
    }
    for (int j = 2; j < q; ++j)
      for (int i = 0; i < n; ++i) zn[j][i] -= l[j] * zn[q][i];
  }
  void adjust_order(int deltaq) {
    if (q == 2 && deltaq != 1) return;   // order 2 -> 1 needs no history change
    if (deltaq == 1) increase_bdf(); else decrease_bdf();
  }
  void adjust_params() {
    if (qprime != q) {
      adjust_order(qprime - q);
      q = qprime;
      L = q + 1;
      qwait = L;
    }
    rescale();
  }

  // ---- BDF coefficients (cvSetBDF / cvSetTqBDF) ----
  void set_bdf() {
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
    // test quantities
    const double A1 = 1.0 - alpha0_hat + alpha0;
    const double A2 = 1.0 + q * A1;
    tq[2] = std::fabs(A1 / (alpha0 * A2));
    tq[5] = std::fabs(A2 * xistar_inv / (l[q] * xi_inv));
    if (qwait == 1) {
      if (q > 1) {
        const double C = xistar_inv / l[q];
        const double A3 = alpha0 + 1.0 / q;
        const double A4 = alpha0_hat + xi_inv;
        const double Cpinv = (1.0 - A4 + A3) / A3;
        tq[1] = std::fabs(C * Cpinv);
      } else {
        tq[1] = 1.0;
      }
      hsum += tau[q];
      xi_inv = h / hsum;
      const double A5 = alpha0 - 1.0 / (q + 1);
      const double A6 = alpha0_hat - xi_inv;
      const double Cppinv = (1.0 - A6 + A5) / A2;
      tq[3] = std::fabs(Cppinv / (xi_inv * (q + 2) * A5));
    }
    tq[4] = NLSCOEF / tq[2];
  }
  void set_coeffs() {
    set_bdf();
    rl1 = 1.0 / l[1];
    gamma = h * rl1;
    if (nst == 0) gammap = gamma;
    gamrat = (nst > 0) ? gamma / gammap : 1.0;
  }

  // ---- Newton iteration (cvNls + SUNNonlinSol_Newton + cvNlsConvTest) ----
  // returns 0 converged, 1 recoverable failure
  int nls(int nflag) {
    int convfail = (nflag == FIRST_CALL || nflag == PREV_ERR_FAIL) ? NO_FAILURES : FAIL_OTHER;
    bool callSetup = (nflag == PREV_CONV_FAIL) || (nflag == PREV_ERR_FAIL) || (nst == 0) ||
                     (nst >= nstlp + MSBP) || (std::fabs(gamrat - 1.0) > DGMAX);
    if (gmres) {   // no set-up function: no matrix to refresh, the rate estimate restarts
      callSetup = false;
      crate = 1.0;
    }
    for (;;) {
      for (int i = 0; i < n; ++i) { acor[i] = 0.0; y[i] = zn[0][i]; }
      F(tn, y, ftemp);
      if (callSetup) {
        const int ier = lsetup(convfail);
        ++nsetups;
        gamrat = 1.0;
        gammap = gamma;
        crate = 1.0;
        nstlp = nst;
        if (ier != 0) return 1;
      }
      int m = 0;
      int ret = 1;
      for (;;) {
        ++nni;
        for (int i = 0; i < n; ++i) tempv[i] = gamma * ftemp[i] - rl1 * zn[1][i] - acor[i];
        if (gmres) {
          if (gmres_solve(tempv, m) != 0) { ret = 1; break; }
        } else {
          lu_solve(tempv);
        }
        if (!gmres && gamrat != 1.0) {
          const double s = 2.0 / (1.0 + gamrat);
          for (int i = 0; i < n; ++i) tempv[i] *= s;
        }
        for (int i = 0; i < n; ++i) acor[i] += tempv[i];
        const double del = wrms(tempv);
        if (m > 0) crate = std::max(CRDOWN * crate, del / delp);
        const double dcon = del * std::min(1.0, crate) / tq[4];
        if (dcon <= 1.0) {
          acnrm = (m == 0) ? del : wrms(acor);
          jcur = false;
          for (int i = 0; i < n; ++i) y[i] = zn[0][i] + acor[i];
          return 0;
        }
        if (m >= 1 && del > RDIV * delp) { ret = 1; break; }
        delp = del;
        ++m;
        if (m >= MAXCOR) { ret = 1; break; }
        for (int i = 0; i < n; ++i) y[i] = zn[0][i] + acor[i];
        F(tn, y, ftemp);
      }
      if (ret > 0 && !jcur && !gmres) {
        callSetup = true;
        convfail = FAIL_BAD_J;
        continue;
      }
      return ret;
    }
  }

  int handle_nflag(int* nflag, double saved_t, int* ncf) {
    ++ncfn;
    restore(saved_t);
    ++(*ncf);
    etamax = 1.0;
    if (*ncf == MXNCF) return ST_CONV_FAIL;
    eta = ETACF;
    *nflag = PREV_CONV_FAIL;
    rescale();
    return PREDICT_AGAIN;
  }

  int error_test(int* nflag, double saved_t, int* nef, double* dsm) {
    *dsm = acnrm * tq[2];
    if (*dsm <= 1.0) return ST_OK;
    ++(*nef);
    ++netf;
    *nflag = PREV_ERR_FAIL;
    restore(saved_t);
    if (*nef == MXNEF) return ST_ERR_FAIL;
    etamax = 1.0;
    if (*nef <= MXNEF1) {
      eta = 1.0 / (std::pow(BIAS2 * (*dsm), 1.0 / L) + ADDON);
      eta = std::max(ETAMIN, eta);
      if (*nef >= SMALL_NEF) eta = std::min(eta, ETAMXF);
      rescale();
      return TRY_AGAIN;
    }
    if (q > 1) {
      eta = ETAMIN;
      adjust_order(-1);
      L = q;
      --q;
      qwait = L;
      rescale();
      return TRY_AGAIN;
    }
    eta = ETAMIN;
    h *= eta;
    hscale = h;
    qwait = LONG_WAIT;
    F(tn, zn[0], tempv);
    for (int i = 0; i < n; ++i) zn[1][i] = h * tempv[i];
    return TRY_AGAIN;
  }

  void complete_step() {
    ++nst;
    for (int i = q; i >= 2; --i) tau[i] = tau[i - 1];
    if (q == 1 && nst > 1) tau[2] = tau[1];
    tau[1] = h;
    for (int j = 0; j <= q; ++j)
      for (int i = 0; i < n; ++i) zn[j][i] += l[j] * acor[i];
    --qwait;
    if (qwait == 1 && q != qmax) {
      for (int i = 0; i < n; ++i) zn[qmax][i] = acor[i];
      saved_tq5 = tq[5];
    }
  }

  void set_eta() {
    if (eta < THRESH) {
      eta = 1.0;
      hprime = h;
    } else {
      eta = std::min(eta, etamax);
      hprime = h * eta;
    }
  }
  double compute_etaqm1() {
    etaqm1 = 0.0;
    if (q > 1) {
      const double ddn = wrms(zn[q]) * tq[1];
      etaqm1 = 1.0 / (std::pow(BIAS1 * ddn, 1.0 / q) + ADDON);
    }
    return etaqm1;
  }
  double compute_etaqp1() {
    etaqp1 = 0.0;
    if (q != qmax) {
      if (saved_tq5 == 0.0) return etaqp1;
      const double cquot = (tq[5] / saved_tq5) * std::pow(h / tau[2], double(L));
      for (int i = 0; i < n; ++i) tempv[i] = acor[i] - cquot * zn[qmax][i];
      const double dup = wrms(tempv) * tq[3];
      etaqp1 = 1.0 / (std::pow(BIAS3 * dup, 1.0 / (L + 1)) + ADDON);
    }
    return etaqp1;
  }
  void choose_eta() {
    const double etam = std::max(etaqm1, std::max(etaq, etaqp1));
    if (etam < THRESH) {
      eta = 1.0;
      qprime = q;
      return;
    }
    if (etam == etaq) {
      eta = etaq;
      qprime = q;
    } else if (etam == etaqm1) {
      eta = etaqm1;
      qprime = q - 1;
    } else {
      eta = etaqp1;
      qprime = q + 1;
      for (int i = 0; i < n; ++i) zn[qmax][i] = acor[i];
    }
  }
  void prepare_next_step(double dsm) {
    if (etamax == 1.0) {
      qwait = std::max(qwait, 2);
      qprime = q;
      hprime = h;
      eta = 1.0;
      return;
    }
    etaq = 1.0 / (std::pow(BIAS2 * dsm, 1.0 / L) + ADDON);
    if (qwait != 0) {
      eta = etaq;
      qprime = q;
      set_eta();
      return;
    }
    qwait = 2;
    compute_etaqm1();
    compute_etaqp1();
    choose_eta();
    set_eta();
  }

  int step() {
    const double saved_t = tn;
    int ncf = 0, nef = 0, nflag = FIRST_CALL;
    double dsm = 0.0;
    if (nst > 0 && hprime != h) adjust_params();
    for (;;) {
      predict();
      set_coeffs();
      const int nret = nls(nflag);
      if (nret != 0) {
        const int k = handle_nflag(&nflag, saved_t, &ncf);
        if (k == PREDICT_AGAIN) continue;
        return k;
      }
      const int e = error_test(&nflag, saved_t, &nef, &dsm);
      if (e == TRY_AGAIN) continue;
      if (e != ST_OK) return e;
      break;
    }
    complete_step();
    prepare_next_step(dsm);
    etamax = (nst <= SMALL_NST) ? ETAMX2 : ETAMX3;
    return ST_OK;
  }

  // integrate from 0 to tout in CV_NORMAL mode; result in yout
  int integrate(const double* y0, double tout, double* yout) {
    tn = 0.0;
    q = 1; L = 2; qwait = L; etamax = ETAMX1;
    nst = nfe = nje = nsetups = nni = ncfn = netf = 0;
    nli = nfe_ls = 0;
    nstlp = nstlj = 0;
    saved_tq5 = 0.0;
    crate = 1.0;
    for (int j = 0; j < LMAX; ++j) for (int i = 0; i < n; ++i) zn[j][i] = 0.0;
    for (int i = 0; i <= LMAX; ++i) tau[i] = 0.0;
    for (int i = 0; i < n; ++i) zn[0][i] = y0[i];
    ewtset(zn[0]);
    F(tn, zn[0], zn[1]);
    int r = hin(tout);
    if (r != ST_OK) return r;
    hscale = h;
    hprime = h;
    for (int i = 0; i < n; ++i) zn[1][i] *= h;
    long nstloc = 0;
    for (;;) {
      if (nst > 0) ewtset(zn[0]);
      if (nstloc >= mxstep) return ST_TOO_MUCH_WORK;
      const double tolsf = UROUND * wrms(zn[0]);
      if (tolsf > 1.0) return ST_TOO_MUCH_ACC;
      r = step();
      if (r != ST_OK) return r;
      ++nstloc;
      if ((tn - tout) * h >= 0.0) {
        const double s = (tout - tn) / h;
        for (int i = 0; i < n; ++i) {
          double v = zn[q][i];
          for (int j = q - 1; j >= 0; --j) v = zn[j][i] + s * v;
          yout[i] = v;
        }
        for (int i = 0; i < n; ++i)
          if (!std::isfinite(yout[i])) return ST_NONFINITE;
        return ST_OK;
      }
    }
  }
};

struct ChemProblem : Problem {
  const orc_mech* m;
  const double* fY;
  double rhoH0, fH;
  void f(double t, const double* y, double* yd) override { orc_rhs(m, t, y, fY, rhoH0, fH, yd); }
  void jac(double t, const double* y, const double*, double* J) override {
    // analytic Jacobian at the temperature the right-hand side uses
    double yy[ORC_NMAX];
    const int N = m->ns;
    for (int i = 0; i < N; ++i) yy[i] = y[i];
    yy[N] = orc_cell_T(m, t, y, rhoH0, fH);
    orc_jac(m, yy, fY, fH, J);
  }
};

struct LinearProblem : Problem {
  int n;
  const double* A;
  void f(double, const double* y, double* yd) override {
    for (int i = 0; i < n; ++i) {
      double s = 0.0;
      for (int j = 0; j < n; ++j) s += A[i * n + j] * y[j];
      yd[i] = s;
    }
  }
  void jac(double, const double*, const double*, double* J) override {
    for (int i = 0; i < n * n; ++i) J[i] = A[i];
  }
};

void fill_stats(const Bdf& b, int status, orc_cellstats* st) {
  if (!st) return;
  st->nst = b.nst; st->nfe = b.nfe; st->nje = b.nje; st->nsetups = b.nsetups;
  st->nni = b.nni; st->ncfn = b.ncfn; st->netf = b.netf;
  st->status = status; st->qlast = b.q; st->hlast = b.h;
  st->nli = b.nli;
}

}  // namespace

extern "C" int orc_bdf_linear(int n, const double* A, double* y, double tout, const orc_opts* o,
                              orc_cellstats* st) {
  LinearProblem P;
  P.n = n;
  P.A = A;
  Bdf* b = new Bdf();
  b->P = &P;
  b->n = n;
  b->rtol = o->rtol;
  for (int i = 0; i < n; ++i) b->atol[i] = o->atol_vec ? o->atol_vec[i] : o->atol;
  b->qmax = o->max_order;
  b->mxstep = o->max_steps;
  b->gmres = o->solve_type == 1;
  b->maxl = o->gmres_maxl > 0 ? o->gmres_maxl : 5;
  double yout[ORC_NMAX];
  const int r = b->integrate(y, tout, yout);
  if (r == 0) for (int i = 0; i < n; ++i) y[i] = yout[i];
  fill_stats(*b, r, st);
  delete b;
  return r;
}

extern "C" int orc_react(const orc_mech* m, int ncell, double* rY, const double* rYsrc, double* T,
                         double* rE, const double* rEsrc, double* FC, const int* mask, double dt,
                         const orc_opts* o, int nthreads, orc_cellstats* stats) {
  const int N = m->ns, n = N + 1;
  int nfail = 0;
#ifdef _OPENMP
  if (nthreads <= 0) nthreads = omp_get_max_threads();
#pragma omp parallel num_threads(nthreads) reduction(+ : nfail)
#endif
  {
    Bdf* b = new Bdf();
    ChemProblem P;
    P.m = m;
    b->P = &P;
    b->n = n;
    b->rtol = o->rtol;
    for (int i = 0; i < n; ++i) b->atol[i] = o->atol_vec ? o->atol_vec[i] : o->atol;
    b->qmax = o->max_order;
    b->mxstep = o->max_steps;
    b->gmres = o->solve_type == 1;
    b->maxl = o->gmres_maxl > 0 ? o->gmres_maxl : 5;
    double y0[ORC_NMAX], yout[ORC_NMAX], fY[ORC_NMAX];
#ifdef _OPENMP
#pragma omp for schedule(dynamic, 4)
#endif
    for (int c = 0; c < ncell; ++c) {
      if (mask && mask[c] < 0) {
        if (FC) FC[c] = 0.0;
        if (stats) { std::memset(&stats[c], 0, sizeof(orc_cellstats)); }
        continue;
      }
      for (int k = 0; k < N; ++k) {
        y0[k] = rY[(long)k * ncell + c];
        fY[k] = rYsrc ? rYsrc[(long)k * ncell + c] : 0.0;
      }
      y0[N] = T[c];
      P.fY = fY;
      P.rhoH0 = rE[c];
      P.fH = rEsrc ? rEsrc[c] : 0.0;
      const int r = b->integrate(y0, dt, yout);
      if (r == 0) {
        for (int k = 0; k < N; ++k) rY[(long)k * ncell + c] = yout[k];
        T[c] = yout[N];
        rE[c] = P.rhoH0 + P.fH * dt;
        if (FC) FC[c] = double(b->nfe);
      } else {
        ++nfail;
        if (FC) FC[c] = -double(b->nfe);
      }
      fill_stats(*b, r, stats ? &stats[c] : nullptr);
    }
    delete b;
  }
  return nfail;
}

/* The GMRES core on a dense matrix (row-major n x n A): solves A x = b in the scaling w
   to 2-norm tolerance delta with at most maxl Krylov vectors.  out[0] = final residual
   estimate, out[1] = Krylov iterations.  Returns the SPGMR status (0 converged,
   1 residual reduced, 2 no reduction, 3 QR breakdown). */
extern "C" int orc_spgmr_dense(int n, const double* A, const double* b, const double* w, int maxl, double delta,
                               double* x, double* out) {
  auto atimes = [&](const double* v, double* z) {
    for (int i = 0; i < n; ++i) {
      double s = 0.0;
      for (int j = 0; j < n; ++j) s += A[i * n + j] * v[j];
      z[i] = s;
    }
  };
  double rho = 0.0;
  int it = 0;
  const int r = spgmr(n, maxl, b, w, delta, atimes, x, &rho, &it);
  if (out) { out[0] = rho; out[1] = it; }
  return r;
}
