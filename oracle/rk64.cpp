// Explicit Runge-Kutta reactor (ReactorRK64) -- TEST INFRASTRUCTURE ONLY (see chem_oracle.h).
//
// PelePhysics' ReactorRK64 (selected by `chem_integrator = "ReactorRK64"`,
// Tools/CMake/BuildPeleLMExe.cmake:124-125) is a six-stage, fourth-order explicit
// Runge-Kutta integrator with adaptive sub-steps, for cells whose chemistry is not stiff.
// Its source is not in the reference tree (empty submodule), so this restates the method
// the reactor here implements: the six-stage Cash-Karp pair (Cash & Karp, ACM TOMS 16,
// 1990), the fourth-order solution propagated and its difference from the fifth-order one
// as the local error, in the WRMS norm of the reactor's tolerances
//   sc_i = rtol max(|y_i|, |y_new,i|) + atol_i.
// A step with error <= 1 is accepted and the next grows by min(5, 0.9 err^-1/5); a
// rejected one (or one whose stages left the physical states, a non-finite error) shrinks
// by max(0.2, 0.9 err^-1/4).  Sub-steps stay within
// [dt / nsub_max, dt / nsub_min] and start at dt / nsub_guess; a step that fails the
// error test at the smallest size fails the cell.  The right-hand side is the h-reactor's
// (orc_rhs: T re-solved from rho h at every evaluation).
#include <algorithm>
#include <cmath>

#ifdef _OPENMP
#include <omp.h>
#endif

#include "oracle_internal.hpp"

namespace {

// the Cash-Karp tableau
constexpr double CK_C[6] = {0.0, 1.0 / 5.0, 3.0 / 10.0, 3.0 / 5.0, 1.0, 7.0 / 8.0};
constexpr double CK_A[6][5] = {
    {0, 0, 0, 0, 0},
    {1.0 / 5.0, 0, 0, 0, 0},
    {3.0 / 40.0, 9.0 / 40.0, 0, 0, 0},
    {3.0 / 10.0, -9.0 / 10.0, 6.0 / 5.0, 0, 0},
    {-11.0 / 54.0, 5.0 / 2.0, -70.0 / 27.0, 35.0 / 27.0, 0},
    {1631.0 / 55296.0, 175.0 / 512.0, 575.0 / 13824.0, 44275.0 / 110592.0, 253.0 / 4096.0}};
constexpr double CK_B4[6] = {2825.0 / 27648.0, 0.0, 18575.0 / 48384.0, 13525.0 / 55296.0, 277.0 / 14336.0, 1.0 / 4.0};
constexpr double CK_B5[6] = {37.0 / 378.0, 0.0, 250.0 / 621.0, 125.0 / 594.0, 0.0, 512.0 / 1771.0};

}  // namespace

extern "C" void orc_rk64_tableau(double* out) {
  int o = 0;
  for (int i = 0; i < 6; ++i) out[o++] = CK_C[i];
  for (int i = 0; i < 6; ++i)
    for (int j = 0; j < 5; ++j) out[o++] = CK_A[i][j];
  for (int i = 0; i < 6; ++i) out[o++] = CK_B4[i];
  for (int i = 0; i < 6; ++i) out[o++] = CK_B5[i];
}

namespace {

// y' = f(t, y) of size n from 0 to dt; returns 0 on success; nst / nfe / nrej counts
template <class RHS>
int rk64_integrate(RHS&& f, int n, double* y, double dt, double rtol, const double* atol, int nsub_guess,
                   int nsub_min, int nsub_max, long* nst_out, long* nfe_out, long* nrej_out) {
  const double hmin = dt / nsub_max, hmax = dt / nsub_min;
  double h = dt / nsub_guess;
  double t = 0.0;
  long nst = 0, nfe = 0, nrej = 0;
  double K[6][ORC_NMAX], yt[ORC_NMAX], y4[ORC_NMAX];
  int status = 0;
  const long max_attempts = 4L * nsub_max + 100;
  while (t < dt) {
    if (nst + nrej >= max_attempts) { status = -1; break; }
    bool last = false;
    if (t + h >= dt) { h = dt - t; last = true; }
    for (int s = 0; s < 6; ++s) {
      for (int i = 0; i < n; ++i) {
        double acc = 0.0;
        for (int j = 0; j < s; ++j) acc += CK_A[s][j] * K[j][i];
        yt[i] = y[i] + h * acc;
      }
      f(t + CK_C[s] * h, yt, K[s]);
      ++nfe;
    }
    double en = 0.0;
    for (int i = 0; i < n; ++i) {
      double a4 = 0.0, e = 0.0;
      for (int s = 0; s < 6; ++s) {
        a4 += CK_B4[s] * K[s][i];
        e += (CK_B5[s] - CK_B4[s]) * K[s][i];
      }
      y4[i] = y[i] + h * a4;
      const double sc = rtol * std::max(std::fabs(y[i]), std::fabs(y4[i])) + atol[i];
      const double r = h * e / sc;
      en += r * r;
    }
    en = std::sqrt(en / n);
    if (!std::isfinite(en)) en = HUGE_VAL;   // a stage left the physical states: reject
    if (en <= 1.0) {
      for (int i = 0; i < n; ++i) y[i] = y4[i];
      t = last ? dt : t + h;
      ++nst;
      h = std::min(hmax, h * (en > 0.0 ? std::min(5.0, 0.9 * std::pow(en, -0.2)) : 5.0));
    } else {
      ++nrej;
      if (h <= hmin * (1.0 + 1e-12)) { status = -2; break; }   // the error test fails at the smallest step
      h = std::max(hmin, h * std::max(0.2, 0.9 * std::pow(en, -0.25)));
    }
  }
  *nst_out = nst;
  *nfe_out = nfe;
  *nrej_out = nrej;
  return status;
}

}  // namespace

// y' = A y (n x n, row-major) from 0 to tout with the same integrator (pins it against
// exp(A t)); returns the status, steps and evaluations in st[0..2]
extern "C" int orc_rk64_linear(int n, const double* A, double* y, double tout, double rtol, double atol,
                               int nsub_guess, int nsub_min, int nsub_max, long* st) {
  double at[ORC_NMAX];
  for (int i = 0; i < n; ++i) at[i] = atol;
  auto f = [&](double, const double* x, double* xd) {
    for (int i = 0; i < n; ++i) {
      double s = 0.0;
      for (int j = 0; j < n; ++j) s += A[i * n + j] * x[j];
      xd[i] = s;
    }
  };
  long nst = 0, nfe = 0, nrej = 0;
  const int r = rk64_integrate(f, n, y, tout, rtol, at, nsub_guess, nsub_min, nsub_max, &nst, &nfe, &nrej);
  if (st) { st[0] = nst; st[1] = nfe; st[2] = nrej; }
  return r;
}

extern "C" int orc_rk64_react(const orc_mech* m, int ncell, double* rY, const double* rYsrc, double* T, double* rE,
                              const double* rEsrc, double* FC, const int* mask, double dt, double rtol, double atol,
                              int nsub_guess, int nsub_min, int nsub_max, int nthreads) {
  const int N = m->ns, n = N + 1;
  int nfail = 0;
#ifdef _OPENMP
  if (nthreads <= 0) nthreads = omp_get_max_threads();
#pragma omp parallel for num_threads(nthreads) schedule(dynamic, 4) reduction(+ : nfail)
#endif
  for (int c = 0; c < ncell; ++c) {
    if (mask && mask[c] < 0) {
      if (FC) FC[c] = 0.0;
      continue;
    }
    double y[ORC_NMAX], fY[ORC_NMAX], at[ORC_NMAX];
    for (int k = 0; k < N; ++k) {
      y[k] = rY[(long)k * ncell + c];
      fY[k] = rYsrc ? rYsrc[(long)k * ncell + c] : 0.0;
    }
    y[N] = T[c];
    for (int i = 0; i < n; ++i) at[i] = atol;
    const double rhoH0 = rE[c], fH = rEsrc ? rEsrc[c] : 0.0;
    long nst = 0, nfe = 0, nrej = 0;
    auto f = [&](double t, const double* x, double* xd) { orc_rhs(m, t, x, fY, rhoH0, fH, xd); };
    const int st = rk64_integrate(f, n, y, dt, rtol, at, nsub_guess, nsub_min, nsub_max, &nst, &nfe, &nrej);
    if (st != 0) {
      ++nfail;
      if (FC) FC[c] = -double(nfe);
      continue;
    }
    for (int k = 0; k < N; ++k) rY[(long)k * ncell + c] = y[k];
    T[c] = y[N];
    rE[c] = rhoH0 + fH * dt;
    if (FC) FC[c] = double(nfe);
  }
  return nfail;
}
