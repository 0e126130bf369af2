// Radau IIA of order 5 for the reactor ODE -- TEST INFRASTRUCTURE ONLY (see chem_oracle.h).
//
// A second stiff integrator, independent of the BDF in cvode_bdf.cpp, so that the tight
// FP64 references the tests use are checked by two methods of different families.  The
// method is the three-stage Radau IIA collocation method (Hairer & Wanner, Solving
// Ordinary Differential Equations II, section IV.5 and IV.8):
//
//   nodes c = ((4 - sqrt 6)/10, (4 + sqrt 6)/10, 1), A_ij = integral_0^{c_i} l_j(s) ds with
//   l_j the Lagrange basis on the nodes, y1 = y0 + Z_3 (stiffly accurate);
//
// the stage increments Z_i = h sum_j A_ij f(x0 + c_j h, y0 + Z_j) are found by simplified
// Newton on the full 3n system (I - h A (x) J) dZ = -Z + h (A (x) I) F(Z), with the
// analytic Jacobian at the step's start; the local error is the embedded order-3 estimate
//
//   err = (I - h g0 J)^-1 (g0 h f(x0, y0) + sum_j e_j Z_j),   e = (bhat - b)^T A^-1,
//
// where g0 is the real eigenvalue of A and bhat solves the order conditions
// sum_i bhat_i c_i^(q-1) = 1/q - g0 [q = 1], q = 1..3 (a rejected first estimate is
// refined once with f(x0, y0 + err), as Hairer & Wanner do for stiff components).  Every
// coefficient is computed from these definitions at start-up rather than tabulated.
// Step size: h_new = h min(8, max(0.2, 0.9 err^(-1/4))); the next step's Newton starts
// from the extrapolated collocation polynomial.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#ifdef _OPENMP
#include <omp.h>
#endif

#include "oracle_internal.hpp"

namespace {

struct RadauCoef {
  double c[3], A[3][3], Ai[3][3], e[3], g0;
};

// solve the 3x3 system M x = r (Gaussian elimination with partial pivoting)
void solve3(double M[3][3], double r[3]) {
  for (int k = 0; k < 3; ++k) {
    int p = k;
    for (int i = k + 1; i < 3; ++i)
      if (std::fabs(M[i][k]) > std::fabs(M[p][k])) p = i;
    if (p != k) {
      for (int j = 0; j < 3; ++j) std::swap(M[k][j], M[p][j]);
      std::swap(r[k], r[p]);
    }
    for (int i = k + 1; i < 3; ++i) {
      const double l = M[i][k] / M[k][k];
      for (int j = k; j < 3; ++j) M[i][j] -= l * M[k][j];
      r[i] -= l * r[k];
    }
  }
  for (int k = 2; k >= 0; --k) {
    double s = r[k];
    for (int j = k + 1; j < 3; ++j) s -= M[k][j] * r[j];
    r[k] = s / M[k][k];
  }
}

RadauCoef make_coef() {
  RadauCoef rc{};
  const double s6 = std::sqrt(6.0);
  rc.c[0] = (4.0 - s6) / 10.0;
  rc.c[1] = (4.0 + s6) / 10.0;
  rc.c[2] = 1.0;
  // A_ij = int_0^{c_i} l_j: the collocation conditions sum_j A_ij c_j^(q-1) = c_i^q / q,
  // q = 1..3, give row i from the transposed Vandermonde system
  for (int i = 0; i < 3; ++i) {
    double V[3][3], r[3];
    for (int q = 0; q < 3; ++q) {
      for (int j = 0; j < 3; ++j) V[q][j] = std::pow(rc.c[j], q);
      r[q] = std::pow(rc.c[i], q + 1) / (q + 1);
    }
    solve3(V, r);
    for (int j = 0; j < 3; ++j) rc.A[i][j] = r[j];
  }
  // A^-1 column by column
  for (int j = 0; j < 3; ++j) {
    double M[3][3], r[3] = {0.0, 0.0, 0.0};
    std::memcpy(M, rc.A, sizeof(M));
    r[j] = 1.0;
    solve3(M, r);
    for (int i = 0; i < 3; ++i) rc.Ai[i][j] = r[i];
  }
  // the real eigenvalue of A: root of det(A - g I) = -g^3 + tr g^2 - m2 g + det, by Newton
  // from above (the complex pair has a smaller modulus for this method)
  const double (&a)[3][3] = rc.A;
  const double tr = a[0][0] + a[1][1] + a[2][2];
  const double m2 = a[0][0] * a[1][1] - a[0][1] * a[1][0] + a[0][0] * a[2][2] - a[0][2] * a[2][0] +
                    a[1][1] * a[2][2] - a[1][2] * a[2][1];
  const double det = a[0][0] * (a[1][1] * a[2][2] - a[1][2] * a[2][1]) -
                     a[0][1] * (a[1][0] * a[2][2] - a[1][2] * a[2][0]) +
                     a[0][2] * (a[1][0] * a[2][1] - a[1][1] * a[2][0]);
  double g = 1.0;
  for (int it = 0; it < 100; ++it) {
    const double p = g * g * g - tr * g * g + m2 * g - det;
    const double dp = 3.0 * g * g - 2.0 * tr * g + m2;
    const double dg = p / dp;
    g -= dg;
    if (std::fabs(dg) <= 1e-16 * std::fabs(g)) break;
  }
  rc.g0 = g;
  // embedded weights: sum_i bhat_i c_i^(q-1) = 1/q - g0 [q = 1]
  {
    double V[3][3], r[3];
    for (int q = 0; q < 3; ++q) {
      for (int i = 0; i < 3; ++i) V[q][i] = std::pow(rc.c[i], q);
      r[q] = 1.0 / (q + 1) - (q == 0 ? g : 0.0);
    }
    solve3(V, r);
    for (int j = 0; j < 3; ++j) {
      double s = 0.0;
      for (int i = 0; i < 3; ++i) s += (r[i] - rc.A[2][i]) * rc.Ai[i][j];
      rc.e[j] = s;
    }
  }
  return rc;
}

const RadauCoef& coef() {
  static const RadauCoef rc = make_coef();
  return rc;
}

// dense LU with partial pivoting, row-major n x n, in place; returns false if singular
bool lu(int n, double* a, int* piv) {
  for (int k = 0; k < n; ++k) {
    int p = k;
    double mx = std::fabs(a[(long)k * n + k]);
    for (int i = k + 1; i < n; ++i) {
      const double v = std::fabs(a[(long)i * n + k]);
      if (v > mx) { mx = v; p = i; }
    }
    piv[k] = p;
    if (mx == 0.0) return false;
    if (p != k)
      for (int j = 0; j < n; ++j) std::swap(a[(long)k * n + j], a[(long)p * n + j]);
    const double inv = 1.0 / a[(long)k * n + k];
    for (int i = k + 1; i < n; ++i) {
      const double l = a[(long)i * n + k] * inv;
      a[(long)i * n + k] = l;
      if (l != 0.0)
        for (int j = k + 1; j < n; ++j) a[(long)i * n + j] -= l * a[(long)k * n + j];
    }
  }
  return true;
}

// the factorisation swaps whole rows (multipliers included), so every interchange is
// applied to b before the forward substitution
void lu_solve(int n, const double* a, const int* piv, double* b) {
  for (int k = 0; k < n; ++k)
    if (piv[k] != k) std::swap(b[k], b[piv[k]]);
  for (int k = 0; k < n; ++k)
    for (int i = k + 1; i < n; ++i) b[i] -= a[(long)i * n + k] * b[k];
  for (int k = n - 1; k >= 0; --k) {
    double s = b[k];
    for (int j = k + 1; j < n; ++j) s -= a[(long)k * n + j] * b[j];
    b[k] = s / a[(long)k * n + k];
  }
}

struct Radau {
  const orc_mech* m;
  const double* fY;
  double rhoH0, fH, rtol;
  const double* atol;   // n values
  int n;
  long nst = 0, nfe = 0, nje = 0, nrej = 0;
  std::vector<double> W, E, J;
  std::vector<int> pw, pe;

  void f(double t, const double* y, double* yd) {
    orc_rhs(m, t, y, fY, rhoH0, fH, yd);
    ++nfe;
  }
  void jac(double t, const double* y) {
    double yy[ORC_NMAX];
    const int N = n - 1;
    for (int i = 0; i < N; ++i) yy[i] = y[i];
    yy[N] = orc_cell_T(m, t, y, rhoH0, fH);
    orc_jac(m, yy, fY, fH, J.data());
    ++nje;
  }
  double wrms(const double* v, const double* sc, int len) const {
    double s = 0.0;
    for (int i = 0; i < len; ++i) {
      const double x = v[i] / sc[i % n];
      s += x * x;
    }
    return std::sqrt(s / len);
  }

  // integrate y from 0 to tout; returns 0 on success
  int integrate(double* y, double tout, long max_steps) {
    const RadauCoef& rc = coef();
    const int n3 = 3 * n;
    W.assign((size_t)n3 * n3, 0.0);
    E.assign((size_t)n * n, 0.0);
    J.assign((size_t)n * n, 0.0);
    pw.assign(n3, 0);
    pe.assign(n, 0);
    std::vector<double> Z(n3, 0.0), Zp(n3, 0.0), F(n3), R(n3), y1(n), f0(n), sc(n), tmp(n), err(n), ye(n);
    double x = 0.0, h = 1e-6 * tout, hprev = 0.0, eta_old = 1.0;
    bool have_prev = false, first = true, last_rejected = false;
    while (x < tout) {
      if (nst >= max_steps) return -1;
      if (x + h >= tout || x + 1.0001 * h >= tout) h = tout - x;
      if (!(h > 1e-14 * tout)) return -2;
      for (int i = 0; i < n; ++i) sc[i] = atol[i] + rtol * std::fabs(y[i]);
      jac(x, y);
      f(x, y, f0.data());
      // W = I - h A (x) J
      for (int i = 0; i < 3; ++i)
        for (int a = 0; a < n; ++a)
          for (int j = 0; j < 3; ++j)
            for (int b = 0; b < n; ++b)
              W[(size_t)(i * n + a) * n3 + (j * n + b)] = (i == j && a == b ? 1.0 : 0.0) - h * rc.A[i][j] * J[(size_t)a * n + b];
      if (!lu(n3, W.data(), pw.data())) { h *= 0.5; ++nrej; continue; }
      // starting stages: the previous step's collocation polynomial, extrapolated
      if (have_prev) {
        const double ratio = h / hprev;
        for (int i = 0; i < 3; ++i) {
          const double s = 1.0 + rc.c[i] * ratio;   // in units of the previous step
          // Lagrange on nodes {0, c1, c2, c3} with values {0, Z1, Z2, Z3} (Z from y_prev)
          double L[3];
          for (int j = 0; j < 3; ++j) {
            double v = s / rc.c[j];
            for (int k = 0; k < 3; ++k)
              if (k != j) v *= (s - rc.c[k]) / (rc.c[j] - rc.c[k]);
            L[j] = v;
          }
          for (int a = 0; a < n; ++a)
            Z[i * n + a] = L[0] * Zp[a] + L[1] * Zp[n + a] + L[2] * Zp[2 * n + a] - Zp[2 * n + a];
        }
      } else {
        std::fill(Z.begin(), Z.end(), 0.0);
      }
      // simplified Newton
      bool conv = false;
      double dn_old = 0.0, eta = std::pow(std::max(eta_old, 1e-16), 0.8);
      for (int it = 0; it < 10; ++it) {
        for (int i = 0; i < 3; ++i) {
          for (int a = 0; a < n; ++a) tmp[a] = y[a] + Z[i * n + a];
          f(x + rc.c[i] * h, tmp.data(), &F[i * n]);
        }
        for (int i = 0; i < 3; ++i)
          for (int a = 0; a < n; ++a) {
            double s = 0.0;
            for (int j = 0; j < 3; ++j) s += rc.A[i][j] * F[j * n + a];
            R[i * n + a] = -Z[i * n + a] + h * s;
          }
        lu_solve(n3, W.data(), pw.data(), R.data());
        bool finite = true;
        for (int k = 0; k < n3; ++k) {
          Z[k] += R[k];
          finite = finite && std::isfinite(Z[k]);
        }
        if (!finite) break;
        const double dn = wrms(R.data(), sc.data(), n3);
        if (it > 0) {
          const double theta = dn / dn_old;
          if (theta >= 0.99) break;   // diverging
          eta = theta / (1.0 - theta);
        }
        dn_old = dn;
        if (eta * dn <= 1e-2 || dn <= 1e-3 * 1e-16 / rtol) {
          conv = true;
          eta_old = eta;
          break;
        }
      }
      if (!conv) {
        h *= 0.5;
        ++nrej;
        have_prev = false;
        last_rejected = true;
        continue;
      }
      // local error
      for (int a = 0; a < n; ++a) {
        y1[a] = y[a] + Z[2 * n + a];
        sc[a] = atol[a] + rtol * std::max(std::fabs(y[a]), std::fabs(y1[a]));
      }
      for (int a = 0; a < n; ++a)
        for (int b = 0; b < n; ++b)
          E[(size_t)a * n + b] = (a == b ? 1.0 : 0.0) - h * rc.g0 * J[(size_t)a * n + b];
      if (!lu(n, E.data(), pe.data())) { h *= 0.5; ++nrej; continue; }
      auto estimate = [&](const double* fx) {
        for (int a = 0; a < n; ++a)
          err[a] = rc.g0 * h * fx[a] + rc.e[0] * Z[a] + rc.e[1] * Z[n + a] + rc.e[2] * Z[2 * n + a];
        lu_solve(n, E.data(), pe.data(), err.data());
        return wrms(err.data(), sc.data(), n);
      };
      double en = estimate(f0.data());
      if (en >= 1.0 && (first || last_rejected)) {
        for (int a = 0; a < n; ++a) ye[a] = y[a] + err[a];
        f(x, ye.data(), tmp.data());
        en = estimate(tmp.data());
      }
      const double fac = std::min(8.0, std::max(0.2, 0.9 * std::pow(std::max(en, 1e-12), -0.25)));
      if (en < 1.0) {
        x = (h == tout - x) ? tout : x + h;
        for (int a = 0; a < n; ++a) y[a] = y1[a];
        Zp = Z;
        hprev = h;
        have_prev = true;
        ++nst;
        first = false;
        h *= last_rejected ? std::min(1.0, fac) : fac;   // no growth right after a rejection
        last_rejected = false;
      } else {
        h *= std::min(1.0, fac);
        ++nrej;
        last_rejected = true;
      }
    }
    for (int a = 0; a < n; ++a)
      if (!std::isfinite(y[a])) return -3;
    return 0;
  }
};

}  // namespace

// the method's coefficients as start-up derived them: c (3), A (3 x 3 row-major), the
// error weights e (3) and g0, 16 values
extern "C" void orc_radau_coefficients(double* out) {
  const RadauCoef& rc = coef();
  for (int i = 0; i < 3; ++i) out[i] = rc.c[i];
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) out[3 + 3 * i + j] = rc.A[i][j];
  for (int i = 0; i < 3; ++i) out[12 + i] = rc.e[i];
  out[15] = rc.g0;
}

extern "C" int orc_radau_react(const orc_mech* m, int ncell, double* rY, const double* rYsrc, double* T,
                               double* rE, const double* rEsrc, double dt, double rtol, double atol,
                               int nthreads, long* nsteps) {
  const int N = m->ns, n = N + 1;
  int nfail = 0;
#ifdef _OPENMP
  if (nthreads <= 0) nthreads = omp_get_max_threads();
#pragma omp parallel for num_threads(nthreads) schedule(dynamic, 1) reduction(+ : nfail)
#endif
  for (int c = 0; c < ncell; ++c) {
    double y[ORC_NMAX], fY[ORC_NMAX], at[ORC_NMAX];
    for (int k = 0; k < N; ++k) {
      y[k] = rY[(long)k * ncell + c];
      fY[k] = rYsrc ? rYsrc[(long)k * ncell + c] : 0.0;
    }
    y[N] = T[c];
    for (int i = 0; i < n; ++i) at[i] = atol;
    Radau r;
    r.m = m;
    r.fY = fY;
    r.rhoH0 = rE[c];
    r.fH = rEsrc ? rEsrc[c] : 0.0;
    r.rtol = rtol;
    r.atol = at;
    r.n = n;
    const int st = r.integrate(y, dt, 1000000);
    if (nsteps) nsteps[c] = st == 0 ? r.nst : -1;
    if (st != 0) {
      ++nfail;
      continue;
    }
    for (int k = 0; k < N; ++k) rY[(long)k * ncell + c] = y[k];
    T[c] = y[N];
    rE[c] = r.rhoH0 + r.fH * dt;
  }
  return nfail;
}
