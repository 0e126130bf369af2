// Scaled GMRES for the Newton systems of the BDF (cvode.solve_type = GMRES) --
// TEST INFRASTRUCTURE ONLY (see chem_oracle.h).
//
// The reference's GPU decks select it (Exec/RegTests/FlameSheet/inputs.3d-regt_GPU:89,
// inputs.3d-regt_GPUdodecane:91, inputs.3d-regt_GPUheptane:90: cvode.solve_type = GMRES).
// SUNDIALS is not vendored in the reference (Submodules/sundials is empty), so this
// restates the published algorithm of SUNLinSol_SPGMR with CVODE's defaults (SUNDIALS
// documentation, "The SUNLinSol_SPGMR module"; Saad & Schultz, SIAM J. Sci. Stat.
// Comput. 7, 1986): no preconditioner, the scaled system S A S^-1 (S x) = S b with
// S = diag(w) (CVODE passes its error weights as both scalings), zero initial guess,
// at most maxl Krylov vectors (CVODE's default 5) and no restarts, modified Gram-Schmidt with one
// reorthogonalisation pass when the new vector's norm falls below 1e-3 of its norm before
// orthogonalisation (|v| + 1000 |v_in| == 1000 |v_in| in floating point), the Hessenberg
// matrix reduced by Givens rotations as it grows (c, s from the larger of the two
// entries), and the residual estimate |prod s_i| beta.
#pragma once
#include <cmath>

#include "oracle_internal.hpp"

constexpr int SPGMR_MAXL = ORC_NMAX;   // largest Krylov dimension (CVODE's default is 5)

enum { SPGMR_SUCCESS = 0, SPGMR_RES_REDUCED = 1, SPGMR_CONV_FAIL = 2, SPGMR_QR_FAIL = 3 };

inline double spgmr_dot(int n, const double* a, const double* b) {
  double s = 0.0;
  for (int i = 0; i < n; ++i) s += a[i] * b[i];
  return s;
}

// Solves A x = b, A given by atimes(v, z): z = A v.  w: the scaling (S = diag(w)); delta:
// the tolerance on || S (b - A x) ||_2.  Returns SPGMR_SUCCESS (x solves to delta),
// SPGMR_RES_REDUCED (the residual fell below ||S b||_2 but not to delta; x is the
// Krylov solution), SPGMR_CONV_FAIL (no reduction; x = 0) or SPGMR_QR_FAIL (a zero
// diagonal in the Givens-reduced Hessenberg matrix; x = 0).  rho: the final residual
// estimate, nli: Krylov iterations (one A v product each).
template <class ATimes>
int spgmr(int n, int maxl, const double* b, const double* w, double delta, ATimes&& atimes, double* x,
          double* rho_out, int* nli_out) {
  double V[SPGMR_MAXL + 1][ORC_NMAX];
  double H[SPGMR_MAXL + 1][SPGMR_MAXL] = {};
  double giv[2 * SPGMR_MAXL], yg[SPGMR_MAXL + 1];
  double vt[ORC_NMAX], z[ORC_NMAX];
  if (maxl > SPGMR_MAXL) maxl = SPGMR_MAXL;
  for (int i = 0; i < n; ++i) {
    x[i] = 0.0;
    V[0][i] = w[i] * b[i];
  }
  const double beta = std::sqrt(spgmr_dot(n, V[0], V[0]));
  *nli_out = 0;
  *rho_out = beta;
  if (beta <= delta) return SPGMR_SUCCESS;
  double rho = beta, rotp = 1.0;
  const double rb = 1.0 / beta;
  for (int i = 0; i < n; ++i) V[0][i] *= rb;
  int krydim = 0;
  bool converged = false;
  for (int l = 0; l < maxl; ++l) {
    ++*nli_out;
    krydim = l + 1;
    // V[l+1] = S A S^-1 V[l]
    for (int i = 0; i < n; ++i) vt[i] = V[l][i] / w[i];
    atimes(vt, z);
    double* vk = V[l + 1];
    for (int i = 0; i < n; ++i) vk[i] = w[i] * z[i];
    // modified Gram-Schmidt against V[0..l], one reorthogonalisation pass if needed
    const double vk_norm = std::sqrt(spgmr_dot(n, vk, vk));
    for (int i = 0; i <= l; ++i) {
      const double h = spgmr_dot(n, V[i], vk);
      H[i][l] = h;
      for (int j = 0; j < n; ++j) vk[j] = vk[j] - h * V[i][j];
    }
    double nrm = std::sqrt(spgmr_dot(n, vk, vk));
    const double big = 1000.0 * vk_norm;
    if (big + nrm == big) {
      double newp = 0.0;
      for (int i = 0; i <= l; ++i) {
        const double t = spgmr_dot(n, V[i], vk);
        newp += t * t;
        H[i][l] += t;
        for (int j = 0; j < n; ++j) vk[j] = vk[j] - t * V[i][j];
      }
      if (newp != 0.0) nrm = std::sqrt(nrm * nrm - newp);
    }
    H[l + 1][l] = nrm;
    // the earlier rotations on the new column, then the rotation that zeroes H[l+1][l]
    for (int i = 0; i < l; ++i) {
      const double c = giv[2 * i], s = giv[2 * i + 1];
      const double t1 = H[i][l], t2 = H[i + 1][l];
      H[i][l] = c * t1 - s * t2;
      H[i + 1][l] = s * t1 + c * t2;
    }
    const double t1 = H[l][l], t2 = H[l + 1][l];
    double c, s;
    if (t2 == 0.0) {
      c = 1.0;
      s = 0.0;
    } else if (std::fabs(t2) >= std::fabs(t1)) {
      const double t3 = t1 / t2;
      s = -1.0 / std::sqrt(1.0 + t3 * t3);
      c = -s * t3;
    } else {
      const double t3 = t2 / t1;
      c = 1.0 / std::sqrt(1.0 + t3 * t3);
      s = -c * t3;
    }
    giv[2 * l] = c;
    giv[2 * l + 1] = s;
    H[l][l] = c * t1 - s * t2;
    if (H[l][l] == 0.0) return SPGMR_QR_FAIL;
    rotp *= s;
    rho = std::fabs(rotp * beta);
    *rho_out = rho;
    if (rho <= delta) {
      converged = true;
      break;
    }
    const double rinv = 1.0 / H[l + 1][l];
    for (int i = 0; i < n; ++i) vk[i] *= rinv;
  }
  if (!converged && !(rho < beta)) return SPGMR_CONV_FAIL;
  // y = R^-1 Q^T (beta e1), x = S^-1 V y
  yg[0] = beta;
  for (int i = 1; i <= krydim; ++i) yg[i] = 0.0;
  for (int k = 0; k < krydim; ++k) {
    const double c = giv[2 * k], s = giv[2 * k + 1];
    const double t1 = yg[k], t2 = yg[k + 1];
    yg[k] = c * t1 - s * t2;
    yg[k + 1] = s * t1 + c * t2;
  }
  for (int k = krydim - 1; k >= 0; --k) {
    if (H[k][k] == 0.0) return SPGMR_QR_FAIL;
    yg[k] /= H[k][k];
    for (int i = 0; i < k; ++i) yg[i] -= yg[k] * H[i][k];
  }
  for (int i = 0; i < n; ++i) {
    double s = yg[0] * V[0][i];
    for (int k = 1; k < krydim; ++k) s += yg[k] * V[k][i];
    x[i] = s / w[i];
  }
  return converged ? SPGMR_SUCCESS : SPGMR_RES_REDUCED;
}
