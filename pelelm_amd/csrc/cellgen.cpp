// Seeded synthetic reactor inputs: fresh fuel/air blended with its adiabatic
// (HP) equilibrium products, the construction the chemistry benchmarks use.
//
// Every cell is generated from Philox4x32-10 keyed by (seed, cell index), so a cell's
// state does not depend on the thread count or on which other cells are generated.
// Per cell: c (progress variable, distribution set by `mode`), phi ~ U[phi_lo, phi_hi];
// fresh mixture at 300 K and its HP equilibrium at 1 atm (Gordon & McBride, NASA
// RP-1311, eqs. 2.24-2.26, Newton in ln n with the CEA step control); Y = (1-c) Y_fresh
// + c Y_eq with a 1e-8 floor and renormalisation; T = 300 + c (T_ad - 300) + N(0, 20 K);
// rho from the ideal-gas law, rho h = rho h(T, Y).  Output is CGS and SoA
// (component-major, cell fastest), ready for the reactor's flat entry point; the
// species order is the mechanism's (Source/IndexDefines.H:11-15 state order).
//
// Host-only, not part of the reactor hot path.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <fstream>
#include <sstream>
#include <string>
#include <vector>

#ifdef _OPENMP
#include <omp.h>
#endif

#include "pelelm_cells.h"

namespace {

// ---------------------------------------------------------------------------
// Philox4x32-10 (Salmon, Moraes, Dror & Shaw, SC'11)
// ---------------------------------------------------------------------------
inline void mulhilo(uint32_t a, uint32_t b, uint32_t& hi, uint32_t& lo) {
  const uint64_t p = uint64_t(a) * uint64_t(b);
  hi = uint32_t(p >> 32);
  lo = uint32_t(p);
}

void philox4x32_10(const uint32_t ctr_in[4], const uint32_t key_in[2], uint32_t out[4]) {
  uint32_t c0 = ctr_in[0], c1 = ctr_in[1], c2 = ctr_in[2], c3 = ctr_in[3];
  uint32_t k0 = key_in[0], k1 = key_in[1];
  for (int r = 0; r < 10; ++r) {
    uint32_t hi0, lo0, hi1, lo1;
    mulhilo(0xD2511F53u, c0, hi0, lo0);
    mulhilo(0xCD9E8D57u, c2, hi1, lo1);
    const uint32_t n0 = hi1 ^ c1 ^ k0, n1 = lo1, n2 = hi0 ^ c3 ^ k1, n3 = lo0;
    c0 = n0; c1 = n1; c2 = n2; c3 = n3;
    k0 += 0x9E3779B9u;
    k1 += 0xBB67AE85u;
  }
  out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

inline double u01(uint32_t x) { return (double(x) + 0.5) * (1.0 / 4294967296.0); }

// ---------------------------------------------------------------------------
// mechanism table (same flat format the reactor's host code and oracle read)
// ---------------------------------------------------------------------------
struct Table {
  int ne = 0, ns = 0;
  double Ru = 0, Patm = 0;
  std::vector<std::string> names, elems;
  std::vector<double> W, Tmid, lo, hi;
  std::vector<int> ecomp;
  bool load(const char* path) {
    std::ifstream in(path);
    if (!in) return false;
    std::string line;
    int nr = 0;
    while (std::getline(in, line)) {
      std::istringstream ss(line);
      std::string tag;
      ss >> tag;
      if (tag == "mech") {
        std::string nm;
        ss >> nm >> ne >> ns >> nr >> Ru >> Patm;
      } else if (tag == "elem") {
        std::string e;
        ss >> e;
        elems.push_back(e);
      } else if (tag == "spec") {
        std::string nm;
        double w, a, b, c;
        ss >> nm >> w >> a >> b >> c;
        names.push_back(nm);
        W.push_back(w);
        Tmid.push_back(b);
        for (int i = 0; i < 7; ++i) { double x; ss >> x; lo.push_back(x); }
        for (int i = 0; i < 7; ++i) { double x; ss >> x; hi.push_back(x); }
        for (int e = 0; e < ne; ++e) { int x; ss >> x; ecomp.push_back(x); }
      }
    }
    return (int)names.size() == ns && ns > 0;
  }
  int index(const char* nm) const {
    for (int k = 0; k < ns; ++k)
      if (names[k] == nm) return k;
    return -1;
  }
  void thermo(double T, double* cpR, double* hRT, double* gRT) const {
    const double logT = std::log(T);
    for (int k = 0; k < ns; ++k) {
      const double* a = (T < Tmid[k]) ? &lo[7 * k] : &hi[7 * k];
      const double cp = a[0] + T * (a[1] + T * (a[2] + T * (a[3] + T * a[4])));
      const double h = a[0] + T * (a[1] / 2 + T * (a[2] / 3 + T * (a[3] / 4 + T * a[4] / 5))) + a[5] / T;
      const double s = a[0] * logT + T * (a[1] + T * (a[2] / 2 + T * (a[3] / 3 + T * a[4] / 4))) + a[6];
      if (cpR) cpR[k] = cp;
      if (hRT) hRT[k] = h;
      if (gRT) gRT[k] = h - s;
    }
  }
  double h_mass(double T, const double* Y) const {   // erg/g
    std::vector<double> hRT(ns);
    thermo(T, nullptr, hRT.data(), nullptr);
    double h = 0;
    for (int k = 0; k < ns; ++k) h += Y[k] * hRT[k] / W[k];
    return h * Ru * T;
  }
};

// Solve A x = b in place (small dense system, partial pivoting). Returns false if singular.
bool solve_small(int n, double* A, double* b) {
  for (int k = 0; k < n; ++k) {
    int p = k;
    for (int i = k + 1; i < n; ++i)
      if (std::fabs(A[i * n + k]) > std::fabs(A[p * n + k])) p = i;
    if (A[p * n + k] == 0.0) return false;
    if (p != k) {
      for (int j = 0; j < n; ++j) std::swap(A[k * n + j], A[p * n + j]);
      std::swap(b[k], b[p]);
    }
    for (int i = k + 1; i < n; ++i) {
      const double f = A[i * n + k] / A[k * n + k];
      for (int j = k; j < n; ++j) A[i * n + j] -= f * A[k * n + j];
      b[i] -= f * b[k];
    }
  }
  for (int k = n - 1; k >= 0; --k) {
    double s = b[k];
    for (int j = k + 1; j < n; ++j) s -= A[k * n + j] * b[j];
    b[k] = s / A[k * n + k];
  }
  return true;
}

// HP equilibrium at P = P_atm for a mixture of moles X0 (sum 1) at T0.
// Gordon-McBride iteration in (pi_e, dln n, dln T); returns 0 on convergence.
int equilibrium_hp(const Table& t, const double* X0, double T0, double* Xeq, double* Teq) {
  const int ns = t.ns;
  // active elements: nonzero abundance; active species: only active elements
  std::vector<double> b(t.ne, 0.0);
  for (int k = 0; k < ns; ++k)
    for (int e = 0; e < t.ne; ++e) b[e] += X0[k] * t.ecomp[k * t.ne + e];
  std::vector<int> el, sp;
  for (int e = 0; e < t.ne; ++e)
    if (b[e] > 0) el.push_back(e);
  for (int k = 0; k < ns; ++k) {
    bool ok = true;
    for (int e = 0; e < t.ne; ++e)
      if (t.ecomp[k * t.ne + e] && b[e] <= 0) ok = false;
    if (ok) sp.push_back(k);
  }
  const int E = (int)el.size(), S = (int)sp.size();
  std::vector<double> hRT0(ns);
  t.thermo(T0, nullptr, hRT0.data(), nullptr);
  double h0 = 0.0;   // assigned enthalpy / R  [K per mole of fresh mixture]
  for (int k = 0; k < ns; ++k) h0 += X0[k] * hRT0[k] * T0;
  std::vector<double> n(S, 1.0 / S), cpR(ns), hRT(ns), gRT(ns), mu(S), dln(S);
  double ntot = 1.0, T = 2000.0;
  const int M = E + 2;
  std::vector<double> A(M * M), rhs(M);
  for (int it = 0; it < 500; ++it) {
    t.thermo(T, cpR.data(), hRT.data(), gRT.data());
    double sumn = 0.0;
    for (int j = 0; j < S; ++j) sumn += n[j];
    for (int j = 0; j < S; ++j) mu[j] = gRT[sp[j]] + std::log(n[j] / ntot);
    std::fill(A.begin(), A.end(), 0.0);
    std::fill(rhs.begin(), rhs.end(), 0.0);
    // rows 0..E-1: element balances; row E: total moles; row E+1: energy
    for (int j = 0; j < S; ++j) {
      const int k = sp[j];
      const double H = hRT[k];
      for (int r = 0; r < E; ++r) {
        const double ar = t.ecomp[k * t.ne + el[r]] * n[j];
        if (ar == 0.0) continue;
        for (int c = 0; c < E; ++c) A[r * M + c] += ar * t.ecomp[k * t.ne + el[c]];
        A[r * M + E] += ar;
        A[r * M + E + 1] += ar * H;
        rhs[r] += ar * mu[j];
      }
      for (int c = 0; c < E; ++c) {
        A[E * M + c] += t.ecomp[k * t.ne + el[c]] * n[j];
        A[(E + 1) * M + c] += t.ecomp[k * t.ne + el[c]] * n[j] * H;
      }
      A[E * M + E + 1] += n[j] * H;
      rhs[E] += n[j] * mu[j];
      A[(E + 1) * M + E] += n[j] * H;
      A[(E + 1) * M + E + 1] += n[j] * (cpR[k] + H * H);
      rhs[E + 1] += n[j] * H * mu[j] - n[j] * H;
    }
    for (int r = 0; r < E; ++r) {
      double bj = 0.0;
      for (int j = 0; j < S; ++j) bj += t.ecomp[sp[j] * t.ne + el[r]] * n[j];
      rhs[r] += b[el[r]] - bj;
    }
    A[E * M + E] = sumn - ntot;
    rhs[E] += ntot - sumn;
    rhs[E + 1] += h0 / T;
    if (!solve_small(M, A.data(), rhs.data())) return -1;
    const double dlnn = rhs[E], dlnT = rhs[E + 1];
    double dmax = std::max(5.0 * std::fabs(dlnT), 5.0 * std::fabs(dlnn));
    for (int j = 0; j < S; ++j) {
      double s = -mu[j] + dlnn + hRT[sp[j]] * dlnT;
      for (int r = 0; r < E; ++r) s += t.ecomp[sp[j] * t.ne + el[r]] * rhs[r];
      dln[j] = s;
      if (n[j] / ntot > 1e-8) dmax = std::max(dmax, std::fabs(s));
    }
    double lam = (dmax > 2.0) ? 2.0 / dmax : 1.0;
    // trace species may not rise above 1e-4 of the total in one step
    for (int j = 0; j < S; ++j) {
      if (n[j] / ntot <= 1e-8 && dln[j] >= 0.0) {
        const double lim = std::fabs((-std::log(n[j] / ntot) - 9.2103404) / (dln[j] - dlnn + 1e-300));
        lam = std::min(lam, lim);
      }
    }
    lam = std::min(1.0, lam);
    for (int j = 0; j < S; ++j) n[j] = std::exp(std::max(-700.0, std::log(n[j]) + lam * dln[j]));
    ntot *= std::exp(lam * dlnn);
    T *= std::exp(lam * dlnT);
    double conv = std::fabs(dlnn) + std::fabs(dlnT);
    for (int j = 0; j < S; ++j) conv = std::max(conv, n[j] / sumn * std::fabs(dln[j]));
    if (lam == 1.0 && conv < 1e-12 && it > 2) {
      std::fill(Xeq, Xeq + ns, 0.0);
      double s = 0.0;
      for (int j = 0; j < S; ++j) s += n[j];
      for (int j = 0; j < S; ++j) Xeq[sp[j]] = n[j] / s;
      *Teq = T;
      return 0;
    }
  }
  return -2;
}

void x2y(const Table& t, const double* X, double* Y) {
  double w = 0;
  for (int k = 0; k < t.ns; ++k) w += X[k] * t.W[k];
  for (int k = 0; k < t.ns; ++k) Y[k] = X[k] * t.W[k] / w;
}

}  // namespace

extern "C" {

void mi_cells_philox(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]) {
  philox4x32_10(ctr, key, out);
}

int mi_cells_equilibrium_hp(const char* tbl, const double* X0, double T0, double* Yeq, double* Teq) {
  Table t;
  if (!t.load(tbl)) return -10;
  std::vector<double> Xeq(t.ns);
  const int r = equilibrium_hp(t, X0, T0, Xeq.data(), Teq);
  if (r == 0) x2y(t, Xeq.data(), Yeq);
  return r;
}

int mi_cells_generate(const char* tbl, const mi_cells_spec* spec, long ncell, double* rY, double* T,
                      double* rE, double* c_out, int* class_out, int nthreads) {
  Table t;
  if (!t.load(tbl)) return -10;
  const int ifuel = t.index(spec->fuel), io2 = t.index("O2"), in2 = t.index("N2");
  if (ifuel < 0 || io2 < 0 || in2 < 0) return -11;
  const int ns = t.ns;
  const uint32_t key[2] = {uint32_t(spec->seed), uint32_t(spec->seed >> 32)};

  // class assignment (mode 1): exact split, then a seeded Fisher-Yates shuffle
  std::vector<int> cls;
  if (spec->mode == 1) {
    const long ncold = std::lround(0.6 * double(ncell));
    const long neq = std::lround(0.3 * double(ncell));
    cls.resize(ncell);
    for (long i = 0; i < ncell; ++i) cls[i] = (i < ncold) ? 0 : (i < ncold + neq ? 1 : 2);
    for (long i = ncell - 1; i >= 1; --i) {
      const uint32_t ctr[4] = {uint32_t(i), uint32_t(uint64_t(i) >> 32), 1u, 0u};
      uint32_t r[4];
      philox4x32_10(ctr, key, r);
      const uint64_t x = (uint64_t(r[0]) << 32) | r[1];
      const long j = long(x % uint64_t(i + 1));
      std::swap(cls[i], cls[j]);
    }
  }

  int nfail = 0;
#ifdef _OPENMP
  if (nthreads <= 0) nthreads = omp_get_max_threads();
#pragma omp parallel for schedule(dynamic, 256) num_threads(nthreads) reduction(+ : nfail)
#endif
  for (long i = 0; i < ncell; ++i) {
    const uint32_t ctr[4] = {uint32_t(i), uint32_t(uint64_t(i) >> 32), 0u, 0u};
    uint32_t r[4];
    philox4x32_10(ctr, key, r);
    double c;
    int cl = 0;
    if (spec->mode == 1) {
      cl = cls[i];
      const double u = u01(r[0]);
      c = (cl == 0) ? 0.05 * u : (cl == 1 ? 1.0 - 0.05 * u : 0.3 + 0.4 * u);
    } else {
      c = u01(r[0]);
    }
    const double phi = spec->phi_lo + (spec->phi_hi - spec->phi_lo) * u01(r[1]);
    const double z = std::sqrt(-2.0 * std::log(u01(r[2]))) * std::cos(2.0 * M_PI * u01(r[3]));
    std::vector<double> X0(ns, 0.0), Xeq(ns), Y0(ns), Yeq(ns), Y(ns);
    const double nO2 = spec->o2_per_fuel / phi;
    X0[ifuel] = 1.0;
    X0[io2] = nO2;
    X0[in2] = nO2 * 0.79 / 0.21;
    const double s = 1.0 + nO2 / 0.21;
    for (int k = 0; k < ns; ++k) X0[k] /= s;
    double Tad = spec->T_fresh;
    if (equilibrium_hp(t, X0.data(), spec->T_fresh, Xeq.data(), &Tad) != 0) {
      ++nfail;
      Xeq = X0;
      Tad = spec->T_fresh;
    }
    x2y(t, X0.data(), Y0.data());
    x2y(t, Xeq.data(), Yeq.data());
    double sum = 0.0;
    for (int k = 0; k < ns; ++k) {
      Y[k] = std::max((1.0 - c) * Y0[k] + c * Yeq[k], 1e-8);
      sum += Y[k];
    }
    for (int k = 0; k < ns; ++k) Y[k] /= sum;
    const double Tc = spec->T_fresh + c * (Tad - spec->T_fresh) + spec->T_noise * z;
    double invW = 0.0;
    for (int k = 0; k < ns; ++k) invW += Y[k] / t.W[k];
    const double rho = spec->pressure / (t.Ru * Tc * invW);
    for (int k = 0; k < ns; ++k) rY[(long)k * ncell + i] = rho * Y[k];
    T[i] = Tc;
    rE[i] = rho * t.h_mass(Tc, Y.data());
    if (c_out) c_out[i] = c;
    if (class_out) class_out[i] = cl;
  }
  return nfail;
}

}  // extern "C"
