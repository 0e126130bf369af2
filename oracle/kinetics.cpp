// CPU oracle: mechanism tables, thermo, production rates, reactor RHS and its
// analytic Jacobian.  TEST INFRASTRUCTURE ONLY (see chem_oracle.h).
//
// Rate law (Docs/source/manual/Model.rst:227-264): k_f = A T^beta exp(-Ta/T);
// third-body [M] = sum_k eps_k [X_k]; fall-off k = k_inf Pr/(1+Pr) F with
// Pr = k_0 [M]/k_inf and the Troe broadening (Gilbert, Luther & Troe 1983);
// k_r = k_f / K_c, K_c = exp(-sum nu g/RT) (P_atm/RT)^(sum nu); wdot_k = sum_j nu_kj q_j.
// Thermo: NASA-7 two-range polynomials (Docs/source/manual/Model.rst:271-296).
#include "oracle_internal.hpp"

#include <cmath>
#include <complex>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <sstream>
#include <string>

using cplx = std::complex<double>;

static inline double re(double x) { return x; }
static inline double re(const cplx& x) { return x.real(); }

// ---------------------------------------------------------------------------
// table loading (format written by pelelm_amd/mechanisms/codegen.py)
// ---------------------------------------------------------------------------
orc_mech* orc_load(const char* path) {
  std::ifstream in(path);
  if (!in) return nullptr;
  auto* m = new orc_mech();
  std::string line;
  while (std::getline(in, line)) {
    std::istringstream ss(line);
    std::string tag;
    ss >> tag;
    if (tag == "mech") {
      ss >> m->name >> m->ne >> m->ns >> m->nr >> m->Ru >> m->Patm;
    } else if (tag == "elem") {
      std::string e;
      ss >> e;
      m->elems.push_back(e);
    } else if (tag == "spec") {
      std::string name;
      double W, Tlo, Tmid, Thi;
      ss >> name >> W >> Tlo >> Tmid >> Thi;
      m->names.push_back(name);
      m->W.push_back(W);
      m->Tmid.push_back(Tmid);
      for (int i = 0; i < 7; ++i) { double a; ss >> a; m->lo.push_back(a); }
      for (int i = 0; i < 7; ++i) { double a; ss >> a; m->hi.push_back(a); }
      for (int e = 0; e < m->ne; ++e) { int c; ss >> c; m->ecomp.push_back(c); }
    } else if (tag == "reac") {
      Rxn r{};
      ss >> r.kind >> r.rev >> r.A >> r.b >> r.Ta;
      ss >> r.nr;
      for (int i = 0; i < r.nr; ++i) ss >> r.ri[i] >> r.rc[i];
      ss >> r.np;
      for (int i = 0; i < r.np; ++i) ss >> r.pi[i] >> r.pc[i];
      ss >> r.A0 >> r.b0 >> r.Ta0 >> r.ntroe;
      for (int i = 0; i < r.ntroe; ++i) ss >> r.troe[i];
      int neff;
      ss >> neff;
      r.eff.assign(m->ns, 1.0);
      for (int i = 0; i < neff; ++i) { int k; double e; ss >> k >> e; r.eff[k] = e; }
      r.dnu = 0;
      for (int i = 0; i < r.np; ++i) r.dnu += r.pc[i];
      for (int i = 0; i < r.nr; ++i) r.dnu -= r.rc[i];
      m->rx.push_back(r);
    }
  }
  if ((int)m->names.size() != m->ns || (int)m->rx.size() != m->nr || m->ns + 1 > ORC_NMAX) {
    delete m;
    return nullptr;
  }
  return m;
}

void orc_free(orc_mech* m) { delete m; }
int orc_nspec(const orc_mech* m) { return m->ns; }
int orc_nreac(const orc_mech* m) { return m->nr; }
int orc_nelem(const orc_mech* m) { return m->ne; }
double orc_gas_constant(const orc_mech* m) { return m->Ru; }
void orc_molwt(const orc_mech* m, double* W) { for (int k = 0; k < m->ns; ++k) W[k] = m->W[k]; }
void orc_elem_comp(const orc_mech* m, int* ec) {
  for (size_t i = 0; i < m->ecomp.size(); ++i) ec[i] = m->ecomp[i];
}

// ---------------------------------------------------------------------------
// thermo
// ---------------------------------------------------------------------------
template <class S>
static void thermo_t(const orc_mech* m, S T, S* cpR, S* hRT, S* sR) {
  const S logT = std::log(T);
  for (int k = 0; k < m->ns; ++k) {
    const double* a = (re(T) < m->Tmid[k]) ? &m->lo[7 * k] : &m->hi[7 * k];
    if (cpR) cpR[k] = a[0] + T * (a[1] + T * (a[2] + T * (a[3] + T * a[4])));
    if (hRT) hRT[k] = a[0] + T * (a[1] / 2.0 + T * (a[2] / 3.0 + T * (a[3] / 4.0 + T * a[4] / 5.0))) + a[5] / T;
    if (sR) sR[k] = a[0] * logT + T * (a[1] + T * (a[2] / 2.0 + T * (a[3] / 3.0 + T * a[4] / 4.0))) + a[6];
  }
}

// d(cp_k/R)/dT
static void dcpR_dT(const orc_mech* m, double T, double* d) {
  for (int k = 0; k < m->ns; ++k) {
    const double* a = (T < m->Tmid[k]) ? &m->lo[7 * k] : &m->hi[7 * k];
    d[k] = a[1] + T * (2.0 * a[2] + T * (3.0 * a[3] + T * 4.0 * a[4]));
  }
}

void orc_thermo(const orc_mech* m, double T, double* cpR, double* hRT, double* sR) {
  thermo_t<double>(m, T, cpR, hRT, sR);
}

// ---------------------------------------------------------------------------
// production rates (templated for the complex-step check)
// ---------------------------------------------------------------------------
template <class S>
static inline S ipow(S x, int n) {
  S r = x;
  for (int i = 1; i < n; ++i) r *= x;
  return r;
}

template <class S>
static S troe_F(const Rxn& r, S T, S Pr) {
  const double a = r.troe[0];
  S Fc = (1.0 - a) * std::exp(-T / r.troe[1]) + a * std::exp(-T / r.troe[2]);
  if (r.ntroe == 4) Fc += std::exp(-r.troe[3] / T);
  if (re(Fc) < 1e-300) Fc = S(1e-300);
  if (re(Pr) < 1e-300) Pr = S(1e-300);
  const S logFc = std::log10(Fc);
  const S c = -0.4 - 0.67 * logFc;
  const S n = 0.75 - 1.27 * logFc;
  const S u = std::log10(Pr) + c;
  const S f1 = u / (n - 0.14 * u);
  const S logF = logFc / (1.0 + f1 * f1);
  return std::exp(logF * M_LN10);
}

template <class S>
static void wdot_t(const orc_mech* m, const S* C, S T, S* wd, S* qnet, S* qfwd = nullptr,
                   S* qrev = nullptr) {
  S gRT[ORC_NMAX], hRT[ORC_NMAX], sR[ORC_NMAX];
  thermo_t<S>(m, T, nullptr, hRT, sR);
  for (int k = 0; k < m->ns; ++k) { gRT[k] = hRT[k] - sR[k]; wd[k] = 0.0; }
  const S logT = std::log(T);
  const S invT = 1.0 / T;
  const S refC = m->Patm / (m->Ru * T);
  for (int j = 0; j < m->nr; ++j) {
    const Rxn& r = m->rx[j];
    const S kf = r.A * std::exp(r.b * logT - r.Ta * invT);
    S keff = kf;
    if (r.kind != 0) {
      S Mc = 0.0;
      for (int k = 0; k < m->ns; ++k) Mc += r.eff[k] * C[k];
      if (r.kind == 1) {
        keff = kf * Mc;
      } else {
        const S k0 = r.A0 * std::exp(r.b0 * logT - r.Ta0 * invT);
        const S Pr = k0 * Mc / kf;
        S F = 1.0;
        if (r.ntroe > 0) F = troe_F<S>(r, T, Pr);
        keff = kf * Pr / (1.0 + Pr) * F;
      }
    }
    S pf = 1.0, pr = 1.0;
    for (int i = 0; i < r.nr; ++i) pf *= ipow<S>(C[r.ri[i]], r.rc[i]);
    for (int i = 0; i < r.np; ++i) pr *= ipow<S>(C[r.pi[i]], r.pc[i]);
    S q;
    if (r.rev) {
      S dG = 0.0;
      for (int i = 0; i < r.np; ++i) dG += double(r.pc[i]) * gRT[r.pi[i]];
      for (int i = 0; i < r.nr; ++i) dG -= double(r.rc[i]) * gRT[r.ri[i]];
      // 1/Kc = exp(dG) * refC^(-dnu)
      S invKc = std::exp(dG);
      if (r.dnu != 0) invKc *= std::pow(refC, -double(r.dnu));
      q = keff * (pf - pr * invKc);
      if (qrev) qrev[j] = keff * pr * invKc;
    } else {
      q = keff * pf;
      if (qrev) qrev[j] = 0.0;
    }
    if (qfwd) qfwd[j] = keff * pf;
    if (qnet) qnet[j] = q;
    for (int i = 0; i < r.nr; ++i) wd[r.ri[i]] -= double(r.rc[i]) * q;
    for (int i = 0; i < r.np; ++i) wd[r.pi[i]] += double(r.pc[i]) * q;
  }
}

void orc_wdot(const orc_mech* m, const double* C, double T, double* wd, double* qnet) {
  wdot_t<double>(m, C, T, wd, qnet);
}

void orc_rop(const orc_mech* m, const double* C, double T, double* qf, double* qr) {
  double wd[ORC_NMAX];
  wdot_t<double>(m, C, T, wd, (double*)nullptr, qf, qr);
}

// ---------------------------------------------------------------------------
// HY2T (Newton on h(T, Y) = h), EOS::HY2T as used at Source/PeleLM_K.H:89,527
// ---------------------------------------------------------------------------
double orc_hy2t(const orc_mech* m, double h, const double* Y, double T, int* iters) {
  double hRT[ORC_NMAX], cpR[ORC_NMAX];
  int it = 0;
  for (; it < ORC_HY2T_MAXIT; ++it) {
    thermo_t<double>(m, T, cpR, hRT, nullptr);
    double hT = 0.0, cp = 0.0;
    for (int k = 0; k < m->ns; ++k) {
      hT += Y[k] * hRT[k] / m->W[k];
      cp += Y[k] * cpR[k] / m->W[k];
    }
    hT *= m->Ru * T;
    cp *= m->Ru;
    const double dT = (h - hT) / cp;
    T += dT;
    if (std::fabs(dT) <= ORC_HY2T_RTOL * T) { ++it; break; }
  }
  if (iters) *iters = it;
  return T;
}

// ---------------------------------------------------------------------------
// reactor right-hand side
// ---------------------------------------------------------------------------
template <class S>
static void rhs_at_T(const orc_mech* m, const S* y, S T, const double* fY, double fH, S* yd) {
  const int N = m->ns;
  S C[ORC_NMAX], wd[ORC_NMAX], hRT[ORC_NMAX], cpR[ORC_NMAX];
  for (int k = 0; k < N; ++k) C[k] = y[k] / m->W[k];
  wdot_t<S>(m, C, T, wd, (S*)nullptr);
  thermo_t<S>(m, T, cpR, hRT, nullptr);
  // energy closure from rho h = sum_k rhoY_k h_k(T): rho cp dT/dt = F_h - sum_k h_k d(rhoY_k)/dt,
  // where d(rhoY_k)/dt = W_k wdot_k + F_Y,k (the species forcing carries its enthalpy)
  S rhocp = 0.0, hw = 0.0;
  for (int k = 0; k < N; ++k) {
    yd[k] = wd[k] * m->W[k] + (fY ? fY[k] : 0.0);
    rhocp += y[k] * cpR[k] / m->W[k];
    hw += (wd[k] + (fY ? fY[k] / m->W[k] : 0.0)) * hRT[k];
  }
  rhocp *= m->Ru;
  hw *= m->Ru * T;
  yd[N] = (fH - hw) / rhocp;
}

double orc_cell_T(const orc_mech* m, double t, const double* y, double rhoH0, double fH) {
  const int N = m->ns;
  double rho = 0.0, Y[ORC_NMAX];
  for (int k = 0; k < N; ++k) rho += y[k];
  for (int k = 0; k < N; ++k) Y[k] = y[k] / rho;
  return orc_hy2t(m, (rhoH0 + fH * t) / rho, Y, y[N], nullptr);
}

void orc_rhs(const orc_mech* m, double t, const double* y, const double* fY, double rhoH0, double fH,
             double* yd) {
  const double T = orc_cell_T(m, t, y, rhoH0, fH);
  rhs_at_T<double>(m, y, T, fY, fH, yd);
}

void orc_rhs_fixedT(const orc_mech* m, const double* y, const double* fY, double fH, double* yd) {
  rhs_at_T<double>(m, y, y[m->ns], fY, fH, yd);
}

void orc_jac_cs(const orc_mech* m, const double* y, const double* fY, double fH, double* J) {
  const int n = m->ns + 1;
  const double hstep = 1e-30;
  cplx yc[ORC_NMAX], yd[ORC_NMAX];
  for (int c = 0; c < n; ++c) {
    for (int i = 0; i < n; ++i) yc[i] = cplx(y[i], 0.0);
    yc[c] = cplx(y[c], hstep);
    rhs_at_T<cplx>(m, yc, yc[n - 1], fY, fH, yd);
    for (int i = 0; i < n; ++i) J[i * n + c] = yd[i].imag() / hstep;
  }
}

// ---------------------------------------------------------------------------
// analytic Jacobian of rhs_at_T (T independent)
// ---------------------------------------------------------------------------
void orc_jac(const orc_mech* m, const double* y, const double* fY, double fH, double* J) {
  const int N = m->ns, n = N + 1;
  const double T = y[N];
  double C[ORC_NMAX], gRT[ORC_NMAX], hRT[ORC_NMAX], sR[ORC_NMAX], cpR[ORC_NMAX], dcp[ORC_NMAX];
  // dwdot/dC (N x N, row-major) and dwdot/dT
  static thread_local double dWdC[ORC_NMAX * ORC_NMAX];
  double dWdT[ORC_NMAX], wd[ORC_NMAX];
  for (int k = 0; k < N; ++k) C[k] = y[k] / m->W[k];
  thermo_t<double>(m, T, cpR, hRT, sR);
  dcpR_dT(m, T, dcp);
  for (int k = 0; k < N; ++k) { gRT[k] = hRT[k] - sR[k]; wd[k] = 0.0; dWdT[k] = 0.0; }
  std::memset(dWdC, 0, sizeof(double) * N * N);
  const double logT = std::log(T), invT = 1.0 / T;
  const double refC = m->Patm / (m->Ru * T);
  double dq[ORC_NMAX];
  for (int j = 0; j < m->nr; ++j) {
    const Rxn& r = m->rx[j];
    const double kf = r.A * std::exp(r.b * logT - r.Ta * invT);
    const double dlnkf = (r.b + r.Ta * invT) * invT;   // d ln kf / dT
    double keff = kf, dkeff_dT = kf * dlnkf, dkeff_dM = 0.0, Mc = 0.0;
    if (r.kind != 0) {
      for (int k = 0; k < N; ++k) Mc += r.eff[k] * C[k];
      if (r.kind == 1) {
        keff = kf * Mc;
        dkeff_dT = kf * dlnkf * Mc;
        dkeff_dM = kf;
      } else {
        const double k0 = r.A0 * std::exp(r.b0 * logT - r.Ta0 * invT);
        const double dlnk0 = (r.b0 + r.Ta0 * invT) * invT;
        const double Pr = k0 * Mc / kf;
        const double dPr_dM = k0 / kf;
        const double dPr_dT = Pr * (dlnk0 - dlnkf);
        double F = 1.0, dF_dPr = 0.0, dF_dT = 0.0;
        if (r.ntroe > 0 && Pr > 1e-300) {
          const double a = r.troe[0];
          const double e3 = std::exp(-T / r.troe[1]), e1 = std::exp(-T / r.troe[2]);
          double Fc = (1.0 - a) * e3 + a * e1;
          double dFc = -(1.0 - a) / r.troe[1] * e3 - a / r.troe[2] * e1;
          if (r.ntroe == 4) {
            const double e2 = std::exp(-r.troe[3] / T);
            Fc += e2;
            dFc += r.troe[3] / (T * T) * e2;
          }
          if (Fc < 1e-300) { Fc = 1e-300; dFc = 0.0; }
          const double logFc = std::log10(Fc);
          const double c = -0.4 - 0.67 * logFc, nn = 0.75 - 1.27 * logFc;
          const double u = std::log10(Pr) + c;
          const double d = nn - 0.14 * u;
          const double f1 = u / d;
          const double den = 1.0 + f1 * f1;
          const double logF = logFc / den;
          F = std::exp(logF * M_LN10);
          const double df1_du = nn / (d * d);                     // at fixed logFc
          const double df1_dlFc = (-0.67 * d + 1.1762 * u) / (d * d);
          const double dlogF_dlPr = -logFc * 2.0 * f1 / (den * den) * df1_du;
          const double dlogF_dlFc = 1.0 / den - logFc * 2.0 * f1 / (den * den) * df1_dlFc;
          dF_dPr = F * dlogF_dlPr / Pr;                           // F ln10 dlogF/dlogPr * 1/(Pr ln10)
          dF_dT = F * M_LN10 * dlogF_dlFc * dFc / (Fc * M_LN10);
        }
        const double g = Pr / (1.0 + Pr);
        const double dg = 1.0 / ((1.0 + Pr) * (1.0 + Pr));
        keff = kf * g * F;
        dkeff_dM = kf * (dg * F + g * dF_dPr) * dPr_dM;
        dkeff_dT = kf * dlnkf * g * F + kf * (dg * F + g * dF_dPr) * dPr_dT + kf * g * dF_dT;
      }
    }
    // concentration products and their partial derivatives
    double pf = 1.0, pr = 1.0;
    for (int i = 0; i < r.nr; ++i) pf *= ipow<double>(C[r.ri[i]], r.rc[i]);
    for (int i = 0; i < r.np; ++i) pr *= ipow<double>(C[r.pi[i]], r.pc[i]);
    double invKc = 0.0, dlnKc = 0.0;
    if (r.rev) {
      double dG = 0.0, dH = 0.0;
      for (int i = 0; i < r.np; ++i) { dG += r.pc[i] * gRT[r.pi[i]]; dH += r.pc[i] * hRT[r.pi[i]]; }
      for (int i = 0; i < r.nr; ++i) { dG -= r.rc[i] * gRT[r.ri[i]]; dH -= r.rc[i] * hRT[r.ri[i]]; }
      invKc = std::exp(dG);
      if (r.dnu != 0) invKc *= std::pow(refC, -double(r.dnu));
      dlnKc = (dH - r.dnu) * invT;
    }
    const double net = pf - pr * invKc;
    const double q = keff * net;
    for (int k = 0; k < N; ++k) dq[k] = (r.kind != 0) ? dkeff_dM * r.eff[k] * net : 0.0;
    // d pf / d C_s for each reactant s: rc * C_s^(rc-1) * prod_{other}
    for (int i = 0; i < r.nr; ++i) {
      double d = r.rc[i] * ipow<double>(C[r.ri[i]], r.rc[i] - 1);
      if (r.rc[i] == 1) d = 1.0;
      for (int i2 = 0; i2 < r.nr; ++i2)
        if (i2 != i) d *= ipow<double>(C[r.ri[i2]], r.rc[i2]);
      dq[r.ri[i]] += keff * d;
    }
    if (r.rev) {
      for (int i = 0; i < r.np; ++i) {
        double d = r.pc[i] * ipow<double>(C[r.pi[i]], r.pc[i] - 1);
        if (r.pc[i] == 1) d = 1.0;
        for (int i2 = 0; i2 < r.np; ++i2)
          if (i2 != i) d *= ipow<double>(C[r.pi[i2]], r.pc[i2]);
        dq[r.pi[i]] -= keff * invKc * d;
      }
    }
    const double dqdT = dkeff_dT * net + keff * pr * invKc * dlnKc;
    auto scatter = [&](int s, double nu) {
      wd[s] += nu * q;
      dWdT[s] += nu * dqdT;
      double* row = &dWdC[s * N];
      for (int k = 0; k < N; ++k) row[k] += nu * dq[k];
    };
    for (int i = 0; i < r.nr; ++i) scatter(r.ri[i], -double(r.rc[i]));
    for (int i = 0; i < r.np; ++i) scatter(r.pi[i], double(r.pc[i]));
  }
  // assemble J of rhs_at_T
  // molar species source wdot_k + F_Y,k / W_k (what the energy equation weighs by h_k)
  double src[ORC_NMAX];
  for (int k = 0; k < N; ++k) src[k] = wd[k] + (fY ? fY[k] / m->W[k] : 0.0);
  double rhocp = 0.0, drhocp = 0.0, hw = 0.0;
  for (int k = 0; k < N; ++k) {
    rhocp += y[k] * cpR[k] / m->W[k];
    drhocp += y[k] * dcp[k] / m->W[k];
    hw += src[k] * hRT[k];
  }
  rhocp *= m->Ru;
  drhocp *= m->Ru;
  hw *= m->Ru * T;
  const double fT = (fH - hw) / rhocp;
  for (int i = 0; i < N; ++i) {
    for (int k = 0; k < N; ++k) J[i * n + k] = m->W[i] / m->W[k] * dWdC[i * N + k];
    J[i * n + N] = m->W[i] * dWdT[i];
  }
  // T row: fT = (fH - sum_i H_i (wdot_i + F_Y,i / W_i))/rhocp with H_i = hRT_i R T (molar);
  // the forcing is constant, so only wdot_i depends on rhoY_k
  for (int k = 0; k < N; ++k) {
    double s = 0.0;
    for (int i = 0; i < N; ++i) s += hRT[i] * dWdC[i * N + k];
    s *= m->Ru * T / m->W[k];
    const double cpk = m->Ru * cpR[k] / m->W[k];
    J[N * n + k] = (-s - fT * cpk) / rhocp;
  }
  {
    double s = 0.0;
    for (int i = 0; i < N; ++i) s += cpR[i] * src[i] + hRT[i] * T * dWdT[i];
    s *= m->Ru;
    J[N * n + N] = (-s - fT * drhocp) / rhocp;
  }
}
