// Internal declarations shared by the oracle's translation units.
// TEST INFRASTRUCTURE ONLY (see chem_oracle.h).
#pragma once
#include <string>
#include <vector>

#include "chem_oracle.h"

#define ORC_NMAX 96            // largest ODE size (species + 1) the oracle handles
#define ORC_HY2T_MAXIT 50
#define ORC_HY2T_RTOL 1e-10    // Newton on T stops once |dT| <= 1e-10 T

struct Rxn {
  int kind, rev;               // kind: 0 elementary, 1 three-body, 2 fall-off
  double A, b, Ta;
  int nr, np;
  int ri[4], rc[4], pi[4], pc[4];
  double A0, b0, Ta0;
  int ntroe;
  double troe[4];
  std::vector<double> eff;     // dense third-body efficiencies (default 1)
  int dnu;                     // sum nu'' - sum nu'
};

struct orc_mech {
  std::string name;
  int ne = 0, ns = 0, nr = 0;
  double Ru = 0.0, Patm = 0.0;
  std::vector<std::string> elems, names;
  std::vector<double> W, Tmid, lo, hi;
  std::vector<int> ecomp;
  std::vector<Rxn> rx;
};

double orc_cell_T(const orc_mech* m, double t, const double* y, double rhoH0, double fH);
