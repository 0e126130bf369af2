/* Seeded synthetic reactor inputs (host-only helper library, libpelelm_cells.so).
 * Not part of the reactor's drop-in interface: it builds the cell states the
 * benchmarks and tests feed through it. */
#ifndef PELELM_CELLS_H
#define PELELM_CELLS_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
  const char* fuel;        /* fuel species name, e.g. "CH4", "H2" */
  double o2_per_fuel;      /* stoichiometric O2 moles per fuel mole (CH4 2, H2 0.5) */
  double phi_lo, phi_hi;   /* equivalence ratio ~ U[phi_lo, phi_hi] */
  int mode;                /* 0: c ~ U[0,1]; 1: 60 % c<0.05 / 30 % c>0.95 / 10 % c in [0.3,0.7], shuffled */
  uint64_t seed;
  double T_fresh;          /* 300 K */
  double T_noise;          /* 20 K standard deviation */
  double pressure;         /* dyn/cm^2 */
} mi_cells_spec;

/* Fill SoA arrays rY[nspec*ncell], T[ncell], rE[ncell] (CGS).  c_out / class_out nullable.
   Returns the number of cells whose equilibrium solve failed (they fall back to fresh gas),
   or a negative code if the table cannot be read or a species is missing. */
int mi_cells_generate(const char* tbl_path, const mi_cells_spec* spec, long ncell, double* rY, double* T,
                      double* rE, double* c_out, int* class_out, int nthreads);

/* HP equilibrium at 1 atm of mole fractions X0 at T0: mass fractions Yeq, temperature Teq. */
int mi_cells_equilibrium_hp(const char* tbl_path, const double* X0, double T0, double* Yeq, double* Teq);

/* Philox4x32-10 block (exposed for known-answer tests). */
void mi_cells_philox(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]);

#ifdef __cplusplus
}
#endif
#endif
