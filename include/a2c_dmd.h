/*
 * a2c_dmd.h -- diffusive-MD time step around the hot path (host C++).
 *
 * The reference's DMD driver is absent from its tree; the step order below is
 * the one its functions imply (findNeighbors -> enforceBoundaryConditions ->
 * minimizeFreeEntropy -> calculateChemicalPotential -> integrateDiffusion ->
 * calculateMacroResults).  The helpers restate boundaryconditions.cpp:30-55,
 * calculateresults.cpp:8-78 and readresults.cpp:10-55 on the host; the
 * restart writer has no counterpart in the reference (it writes the format
 * readresults.cpp reads).
 */
#ifndef A2C_DMD_H
#define A2C_DMD_H
#include <stdint.h>

#include "a2c_fe.h"

#ifdef __cplusplus
extern "C" {
#endif

/* enforceBoundaryConditions (boundaryconditions.cpp:30-55): subsurface H
 * sites become reservoirs: gamma = gamma_bd, full_H = 1, x0 = xUb. */
int a2c_enforce_boundary_conditions(int32_t n_sub, const int32_t* h_subsurf, double gamma_bd, double xUb,
                                    double* x0, double* gamma, int32_t* full_h);

/* calculateMacroResults (calculateresults.cpp:8-78): mean H fraction over all
 * H sites and over the sites not listed in h_surf. */
int a2c_macro_results(int32_t n_h, const double* x, int32_t n_surf, const int32_t* h_surf, double* xH,
                      double* xH_interior);

/* readResults (readresults.cpp:10-55): skip 9 header lines, then 9 columns per
 * site: Pd rows "? ? qx qy qz sigma ? ? ?", H rows "? ? qx qy qz sigma x gamma
 * full_H"; x is clamped to [xLb, xUb].  Returns A2C_EINVAL if the file is
 * missing or short (the reference exits). */
int a2c_read_results(const char* path, int32_t n_pd, int32_t n_h, double xLb, double xUb, double* X, double* x,
                     double* gamma, int32_t* full_h);
/* Writer of that format: 9 header lines, then "index species qx qy qz sigma x
 * gamma full_H" per site (Pd rows carry x = 1, gamma = 0, full_H = 0). */
int a2c_write_results(const char* path, int32_t n_pd, int32_t n_h, const double* X, const double* x,
                      const double* gamma, const int32_t* full_h, double time, int32_t step);

/* Extended-XYZ snapshot in the layout of Output::OutputSolutionOnLatticeXYZ
 * (Output.cpp:332-494): "Lattice", "Origin", "pbc" and "Properties" header,
 * then per site: species, position q, displacement q - q0, type (0 Pd, 1 H),
 * sigma, tag, molar fraction x (1 for Pd) and chemical potential gamma. */
int a2c_write_xyz(const char* path, int32_t n_pd, int32_t n_h, const double* X, const double* X0, const double* x,
                  const double* gamma, const double* box, const int32_t* periodic, double time);

typedef struct a2c_dmd_params {
  double dt, nu, Qm;        /* time step, attempt frequency, activation energy */
  double xLb, xUb;          /* occupancy clamp (integrate.cpp:79-81) */
  double gamma_bd;          /* chemical-potential cap (minimizer.cpp:796) */
  double coe_pd[4], coe_h[4];
  int32_t rebuild_neighbors;  /* rebuild the lists from q at the step start */
  a2c_lmvm_opts lmvm;
} a2c_dmd_params;

typedef struct a2c_dmd_report {
  double f;               /* free entropy after the minimization */
  int32_t lmvm_reason, lmvm_iterations, lmvm_evals;
  double total_h_before;  /* sum of x before the diffusion update (compensated sum) */
  double total_h_after;   /* and after */
  double xH, xH_interior;
  double flux_sum;        /* the diffusion step's balance (a2c_fe_diffusion_balance): */
  double flux_gross;      /*   applied increments, their scale, */
  double clamp_lift;      /*   the H added by the clamp to [xLb, xUb] */
  double reservoir_supply;  /* and the H the full_H reservoirs supplied */
} a2c_dmd_report;

void a2c_dmd_params_default(a2c_dmd_params* p);

/* One DMD step on a context: [rebuild lists] -> BC -> minimize X ->
 * chemical potential -> diffusion (x updated, also in the context) -> macro
 * results.  X (4N), x (n_h), gamma (n_h), full_h (n_h) are host arrays
 * updated in place. */
int a2c_dmd_step(a2c_fe_ctx* ctx, const a2c_dmd_params* p, int32_t n_sub, const int32_t* h_subsurf, int32_t n_surf,
                 const int32_t* h_surf, double* X, double* x, double* gamma, int32_t* full_h, a2c_dmd_report* rep);

#ifdef __cplusplus
}
#endif
#endif
