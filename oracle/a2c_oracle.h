/*
 * a2c_oracle.h -- FP64 CPU restatement of the a2c free-entropy hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker for the HIP path:
 * only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
 * load it, and never as the thing measured or shipped.  The product library
 * (a2c_amd/liba2c_fe.so) does not link it and has no CPU fallback.
 *
 * Every function restates the reference's algorithm from the equations in
 * the reference sources (cited per function in a2c_oracle.cpp):
 *   potential.cpp            EAM Pd-H closed forms + central-FD derivatives
 *   quadrature.cpp:43-75     12-point phase-packet rule (weight 0.5/dim, :39)
 *   electrondensity.cpp      per-site averaged electron density
 *   minimizer.cpp:83-369     objective + gradient (TAO callback)
 *   minimizer.cpp:588-802    chemical potential (occupancy derivative)
 *   neighbors.cpp:8-113,
 *   minimizer.cpp:833-1144   neighbor sets, Pd shell counts, first H shell
 *   integrate.cpp:9-87       mean-field H diffusion step
 *
 * Pinning status per function (what checks it besides its own formula):
 *   or_eam*         constants == values extracted from potential.cpp text
 *                   (tests/golden/eam_constants.json); value continuity at the
 *                   switch radii; FCC Pd E(a) minimum near a=3.89; FD vs
 *                   analytic derivatives.
 *   or_quadrature   bit-exact vs the reference's gaussianQuadrature() built
 *                   from quadrature.cpp (oracle/_ref); Gaussian moments.
 *   GH3 rule        degree-4 test potential against the closed-form Gaussian
 *                   moment |d|^4 + 10 s^2 |d|^2 + 15 s^4 (which the 12-point
 *                   rule misses); FD gradient; FD chemical potential.
 *   or_nbr_*        brute-force all-pairs; FCC shell counts; list symmetry.
 *   or_density,
 *   or_eval         sigma->0 limit equals static EAM energy; degree-2 test
 *                   potential against the closed-form Gaussian moment;
 *                   gradient vs central FD of the objective; translation
 *                   invariance; sum of q-gradients == 0.
 *   or_chem_potential  equals central FD of V w.r.t. occupancy (coe = 0);
 *                   correction terms against closed-form FCC neighbor counts.
 *   or_diffuse      two-site closed form; conservation; zero-flux fixed point.
 */
#ifndef A2C_ORACLE_H
#define A2C_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum { OR_F_PD = 0, OR_F_H = 1, OR_RHO_PD = 2, OR_RHO_H = 3,
       OR_PHI_PD = 4, OR_PHI_PDH = 5, OR_PHI_H = 6, OR_NFUNC = 7 };
enum { OR_DERIV_ANALYTIC = 0, OR_DERIV_FD = 1 };
/* OR_POT_TEST replaces every EAM function by g(r)=r^2 (pairs and density) and
 * F(rho)=rho, so the averaged energy has a closed-form Gaussian moment;
 * OR_POT_TEST4 uses g(r)=r^4 (exact for GH3, not for the 12-point rule). */
enum { OR_POT_EAM = 0, OR_POT_TEST = 1, OR_POT_TEST4 = 2, OR_POT_TAB = 3 };
/* Phase-packet rules: the reference's 12-point rule (quadrature.cpp:43-75), or
 * a 3-point-per-axis Gauss-Hermite product rule on the relative coordinate
 * (27 nodes; builder addition, not in the reference): every pair function
 * depends only on r = d + sqrt(2)(sigma_i xi_1 - sigma_j xi_2) ~ N(d, s^2 I)
 * with s^2 = sigma_i^2 + sigma_j^2, so E g(|r|) = sum_k w_k g(|d + s z_k|). */
enum { OR_QUAD_R12 = 0, OR_QUAD_GH3 = 1, OR_QUAD_TABLE = 2 };

typedef struct or_params {
  int32_t n_pd, n_h;
  double box[3];
  int32_t periodic[3];
  double rc;          /* neighbor cutoff (Angstrom) */
  double a_pd;        /* lattice constant used by the shell classification */
  double T;           /* temperature (K) */
  int32_t deriv_mode; /* OR_DERIV_* */
  double eps;         /* FD step (reference: input.file.eps) */
  int32_t potential;  /* OR_POT_* */
  int32_t threads;    /* OpenMP threads for the per-site loops (<=0: 1) */
  int32_t quad_rule;  /* OR_QUAD_* */
} or_params;

double or_eam(int fn, double x);
double or_eam_d(int fn, double x, int deriv_mode, double eps);
/* Constants of function fn in order of appearance in potential.cpp. */
int or_eam_constants(int fn, double* out, int cap);

/* Reference rule: K=12 nodes xi[k][0..5] (site i: 0..2, site j: 3..5), w[k]. */
int or_quadrature(double* xi, double* w);

/* OR_POT_TAB tables (process-wide): tables[f] for f in OR_F_PD..OR_PHI_H,
 * node k at k*d_rho (F) or k*d_r (the others; pair functions as r*phi). */
int or_set_tabulated(int n_rho, double d_rho, int n_r, double d_r, const double* const* tables);
/* OR_QUAD_TABLE nodes (process-wide): xi K x 6, w K. */
int or_set_quad_table(int K, const double* xi, const double* w);
/* One table on its own: value and slope at x (emb: continue past the end). */
int or_tab_eval(int n, double h, const double* y, int emb, double x, double* v, double* d);

typedef struct or_nbr or_nbr;
or_nbr* or_nbr_build(const or_params* p, const double* X);
/* which: 0 PdPd, 1 PdH, 2 HPd, 3 HH, 4 HH1.  Returns row count. */
int or_nbr_get(const or_nbr* nb, int which, const int32_t** off, const int32_t** idx);
const int32_t* or_nbr_pdpdnum(const or_nbr* nb);
int or_nbr_max(const or_nbr* nb);
void or_nbr_free(or_nbr* nb);

int or_density(const or_params* p, const or_nbr* nb, const double* X,
               const double* x, double* rho);
int or_eval(const or_params* p, const or_nbr* nb, const double* X,
            const double* x, double* f, double* G);
int or_chem_potential(const or_params* p, const or_nbr* nb, const double* X,
                      const double* x0, const double* coe_pd, const double* coe_h,
                      double gamma_bd, double* gamma, int32_t* full_h);
int or_diffuse(const or_params* p, const or_nbr* nb, double dt, double nu, double Qm,
               double xLb, double xUb, const double* gamma, const int32_t* full_h,
               const double* x0, double* x);

#ifdef __cplusplus
}
#endif
#endif
