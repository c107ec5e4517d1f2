/*
 * a2c_fe.h -- C-ABI of the MI355X free-entropy engine for Gaussian phase
 * packets (Pd-H EAM).  Plain pointers and sizes; no C++ or torch types.
 *
 * Host data layout (identical to the reference's TAO vector):
 *   X, G : 4*N doubles, site-interleaved [qx, qy, qz, sigma], all Pd sites
 *          first, then all H sites (minimizer.cpp:398-418, 464-473).
 *   x    : n_h doubles, H site occupancy (minimizer.h:18).
 * N = n_pd + n_h.  Pd index i is site i; H index a is site n_pd + a.
 *
 * Status codes: 0 = OK, negative = error (A2C_E*); the message of the last
 * error on a context is returned by a2c_fe_last_error().  No entry point exits
 * the process (the reference exits or aborts: electrondensity.cpp:50-53,
 * neighbors.cpp:52-56, minimizer.cpp:420-423).
 *
 * Threading: one context per host thread.  Calls are synchronous on the
 * context's HIP stream unless named *_device / *_async.
 *
 * Memory: the device buffers of contexts and batches come from a
 * process-wide cache; destroying a context returns them to the cache for
 * the next context, not to the device (large cached blocks go back to the
 * device when an allocation would otherwise fail).  a2c_dev_alloc memory is
 * allocated and freed directly.
 */
#ifndef A2C_FE_H
#define A2C_FE_H
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define A2C_OK 0
#define A2C_EINVAL (-1)         /* bad argument or descriptor */
#define A2C_ENOMEM (-2)         /* device or host allocation failed */
#define A2C_EHIP (-3)           /* HIP runtime error */
#define A2C_ENBR_OVERFLOW (-4)  /* a site has more neighbors than max_neighbors */
#define A2C_ENONFINITE (-5)     /* NaN/Inf objective or gradient */
#define A2C_ESTATE (-6)         /* call order: e.g. eval before build_neighbors */
#define A2C_ENOTCONV (-7)       /* optimizer stopped without converging */

enum a2c_deriv_mode { A2C_DERIV_ANALYTIC = 0, A2C_DERIV_FD = 1 };
enum a2c_precision { A2C_PREC_FP64 = 0, A2C_PREC_MIXED = 1 };
enum a2c_quad_rule {
  A2C_QUAD_R12 = 0, /* the reference's 12-point rule (quadrature.cpp:43-75), weight 1/12 */
  A2C_QUAD_GH3 = 1  /* 3-point-per-axis Gauss-Hermite rule on the relative coordinate
                       r = d + s z, s^2 = sigma_i^2 + sigma_j^2 (27 nodes; not in the
                       reference: exact to degree 5 per axis where the 12-point rule
                       is exact to degree 3) */
};

typedef struct a2c_fe_desc {
  int32_t n_pd, n_h;
  double box[3];          /* box edge lengths; used on periodic axes */
  int32_t periodic[3];    /* 1 = minimum image on that axis (needs box >= 2 rc) */
  double rc;              /* neighbor cutoff, Angstrom (input.file.rc; 5.35) */
  double a_pd;            /* Pd lattice constant for shell classification (input.file.a_Pd) */
  double temperature;     /* T (input.file.T0), K */
  int32_t quad_rule;      /* A2C_QUAD_R12 (default) or A2C_QUAD_GH3 */
  int32_t deriv_mode;     /* A2C_DERIV_ANALYTIC or A2C_DERIV_FD (reference) */
  double fd_eps;          /* FD step for A2C_DERIV_FD (input.file.eps; 1e-5) */
  int32_t precision;      /* A2C_PREC_FP64 or A2C_PREC_MIXED */
  int32_t device;         /* HIP device ordinal */
  int32_t max_neighbors;  /* per-site cap over all lists (0: 256) */
} a2c_fe_desc;

typedef struct a2c_fe_ctx a2c_fe_ctx;

typedef struct a2c_nbr_stats {
  int32_t max_neib;       /* max neighbors of any site (findNeighbors' return) */
  int32_t width[5];       /* max row length of PdPd, PdH, HPd, HH, HH1 */
  int64_t total[5];       /* entries in PdPd, PdH, HPd, HH, HH1 */
  int32_t refreshed;      /* 1: lists filtered from the context's skin lists (below) */
  double max_displacement;  /* largest site displacement since the skin lists were built (A) */
} a2c_nbr_stats;

/* Fill a descriptor with the reference defaults (rc 5.35, eps 1e-5, R12,
 * analytic derivatives, FP64, device 0, non-periodic). */
void a2c_fe_desc_default(a2c_fe_desc* d);

/* Replaces Minimizer::initialize (minimizer.cpp:484-546): sizes the problem
 * and allocates device state. */
int a2c_fe_create(const a2c_fe_desc* desc, a2c_fe_ctx** out);
void a2c_fe_destroy(a2c_fe_ctx* ctx);
const char* a2c_fe_last_error(const a2c_fe_ctx* ctx);
int a2c_fe_set_precision(a2c_fe_ctx* ctx, int32_t precision);

/* Occupancy x0 held fixed during minimization (AppCtx::x0, minimizer.h:18;
 * updateApplicationContext, minimizer.cpp:372-378). */
int a2c_fe_set_occupancy(a2c_fe_ctx* ctx, const double* x);

/* Replaces findNeighbors (neighbors.cpp:8-113) + Minimizer::findLocalNeighbors
 * (minimizer.cpp:833-1144): neighbor sets at rc from the positions in X
 * (sigma ignored), PdPdnum shell counts and the first H shell HH1.  Lists are
 * frozen until the next call.  From the second call on a context also keeps
 * lists at rc + skin (skin 0.2 A; environment A2C_NBR_SKIN, 0 to disable)
 * and, while no site has moved by skin/2 since they were built, filters them
 * instead of walking the cells: the same sets, rows and shell counts. */
int a2c_fe_build_neighbors(a2c_fe_ctx* ctx, const double* X, a2c_nbr_stats* stats);
/* The same from device-resident positions (dX: 4N doubles in HBM, e.g. the
 * state an optimizer left there); the context state becomes dX. */
int a2c_fe_build_neighbors_device(a2c_fe_ctx* ctx, const double* dX, a2c_nbr_stats* stats);

/* CSR export of one list (0 PdPd, 1 PdH, 2 HPd, 3 HH, 4 HH1), rows sorted by
 * neighbor index.  offsets: rows+1 entries; indices: total[which] entries.
 * Indices are species-local (Pd index or H index) like the reference. */
int a2c_fe_get_neighbors(a2c_fe_ctx* ctx, int32_t which, int32_t* offsets, int32_t* indices);
/* PdPdnum: n_pd x 3 shell counts (minimizer.cpp:935-947). */
int a2c_fe_get_pdpdnum(a2c_fe_ctx* ctx, int32_t* pdpdnum);

/* Per-site averaged electron density rho (N doubles, Pd then H):
 * electronDensity (electrondensity.cpp:12-56) for every site. */
int a2c_fe_density(a2c_fe_ctx* ctx, const double* X, double* rho);

/* THE HOT PATH.  Replaces Minimizer::formFunctionAndGradient
 * (minimizer.cpp:83-369): f = V/T - 3 kB sum x ln sigma and G = df/dX. */
int a2c_fe_eval(a2c_fe_ctx* ctx, const double* X, double* f, double* G);

/* Same on device-resident buffers (HBM), asynchronous on the context stream:
 * dX, dG are device pointers of 4N doubles, df a device pointer to 1 double. */
int a2c_fe_eval_device(a2c_fe_ctx* ctx, const double* dX, double* dG, double* df);

/* TAO-callback-shaped adapter, PetscErrorCode (*)(Tao, Vec X, PetscReal* f,
 * Vec G, void* ctx) (minimizer.h:77-78): solver is ignored, ctx is an
 * a2c_fe_ctx*, X and G are host arrays of 4N doubles. */
int a2c_tao_objgrad(void* solver, const double* X, double* f, double* G, void* ctx);

/* Replaces Minimizer::calculateChemicalPotential (minimizer.cpp:588-802):
 * gamma_a = ln(x/(1-x)) + dV/dx_a/(kB T) + corrections + 1.5, capped at
 * gamma_bd; full_h set when capped and x > 0.9999.  Sites with full_h == 1 on
 * entry keep gamma.  x0 is the context occupancy.  coe_pd/coe_h: 4 each. */
int a2c_fe_chem_potential(a2c_fe_ctx* ctx, const double* X, const double* coe_pd,
                          const double* coe_h, double gamma_bd, double* gamma,
                          int32_t* full_h);

/* ---- general potentials and quadrature rules -------------------------
 * The reference parses InteratomicPotential.File (IoData.cpp:523-534,
 * IoData.h:289-300) but evaluates only its compiled-in Pd-H functions
 * (potential.cpp), and uses only its 12-point rule (quadrature.cpp:43-75).
 * These calls give a context a tabulated potential and/or a node table; the
 * context then evaluates (a2c_fe_eval, _density, _chem_potential, the
 * optimizer) with per-site kernels over the full lists, FP64 throughout
 * (A2C_PREC_FP64) or FP32 pair terms with FP64 site sums (A2C_PREC_MIXED),
 * analytic or central-FD derivatives.  Batches do not take such contexts.
 *
 * Tabulated functions are uniform grids from 0 (value k at k*d), interpolated
 * by cubic Hermite pieces whose node slopes are fourth-order central
 * differences (second-order and one-sided at the ends) -- the interpolation
 * the DYNAMO setfl format is written for.  Pair functions are given as
 * r*phi(r).  Past the last node a density or pair function is 0; an
 * embedding function continues along its end slope. */
typedef struct a2c_eam_tabulated {
  int32_t n_rho;      /* nodes of F_pd and F_h, on rho_k = k * d_rho */
  double d_rho;
  int32_t n_r;        /* nodes of the density and pair functions, on r_k = k * d_r */
  double d_r;
  const double* F_pd;      /* embedding energy of a Pd site, n_rho values */
  const double* F_h;       /* embedding energy of an H site */
  const double* f_pd;      /* density contributed by a Pd neighbor, n_r values */
  const double* f_h;       /* density contributed by an H neighbor */
  const double* rphi_pdpd; /* r * phi(r) of Pd-Pd, Pd-H and H-H pairs, n_r values */
  const double* rphi_pdh;
  const double* rphi_hh;
} a2c_eam_tabulated;
int a2c_fe_set_potential_tables(a2c_fe_ctx* ctx, const a2c_eam_tabulated* t);
/* The same from a DYNAMO setfl file (the eam/alloy layout: 3 comment lines;
 * "n el_1 .. el_n"; "n_rho d_rho n_r d_r cutoff"; per element a header line,
 * F(rho) and f(r); then r*phi for every pair i >= j), taking the elements
 * named host (the Pd sites) and solute (the H sites). */
int a2c_fe_set_potential_file(a2c_fe_ctx* ctx, const char* path, const char* host, const char* solute);
/* Host-only reader behind it: the grid (any output may be null) and the seven
 * tables a2c_fe_set_potential_file installs, copied into caller arrays of
 * n_rho or n_r doubles (null arrays skipped: a first call with null arrays
 * returns the sizes). */
int a2c_eam_setfl_read(const char* path, const char* host, const char* solute, int32_t* n_rho, double* d_rho,
                       int32_t* n_r, double* d_r, double* cutoff, double* F_pd, double* F_h, double* f_pd,
                       double* f_h, double* rphi_pdpd, double* rphi_pdh, double* rphi_hh);
/* Back to the compiled-in Pd-H functions. */
int a2c_fe_clear_potential(a2c_fe_ctx* ctx);

/* A node table for the phase-packet average: node k displaces site i by
 * sqrt(2) sigma_i xi1_k and site j by sqrt(2) sigma_j xi2_k, so the pair
 * vector is r_k = q_i - q_j + sqrt(2) (sigma_i xi1_k - sigma_j xi2_k), with
 * weight w_k (the reference's rule is xi = +-sqrt(3) e_m, w = 1/12).
 * xi: K x 6 doubles (xi1 then xi2 per node), w: K doubles.  The table must be
 * swap-symmetric -- (xi2_k, xi1_k, w_k) is also a node -- so a pair's average
 * is the same from either end, which the reference's gradient assumes.  K in
 * 1..A2C_QUAD_MAX_NODES; K = 0 restores the descriptor's rule. */
#define A2C_QUAD_MAX_NODES 512
int a2c_fe_set_quadrature(a2c_fe_ctx* ctx, int32_t K, const double* xi, const double* w);

/* Replaces integrateDiffusion (integrate.cpp:9-87): one explicit step of the
 * mean-field master equation on HH1 with rate nu*exp(-Qm/(kB T)); full_h sites
 * keep x; result clamped to [xLb, xUb].  x0 is the context occupancy; the new
 * occupancy is written to x (n_h) and becomes the context occupancy. */
int a2c_fe_diffuse(a2c_fe_ctx* ctx, double dt, double nu, double Qm, double xLb,
                   double xUb, const double* gamma, const int32_t* full_h, double* x);

/* Balance of total H over one diffusion step (no counterpart in the
 * reference, which sums x after the step: integrate.cpp:76-86).  The edge
 * terms are exactly antisymmetric, so the increments applied to the sites
 * that diffuse (flux_sum) equal the H the full_h reservoirs supply (minus the
 * increments they would have received; they keep x) up to the rounding of the
 * per-site sums; flux_gross (the sum of |edge term| x rate x dt over both
 * ends of every edge) is the scale that rounding is relative to.  The clamp
 * to [xLb, xUb] (integrate.cpp:79-81) adds clamp_lift.  Total H changes by
 * flux_sum + clamp_lift (plus the rounding of x + increment). */
typedef struct a2c_diffusion_balance {
  double flux_sum;
  double flux_gross;
  double clamp_lift;
  double reservoir_supply;
} a2c_diffusion_balance;
/* The balance of the context's last a2c_fe_diffuse call (zeros before one). */
int a2c_fe_diffusion_balance(const a2c_fe_ctx* ctx, a2c_diffusion_balance* out);

/* No counterpart in the reference.  For callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc): a2c_fe_diffuse on device-resident arrays that calls no
 * allocator.  Its scratch (two edge-term factors per H site, from
 * f = gamma - ln(x/(1-x)), and the balance partials, ~17 MB) is a library-static device array, one per GPU for the
 * life of the process, so n_h is limited to A2C_DIFFUSE_STATIC_MAX_H
 * (A2C_EINVAL beyond it; a2c_fe_diffuse has no limit).  d_gamma (n_h
 * doubles) and d_full_h (n_h int32) are inputs; d_x (n_h doubles, not the
 * context's own occupancy) receives the new occupancy, which also becomes the
 * context occupancy; d_balance (4 doubles: the fields of
 * a2c_diffusion_balance in order) may be null.  Asynchronous on the context stream; results are bitwise those
 * of a2c_fe_diffuse.  Calls that overlap on one GPU share the scratch and
 * race: serialise them (one stream, or synchronise between calls). */
#define A2C_DIFFUSE_STATIC_MAX_H 1048576
int a2c_fe_diffuse_device_noalloc(a2c_fe_ctx* ctx, double dt, double nu, double Qm, double xLb,
                                  double xUb, const double* d_gamma, const int32_t* d_full_h,
                                  double* d_x, double* d_balance);

/* ---- bounded limited-memory variable metric (replaces TAOBLMVM, -------
 * ---- minimizer.cpp:535, TaoSolve :569, bounds :804-830) ---------------- */
typedef struct a2c_lmvm_opts {
  int32_t m;            /* stored pairs (-tao_lmm_vectors; 5) */
  int32_t max_it;       /* -tao_max_it (2000) */
  int32_t max_funcs;    /* -tao_max_funcs (4000) */
  double gatol;         /* -tao_gatol: ||proj g||_2 <= gatol (1e-8) */
  double grtol;         /* -tao_grtol: ||proj g||/|f| <= grtol (1e-8) */
  double gttol;         /* -tao_gttol: ||proj g||/||proj g0|| <= gttol (0) */
  double sigma_lb;      /* lower bound on sigma (1e-6); q unbounded */
  int32_t mixed_switch; /* MIXED precision: switch to FP64 evals once
                           ||proj g|| <= mixed_gtol (1) */
  double mixed_gtol;    /* 0 (default): automatic, 2e-9 sqrt(N sites) --
                           about twice the measured difference of the
                           mixed and FP64 gradients at relaxed states */
  double sigma_ub;      /* upper bound on sigma (+inf, as formVariableBounds,
                           minimizer.cpp:804-830).  The functional is
                           unbounded below as sigma -> inf (the entropy term
                           -3 kB x ln sigma outgrows the vanishing pair terms
                           beyond a few A), so long relaxations of large
                           lattices may set a finite cap. */
  int32_t precond;      /* 1: initial inverse Hessian diag(gamma_g) diag(1/w),
                           w = 1 for Pd sites and the occupancy x for H sites
                           (every term of an H packet carries x, so its
                           curvature does); 0: w = 1 (1).  gamma_g = s.y /
                           y.diag(1/w) y of the newest pair over coordinate
                           group g -- Pd positions, Pd widths, H positions, H
                           widths (TAO's BLMVM scales by one gamma) */
} a2c_lmvm_opts;

typedef struct a2c_lmvm_report {
  int32_t reason;       /* >0 converged (1 gatol, 2 grtol, 3 gttol); <0 stopped */
  int32_t iterations;
  int32_t evals;        /* total objective+gradient evaluations */
  int32_t evals_fp64;   /* of which in FP64 */
  double f;             /* final objective */
  double gnorm;         /* final projected-gradient 2-norm */
} a2c_lmvm_report;

void a2c_lmvm_opts_default(a2c_lmvm_opts* o);

/* Replaces Minimizer::minimizeFreeEntropy (minimizer.cpp:548-586): minimizes
 * f over X (sigma_lb <= sigma <= sigma_ub), starting from X0; neighbor lists and x0 stay
 * frozen.  X_out may alias X0 (host arrays of 4N doubles). */
int a2c_fe_minimize(a2c_fe_ctx* ctx, const double* X0, const a2c_lmvm_opts* opts,
                    double* X_out, a2c_lmvm_report* report);

/* ---- device memory helpers (so callers need no other HIP runtime) ------- */
int a2c_dev_alloc(a2c_fe_ctx* ctx, size_t bytes, void** ptr);
int a2c_dev_free(a2c_fe_ctx* ctx, void* ptr);
int a2c_dev_upload(a2c_fe_ctx* ctx, void* dst, const void* src, size_t bytes);
int a2c_dev_download(a2c_fe_ctx* ctx, void* dst, const void* src, size_t bytes);
int a2c_fe_sync(a2c_fe_ctx* ctx);

/* ---- measurement ------------------------------------------------------- */
/* Plain device-to-device copy kernel (16-byte aligned), synchronous; used to
 * calibrate HBM byte counters against a transfer of known size. */
int a2c_dev_copy(a2c_fe_ctx* ctx, void* dst, const void* src, size_t bytes);

typedef struct a2c_timing {
  int32_t n_kernels;
  char names[8][32];    /* kernel names in launch order */
  double avg_ms[8];     /* average duration per launch (HIP events, ctx stream) */
  double total_ms;      /* wall of the whole timed region (HIP events) */
} a2c_timing;

/* Runs `iters` device-resident evaluations of dX on the context stream,
 * timing each kernel with HIP events recorded on that stream. */
int a2c_fe_time_eval(a2c_fe_ctx* ctx, const double* dX, double* dG, double* df,
                     int32_t iters, a2c_timing* out);

/* No counterpart in the reference (measurement).  Runs `iters` diffusion
 * steps of a2c_fe_diffuse on device-resident arrays (d_gamma and d_full_h
 * n_h each; d_x, n_h doubles and not the context occupancy, receives the new
 * occupancy), each from the context occupancy, which stays unchanged; times
 * each kernel with HIP events recorded on the context stream. */
int a2c_fe_time_diffuse(a2c_fe_ctx* ctx, double dt, double nu, double Qm, double xLb, double xUb,
                        const double* d_gamma, const int32_t* d_full_h, double* d_x, int32_t iters,
                        a2c_timing* out);

/* Host-only (no GPU): the FP32 EAM tables of the mixed path, 5 tables
 * (f_Pd, f_H, phi_Pd, phi_PdH, phi_H) x 500 intervals x 4 cubic coefficients;
 * interval k covers r in [1 + 0.01k, 1 + 0.01(k+1)) A.  Returns the float count
 * (10000); fills out when cap is large enough. */
int a2c_eam_tables(float* out, int32_t cap);

/* Host-only (no GPU): the FP64 tables of the FP64 pair-owner passes.  Part
 * one: 3 functions (f_H, phi_PdH, phi_H) x 100 intervals x 10 doubles
 * (polynomial coefficients c0..c8 and one pad); part two, after it: 2
 * functions (f_Pd, phi_Pd) x 100 intervals x 12 doubles (c0..c11).  Interval
 * k covers r in [1 + 0.05k, 1 + 0.05(k+1)) A, value sum_j c_j u^j with
 * u = (r - 1)/0.05 - k.  Returns the double count (5400); fills out when cap
 * is large enough.  No counterpart in the reference (it evaluates the closed
 * forms of potential.cpp:64-200). */
int a2c_eam_tables64(double* out, int32_t cap);

#ifdef __cplusplus
}
#endif
#endif
