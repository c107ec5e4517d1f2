/*
 * a2c_tune.h -- launch-shape table of the pair-owner passes and the
 * autotuner that fills it.  No counterpart in the reference (its CPU loops
 * have no launch shapes).
 *
 * The pair-owner kernels give each site K lanes (K in {1, 2, 4, 8, 16}); lane p
 * of a site takes its owned neighbor slots p, p + K, p + 2K, ... and the K
 * partial sums are combined by a fixed shuffle tree.  K is chosen per
 * lattice from its site count and the evaluation precision through a
 * process-wide table of size classes.  Results depend on K only through the
 * order of a site's sums: a lattice gives the same bits alone and in any
 * batch under one table, and within the stated tolerances under any table.
 *
 * The table is process state: set it before evaluations start; changing it
 * while another thread evaluates is not supported.
 */
#ifndef A2C_TUNE_H
#define A2C_TUNE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define A2C_TUNE_MAX_CLASSES 8

typedef struct a2c_tune_table {
  int32_t n_classes;                         /* 1..A2C_TUNE_MAX_CLASSES */
  int32_t max_sites[A2C_TUNE_MAX_CLASSES];   /* class c: sites <= max_sites[c] (increasing; the last
                                                class also takes every larger lattice) */
  int32_t lanes_mixed[A2C_TUNE_MAX_CLASSES]; /* K of the mixed-precision passes, 1, 2, 4, 8 or 16 */
  int32_t lanes_fp64[A2C_TUNE_MAX_CLASSES];  /* K of the FP64 passes */
} a2c_tune_table;

/* The table in effect; the built-in default is the result of a2c_autotune
 * on MI355X. */
void a2c_tune_get(a2c_tune_table* out);
/* Replace the table (A2C_EINVAL if malformed); null restores the default. */
int a2c_tune_set(const a2c_tune_table* t);

typedef struct a2c_autotune_opts {
  int32_t device;
  int32_t n_sizes;                       /* representative lattices, 1..A2C_TUNE_MAX_CLASSES */
  int32_t cells[A2C_TUNE_MAX_CLASSES];   /* cubic lattices of cells^3 FCC cells (8 cells^3 sites) */
  int64_t target_sites;                  /* copies of each lattice so one fused launch holds about
                                            this many sites (at most 256 copies) */
  int32_t reps;                          /* timed evaluations per candidate */
  int32_t apply;                         /* 1: install the tuned table on success */
} a2c_autotune_opts;

typedef struct a2c_autotune_report {
  int32_t sites[A2C_TUNE_MAX_CLASSES];          /* sites of each representative lattice */
  int32_t copies[A2C_TUNE_MAX_CLASSES];         /* lattices per fused launch */
  double ns_per_site[A2C_TUNE_MAX_CLASSES][2][5]; /* [size][mixed, fp64][K = 1, 2, 4, 8, 16]: wall time of
                                                     one fused evaluation per site */
  double us_alone[A2C_TUNE_MAX_CLASSES][2][5];    /* the same, one lattice per launch: microseconds
                                                     per evaluation */
  int32_t latency[A2C_TUNE_MAX_CLASSES][2];       /* 1: the lone-lattice time chose K */
} a2c_autotune_report;

void a2c_autotune_opts_default(a2c_autotune_opts* o);
/* Time evaluations (objective and gradient, synthetic Pd-H lattices of the
 * batched workload's recipe) of every representative size under every K and
 * both precisions, in fused launches of many copies (throughput) and of one
 * lattice alone (latency).  Per size and precision the K that makes one
 * lattice finish soonest alone wins among those whose throughput is within
 * 5% of the best -- within 25% if the lattice alone, under the best-
 * throughput K, gives the device fewer waves than it has SIMDs (it then
 * leaves SIMDs idle; more lanes cost per site when many lattices share a
 * launch).  Class boundaries lie at the geometric means of neighboring
 * sizes.  out receives the table; report (may be null) the measurements. */
int a2c_autotune(const a2c_autotune_opts* o, a2c_tune_table* out, a2c_autotune_report* report);

#ifdef __cplusplus
}
#endif
#endif
