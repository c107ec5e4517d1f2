/*
 * a2c_batch.h -- many lattices on one GPU: fused multi-lattice evaluation and
 * a lockstep bounded LMVM that can admit and retire lattices between
 * iterations (continuous batching).  No counterpart in the reference, whose
 * Minimizer holds exactly one lattice per MPI job (minimizer.cpp:484-546); each
 * slot here behaves as one reference Minimizer with its own TaoSolve.
 *
 * Determinism: a slot's objective, gradient and optimizer trajectory are
 * bitwise the same whatever other slots share the launch.
 */
#ifndef A2C_BATCH_H
#define A2C_BATCH_H
#include <stddef.h>
#include <stdint.h>

#include "a2c_fe.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct a2c_batch a2c_batch;

typedef struct a2c_batch_desc {
  int32_t device;
  int32_t precision;        /* A2C_PREC_MIXED (default) or A2C_PREC_FP64 */
  int32_t deriv_mode;       /* A2C_DERIV_ANALYTIC or A2C_DERIV_FD */
  double fd_eps;
  int32_t max_lattices;     /* slot capacity */
  size_t hbm_budget_bytes;  /* admission limit (sum of a2c_batch_footprint) */
  int32_t max_neighbors;
  int32_t graphs;           /* 1: a launch sequence that recurs between host synchronisations is
                               captured into a hipGraph and replayed; 0 (default): every launch is
                               issued directly.  Results are bitwise identical either way.  The
                               environment variable A2C_GRAPHS=1 makes 1 the default. */
} a2c_batch_desc;

void a2c_batch_desc_default(a2c_batch_desc* d);
int a2c_batch_create(const a2c_batch_desc* desc, a2c_batch** out);
void a2c_batch_destroy(a2c_batch* b);
const char* a2c_batch_last_error(const a2c_batch* b);

/* Device bytes one lattice occupies (state, lists, work arrays and an LMVM
 * with m pairs, with the device cache's size-class rounding); admission
 * charges a lattice with m = 5 against hbm_budget_bytes and
 * a2c_batch_lmvm_start re-charges it for the m it asks for (A2C_ENOMEM when
 * that exceeds the budget); a2c_batch_remove releases the charge. */
size_t a2c_batch_footprint(int32_t n_pd, int32_t n_h, int32_t lmvm_m);
size_t a2c_batch_mem_used(const a2c_batch* b);

/* Launch sequences issued so far: replayed from a captured hipGraph,
 * captured (and run), issued directly.  Any pointer may be null. */
int a2c_batch_graph_stats(const a2c_batch* b, int64_t* replays, int64_t* captures, int64_t* direct);

/* Admit one lattice: lat describes it like a single context (its precision,
 * derivative mode and device are the batch's); X (4N) and x (n_h) are its
 * state and occupancy; neighbor lists are built from X.  A2C_ENOMEM when no
 * slot is free or the HBM budget would be exceeded. */
int a2c_batch_add(a2c_batch* b, const a2c_fe_desc* lat, const double* X, const double* x, int32_t* slot);
int a2c_batch_remove(a2c_batch* b, int32_t slot);
int a2c_batch_set_state(a2c_batch* b, int32_t slot, const double* X);
/* Current state (the optimizer's iterate once started) and last gradient. */
int a2c_batch_get_state(a2c_batch* b, int32_t slot, double* X, double* G);

/* One fused launch evaluating the listed slots at their states; f per slot. */
int a2c_batch_eval(a2c_batch* b, int32_t n, const int32_t* slots, int32_t precision, double* f);

/* Lockstep minimization: start a slot (from its current state), then advance
 * every running slot by up to max_iters iterations per call. */
int a2c_batch_lmvm_start(a2c_batch* b, int32_t slot, const a2c_lmvm_opts* opts);
int a2c_batch_lmvm_step(a2c_batch* b, int32_t max_iters, int32_t* n_running);
/* The same for the listed slots only (those not running are skipped); a
 * scheduler can step small lattices more often than large ones.  A slot's
 * trajectory does not depend on which other slots share its steps.
 * n_running counts every running slot of the batch. */
int a2c_batch_lmvm_step_slots(a2c_batch* b, int32_t n, const int32_t* slots, int32_t max_iters,
                              int32_t* n_running);
/* reason 0 while the slot is still running */
int a2c_batch_lmvm_report(a2c_batch* b, int32_t slot, a2c_lmvm_report* rep);

#ifdef __cplusplus
}
#endif
#endif
