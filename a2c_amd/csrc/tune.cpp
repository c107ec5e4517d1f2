// Launch-shape table of the pair-owner passes (lanes per site by lattice
// size and precision) and the autotuner that measures it.
//
// The built-in table is the outcome of a2c_autotune on one MI355X with the
// default options (evaluations of 6^3..40^3-cell lattices, in fused launches
// filling about one million sites and alone).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <climits>
#include <cmath>
#include <cstring>
#include <vector>

#include "a2c_batch.h"
#include "a2c_lattice.h"
#include "a2c_tune.h"
#include "fe_internal.hpp"

namespace a2c {
namespace {

a2c_tune_table default_table() {
  // a2c_autotune with the default options on one MI355X.  Up to 11^3 cells a
  // lattice alone fills fewer SIMDs than the device has and takes the K that
  // finishes it soonest within 25% of the best throughput: 6^3 cells, K =
  // 1/2/4/8/16, alone mixed 294/143/98/80/67 us and FP64 687/270/165/124/96
  // us per evaluation, in launches of 256 copies mixed 1.99/1.75/1.67/1.85/
  // 2.30 and FP64 5.39/4.72/4.70/5.11/6.04 ns per site (K = 16 costs 37%
  // there).  Larger lattices take the fastest K alone within 5% of the best
  // throughput (24^3 cells FP64: 4.13/4.18 ns per site at K = 2/4 in a full
  // launch, 602/528 us alone; 40^3 cells mixed 1.74/1.58/1.60/1.79/2.33 ns
  // per site, 871/807/818/915/1193 us alone).
  static const int32_t max_sites[] = {2660, 6604, 15288, 32003, 71831, 170267, 366357, INT_MAX};
  static const int32_t lanes_mixed[] = {8, 8, 8, 4, 4, 4, 2, 2};
  static const int32_t lanes_fp64[] = {8, 8, 8, 4, 4, 4, 4, 2};
  a2c_tune_table t;
  std::memset(&t, 0, sizeof t);
  t.n_classes = 8;
  for (int c = 0; c < t.n_classes; ++c) {
    t.max_sites[c] = max_sites[c];
    t.lanes_mixed[c] = lanes_mixed[c];
    t.lanes_fp64[c] = lanes_fp64[c];
  }
  return t;
}

a2c_tune_table g_table = default_table();

constexpr int NLANES = 5;
// Largest per-site cost in a full launch that a choice for the lone-lattice
// time may add: for a lattice that alone fills fewer SIMDs than the device
// has, and for one that fills them (where near-equal throughputs differ
// alone: at 24^3 cells FP64, K = 2 beat K = 4 by 1% per site in a full
// launch and lost by 14% alone, in the tail of a lockstep run).
constexpr double LATENCY_COST = 1.25, THROUGHPUT_SLACK = 1.05;
const int LANES[NLANES] = {1, 2, 4, 8, 16};
bool valid_lanes(int k) { return k == 1 || k == 2 || k == 4 || k == 8 || k == 16; }
static_assert(MAX_LANES_PER_SITE == 16, "the table allows up to 16 lanes per site");

bool valid(const a2c_tune_table& t) {
  if (t.n_classes < 1 || t.n_classes > A2C_TUNE_MAX_CLASSES) return false;
  for (int c = 0; c < t.n_classes; ++c) {
    if (!valid_lanes(t.lanes_mixed[c]) || !valid_lanes(t.lanes_fp64[c])) return false;
    if (t.max_sites[c] < 1 || (c > 0 && t.max_sites[c] <= t.max_sites[c - 1])) return false;
  }
  return true;
}

// a table with one class: K lanes for every lattice in both precisions
a2c_tune_table uniform(int k) {
  a2c_tune_table t;
  std::memset(&t, 0, sizeof t);
  t.n_classes = 1;
  t.max_sites[0] = INT_MAX;
  t.lanes_mixed[0] = t.lanes_fp64[0] = k;
  return t;
}

}  // namespace

int lanes_per_site(int n_sites, bool fp64) {
  const a2c_tune_table& t = g_table;
  int c = 0;
  while (c + 1 < t.n_classes && n_sites > t.max_sites[c]) ++c;
  return fp64 ? t.lanes_fp64[c] : t.lanes_mixed[c];
}

}  // namespace a2c

extern "C" {

void a2c_tune_get(a2c_tune_table* out) {
  if (out) *out = a2c::g_table;
}

int a2c_tune_set(const a2c_tune_table* t) {
  if (!t) {
    a2c::g_table = a2c::default_table();
    return A2C_OK;
  }
  if (!a2c::valid(*t)) return A2C_EINVAL;
  a2c::g_table = *t;
  return A2C_OK;
}

void a2c_autotune_opts_default(a2c_autotune_opts* o) {
  std::memset(o, 0, sizeof(*o));
  const int cells[] = {6, 8, 11, 14, 18, 24, 32, 40};
  o->n_sizes = 8;
  for (int i = 0; i < 8; ++i) o->cells[i] = cells[i];
  o->target_sites = 1000000;
  o->reps = 10;
  o->apply = 0;
}

int a2c_autotune(const a2c_autotune_opts* o, a2c_tune_table* out, a2c_autotune_report* report) {
  if (!o || !out || o->n_sizes < 1 || o->n_sizes > A2C_TUNE_MAX_CLASSES || o->reps < 1 || o->target_sites < 1)
    return A2C_EINVAL;
  for (int i = 0; i < o->n_sizes; ++i)
    if (o->cells[i] < 3 || o->cells[i] > 128 || (i > 0 && o->cells[i] <= o->cells[i - 1])) return A2C_EINVAL;
  a2c_autotune_report rep;
  std::memset(&rep, 0, sizeof rep);
  a2c_tune_table res;
  std::memset(&res, 0, sizeof res);
  res.n_classes = o->n_sizes;
  const a2c_tune_table saved = a2c::g_table;
  int ncu = 0;
  if (hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, o->device) != hipSuccess || ncu < 1)
    return A2C_EHIP;
  const int64_t simds = 4 * (int64_t)ncu;
  int rc = A2C_OK;
  for (int i = 0; i < o->n_sizes && rc == A2C_OK; ++i) {
    const int n = o->cells[i];
    const int npd = 4 * n * n * n, N = 2 * npd;
    const int copies = (int)std::max<int64_t>(1, std::min<int64_t>(256, o->target_sites / N));
    rep.sites[i] = N;
    rep.copies[i] = copies;
    a2c_batch_desc bd;
    a2c_batch_desc_default(&bd);
    bd.device = o->device;
    bd.precision = A2C_PREC_MIXED;
    bd.max_lattices = copies;
    bd.graphs = 0;
    a2c_batch* b = nullptr;
    if ((rc = a2c_batch_create(&bd, &b))) break;
    std::vector<int32_t> slots(copies);
    std::vector<double> X(8 * (size_t)N), x(npd), box(3), f(copies);
    for (int k = 0; k < copies && rc == A2C_OK; ++k) {
      // the batched workload's recipe: displacement 0.02 A, sigma 0.10 A, mean H fraction in [0.01, 0.6]
      a2c_lattice_desc L;
      a2c_lattice_desc_default(&L);
      L.nx = L.ny = L.nz = n;
      L.seed = 7000 + 97 * (uint64_t)i + k;
      L.disp_std = 0.02;
      L.sigma_lo = L.sigma_hi = 0.10;
      L.x_lo = 0.01 + 0.59 * (k % 8) / 7.0;
      L.x_hi = L.x_lo;
      if ((rc = a2c_lattice_fcc_pdh(&L, X.data(), x.data(), box.data()))) break;
      a2c_fe_desc d;
      a2c_fe_desc_default(&d);
      d.n_pd = d.n_h = npd;
      for (int c = 0; c < 3; ++c) {
        d.box[c] = box[c];
        d.periodic[c] = 1;
      }
      d.temperature = 500.0;
      rc = a2c_batch_add(b, &d, X.data(), x.data(), &slots[k]);
    }
    for (int p = 0; p < 2 && rc == A2C_OK; ++p) {
      const int prec = p == 0 ? A2C_PREC_MIXED : A2C_PREC_FP64;
      double best = HUGE_VAL;
      int bestk = 1;
      for (int li = 0; li < a2c::NLANES && rc == A2C_OK; ++li) {
        a2c::g_table = a2c::uniform(a2c::LANES[li]);
        for (int alone = 0; alone < 2 && rc == A2C_OK; ++alone) {
          const int n = alone ? 1 : copies, reps = alone ? 4 * o->reps : o->reps;
          for (int w = 0; w < 2 && rc == A2C_OK; ++w) rc = a2c_batch_eval(b, n, slots.data(), prec, f.data());
          const auto t0 = std::chrono::steady_clock::now();
          for (int r = 0; r < reps && rc == A2C_OK; ++r) rc = a2c_batch_eval(b, n, slots.data(), prec, f.data());
          const double s = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count() / reps;
          if (alone) {
            rep.us_alone[i][p][li] = s * 1e6;
          } else {
            const double ns = s * 1e9 / ((double)copies * N);
            rep.ns_per_site[i][p][li] = ns;
            if (ns < best) {
              best = ns;
              bestk = a2c::LANES[li];
            }
          }
        }
      }
      // The K that finishes the lattice soonest alone, among those within a
      // slack of the best throughput: LATENCY_COST for a lattice that alone
      // under the best-throughput K gives fewer waves (64 lanes) than the
      // device has SIMDs, THROUGHPUT_SLACK otherwise.
      const double slack = (int64_t)N * bestk < 64 * simds ? a2c::LATENCY_COST : a2c::THROUGHPUT_SLACK;
      int k = bestk;
      double t = HUGE_VAL;
      for (int li = 0; li < a2c::NLANES; ++li)
        if (rep.ns_per_site[i][p][li] <= slack * best && rep.us_alone[i][p][li] < t) {
          t = rep.us_alone[i][p][li];
          k = a2c::LANES[li];
        }
      rep.latency[i][p] = k != bestk;
      (p == 0 ? res.lanes_mixed : res.lanes_fp64)[i] = k;
    }
    a2c_batch_destroy(b);
  }
  a2c::g_table = saved;
  if (rc) return rc;
  for (int i = 0; i + 1 < o->n_sizes; ++i)  // boundary at the geometric mean of neighboring sizes
    res.max_sites[i] = (int32_t)std::sqrt((double)rep.sites[i] * rep.sites[i + 1]);
  res.max_sites[o->n_sizes - 1] = INT_MAX;
  *out = res;
  if (report) *report = rep;
  if (o->apply) a2c::g_table = res;
  return A2C_OK;
}

}  // extern "C"
