// Batched engine: many lattices of uneven size evaluated by one fused launch,
// and a lockstep bounded LMVM over all of them.
//
// Layout.  Every lattice (slot) keeps its own device buffers (state,
// gradient, neighbor lists, per-site work arrays) in an a2c_fe_ctx.  A launch
// over a set of slots uploads one LatDev per slot and a block -> slot map; the
// site blocks of a slot always start at its site 0 and its reductions
// (objective, dot products) have a fixed shape that depends only on its own
// size, so every slot's result is bitwise the same alone or in any mix.
//
// Optimizer.  The single-lattice minimizer (a2c_fe_minimize) runs on this
// code with a batch of one, so it is the same algorithm: L-BFGS two-loop
// recursion (m pairs; initial inverse Hessian diag(gamma_g) D, D the
// occupancy metric and gamma_g = s.y / y.D y of the newest pair over each
// coordinate group), active-set handling of the sigma bounds, Moré-Thuente
// line search with approximate Wolfe acceptance in FP64, steepest-descent
// restart on failure, TAO's gatol/grtol/gttol tests, and the mixed -> FP64
// switch.  In a batch every step is one launch over all slots that need it
// (segmented kernels; the recursion of a small slot in one workgroup), and
// each scalar decision is taken per slot on the host after one
// synchronisation per round.  Slots can be added and started between
// iterations (continuous batching).
//
// Launch sequences.  Every operation between two host synchronisations is
// queued with a key made of its kernel and all its argument values; the
// descriptors those kernels read are staged in one pinned arena, handed out
// in the same order every time, so a steady-state optimizer iteration repeats
// its keys exactly.  With a2c_batch_desc.graphs set, a sequence seen for the
// second time is captured once into a hipGraph and from then on replayed: the
// descriptors are copied, then the whole sequence (kernels and result copies)
// is one graph launch.  Graph and direct issue run the same kernels in the
// same order with the same arguments, so results are bitwise identical either
// way.  Replay is off by default: on this runtime a replayed iteration of a
// small lattice takes as long as direct issue (the per-kernel cost is on the
// device side), and capturing new sequences costs time in a changing batch.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <functional>
#include <new>
#include <string>
#include <unordered_map>
#include <vector>

#include "a2c_batch.h"
#include "fe_internal.hpp"
#include "linesearch.hpp"

namespace a2c {
namespace {

constexpr int TB = SITE_BLOCK;

// ---------------------------------------------------------------- kernels
// Diagonal metric of the optimizer: every (q, sigma) component of a site is
// scaled by 1/w, w = 1 for Pd sites and the occupancy x (floored at 1e-6) for
// H sites.  Every term of the objective that involves an H packet carries
// its occupancy, so its block of the Hessian scales with x; dividing by x
// equalises the curvature of nearly empty and full sites.
struct Metric {
  const double* occ = nullptr;  // null: identity
  int32_t npd = 0;
};
__device__ __forceinline__ double metric_inv(const Metric& m, int s) {
  return (m.occ && s >= m.npd) ? 1.0 / fmax(m.occ[s - m.npd], 1e-6) : 1.0;
}
// z = A D^k x + B y on 4N doubles (k = 0, or 1 when scaled).  The
// coefficients may come from device scalars, so optimizer steps that depend
// on dot products need no host round trip:
//   amode 0: A = a;  1: A = a * p1;  2: A = a * p1 / p2;
//         3: A = -min(a / p1, 0.05 / sqrt(p2))   (scaled steepest descent)
//   bmode 0: B = b;  1: B = b * q1;  2: B = b * (q1 - q2)
struct SegAx {
  double* z;
  const double* x;
  const double* y;
  double a, b;
  const double *p1, *p2, *q1, *q2;
  int32_t amode, bmode;
  Metric met;
  int32_t scaled;
  int32_t nsite, blk0;
};
struct SegDot {  // out[0] = a1 . D^k1 b1, out[1] = a2 . D^k2 b2 (a2 may be null)
  const double* a1;
  const double* b1;
  const double* a2;
  const double* b2;
  Metric met;
  int32_t k1, k2;
  double* out;
  int32_t nsite, blk0;
};

// Coordinate groups of the scaled initial inverse Hessian: positions and
// widths of Pd sites, positions and widths of H sites.  Their curvatures
// differ by orders of magnitude, so the two-loop recursion scales each group
// by its own gamma_g = s_g . y_g / y_g . D y_g of the newest pair (a group
// without positive curvature data takes the total's gamma): measured on 5^3
// and 6^3 lattices, half the iterations of one scalar gamma.
constexpr int NGROUP = 4;
__host__ __device__ inline int group_of(int site, int comp, int npd) {
  return (site >= npd ? 2 : 0) + (comp == 3 ? 1 : 0);
}
__host__ __device__ inline double group_gamma(const double* gs, double sy, double yy, int g) {
  const double a = gs[2 * g], b = gs[2 * g + 1];
  return (a > 0.0 && b > 0.0) ? a / b : sy / yy;
}
// Device scalars per slot: [0, m) the first-loop dots s_i.q of the two-loop
// recursion, m + 2 the second-loop dot, m + 3 and m + 4 g.D g and g.D^2 g of
// scaled steepest descent, [m + 5, m + 5 + 2 NGROUP) s.y and y.D y of the
// newest pair per coordinate group (m and m + 1 are unused).
constexpr int SCAL_PER_SLOT = 80;  // >= m + 5 + 2 NGROUP for m <= 64
constexpr double SIGMA_STEP = 4.0;  // largest factor by which one line-search trial may change a width
constexpr double Q_STEP = 0.25;     // largest displacement of a site in one line-search trial (A)
constexpr int MIXED_MAX_TRIALS = 8;  // line-search trials of a mixed-precision iteration before FP64 takes over

struct SegBound {  // sigma-bound operations: mode 0 project x, 1 gp = P g, 2 zero the active part of d,
                   // 3 form a trial point x_old + stp d, projected, and the step direction it realises
  double* z;        // x (0), gp (1), d (2), trial x (3)
  const double* x;  // x (1, 2), x_old (3)
  const double* y;  // g (1), d (3)
  double* w;        // realised direction (3)
  double lb, ub, stp;
  int32_t nsite, blk0;
};

__global__ __launch_bounds__(TB) void k_seg_axpby(const SegAx* __restrict__ ops, const int32_t* __restrict__ b2l) {
  const SegAx o = ops[b2l[blockIdx.x]];
  const int s = (blockIdx.x - o.blk0) * TB + threadIdx.x;
  if (s >= o.nsite) return;
  const size_t i = 4 * (size_t)s;
  double a = o.a, bb = o.b;
  if (o.amode == 1) a *= *o.p1;
  else if (o.amode == 2) a *= *o.p1 / *o.p2;
  else if (o.amode == 3) a = -fmin(o.a / fmax(*o.p1, 1e-300), 0.05 / sqrt(fmax(*o.p2, 1e-300)));
  if (o.bmode == 1) bb *= *o.q1;
  else if (o.bmode == 2) bb *= *o.q1 - *o.q2;
  if (o.scaled) a *= metric_inv(o.met, s);
#pragma unroll
  for (int c = 0; c < 4; ++c) o.z[i + c] = a * o.x[i + c] + bb * o.y[i + c];
}

__global__ __launch_bounds__(TB) void k_seg_dot(const SegDot* __restrict__ ops, const int32_t* __restrict__ b2l,
                                                double* __restrict__ part, int nblk) {
  __shared__ double r1[TB], r2[TB];
  const SegDot o = ops[b2l[blockIdx.x]];
  const int s = (blockIdx.x - o.blk0) * TB + threadIdx.x;
  double v1 = 0.0, v2 = 0.0;
  if (s < o.nsite) {
    const size_t i = 4 * (size_t)s;
    const double di = metric_inv(o.met, s);
#pragma unroll
    for (int c = 0; c < 4; ++c) v1 += o.a1[i + c] * o.b1[i + c];
    if (o.k1) v1 *= o.k1 == 1 ? di : di * di;
    if (o.a2) {
#pragma unroll
      for (int c = 0; c < 4; ++c) v2 += o.a2[i + c] * o.b2[i + c];
      if (o.k2) v2 *= o.k2 == 1 ? di : di * di;
    }
  }
  r1[threadIdx.x] = v1;
  r2[threadIdx.x] = v2;
  __syncthreads();
  for (int w = TB / 2; w > 0; w >>= 1) {
    if (threadIdx.x < w) {
      r1[threadIdx.x] += r1[threadIdx.x + w];
      r2[threadIdx.x] += r2[threadIdx.x + w];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    part[blockIdx.x] = r1[0];
    part[nblk + blockIdx.x] = r2[0];
  }
}

// one block per op: fixed-order sum of the op's block partials
__global__ __launch_bounds__(1024) void k_seg_dot_final(const SegDot* __restrict__ ops, const double* __restrict__ part,
                                                        int nblk) {
  __shared__ double r1[1024], r2[1024];
  const SegDot o = ops[blockIdx.x];
  const int n = site_blocks(o.nsite);
  double v1 = 0.0, v2 = 0.0;
  for (int i = threadIdx.x; i < n; i += 1024) {
    v1 += part[o.blk0 + i];
    v2 += part[nblk + o.blk0 + i];
  }
  r1[threadIdx.x] = v1;
  r2[threadIdx.x] = v2;
  __syncthreads();
  for (int w = 512; w > 0; w >>= 1) {
    if (threadIdx.x < w) {
      r1[threadIdx.x] += r1[threadIdx.x + w];
      r2[threadIdx.x] += r2[threadIdx.x + w];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    o.out[0] = r1[0];
    if (o.a2) o.out[1] = r2[0];  // single dots write one scalar (device scalars are packed)
  }
}

// block sum in a fixed order: wave trees, then the NT / 64 wave partials in order
template <int NT>
__device__ __forceinline__ double bsum(double v, double* red) {
  for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off);
  __syncthreads();  // the previous sum's readers are done with red
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
  __syncthreads();
  double s = 0.0;
#pragma unroll
  for (int w = 0; w < NT / 64; ++w) s += red[w];
  return s;
}

// K block sums at once (one pair of barriers), each in the fixed order of
// bsum; red holds K * NT / 64 doubles
template <int NT, int K>
__device__ __forceinline__ void bsum_n(double (&v)[K], double* red) {
  constexpr int W = NT / 64;
#pragma unroll
  for (int k = 0; k < K; ++k)
    for (int off = 32; off > 0; off >>= 1) v[k] += __shfl_down(v[k], off);
  __syncthreads();  // the previous sum's readers are done with red
  if ((threadIdx.x & 63) == 0)
#pragma unroll
    for (int k = 0; k < K; ++k) red[k * W + (threadIdx.x >> 6)] = v[k];
  __syncthreads();
#pragma unroll
  for (int k = 0; k < K; ++k) {
    double s = 0.0;
#pragma unroll
    for (int w = 0; w < W; ++w) s += red[k * W + w];
    v[k] = s;
  }
}

// Two-loop recursion of slots too large for one workgroup: one grid pass per
// recursion step for all such slots at once.  A pass applies the update of
// its step and forms block partials of the dot product the next step needs
// (k_seg_final reduces them in a fixed order), so the recursion over np
// pairs is 2 np + 1 passes, each reading q and two vectors, instead of a
// dot and an axpy per step.
struct SegTL {
  double* q;
  const double* v;      // pass 0: g (q = g); else the update vector (y_t, then s_t)
  const double* w;      // dot partner of the updated q (pass 0: s_0)
  const double* y0;     // pass 0: y_0 (group sums of the initial scale)
  const double* ca;     // update q += c v, c = rho (*ca - (cb ? *cb : 0))
  const double* cb;
  double rho;
  int32_t mode;         // 0 update; 1 update, then q = diag(gamma_g) D q; 2 final: d = -(updated q),
                        // widths at an active bound not moved, dot = g . d
  const double* gscal;  // mode 1: {s.y, y.D y} of the newest pair per coordinate group
  Metric met;
  int32_t npd;
  double* d;            // mode 2
  const double* g;
  const double* x;
  double lb, ub;
  double* out;          // the reduced dot
  double* out2;         // pass 0: the 2 NGROUP group sums
  int32_t nsite, blk0;
};
constexpr int TL_PART = 1 + 2 * NGROUP;  // partial sums per block of pass 0

__global__ __launch_bounds__(TB) void k_seg_tl0(const SegTL* __restrict__ ops, const int32_t* __restrict__ b2l,
                                                double* __restrict__ part, int nblk) {
  __shared__ double red[TB / 64];
  const SegTL o = ops[b2l[blockIdx.x]];
  const int s = (blockIdx.x - o.blk0) * TB + threadIdx.x;
  double a = 0.0, gs[2 * NGROUP] = {};
  if (s < o.nsite) {
    const size_t i = 4 * (size_t)s;
    const double di = metric_inv(o.met, s);
#pragma unroll
    for (int c = 0; c < 4; ++c) {
      const double gi = o.v[i + c], si = o.w[i + c], yi = o.y0[i + c];
      const int gr = group_of(s, c, o.npd);
      o.q[i + c] = gi;
      a += si * gi;
#pragma unroll
      for (int k = 0; k < NGROUP; ++k) {
        gs[2 * k] += k == gr ? si * yi : 0.0;
        gs[2 * k + 1] += k == gr ? yi * yi * di : 0.0;
      }
    }
  }
  a = bsum<TB>(a, red);
  if (threadIdx.x == 0) part[blockIdx.x] = a;
#pragma unroll
  for (int k = 0; k < 2 * NGROUP; ++k) {
    const double v = bsum<TB>(gs[k], red);
    if (threadIdx.x == 0) part[(size_t)(1 + k) * nblk + blockIdx.x] = v;
  }
}

__global__ __launch_bounds__(TB) void k_seg_tl(const SegTL* __restrict__ ops, const int32_t* __restrict__ b2l,
                                               double* __restrict__ part) {
  __shared__ double red[TB / 64];
  const SegTL o = ops[b2l[blockIdx.x]];
  const int s = (blockIdx.x - o.blk0) * TB + threadIdx.x;
  double acc = 0.0;
  if (s < o.nsite) {
    const size_t i = 4 * (size_t)s;
    const double c = o.rho * (*o.ca - (o.cb ? *o.cb : 0.0));
    if (o.mode == 2) {
#pragma unroll
      for (int k = 0; k < 4; ++k) {
        double di = -(o.q[i + k] + c * o.v[i + k]);
        if (k == 3) {
          const double xs = o.x[i + 3];
          if ((xs <= o.lb && di < 0.0) || (xs >= o.ub && di > 0.0)) di = 0.0;
        }
        o.d[i + k] = di;
        acc += o.g[i + k] * di;
      }
    } else {
      double sy = 0.0, yy = 0.0;
      if (o.mode == 1) {
#pragma unroll
        for (int k = 0; k < NGROUP; ++k) {
          sy += o.gscal[2 * k];
          yy += o.gscal[2 * k + 1];
        }
      }
      const double di = metric_inv(o.met, s);
#pragma unroll
      for (int k = 0; k < 4; ++k) {
        double qi = o.q[i + k] + c * o.v[i + k];
        if (o.mode == 1) qi = group_gamma(o.gscal, sy, yy, group_of(s, k, o.npd)) * di * qi;
        o.q[i + k] = qi;
        acc += o.w[i + k] * qi;
      }
    }
  }
  acc = bsum<TB>(acc, red);
  if (threadIdx.x == 0) part[blockIdx.x] = acc;
}

// The new curvature pair of an accepted step and the projected gradient in
// one pass: s = x - x_old, y = g - g_old, gp = P g; dots s.y and gp.gp.
struct SegPair {
  const double* x;
  const double* xo;
  const double* g;
  const double* go;
  double* S;
  double* Y;
  double* gp;
  double lb, ub;
  double* out;   // s.y
  double* out2;  // gp.gp
  int32_t nsite, blk0;
};

__global__ __launch_bounds__(TB) void k_seg_pair(const SegPair* __restrict__ ops, const int32_t* __restrict__ b2l,
                                                 double* __restrict__ part, int nblk) {
  __shared__ double red[TB / 64];
  const SegPair o = ops[b2l[blockIdx.x]];
  const int s = (blockIdx.x - o.blk0) * TB + threadIdx.x;
  double sy = 0.0, gg = 0.0;
  if (s < o.nsite) {
    const size_t i = 4 * (size_t)s;
#pragma unroll
    for (int c = 0; c < 4; ++c) {
      const double gi = o.g[i + c];
      const double si = o.x[i + c] - o.xo[i + c], yi = gi - o.go[i + c];
      o.S[i + c] = si;
      o.Y[i + c] = yi;
      sy += si * yi;
      double pi = gi;
      if (c == 3) {
        const double xs = o.x[i + 3];
        if ((xs <= o.lb && gi > 0.0) || (xs >= o.ub && gi < 0.0)) pi = 0.0;
      }
      o.gp[i + c] = pi;
      gg += pi * pi;
    }
  }
  sy = bsum<TB>(sy, red);
  gg = bsum<TB>(gg, red);
  if (threadIdx.x == 0) {
    part[blockIdx.x] = sy;
    part[nblk + blockIdx.x] = gg;
  }
}

// one block per op: fixed-order sums of the op's NOUT block partials
// (the first to out, the others to out2)
template <class Op, int NOUT>
__global__ __launch_bounds__(256) void k_seg_final(const Op* __restrict__ ops, const double* __restrict__ part,
                                                   int nblk) {
  __shared__ double red[4];
  const Op o = ops[blockIdx.x];
  const int n = site_blocks(o.nsite);
#pragma unroll
  for (int k = 0; k < NOUT; ++k) {
    double v = 0.0;
    for (int j = threadIdx.x; j < n; j += 256) v += part[(size_t)k * nblk + o.blk0 + j];
    v = bsum<256>(v, red);
    if (threadIdx.x == 0) (k == 0 ? o.out[0] : o.out2[k - 1]) = v;
  }
}

template <int MODE>
__global__ __launch_bounds__(TB) void k_seg_bound(const SegBound* __restrict__ ops, const int32_t* __restrict__ b2l) {
  const SegBound o = ops[b2l[blockIdx.x]];
  const int s = (blockIdx.x - o.blk0) * TB + threadIdx.x;
  if (s >= o.nsite) return;
  const size_t i = 4 * (size_t)s;
  if (MODE == 0) {
    o.z[i + 3] = fmin(fmax(o.z[i + 3], o.lb), o.ub);
  } else if (MODE == 1) {
    o.z[i] = o.y[i];
    o.z[i + 1] = o.y[i + 1];
    o.z[i + 2] = o.y[i + 2];
    const double gs = o.y[i + 3], xs = o.x[i + 3];
    o.z[i + 3] = ((xs <= o.lb && gs > 0.0) || (xs >= o.ub && gs < 0.0)) ? 0.0 : gs;
  } else if (MODE == 2) {
    const double xs = o.x[i + 3], ds = o.z[i + 3];
    if ((xs <= o.lb && ds < 0.0) || (xs >= o.ub && ds > 0.0)) o.z[i + 3] = 0.0;
  } else {
    // Trial safeguards (they bend the search path, not the fixed point): a
    // site's position moves at most Q_STEP, so that a packet whose metric
    // weight is large (nearly empty H) cannot be thrown out of its cage; a
    // width is clipped onto the bounds and onto a factor SIGMA_STEP of its
    // previous value (a width is a scale, and one step must not carry it into
    // the -3 kB x ln(sigma) wall near zero).  The realised direction is built
    // from d and the scale factors, never as x - x_old, which would lose
    // digits to the ~100 A coordinates once steps are small.
    const double d0 = o.y[i], d1 = o.y[i + 1], d2 = o.y[i + 2];
    const double len = o.stp * sqrt(d0 * d0 + d1 * d1 + d2 * d2);
    const double sc = len > Q_STEP ? Q_STEP / len : 1.0;  // a site moves at most Q_STEP per trial
    const double w0 = sc * d0, w1 = sc * d1, w2 = sc * d2;
    o.w[i] = w0;
    o.w[i + 1] = w1;
    o.w[i + 2] = w2;
    o.z[i] = o.x[i] + o.stp * w0;
    o.z[i + 1] = o.x[i + 1] + o.stp * w1;
    o.z[i + 2] = o.x[i + 2] + o.stp * w2;
    const double so = o.x[i + 3], ds = o.y[i + 3];
    const double lo = fmax(o.lb, so * (1.0 / SIGMA_STEP)), hi = fmin(o.ub, so * SIGMA_STEP);
    const double xs = so + o.stp * ds, cs = fmin(fmax(xs, lo), hi);
    o.z[i + 3] = cs;
    o.w[i + 3] = cs == xs ? ds : (cs - so) / o.stp;
  }
}

// Two-loop recursion of one small slot in one workgroup.  The segmented form
// costs ~4 launches per stored pair; a slot of a few thousand sites is
// launch-bound, so one workgroup per slot runs the whole recursion (every dot
// a fixed-order block sum), then d = -H g with the active part of the sigma
// bound zeroed, and writes g . d.  Pass t applies the update of step t and
// forms the dot the next step needs, so the recursion takes 2 np + 1 passes.
constexpr int TL_THREADS = 1024;
constexpr int TL_MAX_PAIRS = 64;
constexpr int TL_MAX_SITES = 8192;  // larger slots take the segmented kernels
struct TwoLoop {
  const double* g;
  double* q;
  double* d;
  const double* x;
  double* out;                       // out[0] = g . d
  const double* S[TL_MAX_PAIRS];     // step t = 0 is the newest pair
  const double* Y[TL_MAX_PAIRS];
  double rho[TL_MAX_PAIRS];
  Metric met;
  double lb, ub;
  int32_t np, nsite, npd;
};

__device__ __forceinline__ double tl_sum(double v, double* red) { return bsum<TL_THREADS>(v, red); }
template <int K>
__device__ __forceinline__ void tl_sum_n(double (&v)[K], double* red) { bsum_n<TL_THREADS>(v, red); }

// Each thread takes elements i0 + u TL_THREADS (u < TL_U) of a pass at a time
// and loads them all before it stores any: a slot's vectors are distinct
// arrays, but a store to q could otherwise hold back the next element's
// loads, and a pass would pay one memory latency per element.  With TL_U = 8
// a slot of up to 2,048 sites makes one round trip to memory per pass.
constexpr int TL_U = 8;

__global__ __launch_bounds__(TL_THREADS) void k_two_loop(const TwoLoop* __restrict__ ops) {
  __shared__ double red[(1 + 2 * NGROUP) * (TL_THREADS / 64)];
  __shared__ double al[TL_MAX_PAIRS];  // A_t = s_t . q of the first loop
  const TwoLoop& o = ops[blockIdx.x];
  const int n = 4 * o.nsite, np = o.np, npd = o.npd;
  const Metric met = o.met;
  double* __restrict__ q = o.q;
  const double* __restrict__ g = o.g;
  // q = g; A_0 = s_0 . q; per coordinate group g the scale
  // gamma_g = s_0 . y_0 / y_0 . D y_0 over the group, H0 = diag(gamma_g) D
  double a = 0.0, gs[2 * NGROUP] = {};
  {
    const double* __restrict__ s0 = o.S[0];
    const double* __restrict__ y0 = o.Y[0];
    for (int i0 = threadIdx.x; i0 < n; i0 += TL_U * TL_THREADS) {
      double gv[TL_U], sv[TL_U], yv[TL_U];
#pragma unroll
      for (int u = 0; u < TL_U; ++u) {
        const int i = i0 + u * TL_THREADS;
        if (i < n) {
          gv[u] = g[i];
          sv[u] = s0[i];
          yv[u] = y0[i];
        }
      }
#pragma unroll
      for (int u = 0; u < TL_U; ++u) {
        const int i = i0 + u * TL_THREADS;
        if (i < n) {
          const double gi = gv[u], si = sv[u], yi = yv[u];
          const int gr = group_of(i >> 2, i & 3, npd);
          q[i] = gi;
          a += si * gi;
#pragma unroll
          for (int k = 0; k < NGROUP; ++k) {
            gs[2 * k] += k == gr ? si * yi : 0.0;
            gs[2 * k + 1] += k == gr ? yi * yi * metric_inv(met, i >> 2) : 0.0;
          }
        }
      }
    }
  }
  double sums[1 + 2 * NGROUP] = {a};
#pragma unroll
  for (int k = 0; k < 2 * NGROUP; ++k) sums[1 + k] = gs[k];
  tl_sum_n(sums, red);
  double A = sums[0];
  double sy = 0.0, yy = 0.0, gam[NGROUP];
#pragma unroll
  for (int k = 0; k < 2 * NGROUP; ++k) gs[k] = sums[1 + k];
#pragma unroll
  for (int k = 0; k < NGROUP; ++k) {
    sy += gs[2 * k];
    yy += gs[2 * k + 1];
  }
#pragma unroll
  for (int k = 0; k < NGROUP; ++k) gam[k] = group_gamma(gs, sy, yy, k);
  if (threadIdx.x == 0) al[0] = A;
  // q += c v, then (scale) q = diag(gamma_g) D q; returns this thread's part of w . q
  auto update = [&](const double* __restrict__ v, const double* __restrict__ w, double c, bool scale) {
    double acc = 0.0;
    for (int i0 = threadIdx.x; i0 < n; i0 += TL_U * TL_THREADS) {
      double qv[TL_U], vv[TL_U], wv[TL_U];
#pragma unroll
      for (int u = 0; u < TL_U; ++u) {
        const int i = i0 + u * TL_THREADS;
        if (i < n) {
          qv[u] = q[i];
          vv[u] = v[i];
          wv[u] = w[i];
        }
      }
#pragma unroll
      for (int u = 0; u < TL_U; ++u) {
        const int i = i0 + u * TL_THREADS;
        if (i < n) {
          double qi = qv[u] + c * vv[u];
          if (scale) {
            const int gr = group_of(i >> 2, i & 3, npd);
            double gm = gam[0];
#pragma unroll
            for (int k = 1; k < NGROUP; ++k) gm = gr == k ? gam[k] : gm;
            qi = gm * metric_inv(met, i >> 2) * qi;
          }
          q[i] = qi;
          acc += wv[u] * qi;
        }
      }
    }
    return acc;
  };
  // first loop, newest to oldest: q -= rho_t A_t y_t; then q = diag(gamma_g) D q
  for (int t = 0; t < np; ++t) {
    const bool last = t + 1 == np;
    // next dot: A_{t+1}, or B_{np-1} = y . q
    A = tl_sum(update(o.Y[t], last ? o.Y[np - 1] : o.S[t + 1], -o.rho[t] * A, last), red);
    if (!last && threadIdx.x == 0) al[t + 1] = A;
  }
  // second loop, oldest to newest: q += rho_t (A_t - B_t) s_t, B_t = y_t . q
  double B = A;
  for (int t = np - 1; t > 0; --t) B = tl_sum(update(o.S[t], o.Y[t - 1], o.rho[t] * (al[t] - B), false), red);
  {  // t = 0: d = -q, sigma components at an active bound zeroed; g . d
    const double c = o.rho[0] * (al[0] - B);
    const double* __restrict__ st = o.S[0];
    const double* __restrict__ x = o.x;
    double* __restrict__ d = o.d;
    const double lb = o.lb, ub = o.ub;
    double acc = 0.0;
    for (int i0 = threadIdx.x; i0 < n; i0 += TL_U * TL_THREADS) {
      double qv[TL_U], sv[TL_U], gv[TL_U], xv[TL_U];
#pragma unroll
      for (int u = 0; u < TL_U; ++u) {
        const int i = i0 + u * TL_THREADS;
        if (i < n) {
          qv[u] = q[i];
          sv[u] = st[i];
          gv[u] = g[i];
          xv[u] = (i & 3) == 3 ? x[i] : 0.0;
        }
      }
#pragma unroll
      for (int u = 0; u < TL_U; ++u) {
        const int i = i0 + u * TL_THREADS;
        if (i < n) {
          double di = -(qv[u] + c * sv[u]);
          if ((i & 3) == 3) {
            const double xs = xv[u];
            if ((xs <= lb && di < 0.0) || (xs >= ub && di > 0.0)) di = 0.0;
          }
          d[i] = di;
          acc += gv[u] * di;
        }
      }
    }
    B = tl_sum(acc, red);
  }
  if (threadIdx.x == 0) o.out[0] = B;
}

// The same recursion in vector-free form, for np <= VF_PAIRS: one pass forms
// every dot product the two loops need -- s_t . g, s_a . y_b for a older
// than b (the first loop pairs each s with the newer y's, the second each y
// with the older s's), and the group sums of the initial scale -- and the
// loops then run on those scalars; a second pass forms r = diag(gamma_g) D
// (g - sum_t alpha_t y_t) and y_t . r, a third d = -(r + sum_t c_t s_t).
// Three passes and three block reductions instead of 2 np + 1 of each.
constexpr int VF_PAIRS = 5;
constexpr int VF_TRI = VF_PAIRS * (VF_PAIRS - 1) / 2;
constexpr int VF_SUMS = VF_PAIRS + VF_TRI + 2 * NGROUP;
__device__ __forceinline__ constexpr int vf_tri(int a, int b) { return VF_PAIRS + a * (a - 1) / 2 + b; }  // a > b

// The loops on the scalars.  acc: the VF_SUMS sums of the first pass.
// First loop, newest to oldest: alpha_t = rho_t s_t . (g - sum_{u<t} alpha_u y_u).
__device__ __forceinline__ void vf_alpha(const double* acc, const double* rho, int np, double (&al)[VF_PAIRS]) {
#pragma unroll
  for (int t = 0; t < VF_PAIRS; ++t) {
    double v = acc[t];
#pragma unroll
    for (int u = 0; u < t; ++u) v -= al[u] * acc[vf_tri(t, u)];
    al[t] = t < np ? rho[t] * v : 0.0;
  }
}
// Scale of each coordinate group from its sums (the totals where it has none).
__device__ __forceinline__ void vf_gamma(const double* acc, double (&gam)[NGROUP]) {
  const double* gs = acc + VF_PAIRS + VF_TRI;
  double sy = 0.0, yy = 0.0;
#pragma unroll
  for (int k = 0; k < NGROUP; ++k) {
    sy += gs[2 * k];
    yy += gs[2 * k + 1];
  }
#pragma unroll
  for (int k = 0; k < NGROUP; ++k) gam[k] = group_gamma(gs, sy, yy, k);
}
__device__ __forceinline__ double vf_pick(const double (&gam)[NGROUP], int gr) {
  double gm = gam[0];
#pragma unroll
  for (int k = 1; k < NGROUP; ++k) gm = gr == k ? gam[k] : gm;
  return gm;
}
// Second loop, oldest to newest: beta_t = rho_t y_t . (r + sum_{k>t} c_k s_k),
// c_t = alpha_t - beta_t; yr: y_t . r.
__device__ __forceinline__ void vf_coef(const double* acc, const double* yr, const double* rho, int np,
                                        const double (&al)[VF_PAIRS], double (&c)[VF_PAIRS]) {
#pragma unroll
  for (int t = VF_PAIRS - 1; t >= 0; --t) {
    double v = yr[t];
#pragma unroll
    for (int k = t + 1; k < VF_PAIRS; ++k)
      if (k < np) v += c[k] * acc[vf_tri(k, t)];
    c[t] = t < np ? al[t] - rho[t] * v : 0.0;
  }
}

__global__ __launch_bounds__(TL_THREADS) void k_two_loop_vf(const TwoLoop* __restrict__ ops) {
  __shared__ double red[VF_SUMS * (TL_THREADS / 64)];
  const TwoLoop& o = ops[blockIdx.x];
  const int n = 4 * o.nsite, np = o.np, npd = o.npd;
  const Metric met = o.met;
  const double* __restrict__ g = o.g;
  double* __restrict__ q = o.q;
  const double* S[VF_PAIRS];
  const double* Y[VF_PAIRS];
  double rho[VF_PAIRS];
#pragma unroll
  for (int t = 0; t < VF_PAIRS; ++t) {
    S[t] = t < np ? o.S[t] : nullptr;
    Y[t] = t < np ? o.Y[t] : nullptr;
    rho[t] = t < np ? o.rho[t] : 0.0;
  }
  // pass 1: [0, np) s_t . g; vf_tri(a, b) s_a . y_b; then s_0 . y_0 and
  // y_0 . D y_0 per coordinate group
  double acc[VF_SUMS] = {};
  for (int i = threadIdx.x; i < n; i += TL_THREADS) {
    double sv[VF_PAIRS], yv[VF_PAIRS];
    const double gi = g[i];
#pragma unroll
    for (int t = 0; t < VF_PAIRS; ++t)
      if (t < np) {
        sv[t] = S[t][i];
        yv[t] = Y[t][i];
      }
#pragma unroll
    for (int a = 0; a < VF_PAIRS; ++a)
      if (a < np) {
        acc[a] += sv[a] * gi;
#pragma unroll
        for (int b = 0; b < a; ++b) acc[vf_tri(a, b)] += sv[a] * yv[b];
      }
    const int gr = group_of(i >> 2, i & 3, npd);
    const double di = metric_inv(met, i >> 2);
#pragma unroll
    for (int k = 0; k < NGROUP; ++k) {
      acc[VF_PAIRS + VF_TRI + 2 * k] += k == gr ? sv[0] * yv[0] : 0.0;
      acc[VF_PAIRS + VF_TRI + 2 * k + 1] += k == gr ? yv[0] * yv[0] * di : 0.0;
    }
  }
  tl_sum_n(acc, red);
  double al[VF_PAIRS], gam[NGROUP];
  vf_alpha(acc, rho, np, al);
  vf_gamma(acc, gam);
  // pass 2: r = diag(gamma_g) D (g - sum_t alpha_t y_t) into q; y_t . r
  double yr[VF_PAIRS] = {};
  for (int i = threadIdx.x; i < n; i += TL_THREADS) {
    double yv[VF_PAIRS];
#pragma unroll
    for (int t = 0; t < VF_PAIRS; ++t)
      if (t < np) yv[t] = Y[t][i];
    double qi = g[i];
#pragma unroll
    for (int t = 0; t < VF_PAIRS; ++t)
      if (t < np) qi -= al[t] * yv[t];
    const double ri = vf_pick(gam, group_of(i >> 2, i & 3, npd)) * metric_inv(met, i >> 2) * qi;
    q[i] = ri;
#pragma unroll
    for (int t = 0; t < VF_PAIRS; ++t)
      if (t < np) yr[t] += yv[t] * ri;
  }
  tl_sum_n(yr, red);
  double c[VF_PAIRS];
  vf_coef(acc, yr, rho, np, al, c);
  // pass 3: d = -(r + sum_t c_t s_t) (oldest first), widths at an active bound not moved; g . d
  const double* __restrict__ x = o.x;
  double* __restrict__ d = o.d;
  const double lb = o.lb, ub = o.ub;
  double gd = 0.0;
  for (int i = threadIdx.x; i < n; i += TL_THREADS) {
    double sv[VF_PAIRS];
#pragma unroll
    for (int t = 0; t < VF_PAIRS; ++t)
      if (t < np) sv[t] = S[t][i];
    const double gi = g[i], xs = (i & 3) == 3 ? x[i] : 0.0;
    double ri = q[i];
#pragma unroll
    for (int t = VF_PAIRS - 1; t >= 0; --t)
      if (t < np) ri += c[t] * sv[t];
    double di = -ri;
    if ((i & 3) == 3 && ((xs <= lb && di < 0.0) || (xs >= ub && di > 0.0))) di = 0.0;
    d[i] = di;
    gd += gi * di;
  }
  gd = tl_sum(gd, red);
  if (threadIdx.x == 0) o.out[0] = gd;
}

// The vector-free recursion for slots too large for one workgroup: the three
// passes as grid launches over all such slots, each followed by a
// fixed-order reduction of its block partials into the slot's device
// scalars (the VF_SUMS first-pass sums, then the y_t . r), which the next
// pass reads to run the loops on the scalars itself.
struct SegVF {
  const double* g;
  const double* S[VF_PAIRS];  // step t = 0 is the newest pair
  const double* Y[VF_PAIRS];
  double rho[VF_PAIRS];
  double* q;                  // r
  double* d;
  const double* x;
  double* sums;               // VF_SUMS + VF_PAIRS device scalars
  double* out;                // g . d
  Metric met;
  double lb, ub;
  int32_t np, npd;
  int32_t nsite, blk0;
};

__global__ __launch_bounds__(TB) void k_seg_vf1(const SegVF* __restrict__ ops, const int32_t* __restrict__ b2l,
                                                double* __restrict__ part, int nblk) {
  __shared__ double red[VF_SUMS * (TB / 64)];
  const SegVF o = ops[b2l[blockIdx.x]];
  const int s = (blockIdx.x - o.blk0) * TB + threadIdx.x;
  const int np = o.np;
  double acc[VF_SUMS] = {};
  if (s < o.nsite) {
    const size_t i0 = 4 * (size_t)s;
    const double di = metric_inv(o.met, s);
#pragma unroll
    for (int c = 0; c < 4; ++c) {
      const size_t i = i0 + c;
      double sv[VF_PAIRS], yv[VF_PAIRS];
      const double gi = o.g[i];
#pragma unroll
      for (int t = 0; t < VF_PAIRS; ++t)
        if (t < np) {
          sv[t] = o.S[t][i];
          yv[t] = o.Y[t][i];
        }
#pragma unroll
      for (int a = 0; a < VF_PAIRS; ++a)
        if (a < np) {
          acc[a] += sv[a] * gi;
#pragma unroll
          for (int b = 0; b < a; ++b) acc[vf_tri(a, b)] += sv[a] * yv[b];
        }
      const int gr = group_of(s, c, o.npd);
#pragma unroll
      for (int k = 0; k < NGROUP; ++k) {
        acc[VF_PAIRS + VF_TRI + 2 * k] += k == gr ? sv[0] * yv[0] : 0.0;
        acc[VF_PAIRS + VF_TRI + 2 * k + 1] += k == gr ? yv[0] * yv[0] * di : 0.0;
      }
    }
  }
  bsum_n<TB>(acc, red);
  if (threadIdx.x == 0)
#pragma unroll
    for (int k = 0; k < VF_SUMS; ++k) part[(size_t)k * nblk + blockIdx.x] = acc[k];
}

__global__ __launch_bounds__(TB) void k_seg_vf2(const SegVF* __restrict__ ops, const int32_t* __restrict__ b2l,
                                                double* __restrict__ part, int nblk) {
  __shared__ double red[VF_PAIRS * (TB / 64)];
  const SegVF o = ops[b2l[blockIdx.x]];
  const int s = (blockIdx.x - o.blk0) * TB + threadIdx.x;
  const int np = o.np;
  double al[VF_PAIRS], gam[NGROUP];
  vf_alpha(o.sums, o.rho, np, al);
  vf_gamma(o.sums, gam);
  double yr[VF_PAIRS] = {};
  if (s < o.nsite) {
    const size_t i0 = 4 * (size_t)s;
    const double di = metric_inv(o.met, s);
#pragma unroll
    for (int c = 0; c < 4; ++c) {
      const size_t i = i0 + c;
      double yv[VF_PAIRS];
#pragma unroll
      for (int t = 0; t < VF_PAIRS; ++t)
        if (t < np) yv[t] = o.Y[t][i];
      double qi = o.g[i];
#pragma unroll
      for (int t = 0; t < VF_PAIRS; ++t)
        if (t < np) qi -= al[t] * yv[t];
      const double ri = vf_pick(gam, group_of(s, c, o.npd)) * di * qi;
      o.q[i] = ri;
#pragma unroll
      for (int t = 0; t < VF_PAIRS; ++t)
        if (t < np) yr[t] += yv[t] * ri;
    }
  }
  bsum_n<TB>(yr, red);
  if (threadIdx.x == 0)
#pragma unroll
    for (int k = 0; k < VF_PAIRS; ++k) part[(size_t)k * nblk + blockIdx.x] = yr[k];
}

__global__ __launch_bounds__(TB) void k_seg_vf3(const SegVF* __restrict__ ops, const int32_t* __restrict__ b2l,
                                                double* __restrict__ part) {
  __shared__ double red[TB / 64];
  const SegVF o = ops[b2l[blockIdx.x]];
  const int s = (blockIdx.x - o.blk0) * TB + threadIdx.x;
  const int np = o.np;
  double al[VF_PAIRS], c[VF_PAIRS];
  vf_alpha(o.sums, o.rho, np, al);
  vf_coef(o.sums, o.sums + VF_SUMS, o.rho, np, al, c);
  double gd = 0.0;
  if (s < o.nsite) {
    const size_t i0 = 4 * (size_t)s;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      const size_t i = i0 + k;
      double ri = o.q[i];
#pragma unroll
      for (int t = VF_PAIRS - 1; t >= 0; --t)
        if (t < np) ri += c[t] * o.S[t][i];
      double di = -ri;
      if (k == 3) {
        const double xs = o.x[i];
        if ((xs <= o.lb && di < 0.0) || (xs >= o.ub && di > 0.0)) di = 0.0;
      }
      o.d[i] = di;
      gd += o.g[i] * di;
    }
  }
  gd = bsum<TB>(gd, red);
  if (threadIdx.x == 0) part[blockIdx.x] = gd;
}

// one block per op: fixed-order sums of the K block partials of pass PASS
template <int K, int PASS>
__global__ __launch_bounds__(256) void k_seg_vf_final(const SegVF* __restrict__ ops, const double* __restrict__ part,
                                                      int nblk) {
  __shared__ double red[4];
  const SegVF o = ops[blockIdx.x];
  const int n = site_blocks(o.nsite);
  double* out = PASS == 1 ? o.sums : (PASS == 2 ? o.sums + VF_SUMS : o.out);
#pragma unroll
  for (int k = 0; k < K; ++k) {
    double v = 0.0;
    for (int j = threadIdx.x; j < n; j += 256) v += part[(size_t)k * nblk + o.blk0 + j];
    v = bsum<256>(v, red);
    if (threadIdx.x == 0) out[k] = v;
  }
}

}  // namespace

// ---------------------------------------------------------------- host side
// Pinned staging for per-launch descriptor arrays.  Space is handed out
// bump-style and reclaimed only after the stream has been synchronised.
// Operations are queued (q, with their key) and issued at the next
// synchronisation after one copy of everything staged: an optimizer iteration
// is ~50 small launches, and one host-to-device copy per launch would cost
// more than the kernels of a small lattice.
struct Arena {
  char* h = nullptr;
  char* d = nullptr;
  size_t cap = 0, used = 0;
  std::vector<std::function<void(hipStream_t)>> q;
  std::vector<uint64_t> key;  // kind and argument values of every queued operation, in order
  template <class T>
  T* push(const T* src, size_t n) {
    const size_t bytes = ((n * sizeof(T)) + 255) & ~(size_t)255;
    if (used + bytes > cap) return nullptr;
    std::memcpy(h + used, src, n * sizeof(T));
    T* out = reinterpret_cast<T*>(d + used);
    used += bytes;
    return out;
  }
};

struct Slot {
  a2c_fe_ctx* fe = nullptr;
  bool owned = false;
  bool used = false;
  size_t charged = 0;  // bytes of the HBM budget this slot holds (its footprint)
  // optimizer
  bool running = false, started = false, mixed = false;
  a2c_lmvm_opts opts{};
  a2c_lmvm_report rep{};
  int m = 0, npairs = 0, head = 0;
  double *x = nullptr, *g = nullptr, *gp = nullptr, *d = nullptr, *xo = nullptr, *go = nullptr, *q = nullptr;
  std::vector<double*> S, Y;
  std::vector<double> rho, alpha;
  double f = 0, fold = 0, gnorm = 0, gnorm0 = 0, gd = 0, stp = 0;
  bool steepest = false;
  int lsstat = 0, nfev = 0;
  MoreThuente ls;
  int N() const { return fe->N; }
  Metric metric() const {  // the optimizer's diagonal metric (identity unless opts.precond)
    Metric m;
    if (opts.precond) { m.occ = fe->dx; m.npd = fe->desc.n_pd; }
    return m;
  }
};

}  // namespace a2c

namespace a2c {
// Instantiated launch sequences, by key.
struct GraphCache {
  struct Entry {
    std::vector<uint64_t> key;
    hipGraphExec_t exec = nullptr;
    uint64_t last = 0;  // use stamp (least recently used is evicted)
  };
  static constexpr size_t CAP = 64;
  std::unordered_multimap<uint64_t, Entry> map;
  std::unordered_map<uint64_t, int> seen;  // keys issued directly so far (hash -> count)
  uint64_t clock = 0;
  int64_t replays = 0, captures = 0, direct = 0;
};
}  // namespace a2c

struct a2c_batch {
  a2c_batch_desc desc{};
  hipStream_t stream = nullptr;
  std::string err;
  std::vector<a2c::Slot> slots;
  a2c::Arena arena;
  a2c::GraphCache graphs;
  int qrc = 0;  // first error of a queued operation
  bool fused_two_loop = true;  // small slots run the two-loop recursion in one workgroup
  bool vector_free = true;     // slots of at most VF_PAIRS pairs take the vector-free form
  double* dF = nullptr;    // objective per slot (device)
  double* hF = nullptr;    // pinned mirror
  double* dres = nullptr;  // 2 doubles per slot (device dot results read by the host)
  double* hres = nullptr;  // pinned mirror
  double* dscal = nullptr;  // SCAL_PER_SLOT doubles per slot: optimizer scalars that stay on the device
  double* dpart = nullptr;  // PART_PER_BLOCK * max blocks
  size_t part_cap = 0;      // blocks
  float4* dtab = nullptr;
  double* dtab64 = nullptr;
  size_t mem_used = 0;
};

namespace a2c {
namespace {

int berr(a2c_batch* b, int code, const std::string& m) {
  b->err = m;
  return code;
}
#define B_HIP(b, call)                                                                            \
  do {                                                                                            \
    hipError_t e_ = (call);                                                                       \
    if (e_ != hipSuccess) return berr((b), A2C_EHIP, std::string(#call) + ": " + hipGetErrorString(e_)); \
  } while (0)

inline uint64_t u64(const void* p) { return (uint64_t)(uintptr_t)p; }
inline uint64_t u64(double v) {
  uint64_t r;
  std::memcpy(&r, &v, sizeof r);
  return r;
}
inline uint64_t u64(long long v) { return (uint64_t)v; }

// Queue one operation; key holds its kind and every argument value it passes.
void enqueue(a2c_batch* b, std::initializer_list<uint64_t> key, std::function<void(hipStream_t)> fn) {
  b->arena.key.insert(b->arena.key.end(), key);
  b->arena.q.push_back(std::move(fn));
}
void enqueue_d2h(a2c_batch* b, void* dst, const void* src, size_t bytes) {
  enqueue(b, {7, u64(dst), u64(src), (uint64_t)bytes}, [=](hipStream_t st) {
    (void)hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, st);
  });
}

uint64_t key_hash(const std::vector<uint64_t>& k) {  // FNV-1a over the words
  uint64_t h = 1469598103934665603ull;
  for (uint64_t w : k) h = (h ^ w) * 1099511628211ull;
  return h;
}

// Issue the queued sequence: replay its graph, capture it (second sighting
// of its key), or issue it directly.
int issue(a2c_batch* b) {
  Arena& a = b->arena;
  if (a.q.empty()) return A2C_OK;
  if (a.used) B_HIP(b, hipMemcpyAsync(a.d, a.h, a.used, hipMemcpyHostToDevice, b->stream));
  GraphCache& gc = b->graphs;
  hipGraphExec_t exec = nullptr;
  if (b->desc.graphs) {
    a.key.push_back(a.used);
    const uint64_t h = key_hash(a.key);
    auto range = gc.map.equal_range(h);
    for (auto it = range.first; it != range.second; ++it)
      if (it->second.key == a.key) {
        exec = it->second.exec;
        it->second.last = ++gc.clock;
        gc.replays++;
        break;
      }
    if (!exec && ++gc.seen[h] >= 2) {
      if (gc.map.size() >= GraphCache::CAP) {  // evict the least recently used
        auto old = gc.map.begin();
        for (auto it = gc.map.begin(); it != gc.map.end(); ++it)
          if (it->second.last < old->second.last) old = it;
        (void)hipGraphExecDestroy(old->second.exec);
        gc.map.erase(old);
      }
      hipGraph_t g = nullptr;
      B_HIP(b, hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
      b->qrc = 0;
      for (auto& f : a.q) f(b->stream);
      const hipError_t le = hipGetLastError();
      const hipError_t ce = hipStreamEndCapture(b->stream, &g);
      if (le != hipSuccess || ce != hipSuccess || b->qrc) {
        if (g) (void)hipGraphDestroy(g);
        return berr(b, b->qrc ? b->qrc : A2C_EHIP, std::string("launch sequence capture failed: ") +
                                                       hipGetErrorString(le != hipSuccess ? le : ce));
      }
      const hipError_t ie = hipGraphInstantiate(&exec, g, nullptr, nullptr, 0);
      (void)hipGraphDestroy(g);
      B_HIP(b, ie);
      GraphCache::Entry e;
      e.key = a.key;
      e.exec = exec;
      e.last = ++gc.clock;
      gc.map.emplace(h, std::move(e));
      gc.captures++;
    }
  }
  if (exec) {
    B_HIP(b, hipGraphLaunch(exec, b->stream));
  } else {
    b->qrc = 0;
    for (auto& f : a.q) f(b->stream);
    B_HIP(b, hipGetLastError());
    if (b->qrc) return berr(b, b->qrc, "fused eval launch failed");
    gc.direct++;
  }
  a.q.clear();
  a.key.clear();
  return A2C_OK;
}

int bsync(a2c_batch* b) {
  if (int rc = issue(b)) return rc;
  B_HIP(b, hipStreamSynchronize(b->stream));
  b->arena.used = 0;
  return A2C_OK;
}

// make sure the partial-sum buffer holds nblk blocks (x PART_PER_BLOCK: up
// to that many sums per block)
constexpr int PART_PER_BLOCK = VF_SUMS > TL_PART ? VF_SUMS : TL_PART;
int ensure_part(a2c_batch* b, size_t nblk) {
  if (nblk <= b->part_cap) return A2C_OK;
  if (int rc = bsync(b)) return rc;
  (void)dfree(b->dpart);
  b->dpart = nullptr;
  B_HIP(b, dmalloc(&b->dpart, sizeof(double) * PART_PER_BLOCK * nblk));
  b->part_cap = nblk;
  return A2C_OK;
}

// block -> op map for ops over the listed slots; fills blk0 via setter
template <class Op, class Fill>
int seg_launch_prep(a2c_batch* b, const std::vector<int>& ids, Fill fill, Op** dops, int32_t** db2l, int* nblk) {
  std::vector<Op> ops(ids.size());
  std::vector<int32_t> map;
  int cum = 0;
  for (size_t k = 0; k < ids.size(); ++k) {
    Slot& s = b->slots[ids[k]];
    fill(ops[k], s);
    ops[k].nsite = s.N();
    ops[k].blk0 = cum;
    const int nb = site_blocks(s.N());
    map.insert(map.end(), nb, (int32_t)k);
    cum += nb;
  }
  *nblk = cum;
  if (int rc = ensure_part(b, cum)) return rc;
  size_t need = ((ops.size() * sizeof(Op) + 255) & ~(size_t)255) + ((map.size() * 4 + 255) & ~(size_t)255);
  if (b->arena.used + need > b->arena.cap)
    if (int rc = bsync(b)) return rc;
  *dops = b->arena.push(ops.data(), ops.size());
  *db2l = map.empty() ? nullptr : b->arena.push(map.data(), map.size());
  if (!*dops || (!map.empty() && !*db2l)) return berr(b, A2C_ENOMEM, "staging arena too small");
  return A2C_OK;
}

template <class Fill>
int seg_axpby(a2c_batch* b, const std::vector<int>& ids, Fill fill) {
  if (ids.empty()) return A2C_OK;
  auto fill0 = [&](SegAx& o, Slot& s) {
    o.met = Metric{};
    o.scaled = 0;
    o.amode = o.bmode = 0;
    o.p1 = o.p2 = o.q1 = o.q2 = nullptr;
    fill(o, s);
  };
  SegAx* ops;
  int32_t* map;
  int nblk;
  if (int rc = seg_launch_prep<SegAx>(b, ids, fill0, &ops, &map, &nblk)) return rc;
  if (nblk)
    enqueue(b, {1, u64(ops), u64(map), (uint64_t)nblk},
            [=](hipStream_t st) { hipLaunchKernelGGL(k_seg_axpby, dim3(nblk), dim3(TB), 0, st, ops, map); });
  return A2C_OK;
}

// block partials, then one fixed-order sum per op
void enqueue_dot(a2c_batch* b, const SegDot* ops, const int32_t* map, int nblk, unsigned nops) {
  double* part = b->dpart;
  if (nblk)
    enqueue(b, {2, u64(ops), u64(map), u64(part), (uint64_t)nblk}, [=](hipStream_t st) {
      hipLaunchKernelGGL(k_seg_dot, dim3(nblk), dim3(TB), 0, st, ops, map, part, nblk);
    });
  enqueue(b, {3, u64(ops), u64(part), (uint64_t)nblk, nops}, [=](hipStream_t st) {
    hipLaunchKernelGGL(k_seg_dot_final, dim3(nops), dim3(1024), 0, st, ops, part, nblk);
  });
}

// dot products per slot; results (2 per slot) returned in out[k*2 + 0/1]
template <class Fill>
int seg_dot(a2c_batch* b, const std::vector<int>& ids, Fill fill, std::vector<double>& out) {
  out.assign(2 * ids.size(), 0.0);
  if (ids.empty()) return A2C_OK;
  auto fill2 = [&](SegDot& o, Slot& s) {
    o.a2 = o.b2 = nullptr;
    o.met = Metric{};
    o.k1 = o.k2 = 0;
    fill(o, s);
    o.out = b->dres + 2 * (&s - b->slots.data());
  };
  SegDot* ops;
  int32_t* map;
  int nblk;
  if (int rc = seg_launch_prep<SegDot>(b, ids, fill2, &ops, &map, &nblk)) return rc;
  enqueue_dot(b, ops, map, nblk, (unsigned)ids.size());
  enqueue_d2h(b, b->hres, b->dres, sizeof(double) * 2 * b->slots.size());
  if (int rc = bsync(b)) return rc;
  for (size_t k = 0; k < ids.size(); ++k) {
    out[2 * k] = b->hres[2 * ids[k]];
    out[2 * k + 1] = b->hres[2 * ids[k] + 1];  // meaningful only for paired dots (a2 set)
  }
  return A2C_OK;
}

// dot products per slot left on the device (fill sets o.out); no host sync
template <class Fill>
int seg_dot_dev(a2c_batch* b, const std::vector<int>& ids, Fill fill) {
  if (ids.empty()) return A2C_OK;
  auto fill2 = [&](SegDot& o, Slot& s) {
    o.a2 = o.b2 = nullptr;
    o.met = Metric{};
    o.k1 = o.k2 = 0;
    fill(o, s);
  };
  SegDot* ops;
  int32_t* map;
  int nblk;
  if (int rc = seg_launch_prep<SegDot>(b, ids, fill2, &ops, &map, &nblk)) return rc;
  enqueue_dot(b, ops, map, nblk, (unsigned)ids.size());
  return A2C_OK;
}

// one pass of the segmented two-loop recursion (first: pass 0) and the
// fixed-order reduction of its dots
template <class Fill>
int seg_tl(a2c_batch* b, const std::vector<int>& ids, bool first, Fill fill) {
  if (ids.empty()) return A2C_OK;
  auto fill0 = [&](SegTL& o, Slot& s) {
    o = SegTL{};
    fill(o, s);
  };
  SegTL* ops;
  int32_t* map;
  int nblk;
  if (int rc = seg_launch_prep<SegTL>(b, ids, fill0, &ops, &map, &nblk)) return rc;
  double* part = b->dpart;
  const unsigned nops = (unsigned)ids.size();
  if (first) {
    if (nblk)
      enqueue(b, {8, u64(ops), u64(map), u64(part), (uint64_t)nblk}, [=](hipStream_t st) {
        hipLaunchKernelGGL(k_seg_tl0, dim3(nblk), dim3(TB), 0, st, ops, map, part, nblk);
      });
    enqueue(b, {9, u64(ops), u64(part), (uint64_t)nblk, nops}, [=](hipStream_t st) {
      hipLaunchKernelGGL((k_seg_final<SegTL, TL_PART>), dim3(nops), dim3(256), 0, st, ops, part, nblk);
    });
  } else {
    if (nblk)
      enqueue(b, {10, u64(ops), u64(map), u64(part), (uint64_t)nblk}, [=](hipStream_t st) {
        hipLaunchKernelGGL(k_seg_tl, dim3(nblk), dim3(TB), 0, st, ops, map, part);
      });
    enqueue(b, {11, u64(ops), u64(part), (uint64_t)nblk, nops}, [=](hipStream_t st) {
      hipLaunchKernelGGL((k_seg_final<SegTL, 1>), dim3(nops), dim3(256), 0, st, ops, part, nblk);
    });
  }
  return A2C_OK;
}

// the three passes of the segmented vector-free recursion and their reductions
template <class Fill>
int seg_vf(a2c_batch* b, const std::vector<int>& ids, Fill fill) {
  if (ids.empty()) return A2C_OK;
  auto fill0 = [&](SegVF& o, Slot& s) {
    o = SegVF{};
    fill(o, s);
  };
  SegVF* ops;
  int32_t* map;
  int nblk;
  if (int rc = seg_launch_prep<SegVF>(b, ids, fill0, &ops, &map, &nblk)) return rc;
  double* part = b->dpart;
  const unsigned nops = (unsigned)ids.size();
  if (nblk)
    enqueue(b, {14, u64(ops), u64(map), u64(part), (uint64_t)nblk}, [=](hipStream_t st) {
      hipLaunchKernelGGL(k_seg_vf1, dim3(nblk), dim3(TB), 0, st, ops, map, part, nblk);
    });
  enqueue(b, {15, u64(ops), u64(part), (uint64_t)nblk, nops}, [=](hipStream_t st) {
    hipLaunchKernelGGL((k_seg_vf_final<VF_SUMS, 1>), dim3(nops), dim3(256), 0, st, ops, part, nblk);
  });
  if (nblk)
    enqueue(b, {16, u64(ops), u64(map), u64(part), (uint64_t)nblk}, [=](hipStream_t st) {
      hipLaunchKernelGGL(k_seg_vf2, dim3(nblk), dim3(TB), 0, st, ops, map, part, nblk);
    });
  enqueue(b, {17, u64(ops), u64(part), (uint64_t)nblk, nops}, [=](hipStream_t st) {
    hipLaunchKernelGGL((k_seg_vf_final<VF_PAIRS, 2>), dim3(nops), dim3(256), 0, st, ops, part, nblk);
  });
  if (nblk)
    enqueue(b, {18, u64(ops), u64(map), u64(part), (uint64_t)nblk}, [=](hipStream_t st) {
      hipLaunchKernelGGL(k_seg_vf3, dim3(nblk), dim3(TB), 0, st, ops, map, part);
    });
  enqueue(b, {19, u64(ops), u64(part), (uint64_t)nblk, nops}, [=](hipStream_t st) {
    hipLaunchKernelGGL((k_seg_vf_final<1, 3>), dim3(nops), dim3(256), 0, st, ops, part, nblk);
  });
  return A2C_OK;
}

// s, y, gp of the accepted step and the dots s.y, gp.gp per slot, returned in
// out[2k], out[2k + 1] (one round trip)
int seg_pair(a2c_batch* b, const std::vector<int>& ids, std::vector<double>& out) {
  out.assign(2 * ids.size(), 0.0);
  if (ids.empty()) return A2C_OK;
  auto fill = [&](SegPair& o, Slot& s) {
    o.x = s.x; o.xo = s.xo; o.g = s.g; o.go = s.go;
    o.S = s.S[s.head]; o.Y = s.Y[s.head]; o.gp = s.gp;
    o.lb = s.opts.sigma_lb; o.ub = s.opts.sigma_ub;
    o.out = b->dres + 2 * (&s - b->slots.data());
    o.out2 = o.out + 1;
  };
  SegPair* ops;
  int32_t* map;
  int nblk;
  if (int rc = seg_launch_prep<SegPair>(b, ids, fill, &ops, &map, &nblk)) return rc;
  double* part = b->dpart;
  const unsigned nops = (unsigned)ids.size();
  if (nblk)
    enqueue(b, {12, u64(ops), u64(map), u64(part), (uint64_t)nblk}, [=](hipStream_t st) {
      hipLaunchKernelGGL(k_seg_pair, dim3(nblk), dim3(TB), 0, st, ops, map, part, nblk);
    });
  enqueue(b, {13, u64(ops), u64(part), (uint64_t)nblk, nops}, [=](hipStream_t st) {
    hipLaunchKernelGGL((k_seg_final<SegPair, 2>), dim3(nops), dim3(256), 0, st, ops, part, nblk);
  });
  enqueue_d2h(b, b->hres, b->dres, sizeof(double) * 2 * b->slots.size());
  if (int rc = bsync(b)) return rc;
  for (size_t k = 0; k < ids.size(); ++k) {
    out[2 * k] = b->hres[2 * ids[k]];
    out[2 * k + 1] = b->hres[2 * ids[k] + 1];
  }
  return A2C_OK;
}

template <int MODE, class Fill>
int seg_bound(a2c_batch* b, const std::vector<int>& ids, Fill fill) {
  if (ids.empty()) return A2C_OK;
  SegBound* ops;
  int32_t* map;
  int nblk;
  if (int rc = seg_launch_prep<SegBound>(b, ids, fill, &ops, &map, &nblk)) return rc;
  if (nblk)
    enqueue(b, {4, (uint64_t)MODE, u64(ops), u64(map), (uint64_t)nblk},
            [=](hipStream_t st) { hipLaunchKernelGGL(k_seg_bound<MODE>, dim3(nblk), dim3(TB), 0, st, ops, map); });
  return A2C_OK;
}

// Fused objective + gradient of the listed slots at X[slot] -> G[slot].
// state_of(s) selects the state/gradient buffers.  Slots are evaluated in
// FP64 or mixed precision as `fp64` says.  f values land in hF after sync.
// A slot whose context has a general model (a tabulated potential or a node
// table: a2c_fe_set_potential_*, a2c_fe_set_quadrature) is evaluated by the
// general kernels, on its own; the fused launch takes the others.
template <class StateOf>
int gen_eval(a2c_batch* b, int id, bool fp64, StateOf state_of) {
  Slot& s = b->slots[id];
  LatDev L = s.fe->sys();
  const double* X;
  double* G;
  state_of(s, &X, &G);
  L.X = X;
  L.G = G;
  L.f = b->dF + id;
  if (b->arena.used + 256 + sizeof(LatDev) > b->arena.cap)
    if (int rc = bsync(b)) return rc;
  LatDev* dl = b->arena.push(&L, 1);
  if (!dl) return berr(b, A2C_ENOMEM, "staging arena too small");
  const GenDev g = s.fe->gen;
  const int N = s.N(), dm = b->desc.deriv_mode;
  const double eps = b->desc.fd_eps;
  enqueue(b, {8, (uint64_t)fp64, u64(dl), (uint64_t)N, u64(g.nodes), u64(g.tab), u64(g.tab32), (uint64_t)g.K,
              (uint64_t)g.pot, (uint64_t)dm, u64(eps)},
          [=](hipStream_t st) {
            const int rc = run_eval_gen(dl, N, g, dm, eps, !fp64, st, nullptr);
            if (rc && !b->qrc) b->qrc = rc;
          });
  return A2C_OK;
}

template <class StateOf>
int fused_eval(a2c_batch* b, const std::vector<int>& ids_in, bool fp64, StateOf state_of, bool sync = true) {
  std::vector<int> ids;
  for (int i : ids_in) {
    if (!b->slots[i].fe->gen_on) {
      ids.push_back(i);
    } else if (int rc = gen_eval(b, i, fp64, state_of)) {
      return rc;
    }
  }
  if (ids.empty()) {
    if (!sync || ids_in.empty()) return A2C_OK;
    enqueue_d2h(b, b->hF, b->dF, sizeof(double) * b->slots.size());
    return bsync(b);
  }
  std::vector<LatDev> lats(ids.size());
  std::vector<int32_t> map, omap;  // site block -> lattice; pair-owner workgroup -> lattice
  int cum = 0, ocum = 0, rules = 0;
  for (size_t k = 0; k < ids.size(); ++k) {
    Slot& s = b->slots[ids[k]];
    LatDev L = s.fe->sys();
    rules |= 1 << L.rule;
    const double* X;
    double* G;
    state_of(s, &X, &G);
    L.X = X;
    L.G = G;
    L.f = b->dF + ids[k];
    L.blk0 = cum;
    L.kl32 = lanes_per_site(s.N(), false);
    L.kl64 = lanes_per_site(s.N(), true);
    const int kl = fp64 ? L.kl64 : L.kl32;
    L.owg0_32 = L.owg0_64 = ocum;
    lats[k] = L;
    const int nb = site_blocks(s.N());
    map.insert(map.end(), nb, (int32_t)k);
    omap.insert(omap.end(), (size_t)nb * kl, (int32_t)k);
    cum += nb;
    ocum += nb * kl;
  }
  size_t need = ((lats.size() * sizeof(LatDev) + 255) & ~(size_t)255) + ((map.size() * 4 + 255) & ~(size_t)255) +
                ((omap.size() * 4 + 255) & ~(size_t)255);
  if (b->arena.used + need > b->arena.cap)
    if (int rc = bsync(b)) return rc;
  LatDev* dl = b->arena.push(lats.data(), lats.size());
  int32_t* dm = map.empty() ? nullptr : b->arena.push(map.data(), map.size());
  int32_t* dom = omap.empty() ? nullptr : b->arena.push(omap.data(), omap.size());
  if (!dl || (!map.empty() && (!dm || !dom))) return berr(b, A2C_ENOMEM, "staging arena too small");
  Launch L;
  L.lats = dl;
  L.b2l = dm;
  L.nblk = cum;
  L.ob2l = dom;
  L.onblk = ocum;
  L.nlat = (int)ids.size();
  L.rules = rules;
  L.stream = b->stream;
  L.deriv_mode = b->desc.deriv_mode;
  L.fd_eps = b->desc.fd_eps;
  L.tab = b->dtab;
  L.tab64 = b->dtab64;
  L.mixed_full = mixed_full_pairs();
  L.full_pairs64 = fp64_full_pairs();
  enqueue(b, {5, (uint64_t)fp64, u64(L.lats), u64(L.b2l), u64(L.ob2l), (uint64_t)L.nblk, (uint64_t)L.onblk,
              (uint64_t)L.nlat, (uint64_t)L.rules, (uint64_t)L.deriv_mode, u64(L.fd_eps), u64(L.tab), u64(L.tab64),
              (uint64_t)L.mixed_full, (uint64_t)L.full_pairs64},
          [=](hipStream_t st) {
            Launch l = L;
            l.stream = st;
            const int rc = fp64 ? run_eval64(l, nullptr) : run_eval32(l, nullptr);
            if (rc && !b->qrc) b->qrc = rc;
          });
  if (!sync) return A2C_OK;
  enqueue_d2h(b, b->hF, b->dF, sizeof(double) * b->slots.size());
  return bsync(b);
}

// ---------------------------------------------------------------- optimizer
void lmvm_free(Slot& s) {
  for (double* p : {s.x, s.g, s.gp, s.d, s.xo, s.go, s.q}) (void)dfree(p);
  for (double* p : s.S) (void)dfree(p);
  for (double* p : s.Y) (void)dfree(p);
  s.x = s.g = s.gp = s.d = s.xo = s.go = s.q = nullptr;
  s.S.clear();
  s.Y.clear();
}

int lmvm_alloc(a2c_batch* b, Slot& s) {
  lmvm_free(s);
  const size_t n = std::max<size_t>(4 * (size_t)s.N(), 4);
  for (double** p : {&s.x, &s.g, &s.gp, &s.d, &s.xo, &s.go, &s.q}) B_HIP(b, dmalloc(p, n * sizeof(double)));
  s.S.assign(s.m, nullptr);
  s.Y.assign(s.m, nullptr);
  for (int i = 0; i < s.m; ++i) {
    B_HIP(b, dmalloc(&s.S[i], n * sizeof(double)));
    B_HIP(b, dmalloc(&s.Y[i], n * sizeof(double)));
  }
  s.rho.assign(s.m, 0.0);
  s.alpha.assign(s.m, 0.0);
  return A2C_OK;
}

auto lmvm_state = [](Slot& s, const double** X, double** G) {
  *X = s.x;
  *G = s.g;
};

// Norm of the projected gradient below which a mixed-precision slot continues
// in FP64: mixed_gtol, or automatically 2e-9 sqrt(N) -- about twice the
// difference of the mixed and FP64 gradients at relaxed states (measured
// 9e-10 sqrt(N) for 1,728 to 110,592 sites).  On 64 lattices of 4,096 to
// 110,592 sites, 1.5e-9 / 2e-9 / 3e-9 / 5e-9 / 1e-8 / 2e-8 sqrt(N) took
// 2,870 / 3,042 / 3,317 / 3,414 / 3,850 / 4,423 FP64 evaluations of about
// 8,000; below 2e-9 the mixed phase itself starts to need more of them.
double mixed_switch_gnorm(const Slot& s) {
  return s.opts.mixed_gtol > 0.0 ? s.opts.mixed_gtol : 2e-9 * std::sqrt((double)s.N());
}

int converged_reason(const Slot& s) {
  if (!std::isfinite(s.gnorm) || !std::isfinite(s.f)) return -4;
  if (s.gnorm <= s.opts.gatol) return 1;
  if (s.opts.grtol > 0 && s.gnorm <= s.opts.grtol * std::fabs(s.f)) return 2;
  if (s.opts.gttol > 0 && s.gnorm <= s.opts.gttol * s.gnorm0) return 3;
  return 0;
}

// evaluate f, g at x for the listed slots, each in its own precision; with
// sync = false the objective copy is only enqueued (the caller's next
// synchronising step completes it, then take_f reads it)
int eval_slots(a2c_batch* b, const std::vector<int>& ids, bool sync = true) {
  std::vector<int> m32, m64;
  for (int i : ids) {
    Slot& s = b->slots[i];
    (s.mixed ? m32 : m64).push_back(i);
    s.rep.evals++;
    if (!s.mixed) s.rep.evals_fp64++;
  }
  if (int rc = fused_eval(b, m32, false, lmvm_state, false)) return rc;
  if (int rc = fused_eval(b, m64, true, lmvm_state, false)) return rc;
  enqueue_d2h(b, b->hF, b->dF, sizeof(double) * b->slots.size());
  if (!sync) return A2C_OK;
  if (int rc = bsync(b)) return rc;
  for (int i : ids) b->slots[i].f = b->hF[i];
  return A2C_OK;
}
void take_f(a2c_batch* b, const std::vector<int>& ids) {
  for (int i : ids) b->slots[i].f = b->hF[i];
}

int proj_gnorm(a2c_batch* b, const std::vector<int>& ids) {
  if (int rc = seg_bound<1>(b, ids, [](SegBound& o, Slot& s) {
        o.z = s.gp; o.x = s.x; o.y = s.g; o.lb = s.opts.sigma_lb; o.ub = s.opts.sigma_ub;
      }))
    return rc;
  std::vector<double> v;
  if (int rc = seg_dot(b, ids, [](SegDot& o, Slot& s) { o.a1 = s.gp; o.b1 = s.gp; }, v)) return rc;
  for (size_t k = 0; k < ids.size(); ++k) b->slots[ids[k]].gnorm = std::sqrt(v[2 * k]);
  return A2C_OK;
}

// slots whose mixed phase ended: re-evaluate in FP64 at the same point
int switch_to_fp64(a2c_batch* b, const std::vector<int>& ids) {
  if (ids.empty()) return A2C_OK;
  for (int i : ids) b->slots[i].mixed = false;
  if (int rc = eval_slots(b, ids)) return rc;
  return proj_gnorm(b, ids);
}

// first evaluation of newly started slots
int lmvm_init(a2c_batch* b, const std::vector<int>& ids) {
  if (ids.empty()) return A2C_OK;
  if (int rc = seg_bound<0>(b, ids, [](SegBound& o, Slot& s) {
        o.z = s.x; o.lb = s.opts.sigma_lb; o.ub = s.opts.sigma_ub;
      }))
    return rc;
  if (int rc = eval_slots(b, ids)) return rc;
  if (int rc = proj_gnorm(b, ids)) return rc;
  std::vector<int> sw;
  for (int i : ids) {
    Slot& s = b->slots[i];
    // FP64 from the start when the mixed gradient is already small or the mixed
    // evaluation is non-finite (a pair closer than its tables reach)
    if (s.mixed && s.opts.mixed_switch &&
        (s.gnorm <= mixed_switch_gnorm(s) || !std::isfinite(s.f) || !std::isfinite(s.gnorm)))
      sw.push_back(i);
  }
  if (int rc = switch_to_fp64(b, sw)) return rc;
  for (int i : ids) {
    Slot& s = b->slots[i];
    s.gnorm0 = s.gnorm;
    s.started = true;
    s.rep.reason = converged_reason(s);
    s.rep.f = s.f;
    s.rep.gnorm = s.gnorm;
    if (s.rep.reason != 0) s.running = false;
  }
  return A2C_OK;
}

// d = -H g for the listed slots (two-loop recursion, or scaled steepest
// descent for slots with no pairs or marked steepest)
int directions(a2c_batch* b, const std::vector<int>& ids) {
  std::vector<int> tl, tlvf, sd, small;  // segmented two-loop (per step / vector-free), steepest descent, fused
  int maxp = 0;
  for (int i : ids) {
    Slot& s = b->slots[i];
    if (s.steepest || s.npairs == 0) {
      s.npairs = 0;
      sd.push_back(i);
    } else if (b->fused_two_loop && s.N() <= TL_MAX_SITES) {
      small.push_back(i);
    } else if (b->vector_free && s.npairs <= VF_PAIRS) {
      tlvf.push_back(i);
    } else {
      tl.push_back(i);
      maxp = std::max(maxp, s.npairs);
    }
  }
  // fused recursion: vector-free for up to VF_PAIRS pairs, one pass per step above
  for (int vf = 1; vf >= 0; --vf) {
    std::vector<int> part;
    for (int i : small)
      if ((b->vector_free && b->slots[i].npairs <= VF_PAIRS) == (vf == 1)) part.push_back(i);
    if (part.empty()) continue;
    std::vector<TwoLoop> ops(part.size());
    for (size_t k = 0; k < part.size(); ++k) {
      Slot& s = b->slots[part[k]];
      TwoLoop& o = ops[k];
      o.g = s.g; o.q = s.q; o.d = s.d; o.x = s.x;
      o.out = b->dres + 2 * part[k];
      for (int t = 0; t < s.npairs; ++t) {
        const int p = (s.head - 1 - t + s.m) % s.m;
        o.S[t] = s.S[p]; o.Y[t] = s.Y[p]; o.rho[t] = s.rho[p];
      }
      o.met = s.metric();
      o.lb = s.opts.sigma_lb; o.ub = s.opts.sigma_ub;
      o.np = s.npairs; o.nsite = s.N(); o.npd = s.fe->desc.n_pd;
    }
    const size_t need = (ops.size() * sizeof(TwoLoop) + 255) & ~(size_t)255;
    if (b->arena.used + need > b->arena.cap)
      if (int rc = bsync(b)) return rc;
    const TwoLoop* dops = b->arena.push(ops.data(), ops.size());
    if (!dops) return berr(b, A2C_ENOMEM, "staging arena too small");
    const unsigned nops = (unsigned)ops.size();
    enqueue(b, {6, (uint64_t)vf, u64(dops), nops}, [=](hipStream_t st) {
      if (vf) hipLaunchKernelGGL(k_two_loop_vf, dim3(nops), dim3(TL_THREADS), 0, st, dops);
      else hipLaunchKernelGGL(k_two_loop, dim3(nops), dim3(TL_THREADS), 0, st, dops);
    });
  }
  std::vector<double> v;
  auto scal = [b](const Slot& s) { return b->dscal + SCAL_PER_SLOT * (&s - b->slots.data()); };
  static_assert(VF_SUMS + VF_PAIRS <= SCAL_PER_SLOT, "device scalars of the vector-free recursion");
  if (int rc = seg_vf(b, tlvf, [&](SegVF& o, Slot& s) {
        o.g = s.g; o.q = s.q; o.d = s.d; o.x = s.x;
        for (int t = 0; t < s.npairs; ++t) {
          const int p = (s.head - 1 - t + s.m) % s.m;
          o.S[t] = s.S[p]; o.Y[t] = s.Y[p]; o.rho[t] = s.rho[p];
        }
        o.np = s.npairs; o.npd = s.fe->desc.n_pd; o.met = s.metric();
        o.lb = s.opts.sigma_lb; o.ub = s.opts.sigma_ub;
        o.sums = scal(s); o.out = b->dres + 2 * (&s - b->slots.data());
      }))
    return rc;
  if (!tl.empty()) {
    // Two-loop recursion with every dot product and coefficient on the
    // device (no host round trip inside the recursion): pass 0, then
    // passes 1..np of the first loop and np+1..2np of the second; a slot
    // with fewer pairs than the batch's largest is done after its 2np.
    auto pidx = [](const Slot& s, int t) { return (s.head - 1 - t + s.m) % s.m; };
    if (int rc = seg_tl(b, tl, true, [&](SegTL& o, Slot& s) {
          o.q = s.q; o.v = s.g; o.w = s.S[pidx(s, 0)]; o.y0 = s.Y[pidx(s, 0)];
          o.met = s.metric(); o.npd = s.fe->desc.n_pd;
          o.out = scal(s) + pidx(s, 0); o.out2 = scal(s) + s.m + 5;
        }))
      return rc;
    for (int p = 1; p <= 2 * maxp; ++p) {
      std::vector<int> sub;
      for (int i : tl)
        if (p <= 2 * b->slots[i].npairs) sub.push_back(i);
      if (int rc = seg_tl(b, sub, false, [&](SegTL& o, Slot& s) {
            const int np = s.npairs;
            o.q = s.q;
            if (p <= np) {  // first loop, newest to oldest: q -= rho_t A_t y_t, A_t = s_t . q
              const int t = p - 1;
              o.v = s.Y[pidx(s, t)]; o.rho = -s.rho[pidx(s, t)]; o.ca = scal(s) + pidx(s, t);
              if (t + 1 < np) { o.w = s.S[pidx(s, t + 1)]; o.out = scal(s) + pidx(s, t + 1); }
              else {  // then q = diag(gamma_g) D q; B = y_{np-1} . q
                o.mode = 1; o.gscal = scal(s) + s.m + 5; o.met = s.metric(); o.npd = s.fe->desc.n_pd;
                o.w = s.Y[pidx(s, np - 1)]; o.out = scal(s) + s.m + 2;
              }
            } else {  // second loop, oldest to newest: q += rho_t (A_t - B_t) s_t, B_t = y_t . q
              const int t = 2 * np - p;
              o.v = s.S[pidx(s, t)]; o.rho = s.rho[pidx(s, t)]; o.ca = scal(s) + pidx(s, t); o.cb = scal(s) + s.m + 2;
              if (t > 0) { o.w = s.Y[pidx(s, t - 1)]; o.out = scal(s) + s.m + 2; }
              else {  // d = -q, widths at an active bound not moved; g . d
                o.mode = 2; o.d = s.d; o.g = s.g; o.x = s.x; o.lb = s.opts.sigma_lb; o.ub = s.opts.sigma_ub;
                o.out = b->dres + 2 * (&s - b->slots.data());
              }
            }
          }))
        return rc;
    }
  }
  if (!sd.empty()) {
    // scaled steepest descent d = -delta D g: TAO's initial scale
    // 2|f| / g.D g, capped so that the first trial moves no coordinate
    // group by more than 0.05 A in the 2-norm (both on the device)
    if (int rc = seg_dot_dev(b, sd, [&](SegDot& o, Slot& s) {
          o.a1 = s.g; o.b1 = s.g; o.a2 = s.g; o.b2 = s.g; o.met = s.metric(); o.k1 = 1; o.k2 = 2;
          o.out = scal(s) + s.m + 3;
        }))
      return rc;
    if (int rc = seg_axpby(b, sd, [&](SegAx& o, Slot& s) {
          o.z = s.d; o.x = s.g; o.y = s.g; o.a = s.f != 0.0 ? 2.0 * std::fabs(s.f) : 2.0; o.amode = 3;
          o.p1 = scal(s) + s.m + 3; o.p2 = scal(s) + s.m + 4; o.b = 0.0; o.met = s.metric(); o.scaled = 1;
        }))
      return rc;
  }
  if (int rc = seg_bound<2>(b, sd, [](SegBound& o, Slot& s) {
        o.z = s.d; o.x = s.x; o.lb = s.opts.sigma_lb; o.ub = s.opts.sigma_ub;
      }))
    return rc;
  if (!sd.empty()) {  // its copy of the results also brings the two-loop slots' g . d
    if (int rc = seg_dot(b, sd, [](SegDot& o, Slot& s) { o.a1 = s.g; o.b1 = s.d; }, v)) return rc;
    for (size_t k = 0; k < sd.size(); ++k) b->slots[sd[k]].gd = v[2 * k];
  } else {
    enqueue_d2h(b, b->hres, b->dres, sizeof(double) * 2 * b->slots.size());
    if (int rc = bsync(b)) return rc;
  }
  for (int i : small) b->slots[i].gd = b->hres[2 * i];
  for (int i : tl) b->slots[i].gd = b->hres[2 * i];
  for (int i : tlvf) b->slots[i].gd = b->hres[2 * i];
  return A2C_OK;
}

// One LMVM iteration for every running, started slot.
int lmvm_iteration(a2c_batch* b, std::vector<int>& act) {
  for (int i : act) b->slots[i].steepest = false;
  if (int rc = directions(b, act)) return rc;
  // retry non-descent directions with steepest descent; give up if still bad
  std::vector<int> redo, go;
  for (int i : act) (b->slots[i].gd < 0.0 ? go : redo).push_back(i);
  if (!redo.empty()) {
    for (int i : redo) b->slots[i].steepest = true;
    if (int rc = directions(b, redo)) return rc;
    for (int i : redo) {
      Slot& s = b->slots[i];
      if (s.gd < 0.0) go.push_back(i);
      else { s.rep.reason = -5; s.running = false; }
    }
  }
  std::sort(go.begin(), go.end());
  if (go.empty()) return A2C_OK;
  // line search setup (projected search: trial points are clipped onto the
  // sigma bounds, and the slope is taken along the projected step)
  for (size_t k = 0; k < go.size(); ++k) {
    Slot& s = b->slots[go[k]];
    // the iteration's start point and gradient become x_old, g_old by a swap
    // of buffer roles (every trial overwrites all of x and g)
    std::swap(s.x, s.xo);
    std::swap(s.g, s.go);
    s.fold = s.f;
    s.ls = MoreThuente();
    s.ls.stpmax = 1e20;
    // approximate Wolfe: f may rise by epsf when the slope conditions hold --
    // about 20x the rounding noise of the objective (FP64 ~1e-13 |f|, mixed
    // ~5e-11 |f|, measured), so searches near the noise floor still progress
    s.ls.epsf = (s.mixed ? 1e-9 : 1e-11) * std::fabs(s.f);
    s.stp = 1.0;
    s.ls.start(s.f, s.gd, s.stp);
    s.nfev = 0;
    s.lsstat = 0;
  }
  std::vector<int> srch = go;
  std::vector<double> v;
  while (!srch.empty()) {
    std::vector<int> run, cap;
    for (int i : srch) (b->slots[i].rep.evals >= b->slots[i].opts.max_funcs ? cap : run).push_back(i);
    for (int i : cap) b->slots[i].lsstat = -1;
    if (run.empty()) break;
    if (int rc = seg_bound<3>(b, run, [](SegBound& o, Slot& s) {
          o.z = s.x; o.x = s.xo; o.y = s.d; o.w = s.q;
          o.lb = s.opts.sigma_lb; o.ub = s.opts.sigma_ub; o.stp = s.stp;
        }))
      return rc;
    if (int rc = eval_slots(b, run, false)) return rc;  // f and the slope come back in one round trip
    if (int rc = seg_dot(b, run, [](SegDot& o, Slot& s) { o.a1 = s.g; o.b1 = s.q; }, v)) return rc;
    take_f(b, run);
    std::vector<int> next;
    for (size_t k = 0; k < run.size(); ++k) {
      Slot& s = b->slots[run[k]];
      s.nfev++;
      if (!std::isfinite(s.f)) { s.lsstat = -2; continue; }
      s.lsstat = s.ls.iterate(s.f, v[2 * k], s.stp);  // slope along the realised (projected) direction
      // A mixed search that needs many trials has met the noise of the FP32
      // objective: give up (the slot then continues in FP64 from the
      // iteration's start) instead of spending up to 30 evaluations.
      if (s.lsstat == 0 && s.mixed && s.opts.mixed_switch && s.nfev >= MIXED_MAX_TRIALS) s.lsstat = -3;
      if (s.lsstat == 0 && s.nfev >= 30) s.lsstat = 2;
      if (s.lsstat == 0) next.push_back(run[k]);
    }
    srch = next;
  }
  // outcomes
  std::vector<int> ok, bad, sw;
  for (int i : go) {
    Slot& s = b->slots[i];
    const bool decreased = (s.f < s.fold) || s.lsstat == 1;
    (s.lsstat >= 0 && decreased ? ok : bad).push_back(i);
  }
  if (!bad.empty()) {
    for (int i : bad) {
      Slot& s = b->slots[i];
      std::swap(s.x, s.xo);  // back to the iteration's start
      std::swap(s.g, s.go);
      s.f = s.fold;
      if (s.lsstat == -1) { s.rep.reason = -3; s.running = false; }
      else if (s.mixed && s.opts.mixed_switch) {  // incl. non-finite mixed trials
        sw.push_back(i);
        if (s.lsstat != -3) s.npairs = 0;  // a search stopped by the trial cap keeps its curvature pairs
      }
      else if (s.lsstat == -2) { s.rep.reason = -4; s.running = false; }
      else if (s.npairs == 0) { s.rep.reason = -6; s.running = false; }
      else s.npairs = 0;
    }
  }
  if (int rc = switch_to_fp64(b, sw)) return rc;
  for (int i : sw) {
    Slot& s = b->slots[i];
    s.rep.reason = converged_reason(s);
    if (s.rep.reason) s.running = false;
  }
  if (!ok.empty()) {
    // the new pair, s.y and the projected-gradient norm in one round trip
    if (int rc = seg_pair(b, ok, v)) return rc;
    for (size_t k = 0; k < ok.size(); ++k) {
      Slot& s = b->slots[ok[k]];
      s.gnorm = std::sqrt(v[2 * k + 1]);
      if (v[2 * k] > 1e-300) {  // keep only pairs with positive curvature
        s.rho[s.head] = 1.0 / v[2 * k];
        s.head = (s.head + 1) % s.m;
        s.npairs = std::min(s.npairs + 1, s.m);
      } else if (s.npairs == s.m) {
        // k_seg_pair wrote the rejected s, y over S[head], Y[head], which with
        // a full history held the oldest pair still in use: that pair is gone,
        // so the recursion stops one short of it (the next accepted pair
        // fills the slot again)
        s.npairs = s.m - 1;
      }
    }
    std::vector<int> sw2;
    for (int i : ok) {
      Slot& s = b->slots[i];
      s.rep.iterations++;
      if (s.mixed && s.opts.mixed_switch && s.gnorm <= mixed_switch_gnorm(s)) sw2.push_back(i);
    }
    if (int rc = switch_to_fp64(b, sw2)) return rc;
    for (int i : ok) {
      Slot& s = b->slots[i];
      s.rep.reason = converged_reason(s);
      if (s.rep.reason == 0 && s.rep.iterations >= s.opts.max_it) s.rep.reason = -2;
      if (s.rep.reason) s.running = false;
    }
  }
  for (int i : act) {
    Slot& s = b->slots[i];
    s.rep.f = s.f;
    s.rep.gnorm = s.gnorm;
  }
  return A2C_OK;
}

}  // namespace

// Advance every running slot (or, with a filter, the running slots it
// selects) by up to max_iters iterations.
int batch_step(a2c_batch* b, int max_iters, int* n_running, const std::vector<char>* only = nullptr) {
  auto chosen = [&](size_t i) { return !only || (*only)[i]; };
  std::vector<int> init;
  for (size_t i = 0; i < b->slots.size(); ++i)
    if (chosen(i) && b->slots[i].used && b->slots[i].running && !b->slots[i].started) init.push_back((int)i);
  if (int rc = lmvm_init(b, init)) return rc;
  for (int it = 0; it < max_iters; ++it) {
    std::vector<int> act;
    for (size_t i = 0; i < b->slots.size(); ++i) {
      Slot& s = b->slots[i];
      if (chosen(i) && s.used && s.running && s.started) {
        if (s.rep.iterations >= s.opts.max_it) { s.rep.reason = -2; s.running = false; continue; }
        act.push_back((int)i);
      }
    }
    if (act.empty()) break;
    if (int rc = lmvm_iteration(b, act)) return rc;
  }
  int nr = 0;
  for (auto& s : b->slots) nr += (s.used && s.running) ? 1 : 0;
  if (n_running) *n_running = nr;
  return A2C_OK;
}

int batch_eval_slots(a2c_batch* b, const std::vector<int>& ids, bool fp64) {
  return fused_eval(b, ids, fp64, [](Slot& s, const double** X, double** G) {
    *X = s.fe->dX;
    *G = s.fe->dG;
  });
}

}  // namespace a2c

// =========================================================================
// C-ABI
// =========================================================================
using namespace a2c;

extern "C" {

void a2c_batch_desc_default(a2c_batch_desc* d) {
  std::memset(d, 0, sizeof(*d));
  d->device = 0;
  d->precision = A2C_PREC_MIXED;
  d->deriv_mode = A2C_DERIV_ANALYTIC;
  d->fd_eps = 1e-5;
  d->max_lattices = 64;
  d->hbm_budget_bytes = (size_t)200 << 30;
  d->max_neighbors = NBR_CAP;
  const char* g = std::getenv("A2C_GRAPHS");
  d->graphs = (g && std::string(g) == "1") ? 1 : 0;
}

int a2c_batch_create(const a2c_batch_desc* desc, a2c_batch** out) {
  if (!desc || !out || desc->max_lattices <= 0 || desc->max_lattices > (1 << 20)) return A2C_EINVAL;
  if (desc->precision != A2C_PREC_FP64 && desc->precision != A2C_PREC_MIXED) return A2C_EINVAL;
  if (desc->graphs != 0 && desc->graphs != 1) return A2C_EINVAL;
  *out = nullptr;
  a2c_batch* b = new (std::nothrow) a2c_batch;
  if (!b) return A2C_ENOMEM;
  b->desc = *desc;
  // A2C_TWO_LOOP: "segmented" sends every slot through the segmented kernels,
  // "perstep" every recursion through the pass-per-step form (both may be given,
  // comma-separated)
  const char* tl = std::getenv("A2C_TWO_LOOP");
  const std::string tls = tl ? tl : "";
  b->fused_two_loop = tls.find("segmented") == std::string::npos;
  b->vector_free = tls.find("perstep") == std::string::npos;
  auto fail = [&](int rc) { a2c_batch_destroy(b); return rc; };
  if (hipSetDevice(desc->device) != hipSuccess) return fail(A2C_EHIP);
  if (hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking) != hipSuccess) return fail(A2C_EHIP);
  const size_t cap = desc->max_lattices;
  b->slots.resize(cap);
  b->arena.cap = (size_t)32 << 20;
  bool ok = hmalloc(&b->arena.h, b->arena.cap) == hipSuccess && dmalloc(&b->arena.d, b->arena.cap) == hipSuccess &&
            dmalloc(&b->dF, 8 * cap) == hipSuccess && hmalloc(&b->hF, 8 * cap) == hipSuccess &&
            dmalloc(&b->dres, 16 * cap) == hipSuccess && hmalloc(&b->hres, 16 * cap) == hipSuccess &&
            dmalloc(&b->dscal, sizeof(double) * SCAL_PER_SLOT * cap) == hipSuccess;
  if (!ok) return fail(A2C_ENOMEM);
  DeviceTables tabs;
  if (device_tables(desc->device, &tabs) != hipSuccess) return fail(A2C_EHIP);
  b->dtab = tabs.t32;  // shared, not freed with the batch
  b->dtab64 = tabs.t64;
  *out = b;
  return A2C_OK;
}

void a2c_batch_destroy(a2c_batch* b) {
  if (!b) return;
  if (b->stream) (void)hipStreamSynchronize(b->stream);
  for (auto& e : b->graphs.map) (void)hipGraphExecDestroy(e.second.exec);
  for (auto& s : b->slots) {
    lmvm_free(s);
    if (s.owned) a2c_fe_destroy(s.fe);
  }
  (void)dfree(b->arena.d);
  if (b->arena.h) (void)hfree(b->arena.h);
  (void)dfree(b->dF);
  if (b->hF) (void)hfree(b->hF);
  (void)dfree(b->dres);
  (void)dfree(b->dscal);
  if (b->hres) (void)hfree(b->hres);
  (void)dfree(b->dpart);
  if (b->stream) (void)hipStreamDestroy(b->stream);
  delete b;
}

const char* a2c_batch_last_error(const a2c_batch* b) { return b ? b->err.c_str() : "null batch"; }

size_t a2c_batch_footprint(int32_t n_pd, int32_t n_h, int32_t lmvm_m) {
  const size_t N = (size_t)n_pd + n_h;
  // state/gradient/staging 4N doubles x 4, per-site arrays ~5 doubles, ELL
  // lists ~100 int32 per site, the owned/received halves of two pair lists
  // per site (~28 wide: 2 int32 per owned slot, a 32-byte share per received
  // slot), optimizer (7 + 2m) x 4N doubles; plus a quarter for the device
  // cache's size classes (a block is rounded up by at most a quarter)
  const size_t raw =
      N * (4 * 32 + 40 + 100 * 4 + 2 * 28 * (2 * 4 + 32) + (7 + 2 * (size_t)std::max(lmvm_m, 1)) * 32) +
      ((size_t)1 << 20);
  return raw + raw / 4;
}

size_t a2c_batch_mem_used(const a2c_batch* b) { return b ? b->mem_used : 0; }

int a2c_batch_graph_stats(const a2c_batch* b, int64_t* replays, int64_t* captures, int64_t* direct) {
  if (!b) return A2C_EINVAL;
  if (replays) *replays = b->graphs.replays;
  if (captures) *captures = b->graphs.captures;
  if (direct) *direct = b->graphs.direct;
  return A2C_OK;
}

int a2c_batch_add(a2c_batch* b, const a2c_fe_desc* lat, const double* X, const double* x, int32_t* slot) {
  if (!b) return A2C_EINVAL;
  b->err.clear();
  if (!lat || !X || !slot || (lat->n_h > 0 && !x)) return berr(b, A2C_EINVAL, "null argument");
  int free_slot = -1;
  for (size_t i = 0; i < b->slots.size(); ++i)
    if (!b->slots[i].used) { free_slot = (int)i; break; }
  if (free_slot < 0) return berr(b, A2C_ENOMEM, "no free slot");
  const size_t need = a2c_batch_footprint(lat->n_pd, lat->n_h, 5);
  if (b->mem_used + need > b->desc.hbm_budget_bytes) return berr(b, A2C_ENOMEM, "HBM budget exceeded");
  a2c_fe_desc d = *lat;
  d.device = b->desc.device;
  d.precision = b->desc.precision;
  d.deriv_mode = b->desc.deriv_mode;
  d.fd_eps = b->desc.fd_eps;
  if (d.max_neighbors == 0) d.max_neighbors = b->desc.max_neighbors;
  a2c_fe_ctx* fe = nullptr;
  int rc = fe_create(&d, b->stream, &fe);
  if (rc) return berr(b, rc, "lattice context creation failed");
  if ((rc = a2c_fe_set_occupancy(fe, x)) || (rc = a2c_fe_build_neighbors(fe, X, nullptr))) {
    b->err = a2c_fe_last_error(fe);
    a2c_fe_destroy(fe);
    return rc;
  }
  Slot& s = b->slots[free_slot];
  s = Slot();
  s.fe = fe;
  s.owned = true;
  s.used = true;
  s.charged = need;
  b->mem_used += need;
  *slot = free_slot;
  return A2C_OK;
}

int a2c_batch_remove(a2c_batch* b, int32_t slot) {
  if (!b) return A2C_EINVAL;
  if (slot < 0 || slot >= (int)b->slots.size() || !b->slots[slot].used) return berr(b, A2C_EINVAL, "bad slot");
  if (int rc = bsync(b)) return rc;
  Slot& s = b->slots[slot];
  b->mem_used -= std::min(b->mem_used, s.charged);
  lmvm_free(s);
  if (s.owned) a2c_fe_destroy(s.fe);
  s = Slot();
  return A2C_OK;
}

int a2c_batch_set_state(a2c_batch* b, int32_t slot, const double* X) {
  if (!b || slot < 0 || slot >= (int)b->slots.size() || !b->slots[slot].used || !X) return A2C_EINVAL;
  if (int rc = bsync(b)) return rc;  // no launch may still read the state
  Slot& s = b->slots[slot];
  B_HIP(b, hipMemcpy(s.fe->dX, X, 32 * (size_t)s.N(), hipMemcpyHostToDevice));
  return A2C_OK;
}

int a2c_batch_get_state(a2c_batch* b, int32_t slot, double* X, double* G) {
  if (!b || slot < 0 || slot >= (int)b->slots.size() || !b->slots[slot].used) return A2C_EINVAL;
  if (int rc = bsync(b)) return rc;
  Slot& s = b->slots[slot];
  const double* src = s.x ? s.x : s.fe->dX;
  const double* gsrc = s.x ? s.g : s.fe->dG;
  if (X) B_HIP(b, hipMemcpy(X, src, 32 * (size_t)s.N(), hipMemcpyDeviceToHost));
  if (G) B_HIP(b, hipMemcpy(G, gsrc, 32 * (size_t)s.N(), hipMemcpyDeviceToHost));
  return A2C_OK;
}

int a2c_batch_eval(a2c_batch* b, int32_t n, const int32_t* slots, int32_t precision, double* f) {
  if (!b || n < 0 || (n > 0 && (!slots || !f))) return A2C_EINVAL;
  b->err.clear();
  std::vector<int> ids(slots, slots + n);
  for (int i : ids)
    if (i < 0 || i >= (int)b->slots.size() || !b->slots[i].used) return berr(b, A2C_EINVAL, "bad slot");
  if (int rc = batch_eval_slots(b, ids, precision != A2C_PREC_MIXED)) return rc;
  for (int k = 0; k < n; ++k) f[k] = b->hF[ids[k]];
  return A2C_OK;
}

int a2c_batch_lmvm_start(a2c_batch* b, int32_t slot, const a2c_lmvm_opts* o) {
  if (!b || !o) return A2C_EINVAL;
  b->err.clear();
  if (slot < 0 || slot >= (int)b->slots.size() || !b->slots[slot].used) return berr(b, A2C_EINVAL, "bad slot");
  if (o->m < 1 || o->m > 64 || o->max_it < 0 || o->max_funcs < 1 || !(o->sigma_lb >= 0) ||
      !(o->sigma_ub > o->sigma_lb) || (o->precond != 0 && o->precond != 1) || !(o->mixed_gtol >= 0.0))
    return berr(b, A2C_EINVAL, "bad options");
  Slot& s = b->slots[slot];
  if (int rc = bsync(b)) return rc;  // queued launches may still use the buffers lmvm_alloc replaces
  // admission charged an optimizer of 5 pairs; charge the one asked for
  if (s.owned) {
    const size_t want = a2c_batch_footprint(s.fe->desc.n_pd, s.fe->desc.n_h, o->m);
    if (want > s.charged && b->mem_used - s.charged + want > b->desc.hbm_budget_bytes)
      return berr(b, A2C_ENOMEM, "HBM budget exceeded by the optimizer's memory");
    b->mem_used = b->mem_used - s.charged + want;
    s.charged = want;
  }
  s.opts = *o;
  s.m = o->m;
  if (int rc = lmvm_alloc(b, s)) return rc;
  B_HIP(b, hipMemcpyAsync(s.x, s.fe->dX, 32 * (size_t)s.N(), hipMemcpyDeviceToDevice, b->stream));
  s.rep = a2c_lmvm_report{};
  s.npairs = s.head = 0;
  s.mixed = s.fe->desc.precision == A2C_PREC_MIXED;
  s.running = true;
  s.started = false;
  return A2C_OK;
}

int a2c_batch_lmvm_step(a2c_batch* b, int32_t max_iters, int32_t* n_running) {
  if (!b || max_iters < 0) return A2C_EINVAL;
  b->err.clear();
  return batch_step(b, max_iters, n_running);
}

int a2c_batch_lmvm_step_slots(a2c_batch* b, int32_t n, const int32_t* slots, int32_t max_iters,
                              int32_t* n_running) {
  if (!b || n < 0 || (n > 0 && !slots) || max_iters < 0) return A2C_EINVAL;
  b->err.clear();
  std::vector<char> only(b->slots.size(), 0);
  for (int k = 0; k < n; ++k) {
    if (slots[k] < 0 || slots[k] >= (int)b->slots.size() || !b->slots[slots[k]].used)
      return berr(b, A2C_EINVAL, "bad slot");
    only[slots[k]] = 1;
  }
  return batch_step(b, max_iters, n_running, &only);
}

int a2c_batch_lmvm_report(a2c_batch* b, int32_t slot, a2c_lmvm_report* rep) {
  if (!b || !rep || slot < 0 || slot >= (int)b->slots.size() || !b->slots[slot].used) return A2C_EINVAL;
  *rep = b->slots[slot].rep;
  return A2C_OK;
}

}  // extern "C"

// Single-lattice minimizer on the same code: a batch of one wrapping ctx.
namespace a2c {
int lmvm_minimize(a2c_fe_ctx* c, const double* X0, const a2c_lmvm_opts* o, double* Xout, a2c_lmvm_report* rep) {
  a2c_batch_desc bd;
  a2c_batch_desc_default(&bd);
  bd.device = c->desc.device;
  bd.precision = c->desc.precision;
  bd.deriv_mode = c->desc.deriv_mode;
  bd.fd_eps = c->desc.fd_eps;
  bd.max_lattices = 1;
  // the context's stream may still run work that uses the buffers the
  // optimizer shares with it (state, scratch of the evaluation)
  A2C_HIP(c, hipStreamSynchronize(c->stream));
  a2c_batch* b = nullptr;
  if (int rc = a2c_batch_create(&bd, &b)) return set_err(c, rc, "batch creation failed");
  struct Guard {
    a2c_batch* b;
    ~Guard() { a2c_batch_destroy(b); }
  } guard{b};
  if (c->desc.precision == A2C_PREC_MIXED && !c->daux)
    if (int rc = build_tables(c)) return rc;
  Slot& s = b->slots[0];
  s.fe = c;
  s.owned = false;
  s.used = true;
  A2C_HIP(c, hipMemcpy(c->dX, X0, 32 * (size_t)c->N, hipMemcpyHostToDevice));
  c->lat_dirty = true;
  int rc = a2c_batch_lmvm_start(b, 0, o);
  if (!rc) {
    int nr = 1;
    while (nr > 0 && !rc) rc = batch_step(b, 1 << 20, &nr);
  }
  if (rc) return set_err(c, rc, b->err);
  *rep = s.rep;
  A2C_HIP(c, hipMemcpy(Xout, s.x, 32 * (size_t)c->N, hipMemcpyDeviceToHost));
  return A2C_OK;
}
}  // namespace a2c

extern "C" void a2c_lmvm_opts_default(a2c_lmvm_opts* o) {
  o->m = 5;            // -tao_lmm_vectors
  o->max_it = 2000;    // -tao_max_it
  o->max_funcs = 4000; // -tao_max_funcs
  o->gatol = 1e-8;
  o->grtol = 1e-8;
  o->gttol = 0.0;
  o->sigma_lb = 1e-6;  // formVariableBounds, minimizer.cpp:815,822
  o->mixed_switch = 1;
  o->mixed_gtol = 0.0;  // automatic: 2e-9 sqrt(N)
  o->sigma_ub = HUGE_VAL;  // formVariableBounds leaves sigma unbounded above
  o->precond = 1;
}

extern "C" int a2c_fe_minimize(a2c_fe_ctx* c, const double* X0, const a2c_lmvm_opts* opts, double* X_out,
                               a2c_lmvm_report* report) {
  if (!c) return A2C_EINVAL;
  c->err.clear();
  if (!X0 || !X_out || !opts || !report) return a2c::set_err(c, A2C_EINVAL, "null argument");
  if (!c->have_nbr || !c->have_x) return a2c::set_err(c, A2C_ESTATE, "neighbors/occupancy not set");
  if (opts->m < 1 || opts->max_it < 0 || opts->max_funcs < 1 || !(opts->sigma_lb >= 0) ||
      !(opts->sigma_ub > opts->sigma_lb) || (opts->precond != 0 && opts->precond != 1))
    return a2c::set_err(c, A2C_EINVAL, "bad options");
  return a2c::lmvm_minimize(c, X0, opts, X_out, report);
}
