// FP64 objective + gradient: the plain kernel of
// Minimizer::formFunctionAndGradient (minimizer.cpp:83-369).
//
//   k_density64  phase I  (:140-157; electrondensity.cpp:12-56): rho per
//                site, then F(rho), F'(rho) once per site.
//   k_grad64     phase II (:162-351): pair terms and gradient per site, the
//                per-site objective share, and a fixed-shape block reduction.
//   k_final_sum  per lattice, a fixed-order sum of its block partials -> f.
//
// Analytic mode evaluates each unordered pair once (k_dens_own64 ->
// k_dens_recv64 -> k_grad_own64 -> k_grad_recv64): the pair's owner computes
// both ends' shares and writes the partner's share into the partner's
// received slot, as the mixed path does.  These two owner kernels take EAM
// functions from FP64 polynomial tables in LDS (on 0.05 A intervals,
// interpolating the long-double closed forms): the exponential functions
// (f_H, phi_PdH, phi_H) at degree 8, within 1e-13 of the closed forms in
// value and slope, instead of three FP64 exp per quadrature point; the Pd
// product polynomials f_Pd (density pass) and phi_Pd (gradient pass) at
// degree 11, within the ~2e-11 rounding noise of the FP64 evaluation of
// those polynomials themselves.  f_Pd' keeps its closed form.  1/|r| comes
// from the hardware estimate refined by two Newton steps.  FD mode keeps the per-site
// full-list loop nest of the reference (every pair from both ends), since
// its derivatives are the reference's central differences per site.
//
// One lane per site over ELL rows: a wave holds 64 consecutive sites of one
// species, reads slot t of their rows as one coalesced load, and keeps the
// per-site sums in registers (no atomics, so results are bitwise
// reproducible).  Blocks map to lattices through b2l, so one launch serves a
// batch of uneven lattices.  The 12-point rule of quadrature.cpp:43-75 is used
// in its structured form: point (c, +-) of site i displaces r along axis c by
// +-sqrt(6) sigma_i, point (c, +-) of site j by -+sqrt(6) sigma_j.  The
// Gauss-Hermite rule (A2C_QUAD_GH3) has its own instances of the pair
// kernels; each instance skips the lattices of the other rule, and the
// per-site kernels apply the rule's factors (1/12 and sqrt(6), or 1).
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <string>
#include <type_traits>
#include <vector>

#include "eam64.hpp"
#include "fe_internal.hpp"

namespace a2c {
namespace {

constexpr int TB = SITE_BLOCK;
constexpr double QW = 1.0 / 12.0;             // 0.5/dim, quadrature.cpp:39
constexpr double SQ6 = 2.449489742783178098;  // sqrt(2)*sqrt(dim/2)
constexpr int GH3 = A2C_QUAD_GH3;
// Factors the per-site kernels apply to the pair sums: the 12-point sums are
// unweighted (weight 1/12) and their sigma terms carry r . e_c, a factor
// sqrt(6) short of d/dsigma; the Gauss-Hermite sums are complete.
template <int RULE> constexpr double rule_qw() { return RULE == GH3 ? 1.0 : QW; }
template <int RULE> constexpr double rule_sq() { return RULE == GH3 ? 1.0 : SQ6; }

struct Site {
  double q[3];
  double s;
};

__device__ __forceinline__ Site load_site(const double* __restrict__ X, int s) {
  const double4 v = *reinterpret_cast<const double4*>(X + 4 * (size_t)s);
  return {{v.x, v.y, v.z}, v.w};
}

// d = q_i - image(q_j): periodic axes shift q_j by L*rint((q_j - q_i)/L).  A
// listed pair is within rc < L/2 of an image, far from a rounding tie, so
// (q_j - q_i)*(1/L) picks the same image as the quotient at a fraction of
// the cost of an FP64 divide.
__device__ __forceinline__ void pair_vec(const LatDev& S, const Site& a, const Site& b, double d[3]) {
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    double qj = b.q[c];
    if (S.periodic[c]) qj = qj - S.box[c] * rint((qj - a.q[c]) * S.ibox[c]);
    d[c] = a.q[c] - qj;
  }
}
// the same with the geometry held in registers (pair-owner loops)
__device__ __forceinline__ void pair_vec(const Image& g, const Site& a, const Site& b, double d[3]) {
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    double qj = b.q[c];
    if (g.per[c]) qj = qj - g.box[c] * rint((qj - a.q[c]) * g.ibox[c]);
    d[c] = a.q[c] - qj;
  }
}
__device__ __forceinline__ Site load_site(const A2C_GLOBAL double4* __restrict__ X, int s) {
  const double4 v = X[s];
  return {{v.x, v.y, v.z}, v.w};
}

// ---------------------------------------------------------------- FP64 tables
// Position of r in the tables: interval k, local coordinate u = (r - 1)/0.05 - k;
// `in` is false outside [1, 6) A, where the closed forms are evaluated.
struct Tpos {
  int k;
  double u;
  bool in;
};
__device__ __forceinline__ Tpos tpos64(double r) {
  const double x = fma(r, 20.0, -20.0);
  Tpos p;
  p.in = x >= 0.0 && x < (double)T64_N;
  p.k = p.in ? (int)x : 0;
  p.u = x - (double)p.k;
  return p;
}
__device__ __forceinline__ double tab64_v(const double* c, double u) {
  double v = c[T64_D];
#pragma unroll
  for (int j = T64_D - 1; j >= 0; --j) v = fma(v, u, c[j]);
  return v;
}
__device__ __forceinline__ eam::VD tab64_vd(const double* c, double u) {
  double v = c[T64_D], d = 0.0;
#pragma unroll
  for (int j = T64_D - 1; j >= 0; --j) {
    d = fma(d, u, v);
    v = fma(v, u, c[j]);
  }
  return {v, 20.0 * d};  // d/dr = 20 d/du
}
__device__ __forceinline__ const double* trow(const double* t, int f, int k) {
  return t + ((size_t)f * T64_N + k) * T64_P;
}
// rows of the Pd polynomials (second part of the tables)
__device__ __forceinline__ const double* qrow(const double* t, int f, int k) {
  return t + T64_PART1 + ((size_t)f * T64_N + k) * T64Q_P;
}
template <int D>
__device__ __forceinline__ double tabq_v(const double* c, double u) {
  double v = c[D];
#pragma unroll
  for (int j = D - 1; j >= 0; --j) v = fma(v, u, c[j]);
  return v;
}
template <int D>
__device__ __forceinline__ eam::VD tabq_vd(const double* c, double u) {
  double v = c[D], d = 0.0;
#pragma unroll
  for (int j = D - 1; j >= 0; --j) {
    d = fma(d, u, v);
    v = fma(v, u, c[j]);
  }
  return {v, 20.0 * d};
}
__device__ __forceinline__ void stage64(double* lds, const double* __restrict__ g, int n) {
  for (int i = threadIdx.x; i < n; i += blockDim.x) lds[i] = g[i];
  __syncthreads();
}
// 1/sqrt(x) for x in the normal range: the hardware estimate, then two steps
// of Newton's iteration y <- y (3 - x y^2) / 2 (quadratic convergence; the
// result is within a few ulp).
__device__ __forceinline__ double rsqrt_nr(double x) {
  double y = __builtin_amdgcn_rsq(x);
  const double h = 0.5 * x;
  y = y * fma(-h * y, y, 1.5);
  y = y * fma(-h * y, y, 1.5);
  return y;
}

template <int DM>
__device__ __forceinline__ double dF(bool pd, double rho, double eps) {
  if (DM == A2C_DERIV_FD) {
    return pd ? (eam::F_Pd(rho + eps) - eam::F_Pd(rho - eps)) / (2.0 * eps)
              : (eam::F_H(rho + eps) - eam::F_H(rho - eps)) / (2.0 * eps);
  }
  return pd ? eam::F_Pd_vd(rho).d : eam::F_H_vd(rho).d;
}

// Point p in 0..11: axis c = (p>>1)%3, sign = p&1 ? -1 : +1, owner = p<6 ? i : j.
__device__ __forceinline__ void point_r(int p, const double d[3], double ai, double aj, double r[3]) {
  const int c = (p >> 1) % 3;
  const double sg = (p & 1) ? -1.0 : 1.0;
  const double disp = (p < 6) ? sg * ai : -sg * aj;
  r[0] = d[0]; r[1] = d[1]; r[2] = d[2];
  r[c] += disp;
}

// Sum of g(|r|) over the nodes of one pair: unweighted over the 12 points, or
// weighted over the 27 Gauss-Hermite nodes (rule_qw completes either).
template <int RULE, class Fn>
__device__ __forceinline__ double pair_quad(const double d[3], double si, double sj, Fn g) {
  double acc = 0.0;
  if (RULE == GH3) {
    const double sd = sqrt(si * si + sj * sj);
#pragma unroll 1
    for (int k = 0; k < 27; ++k) {
      double z[3], w;
      gh_node(k, z, w);
      const double r0 = d[0] + sd * z[0], r1 = d[1] + sd * z[1], r2 = d[2] + sd * z[2];
      acc += w * g(sqrt(r0 * r0 + r1 * r1 + r2 * r2));
    }
  } else {
    const double ai = SQ6 * si, aj = SQ6 * sj;
#pragma unroll
    for (int p = 0; p < 12; ++p) {
      double r[3];
      point_r(p, d, ai, aj, r);
      acc += g(sqrt(r[0] * r[0] + r[1] * r[1] + r[2] * r[2]));
    }
  }
  return acc;
}

// ---------------------------------------------------------------- density
template <int DM, int RULE>
__global__ __launch_bounds__(TB) void k_density64(const LatDev* __restrict__ lats, const int32_t* __restrict__ b2l,
                                                  double eps) {
  const int blk = xcd_block();
  const LatDev& S = lats[b2l[blk]];
  if (S.rule != RULE) return;  // the other rule's instance serves this lattice
  const int s = (blk - S.blk0) * TB + threadIdx.x;
  const int N = S.npd + S.nh;
  if (s >= N) return;
  const double* __restrict__ X = S.X;
  const double* __restrict__ xo = S.x;
  const bool pd = s < S.npd;
  const int row = pd ? s : s - S.npd;
  const Ell LP = pd ? S.pdpd : S.hpd;
  const Ell LH = pd ? S.pdh : S.hh;
  const Site a = load_site(X, s);
  constexpr double qw = rule_qw<RULE>();
  double rho = 0.0;
  const int np = LP.cnt[row];
  for (int t = 0; t < np; ++t) {
    const int j = LP.idx[(size_t)t * LP.rows + row];
    const Site b = load_site(X, j);
    double d[3];
    pair_vec(S, a, b, d);
    rho += pair_quad<RULE>(d, a.s, b.s, [](double r) { return eam::f_Pd(r); }) * qw;
  }
  const int nhn = LH.cnt[row];
  for (int t = 0; t < nhn; ++t) {
    const int h = LH.idx[(size_t)t * LH.rows + row];
    const Site b = load_site(X, S.npd + h);
    double d[3];
    pair_vec(S, a, b, d);
    rho += xo[h] * pair_quad<RULE>(d, a.s, b.s, [](double r) { return eam::f_H(r); }) * qw;
  }
  S.rho[s] = rho;
  const double F = pd ? eam::F_Pd(rho) : eam::F_H(rho);
  S.Fp[s] = dF<DM>(pd, rho, eps);
  S.emb[s] = pd ? F : xo[row] * F;
}

// ---------------------------------------------------------------- gradient
// Pair kinds: 0 Pd-Pd, 1 Pd-H (site Pd), 2 H-Pd (site H), 3 H-H.
template <int DM, int KIND>
__device__ __forceinline__ void pair_point(double r, double Fi, double Fj, double eps, double& phi, double& c) {
  if (DM == A2C_DERIV_FD) {
    auto fd = [&](auto g) { return (g(r + eps) - g(r - eps)) / (2.0 * eps); };
    if (KIND == 0) {
      const double fpd = fd([](double x) { return eam::f_Pd(x); });
      phi = eam::phi_Pd(r);
      c = (Fi + Fj) * fpd + fd([](double x) { return eam::phi_Pd(x); });
    } else if (KIND == 1 || KIND == 2) {
      const double fpd = fd([](double x) { return eam::f_Pd(x); });
      const double fh = fd([](double x) { return eam::f_H(x); });
      phi = eam::phi_morse<0>(r);
      const double dph = fd([](double x) { return eam::phi_morse<0>(x); });
      c = (KIND == 1) ? Fi * fh + Fj * fpd + dph : Fi * fpd + Fj * fh + dph;
    } else {
      const double fh = fd([](double x) { return eam::f_H(x); });
      phi = eam::phi_morse<1>(r);
      c = (Fi + Fj) * fh + fd([](double x) { return eam::phi_morse<1>(x); });
    }
  } else {
    if (KIND == 0) {
      const eam::VD f = eam::f_Pd_vd(r);
      const eam::VD ph = eam::phi_Pd_vd(r, f);
      phi = ph.v;
      c = (Fi + Fj) * f.d + ph.d;
    } else if (KIND == 1 || KIND == 2) {
      const eam::VD f = eam::f_Pd_vd(r);
      const eam::VD h = eam::f_H_vd(r);
      const eam::VD ph = eam::phi_morse_vd<0>(r);
      phi = ph.v;
      c = (KIND == 1) ? Fi * h.d + Fj * f.d + ph.d : Fi * f.d + Fj * h.d + ph.d;
    } else {
      const eam::VD h = eam::f_H_vd(r);
      const eam::VD ph = eam::phi_morse_vd<1>(r);
      phi = ph.v;
      c = (Fi + Fj) * h.d + ph.d;
    }
  }
}

// pair_point (analytic) with the exponential functions from the tables t64;
// `out` is set when r lies outside them (the caller then redoes the pair with
// the closed forms; the values returned here are discarded)
template <int KIND>
__device__ __forceinline__ void pair_point_tab(double r, double Fi, double Fj, const double* t64, double& phi,
                                               double& c, bool& out) {
  const Tpos p = tpos64(r);
  out |= !p.in;
  // f_Pd' stays a closed form here: a second 96-byte row per point makes the
  // pass LDS-bound (measured: 1.83 ms instead of 1.45 ms at 512,000 sites),
  // while phi_Pd from the table keeps the pair energy free of the FP64
  // rounding noise of the product polynomial (the relaxation endgame then
  // needs ~20% fewer FP64 evaluations).
  if (KIND == 0) {
    const eam::VD f = eam::f_Pd_vd(r);
    const eam::VD ph = tabq_vd<T64Q_D>(qrow(t64, T64Q_PHPD, p.k), p.u);
    phi = ph.v;
    c = (Fi + Fj) * f.d + ph.d;
    return;
  }
  const eam::VD h = tab64_vd(trow(t64, T64_FH, p.k), p.u);
  const eam::VD ph = tab64_vd(trow(t64, KIND == 3 ? T64_PHH : T64_PHPDH, p.k), p.u);
  phi = ph.v;
  if (KIND == 3) {
    c = (Fi + Fj) * h.d + ph.d;
  } else {
    const eam::VD f = eam::f_Pd_vd(r);
    c = (KIND == 1) ? Fi * h.d + Fj * f.d + ph.d : Fi * f.d + Fj * h.d + ph.d;
  }
}
// f_H and f_Pd from the tables (values); closed forms outside [1, 6) A
__device__ __forceinline__ double f_H_tab(double r, const double* t64) {
  const Tpos p = tpos64(r);
  return p.in ? tab64_v(trow(t64, T64_FH, p.k), p.u) : eam::f_H(r);
}
__device__ __forceinline__ double f_Pd_tab(double r, const double* t64) {
  const Tpos p = tpos64(r);
  return p.in ? tabq_v<T64Q_D>(qrow(t64, T64Q_FPD, p.k), p.u) : eam::f_Pd(r);
}

// Accumulate one neighbor pair: energy (phi summed over points), q-gradient
// and sigma-gradient (12-point rule: r . xi1 is nonzero only on the six
// points of site i; Gauss-Hermite rule: sum of w (c/|r|) (r . z) sigma_i/s).
template <int DM, int KIND, int RULE>
__device__ __forceinline__ void pair_accum(const double d[3], double si, double sj, double Fi, double Fj,
                                           double eps, double& e, double g[3], double& gs) {
  double pe = 0.0, px = 0.0, py = 0.0, pz = 0.0, ps = 0.0;
  if (RULE == GH3) {
    const double sd = sqrt(si * si + sj * sj);
#pragma unroll 1
    for (int k = 0; k < 27; ++k) {
      double z[3], w;
      gh_node(k, z, w);
      const double r[3] = {d[0] + sd * z[0], d[1] + sd * z[1], d[2] + sd * z[2]};
      const double rn = sqrt(r[0] * r[0] + r[1] * r[1] + r[2] * r[2]);
      double phi, c;
      pair_point<DM, KIND>(rn, Fi, Fj, eps, phi, c);
      pe += w * phi;
      const double u = w * c / rn;
      px += u * r[0];
      py += u * r[1];
      pz += u * r[2];
      ps += u * (r[0] * z[0] + r[1] * z[1] + r[2] * z[2]);
    }
    e += pe;
    g[0] += px; g[1] += py; g[2] += pz;
    gs += ps * (si / sd);
    return;
  }
  const double ai = SQ6 * si, aj = SQ6 * sj;
#pragma unroll 1  // rolled, as in pair_both
  for (int p = 0; p < 12; ++p) {
    double r[3];
    point_r(p, d, ai, aj, r);
    const double rn = sqrt(r[0] * r[0] + r[1] * r[1] + r[2] * r[2]);
    double phi, c;
    pair_point<DM, KIND>(rn, Fi, Fj, eps, phi, c);
    pe += phi;
    const double u = c / rn;
    px += u * r[0];
    py += u * r[1];
    pz += u * r[2];
    if (p < 6) {
      const int ax = (p >> 1) % 3;
      ps += ((p & 1) ? -u : u) * r[ax];  // (c/|r|) r . xi1 / sqrt(3)
    }
  }
  e += pe;
  g[0] += px; g[1] += py; g[2] += pz;
  gs += ps;
}

template <int DM, int RULE>
__global__ __launch_bounds__(TB) void k_grad64(const LatDev* __restrict__ lats, const int32_t* __restrict__ b2l,
                                               double eps) {
  __shared__ double red[TB];
  const int blk = xcd_block();
  const LatDev& S = lats[b2l[blk]];
  if (S.rule != RULE) return;  // whole block: the other rule's instance serves this lattice
  const int lb = blk - S.blk0;
  const int s = lb * TB + threadIdx.x;
  const int N = S.npd + S.nh;
  double fs = 0.0;
  if (s < N) {
    const double* __restrict__ X = S.X;
    const double* __restrict__ xo = S.x;
    const double* __restrict__ Fp = S.Fp;
    const bool pd = s < S.npd;
    const int row = pd ? s : s - S.npd;
    const Ell LP = pd ? S.pdpd : S.hpd;
    const Ell LH = pd ? S.pdh : S.hh;
    const Site a = load_site(X, s);
    const double Fi = Fp[s];
    const double xi = pd ? 1.0 : xo[row];
    double ePd = 0.0, eH = 0.0, g[3] = {0.0, 0.0, 0.0}, gs = 0.0;
    const int np = LP.cnt[row];
    for (int t = 0; t < np; ++t) {
      const int j = LP.idx[(size_t)t * LP.rows + row];
      const Site b = load_site(X, j);
      double d[3];
      pair_vec(S, a, b, d);
      if (pd) pair_accum<DM, 0, RULE>(d, a.s, b.s, Fi, Fp[j], eps, ePd, g, gs);
      else pair_accum<DM, 2, RULE>(d, a.s, b.s, Fi, Fp[j], eps, ePd, g, gs);
    }
    const int nhn = LH.cnt[row];
    for (int t = 0; t < nhn; ++t) {
      const int h = LH.idx[(size_t)t * LH.rows + row];
      const Site b = load_site(X, S.npd + h);
      double d[3];
      pair_vec(S, a, b, d);
      const double xh = xo[h];
      double e1 = 0.0, g1[3] = {0.0, 0.0, 0.0}, gs1 = 0.0;
      if (pd) pair_accum<DM, 1, RULE>(d, a.s, b.s, Fi, Fp[S.npd + h], eps, e1, g1, gs1);
      else pair_accum<DM, 3, RULE>(d, a.s, b.s, Fi, Fp[S.npd + h], eps, e1, g1, gs1);
      eH += xh * e1;
      g[0] += xh * g1[0]; g[1] += xh * g1[1]; g[2] += xh * g1[2];
      gs += xh * gs1;
    }
    // V share of site s: embedding + half of every pair (minimizer.cpp:169-286)
    constexpr double qw = rule_qw<RULE>(), sq = rule_sq<RULE>();
    const double V = S.emb[s] + 0.5 * qw * xi * (ePd + eH);
    const double T = S.T;
    const double sc = qw * xi / T;
    double* __restrict__ G = S.G;
    G[4 * (size_t)s + 0] = sc * g[0];
    G[4 * (size_t)s + 1] = sc * g[1];
    G[4 * (size_t)s + 2] = sc * g[2];
    G[4 * (size_t)s + 3] = sq * sc * gs - 3.0 * KB * xi / a.s;
    fs = V / T - 3.0 * KB * xi * log(a.s);
  }
  red[threadIdx.x] = fs;
  __syncthreads();
#pragma unroll
  for (int w = TB / 2; w > 0; w >>= 1) {
    if (threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) S.part[lb] = red[0];
}

// ================================================================ pair-once
__device__ __forceinline__ const Ell& list_of(const LatDev& S, int list) {
  return list == 0 ? S.pdpd : (list == 1 ? S.pdh : (list == 2 ? S.hpd : S.hh));
}
// partner list: PdPd<->PdPd, PdH<->HPd, HH<->HH
__device__ __forceinline__ const Ell& partner_list(const LatDev& S, int list) {
  return list == 0 ? S.pdpd : (list == 1 ? S.hpd : (list == 2 ? S.pdh : S.hh));
}

template <int RULE>
__global__ __launch_bounds__(TB) void k_dens_own64(const LatDev* __restrict__ lats, const int32_t* __restrict__ ob2l,
                                                   const double* __restrict__ gtab64) {
  // the table layout with only the rows this pass reads staged: f_H (part
  // one, first function) and f_Pd (part two, first function)
  __shared__ __align__(16) double t64[T64_PART1 + T64_N * T64Q_P];
  const int w = xcd_block();
  const LatDev& S = lats[ob2l[w]];
  if (S.rule != RULE) return;
  for (int i = threadIdx.x; i < T64_N * T64_P; i += blockDim.x) t64[i] = gtab64[(size_t)T64_FH * T64_N * T64_P + i];
  stage64(t64 + T64_PART1, gtab64 + T64_PART1 + (size_t)T64Q_FPD * T64_N * T64Q_P, T64_N * T64Q_P);
  const int K = S.kl64;  // lanes per site: lane ph takes owned slots ph, ph + K, ...
  const int s = (w - S.owg0_64) * (TB / K) + (int)threadIdx.x / K;
  const int ph = (int)threadIdx.x % K;
  const int npd = S.npd;
  const int N = npd + S.nh;
  if (s >= N) return;  // whole lane groups
  // everything the pair loops use, out of the descriptor and into registers
  const A2C_GLOBAL double4* __restrict__ X = gmem(reinterpret_cast<const double4*>(S.X));
  const A2C_GLOBAL double* __restrict__ xo = gmem(S.x);
  const Image img = image_of(S);
  const bool pd = s < npd;
  const int row = pd ? s : s - npd;
  const Site a = load_site(X, s);
  const double ai = SQ6 * a.s;
  const double wme = pd ? 1.0 : xo[row];  // shares leave weighted by the sender's occupancy
  double rho = 0.0;
#pragma unroll
  for (int side = 0; side < 2; ++side) {  // Pd neighbors, then H neighbors
    const int list = pd ? side : 2 + side;
    const Ell& L = list_of(S, list);
    const Ell& P = partner_list(S, list);
    const A2C_GLOBAL int32_t* __restrict__ oj = gmem(L.oj);
    const A2C_GLOBAL int32_t* __restrict__ orv = gmem(L.orv);
    A2C_GLOBAL double* __restrict__ share = gmem(reinterpret_cast<double*>(P.rbuf));
    const size_t rows = L.rows, prows = P.rows;
    const int n = gmem(L.ocnt)[row];
    for (int t = ph; t < n; t += K) {
      const size_t e = (size_t)t * rows + row;
      const int jl = oj[e];
      const int rv = orv[e];
      const Site b = load_site(X, side == 0 ? jl : npd + jl);
      const double wj = side == 0 ? 1.0 : xo[jl];
      double d[3];
      pair_vec(img, a, b, d);
      double mine = 0.0, theirs = 0.0;  // neighbor-species density at s, s-species density at the neighbor
      if (RULE == GH3) {
        const double sd = sqrt(a.s * a.s + b.s * b.s);
#pragma unroll 1
        for (int k = 0; k < 27; ++k) {
          double z[3], wk;
          gh_node(k, z, wk);
          const double r0 = d[0] + sd * z[0], r1 = d[1] + sd * z[1], r2 = d[2] + sd * z[2];
          const double q2 = r0 * r0 + r1 * r1 + r2 * r2;
          const double rn = q2 * rsqrt_nr(q2);
          const double fpd = (side == 0 || pd) ? f_Pd_tab(rn, t64) : 0.0;
          const double fh = (side == 1 || !pd) ? f_H_tab(rn, t64) : 0.0;
          mine += wk * (side == 0 ? fpd : fh);
          theirs += wk * (pd ? fpd : fh);
        }
      } else {
        const double aj = SQ6 * b.s;
#pragma unroll 1
        for (int p = 0; p < 12; ++p) {
          double r[3];
          point_r(p, d, ai, aj, r);
          const double q2 = r[0] * r[0] + r[1] * r[1] + r[2] * r[2];
          const double rn = q2 * rsqrt_nr(q2);
          const double fpd = (side == 0 || pd) ? f_Pd_tab(rn, t64) : 0.0;
          const double fh = (side == 1 || !pd) ? f_H_tab(rn, t64) : 0.0;
          mine += side == 0 ? fpd : fh;
          theirs += pd ? fpd : fh;
        }
      }
      rho += wj * mine;
      if (rv >= 0) share[(size_t)rv * prows + jl] = wme * theirs;
    }
  }
  rho = site_sum(rho, K);
  if (ph == 0) gmem(S.rho)[s] = rho;
}

__global__ __launch_bounds__(TB) void k_dens_recv64(const LatDev* __restrict__ lats, const int32_t* __restrict__ b2l) {
  const int blk = xcd_block();
  const LatDev& S = lats[b2l[blk]];
  const int s = (blk - S.blk0) * TB + threadIdx.x;
  const int N = S.npd + S.nh;
  if (s >= N) return;
  const double* __restrict__ xo = S.x;
  const bool pd = s < S.npd;
  const int row = pd ? s : s - S.npd;
  double rho = S.rho[s];
#pragma unroll
  for (int side = 0; side < 2; ++side) {
    const Ell& L = list_of(S, pd ? side : 2 + side);
    const int n = L.rcnt[row];
    // shares arrive weighted by their senders
    for_row_shares<RECV_RU_SCALAR>(gmem(reinterpret_cast<const double*>(L.rbuf)), L.rows, row, n, [&](double v) { rho += v; });
  }
  rho *= S.rule == GH3 ? 1.0 : QW;
  S.rho[s] = rho;
  const double xs = pd ? 1.0 : xo[row];
  const eam::VD F = pd ? eam::F_Pd_vd(rho) : eam::F_H_vd(rho);
  S.Fp[s] = F.d;
  S.emb[s] = xs * F.v;
}

// One owned pair: phi summed over the 12 points, the q-gradient g of site s
// (the partner's is -g), and the sigma terms of both ends (site s: its own
// six points; the partner: the six points that displace it, seen from its
// side with r -> -r).
template <int KIND, int RULE>
__device__ __forceinline__ void pair_both_closed(const double d[3], double si, double sj, double Fi, double Fj,
                                          double& pe, double g[3], double& gsi, double& gsj) {
  pe = 0.0;
  g[0] = g[1] = g[2] = 0.0;
  gsi = gsj = 0.0;
  if (RULE == GH3) {  // both ends move every node: dr/dsigma_k = z sigma_k / s
    const double sd = sqrt(si * si + sj * sj);
    double sz = 0.0;
#pragma unroll 1
    for (int k = 0; k < 27; ++k) {
      double z[3], w;
      gh_node(k, z, w);
      const double r[3] = {d[0] + sd * z[0], d[1] + sd * z[1], d[2] + sd * z[2]};
      const double q2 = r[0] * r[0] + r[1] * r[1] + r[2] * r[2];
      const double ri = rsqrt_nr(q2);
      double phi, c;
      pair_point<A2C_DERIV_ANALYTIC, KIND>(q2 * ri, Fi, Fj, 0.0, phi, c);
      pe += w * phi;
      const double u = w * c * ri;
      g[0] += u * r[0];
      g[1] += u * r[1];
      g[2] += u * r[2];
      sz += u * (r[0] * z[0] + r[1] * z[1] + r[2] * z[2]);
    }
    gsi = sz * (si / sd);
    gsj = sz * (sj / sd);
    return;
  }
  const double ai = SQ6 * si, aj = SQ6 * sj;
#pragma unroll 1  // rolled: the FP64 closed forms keep ~140 VGPRs live (3 waves per SIMD) instead of ~260 (1)
  for (int p = 0; p < 12; ++p) {
    double r[3];
    point_r(p, d, ai, aj, r);
    const double q2 = r[0] * r[0] + r[1] * r[1] + r[2] * r[2];
    const double ri = rsqrt_nr(q2);
    double phi, c;
    pair_point<A2C_DERIV_ANALYTIC, KIND>(q2 * ri, Fi, Fj, 0.0, phi, c);
    pe += phi;
    const double u = c * ri;
    g[0] += u * r[0];
    g[1] += u * r[1];
    g[2] += u * r[2];
    const int ax = (p >> 1) % 3;
    if (p < 6) gsi += ((p & 1) ? -u : u) * r[ax];
    else gsj += ((p & 1) ? u : -u) * r[ax];
  }
}

template <int KIND, int RULE>
__device__ __forceinline__ void pair_both(const double d[3], double si, double sj, double Fi, double Fj,
                                          const double* t64, double& pe, double g[3], double& gsi, double& gsj) {
  pe = 0.0;
  g[0] = g[1] = g[2] = 0.0;
  gsi = gsj = 0.0;
  bool out = false;  // some point outside the tables' [1, 6) A
  if (RULE == GH3) {  // both ends move every node: dr/dsigma_k = z sigma_k / s
    const double sd = sqrt(si * si + sj * sj);
    double sz = 0.0;
#pragma unroll 1
    for (int k = 0; k < 27; ++k) {
      double z[3], w;
      gh_node(k, z, w);
      const double r[3] = {d[0] + sd * z[0], d[1] + sd * z[1], d[2] + sd * z[2]};
      const double q2 = r[0] * r[0] + r[1] * r[1] + r[2] * r[2];
      const double ri = rsqrt_nr(q2);
      double phi, c;
      pair_point_tab<KIND>(q2 * ri, Fi, Fj, t64, phi, c, out);
      pe += w * phi;
      const double u = w * c * ri;
      g[0] += u * r[0];
      g[1] += u * r[1];
      g[2] += u * r[2];
      sz += u * (r[0] * z[0] + r[1] * z[1] + r[2] * z[2]);
    }
    gsi = sz * (si / sd);
    gsj = sz * (sj / sd);
    if (out) pair_both_closed<KIND, RULE>(d, si, sj, Fi, Fj, pe, g, gsi, gsj);
    return;
  }
  const double ai = SQ6 * si, aj = SQ6 * sj;
#pragma unroll 1  // rolled: the FP64 closed forms keep ~140 VGPRs live (3 waves per SIMD) instead of ~260 (1)
  for (int p = 0; p < 12; ++p) {
    double r[3];
    point_r(p, d, ai, aj, r);
    const double q2 = r[0] * r[0] + r[1] * r[1] + r[2] * r[2];
    const double ri = rsqrt_nr(q2);
    double phi, c;
    pair_point_tab<KIND>(q2 * ri, Fi, Fj, t64, phi, c, out);
    pe += phi;
    const double u = c * ri;
    g[0] += u * r[0];
    g[1] += u * r[1];
    g[2] += u * r[2];
    const int ax = (p >> 1) % 3;
    if (p < 6) gsi += ((p & 1) ? -u : u) * r[ax];
    else gsj += ((p & 1) ? u : -u) * r[ax];
  }
  if (out) pair_both_closed<KIND, RULE>(d, si, sj, Fi, Fj, pe, g, gsi, gsj);
}


// 3 waves per SIMD (at most 168 VGPRs per lane)
template <int RULE>
__global__ __launch_bounds__(TB) __attribute__((amdgpu_waves_per_eu(3))) void k_grad_own64(
    const LatDev* __restrict__ lats, const int32_t* __restrict__ ob2l, const double* __restrict__ gtab64) {
  __shared__ __align__(16) double t64[T64_SIZE];
  const int w = xcd_block();
  const LatDev& S = lats[ob2l[w]];
  if (S.rule != RULE) return;
  stage64(t64, gtab64, T64_SIZE);
  const int K = S.kl64;  // lanes per site: lane ph takes owned slots ph, ph + K, ...
  const int s = (w - S.owg0_64) * (TB / K) + (int)threadIdx.x / K;
  const int ph = (int)threadIdx.x % K;
  const int npd = S.npd;
  const int N = npd + S.nh;
  if (s >= N) return;  // whole lane groups
  // everything the pair loops use, out of the descriptor and into registers
  const A2C_GLOBAL double4* __restrict__ X = gmem(reinterpret_cast<const double4*>(S.X));
  const A2C_GLOBAL double* __restrict__ xo = gmem(S.x);
  const A2C_GLOBAL double* __restrict__ Fp = gmem(S.Fp);
  const Image img = image_of(S);
  const bool pd = s < npd;
  const int row = pd ? s : s - npd;
  const Site a = load_site(X, s);
  const double Fi = Fp[s];
  const double wme = pd ? 1.0 : xo[row];  // shares leave weighted by the sender's occupancy
  double e = 0.0, g0 = 0.0, g1 = 0.0, g2 = 0.0, gs = 0.0;
#pragma unroll
  for (int side = 0; side < 2; ++side) {
    const int list = pd ? side : 2 + side;
    const Ell& L = list_of(S, list);
    const Ell& P = partner_list(S, list);
    const A2C_GLOBAL int32_t* __restrict__ oj = gmem(L.oj);
    const A2C_GLOBAL int32_t* __restrict__ orv = gmem(L.orv);
    A2C_GLOBAL double4* __restrict__ share = gmem(reinterpret_cast<double4*>(P.rbuf));
    const size_t rows = L.rows, prows = P.rows;
    const int n = gmem(L.ocnt)[row];
    const int off = side == 0 ? 0 : npd;  // site index of neighbor 0 of this list
    for (int t = ph; t < n; t += K) {
      // (no one-pair-ahead prefetch: an FP64 pair is long enough to cover its
      // gathers, and the registers would push the kernel into spills)
      const size_t ei = (size_t)t * rows + row;
      const int jl = oj[ei];
      const int rv = orv[ei];
      const Site b = load_site(X, off + jl);
      const double Fj = Fp[off + jl];
      const double wj = side == 0 ? 1.0 : xo[jl];
      double d[3];
      pair_vec(img, a, b, d);
      double pe, pg[3], psi, psj;
      if (list == 0) pair_both<0, RULE>(d, a.s, b.s, Fi, Fj, t64, pe, pg, psi, psj);
      else if (list == 1) pair_both<1, RULE>(d, a.s, b.s, Fi, Fj, t64, pe, pg, psi, psj);
      else if (list == 2) pair_both<2, RULE>(d, a.s, b.s, Fi, Fj, t64, pe, pg, psi, psj);
      else pair_both<3, RULE>(d, a.s, b.s, Fi, Fj, t64, pe, pg, psi, psj);
      e += wj * pe;
      g0 += wj * pg[0]; g1 += wj * pg[1]; g2 += wj * pg[2];
      gs += wj * psi;
      if (rv >= 0) share[(size_t)rv * prows + jl] = make_double4(-wme * pg[0], -wme * pg[1], -wme * pg[2], wme * psj);
    }
  }
  e = site_sum(e, K);
  g0 = site_sum(g0, K);
  g1 = site_sum(g1, K);
  g2 = site_sum(g2, K);
  gs = site_sum(gs, K);
  if (ph != 0) return;
  gmem(reinterpret_cast<double4*>(S.G))[s] = make_double4(g0, g1, g2, gs);
  gmem(S.epair)[s] = e;
}

__global__ __launch_bounds__(TB) void k_grad_recv64(const LatDev* __restrict__ lats, const int32_t* __restrict__ b2l) {
  __shared__ double red[TB];
  const int blk = xcd_block();
  const LatDev& S = lats[b2l[blk]];
  const int lb = blk - S.blk0;
  const int s = lb * TB + threadIdx.x;
  const int N = S.npd + S.nh;
  double fs = 0.0;
  if (s < N) {
    const double* __restrict__ xo = S.x;
    const bool pd = s < S.npd;
    const int row = pd ? s : s - S.npd;
    double* __restrict__ G = S.G;
    double g0 = G[4 * (size_t)s], g1 = G[4 * (size_t)s + 1], g2 = G[4 * (size_t)s + 2], gs = G[4 * (size_t)s + 3];
#pragma unroll
    for (int side = 0; side < 2; ++side) {
      const Ell& L = list_of(S, pd ? side : 2 + side);
      const int n = L.rcnt[row];
      // shares arrive weighted by their senders
      for_row_shares<RECV_RU_WIDE>(gmem(reinterpret_cast<const double4*>(L.rbuf)), L.rows, row, n, [&](const double4& v) {
        g0 += v.x; g1 += v.y; g2 += v.z;
        gs += v.w;
      });
    }
    const double xi = pd ? 1.0 : xo[row];
    const double sg = S.X[4 * (size_t)s + 3];
    const double T = S.T;
    const bool gh = S.rule == GH3;
    const double qw = gh ? 1.0 : QW, sq = gh ? 1.0 : SQ6;
    const double V = S.emb[s] + qw * xi * S.epair[s];  // owned pairs carry their full energy
    const double sc = qw * xi / T;
    G[4 * (size_t)s + 0] = sc * g0;
    G[4 * (size_t)s + 1] = sc * g1;
    G[4 * (size_t)s + 2] = sc * g2;
    G[4 * (size_t)s + 3] = sq * sc * gs - 3.0 * KB * xi / sg;
    fs = V / T - 3.0 * KB * xi * log(sg);
  }
  red[threadIdx.x] = fs;
  __syncthreads();
#pragma unroll
  for (int w = TB / 2; w > 0; w >>= 1) {
    if (threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) S.part[lb] = red[0];
}

}  // namespace

// One block per lattice: fixed-order sum of its block partials.
__global__ __launch_bounds__(1024) void k_final_sum(const LatDev* __restrict__ lats) {
  __shared__ double red[1024];
  const LatDev& S = lats[blockIdx.x];
  const int n = site_blocks(S.npd + S.nh);
  double acc = 0.0;
  for (int i = threadIdx.x; i < n; i += 1024) acc += S.part[i];
  red[threadIdx.x] = acc;
  __syncthreads();
  for (int w = 512; w > 0; w >>= 1) {
    if (threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) *S.f = red[0];
}

namespace {
// Launch the instance of each quadrature rule present in the launch.
template <template <int> class K, class... A>
int per_rule(const Launch& L, int nblk, A... args) {
  if (nblk == 0) return A2C_OK;
  if (L.rules & (1 << A2C_QUAD_R12))
    hipLaunchKernelGGL(K<A2C_QUAD_R12>::fn, dim3(nblk), dim3(TB), 0, L.stream, args...);
  if (L.rules & (1 << GH3)) hipLaunchKernelGGL(K<GH3>::fn, dim3(nblk), dim3(TB), 0, L.stream, args...);
  return hipGetLastError() == hipSuccess ? A2C_OK : A2C_EHIP;
}
template <int R> struct DensFd { static constexpr auto fn = k_density64<A2C_DERIV_FD, R>; };
template <int R> struct DensAn { static constexpr auto fn = k_density64<A2C_DERIV_ANALYTIC, R>; };
template <int R> struct GradFd { static constexpr auto fn = k_grad64<A2C_DERIV_FD, R>; };
template <int R> struct GradAn { static constexpr auto fn = k_grad64<A2C_DERIV_ANALYTIC, R>; };
template <int R> struct DensOwn { static constexpr auto fn = k_dens_own64<R>; };
template <int R> struct GradOwn { static constexpr auto fn = k_grad_own64<R>; };
}  // namespace

int run_density64(const Launch& L) {
  if (L.deriv_mode == A2C_DERIV_FD) return per_rule<DensFd>(L, L.nblk, L.lats, L.b2l, L.fd_eps);
  return per_rule<DensAn>(L, L.nblk, L.lats, L.b2l, L.fd_eps);
}

int run_eval64(const Launch& L, Timing* tm) {
  if (L.deriv_mode == A2C_DERIV_ANALYTIC && !L.full_pairs64) {
    if (L.nblk > 0) {
      if (tm) tm->mark(L.stream, "dens_own64");
      if (per_rule<DensOwn>(L, L.onblk, L.lats, L.ob2l, L.tab64)) return A2C_EHIP;
      if (tm) tm->mark(L.stream, "dens_recv64");
      hipLaunchKernelGGL(k_dens_recv64, dim3(L.nblk), dim3(TB), 0, L.stream, L.lats, L.b2l);
      if (hipGetLastError() != hipSuccess) return A2C_EHIP;
      if (tm) tm->mark(L.stream, "grad_own64");
      if (per_rule<GradOwn>(L, L.onblk, L.lats, L.ob2l, L.tab64)) return A2C_EHIP;
      if (tm) tm->mark(L.stream, "grad_recv64");
      hipLaunchKernelGGL(k_grad_recv64, dim3(L.nblk), dim3(TB), 0, L.stream, L.lats, L.b2l);
      if (hipGetLastError() != hipSuccess) return A2C_EHIP;
    }
    if (tm) tm->mark(L.stream, "final_sum");
    if (int rc = run_final_sum(L)) return rc;
    if (tm) tm->mark(L.stream, nullptr);
    return A2C_OK;
  }
  if (tm) tm->mark(L.stream, "density64");
  if (int rc = run_density64(L)) return rc;
  if (tm) tm->mark(L.stream, "grad64");
  if (L.deriv_mode == A2C_DERIV_FD) {
    if (per_rule<GradFd>(L, L.nblk, L.lats, L.b2l, L.fd_eps)) return A2C_EHIP;
  } else if (per_rule<GradAn>(L, L.nblk, L.lats, L.b2l, L.fd_eps)) {
    return A2C_EHIP;
  }
  if (tm) tm->mark(L.stream, "final_sum");
  if (int rc = run_final_sum(L)) return rc;
  if (tm) tm->mark(L.stream, nullptr);
  return A2C_OK;
}

int run_final_sum(const Launch& L) {
  if (L.nlat == 0) return A2C_OK;
  hipLaunchKernelGGL(k_final_sum, dim3(L.nlat), dim3(1024), 0, L.stream, L.lats);
  return hipGetLastError() == hipSuccess ? A2C_OK : A2C_EHIP;
}

// FP64 tables: on each 0.05 A interval the polynomial (degree 8 for the
// exponential functions, 11 for the Pd polynomials) interpolating the
// long-double closed form at the Chebyshev points, in Newton form, expanded
// to monomials in u (all in long double); the switch radii 5.0, 5.05 and
// 5.35 A are interval boundaries.
void host_tables64(std::vector<double>& t) {
  static const std::vector<double> built = [] {
    using LD = long double;
    std::vector<double> out((size_t)T64_SIZE, 0.0);
    auto fill = [&](auto deg, size_t base, int pitch, auto g) {
      constexpr int D = decltype(deg)::value;
      const LD H = 0.05L, PI = 3.14159265358979323846264338327950288L;
      for (int k = 0; k < T64_N; ++k) {
        const LD a = 1.0L + H * k;
        LD u[D + 1], c[D + 1], m[D + 1] = {};
        for (int i = 0; i <= D; ++i) {
          u[i] = 0.5L - 0.5L * cosl((2 * i + 1) * PI / (2.0L * (D + 1)));
          c[i] = g(a + u[i] * H);
        }
        for (int j = 1; j <= D; ++j)  // divided differences
          for (int i = D; i >= j; --i) c[i] = (c[i] - c[i - 1]) / (u[i] - u[i - j]);
        for (int i = D; i >= 0; --i) {  // m <- m (u - u_i) + c_i
          LD n[D + 1] = {};
          for (int q = 0; q < D; ++q) {
            n[q + 1] += m[q];
            n[q] -= m[q] * u[i];
          }
          n[0] += c[i];
          for (int q = 0; q <= D; ++q) m[q] = n[q];
        }
        for (int q = 0; q <= D; ++q) out[base + (size_t)k * pitch + q] = (double)m[q];
      }
    };
    using D8 = std::integral_constant<int, T64_D>;
    using D11 = std::integral_constant<int, T64Q_D>;
    fill(D8{}, (size_t)T64_FH * T64_N * T64_P, T64_P, [](LD r) { return eam::f_H<LD>(r); });
    fill(D8{}, (size_t)T64_PHPDH * T64_N * T64_P, T64_P, [](LD r) { return eam::phi_morse<0, LD>(r); });
    fill(D8{}, (size_t)T64_PHH * T64_N * T64_P, T64_P, [](LD r) { return eam::phi_morse<1, LD>(r); });
    fill(D11{}, T64_PART1 + (size_t)T64Q_FPD * T64_N * T64Q_P, T64Q_P, [](LD r) { return eam::f_Pd<LD>(r); });
    fill(D11{}, T64_PART1 + (size_t)T64Q_PHPD * T64_N * T64Q_P, T64Q_P, [](LD r) { return eam::phi_Pd<LD>(r); });
    return out;
  }();
  t = built;
}

bool mixed_full_pairs() {
  static const bool full = [] {
    const char* v = getenv("A2C_MIXED_PAIRS");
    return v && std::string(v) == "full";
  }();
  return full;
}
bool fp64_full_pairs() {
  static const bool full = [] {
    const char* v = getenv("A2C_FP64_PAIRS");
    return v && std::string(v) == "full";
  }();
  return full;
}

static Launch single_launch(a2c_fe_ctx* c, bool fp64) {
  Launch L;
  L.mixed_full = mixed_full_pairs();
  L.full_pairs64 = fp64_full_pairs();
  L.lats = c->dlat;
  L.b2l = c->db2l;
  L.ob2l = c->db2l;
  L.onblk = site_blocks(c->N) * lanes_per_site(c->N, fp64);
  L.nblk = site_blocks(c->N);
  L.nlat = 1;
  L.rules = 1 << c->desc.quad_rule;
  L.stream = c->stream;
  L.deriv_mode = c->desc.deriv_mode;
  L.fd_eps = c->desc.fd_eps;
  L.tab = c->dtab;
  L.tab64 = c->dtab64;
  return L;
}

int launch_density64(a2c_fe_ctx* c, const double* dX) {
  if (c->gen_on) return launch_density_gen(c, dX);
  if (int rc = sync_lattice(c, dX, c->dG, c->dF)) return rc;
  if (run_density64(single_launch(c, true))) return set_err(c, A2C_EHIP, "density64 launch failed");
  return A2C_OK;
}

int launch_eval64(a2c_fe_ctx* c, const double* dX, double* dG, double* dF, bool timed, Timing* tm) {
  if (int rc = sync_lattice(c, dX, dG, dF)) return rc;
  if (run_eval64(single_launch(c, true), timed ? tm : nullptr)) return set_err(c, A2C_EHIP, "eval64 launch failed");
  return A2C_OK;
}

int launch_eval32(a2c_fe_ctx* c, const double* dX, double* dG, double* dF, bool timed, Timing* tm) {
  if (!c->dtab) return set_err(c, A2C_ESTATE, "tables not built");
  if (int rc = sync_lattice(c, dX, dG, dF)) return rc;
  if (run_eval32(single_launch(c, false), timed ? tm : nullptr)) return set_err(c, A2C_EHIP, "eval32 launch failed");
  return A2C_OK;
}

// Refresh the device descriptor of the context's lattice when the state,
// gradient or objective buffers (or the neighbor lists) changed.
int sync_lattice(a2c_fe_ctx* c, const double* dX, double* dG, double* dF) {
  if (!c->hlat) {  // pinned, at the first single-lattice launch (a batch's lattices never need it)
    A2C_HIP(c, hmalloc(&c->hlat, sizeof(LatDev)));
    new (c->hlat) LatDev();
    c->lat_dirty = true;
  }
  LatDev& h = *c->hlat;
  const int kl32 = lanes_per_site(c->N, false), kl64 = lanes_per_site(c->N, true);
  if (!c->lat_dirty && h.X == dX && h.G == dG && h.f == dF && h.kl32 == kl32 && h.kl64 == kl64) return A2C_OK;
  A2C_HIP(c, hipStreamSynchronize(c->stream));  // the pinned copy may still be in flight
  h = c->sys();
  h.blk0 = 0;
  h.kl32 = kl32;  // the launch grid (single_launch) takes K from the same table
  h.kl64 = kl64;
  h.owg0_32 = h.owg0_64 = 0;
  h.X = dX;
  h.G = dG;
  h.f = dF;
  A2C_HIP(c, hipMemcpyAsync(c->dlat, c->hlat, sizeof(LatDev), hipMemcpyHostToDevice, c->stream));
  c->lat_dirty = false;
  return A2C_OK;
}

}  // namespace a2c

// Host-only: the FP64 tables of the FP64 pair-owner passes (layout in a2c_fe.h).
extern "C" int a2c_eam_tables64(double* out, int32_t cap) {
  std::vector<double> t;
  a2c::host_tables64(t);
  const int n = (int)t.size();
  if (out && cap >= n) std::memcpy(out, t.data(), sizeof(double) * n);
  return n;
}
