// Mixed-precision objective + gradient (Minimizer::formFunctionAndGradient,
// minimizer.cpp:83-369): FP32 bulk quadrature, FP64 everywhere the error
// would not average out.
//
//   FP64: positions, the pair vector d = q_i - q_j (then rounded to FP32, so
//         the rounding is relative to |d| <= rc, not to the box size),
//         per-site accumulation of the per-pair sums, rho, F(rho), F'(rho),
//         the per-site objective share and the global reduction.
//   FP32: the 12 quadrature points of a pair: |r|^2 = |d|^2 + a^2 +- 2 a d_c
//         (each point displaces d along one axis by a = sqrt(6) sigma), 1/|r|
//         by v_rsq_f32, and the EAM functions.  Under the Gauss-Hermite rule
//         the 27 nodes r_k = d + s sqrt(3) a_k (a_k in {-1,0,1}^3) have
//         |r_k|^2 = |d|^2 + 3 s^2 |a_k|^2 + 2 sqrt(3) s (a_k . d).
//
// The EAM functions are piecewise cubic Hermite polynomials on a uniform grid
// in r (h = 0.01 A over [1, 6] A, switch radii 5.0, 5.05, 5.35 on grid
// nodes), built on the host in 80-bit long double from the closed forms and
// their analytic derivatives, stored as FP32 coefficients in the local
// coordinate u in [0,1), and staged in LDS (5 tables x 500 intervals x 16 B =
// 40 KB per workgroup).  Beyond 6 A every function is the exponential tail
// of potential.cpp, evaluated directly.  Interpolation error is below the FP32
// rounding of the table values (~1e-7 relative, mean bias ~1e-10).
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "eam64.hpp"
#include "fe_internal.hpp"

namespace a2c {
namespace {

constexpr int TB = SITE_BLOCK;
constexpr int NT = 500;               // intervals per table (h = 0.0125 would fit 32 KB of LDS, but the
                                      // f_Pd derivative error then reaches 1.4e-5 of its maximum)
constexpr float R0 = 1.0f, R1 = 6.0f, INVH = 100.0f;
constexpr double H64 = 0.01;
constexpr double QW = 1.0 / 12.0;     // 0.5/dim, quadrature.cpp:39
constexpr double SQ6 = 2.449489742783178098;
constexpr float SQ6F = 2.449489742783178098f;
constexpr int GH3 = A2C_QUAD_GH3;
constexpr float GHZF = 1.7320508075688772935f;  // Gauss-Hermite node sqrt(3)
__device__ constexpr float ghwf(int a) { return a == 0 ? 2.0f / 3.0f : 1.0f / 6.0f; }
enum { T_FPD = 0, T_FH = 1, T_PHPD = 2, T_PHPDH = 3, T_PHH = 4, NTAB = 5 };

// Exponential tails beyond the switch radii: funRc * exp(k (r - rs)).
__constant__ float TAIL[NTAB][3] = {
    {0.001510569378297f, (float)(-0.077976566126292 / 0.001510569378297), 5.0f},
    {0.009986990526578f, (float)(-0.013075667086408 / 0.009986990526578), 5.35f},
    {-0.016958780164980f, (float)(0.089019679316328 / -0.016958780164980), 5.05f},
    {-3.353001113951400e-04f, (float)(7.147062196813336e-04 / -3.353001113951400e-04), 5.35f},
    {-0.003702722457147f, (float)(0.005465945471933 / -0.003702722457147), 5.35f}};

__device__ __forceinline__ float rsq(float x) { return __builtin_amdgcn_rsqf(x); }

// Position of r in the tables: interval k and local coordinate u in [0, 1),
// computed once per quadrature point and shared by every function evaluated
// there.  Outside [1, 6) A ("far") k is clamped into the table so the reads
// stay in bounds and their results are replaced: from 6 A on by the
// exponential tails, below 1 A by NaN (no physical configuration comes that
// close -- the nearest Pd-H distance is 1.9 A -- and the non-finite
// objective makes the caller fall back to the FP64 path).
struct Pos {
  int k;
  float u;
  bool far;
};
__device__ __forceinline__ Pos tab_pos(float r) {
  const float xx = fmaf(r, INVH, -R0 * INVH);
  Pos p;
  p.k = min(max((int)xx, 0), NT - 1);
  p.u = __builtin_amdgcn_fractf(xx);
  p.far = !(fabsf(xx - 0.5f * NT) < 0.5f * NT);
  return p;
}
// value of a cubic row {c0, c1, c2, c3}
__device__ __forceinline__ float cubic(const float4 c, float u) { return fmaf(u, fmaf(u, fmaf(u, c.w, c.z), c.y), c.x); }
// value and d/du of a cubic row (Horner with derivative: 5 FMAs)
__device__ __forceinline__ void cubic_vd(const float4 c, float u, float& v, float& du) {
  const float v1 = fmaf(u, c.w, c.z);
  const float v2 = fmaf(u, v1, c.y);
  const float d2 = fmaf(u, c.w, v1);
  v = fmaf(u, v2, c.x);
  du = fmaf(u, d2, v2);
}
// value of a derivative row {c1, 2 c2, 3 c3, 0} (d/du of a cubic row)
__device__ __forceinline__ float quad(const float4 c, float u) { return fmaf(u, fmaf(u, c.z, c.y), c.x); }
// exponential tail of function w (its slope is TAIL[w][1] times the value);
// NaN below the table (see tab_pos)
__device__ __forceinline__ float tail(int w, float r) {
  return r >= R1 ? TAIL[w][0] * __expf(TAIL[w][1] * (r - TAIL[w][2])) : __builtin_nanf("");
}

// value of table w at r (tails beyond the table; NaN below it)
__device__ __forceinline__ float tab_v(const float4* __restrict__ tab, int w, float r) {
  const Pos p = tab_pos(r);
  if (!p.far) return cubic(tab[w * NT + p.k], p.u);
  return tail(w, r);
}
// The same without the branch: the row of the clamped interval, with far set
// when r lies outside the table.  A per-point branch makes every point its own
// basic block, so a wave waits out each point's LDS read before the next one
// can issue; the pair passes evaluate all points of a pair this way and redo
// the (rare) pair with a point outside [1, 6) A through the branching form.
__device__ __forceinline__ float tab_vf(const float4* __restrict__ tab, int w, float r, bool& far) {
  const Pos p = tab_pos(r);
  far |= p.far;
  return cubic(tab[w * NT + p.k], p.u);
}

// f_H(r) = 11.0025 exp(-1.30927 r) (potential.cpp:105-119; the exponential
// tail past 5.35 A continues the same function to 1e-13), evaluated without a
// table as 11.0025 * 2^(K r), K = -1.30927 log2(e) split into FP32 hi and lo
// parts, with the rounding of K r carried into a first-order correction: the
// value is within about an ulp of FP32 and costs one transcendental and five
// FMA-class instructions instead of a 16-byte LDS lookup.  NaN below 1 A, as
// the tables (see tab_pos).  Its slope is FH_K times the value.
constexpr double FH_K = -1.30927;
constexpr double LOG2E = 1.4426950408889634074;
constexpr float FH_C = 11.0025f;
constexpr float FH_KHI = (float)(FH_K * LOG2E);
constexpr float FH_KLO = (float)(FH_K * LOG2E - (double)FH_KHI);
constexpr float FH_KF = (float)FH_K;
constexpr float LN2F = 0.69314718055994530942f;
__device__ __forceinline__ float fH_exp(float r) {
  const float e = FH_KHI * r;
  const float el = fmaf(FH_KLO, r, fmaf(FH_KHI, r, -e));
  const float v = FH_C * __builtin_amdgcn_exp2f(e) * fmaf(el, LN2F, 1.0f);
  return r >= R0 ? v : __builtin_nanf("");
}
// density of a species at r: f_H directly, f_Pd from the value table in LDS
// rows [0, NT) of the pair-owner density pass
template <bool H, bool FAST>
__device__ __forceinline__ float dens(const float4* __restrict__ tab, float r, bool& far) {
  if (H) return fH_exp(r);
  return FAST ? tab_vf(tab, T_FPD, r, far) : tab_v(tab, T_FPD, r);
}

struct Site {
  double q[3];
  double s;
};
__device__ __forceinline__ Site load_site(const double* __restrict__ X, int s) {
  const double4 v = *reinterpret_cast<const double4*>(X + 4 * (size_t)s);
  return {{v.x, v.y, v.z}, v.w};
}
__device__ __forceinline__ Site load_site(const A2C_GLOBAL double4* __restrict__ X, int s) {
  const double4 v = X[s];
  return {{v.x, v.y, v.z}, v.w};
}
// d = q_i - image(q_j) in FP64, rounded to FP32 (geometry held in registers)
__device__ __forceinline__ void pair_vec32(const Image& g, const Site& a, const Site& b, float d[3]) {
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    double qj = b.q[c];
    if (g.per[c]) qj = qj - g.box[c] * rint((qj - a.q[c]) * g.ibox[c]);
    d[c] = (float)(a.q[c] - qj);
  }
}
// d = q_i - image(q_j) in FP64, rounded to FP32
__device__ __forceinline__ void pair_vec32(const LatDev& S, const Site& a, const Site& b, float d[3]) {
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    double qj = b.q[c];
    if (S.periodic[c]) qj = qj - S.box[c] * rint((qj - a.q[c]) * S.ibox[c]);
    d[c] = (float)(a.q[c] - qj);
  }
}

__device__ __forceinline__ void stage_tables(float4* lds, const float4* __restrict__ g, int n) {
  for (int i = threadIdx.x; i < n; i += blockDim.x) lds[i] = g[i];
  __syncthreads();
}
// Gradient passes need f_Pd only through f_Pd', staged as derivative rows
// {c1, 2 c2, 3 c3, 0}; the pair tables stay cubic; f_H is evaluated directly.
// LDS layout of the gradient passes: G_FPD (f_Pd'), G_PHPD, G_PHPDH, G_PHH,
// NT rows each (32 KB).
enum { G_FPD = 0, G_PHPD = 1, G_PHPDH = 2, G_PHH = 3, NGTAB = 4 };
__device__ __forceinline__ void stage_grad_tables(float4* lds, const float4* __restrict__ g) {
  for (int i = threadIdx.x; i < NGTAB * NT; i += blockDim.x) {
    const int w = i / NT, k = i - w * NT;
    const int src = w == G_FPD ? T_FPD : w + 1;  // T_PHPD, T_PHPDH, T_PHH follow T_FH in the global tables
    const float4 c = g[src * NT + k];
    lds[i] = w == G_FPD ? make_float4(c.y, 2.0f * c.z, 3.0f * c.w, 0.0f) : c;
  }
  __syncthreads();
}

// ---------------------------------------------------------------- density
// Sum of the table-w density over the 12 points of one pair.
__device__ __forceinline__ float pair_density(const float4* tab, int w, const float d[3], float ai, float aj) {
  const float D2 = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
  const float bi = D2 + ai * ai, bj = D2 + aj * aj;
  float acc = 0.0f;
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    const float ti = 2.0f * ai * d[c], tj = 2.0f * aj * d[c];
    const float r2[4] = {bi + ti, bi - ti, bj - tj, bj + tj};
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const float x = fmaxf(r2[q], 1e-12f);
      acc += tab_v(tab, w, x * rsq(x));
    }
  }
  return acc;
}

__global__ __launch_bounds__(TB) void k_density32(const LatDev* __restrict__ lats, const int32_t* __restrict__ b2l,
                                                  const float4* __restrict__ gtab) {
  __shared__ float4 tab[2 * NT];  // f_Pd, f_H
  stage_tables(tab, gtab, 2 * NT);
  const int blk = xcd_block();
  const LatDev& S = lats[b2l[blk]];
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
  const float ai = SQ6F * (float)a.s;
  double rho = 0.0;
  const int np = LP.cnt[row];
  for (int t = 0; t < np; ++t) {
    const int j = LP.idx[(size_t)t * LP.rows + row];
    const Site b = load_site(X, j);
    float d[3];
    pair_vec32(S, a, b, d);
    rho += (double)pair_density(tab, T_FPD, d, ai, SQ6F * (float)b.s);
  }
  const int nhn = LH.cnt[row];
  for (int t = 0; t < nhn; ++t) {
    const int h = LH.idx[(size_t)t * LH.rows + row];
    const Site b = load_site(X, S.npd + h);
    float d[3];
    pair_vec32(S, a, b, d);
    rho += xo[h] * (double)pair_density(tab, T_FH, d, ai, SQ6F * (float)b.s);
  }
  rho *= QW;
  S.rho[s] = rho;
  const double xs = pd ? 1.0 : xo[row];
  const eam::VD F = pd ? eam::F_Pd_vd(rho) : eam::F_H_vd(rho);
  S.aux[s] = make_float2((float)F.d, (float)xs);
  S.emb[s] = xs * F.v;
}

// ---------------------------------------------------------------- gradient
// Pair kinds: 0 Pd-Pd, 1 Pd-H (site Pd), 2 H-Pd (site H), 3 H-H.
// Per point: the pair energy phi(r) and cu = h dE/dr (h = 0.01 A, the table
// step; callers scale their per-pair sums by 1/h once) of the pair's share
// of the objective, dE/dr = F'_i f_j'(r) + F'_j f_i'(r) + phi'(r) (f_k: the
// density function of site k's species).  One table position serves all
// three functions.
template <int KIND, bool FAST = false>
__device__ __forceinline__ void point_coef(const float4* tab, float r, float Fi, float Fj, float& phi, float& cu,
                                           bool& far) {
  constexpr int WP = KIND == 0 ? G_PHPD : (KIND == 3 ? G_PHH : G_PHPDH);    // LDS table of phi
  constexpr int TP = KIND == 0 ? T_PHPD : (KIND == 3 ? T_PHH : T_PHPDH);    // its tail
  const Pos p = tab_pos(r);
  float dphi;
  cubic_vd(tab[WP * NT + p.k], p.u, phi, dphi);
  // h f_H'(r) (h = 1/INVH, the table step): f_H' = FH_K f_H
  const float hdfh = KIND == 0 ? 0.0f : (FH_KF / INVH) * fH_exp(r);
  if (KIND == 0) {
    cu = fmaf(Fi + Fj, quad(tab[G_FPD * NT + p.k], p.u), dphi);
  } else if (KIND == 3) {
    cu = fmaf(Fi + Fj, hdfh, dphi);
  } else {
    const float Fpd = KIND == 1 ? Fi : Fj;  // F' of the Pd site multiplies f_H'
    const float Fh = KIND == 1 ? Fj : Fi;   // F' of the H site multiplies f_Pd'
    cu = fmaf(Fpd, hdfh, fmaf(Fh, quad(tab[G_FPD * NT + p.k], p.u), dphi));
  }
  if (FAST) {
    far |= p.far;
  } else if (p.far) {
    phi = tail(TP, r);
    const float dp = TAIL[TP][1] * phi;
    float c;
    if (KIND == 0) {
      c = fmaf(Fi + Fj, TAIL[T_FPD][1] * tail(T_FPD, r), dp);
    } else if (KIND == 3) {
      c = fmaf(Fi + Fj, hdfh * INVH, dp);
    } else {
      const float Fpd = KIND == 1 ? Fi : Fj;
      const float Fh = KIND == 1 ? Fj : Fi;
      c = fmaf(Fpd, hdfh * INVH, fmaf(Fh, TAIL[T_FPD][1] * tail(T_FPD, r), dp));
    }
    cu = c * (1.0f / INVH);
  }
}

struct PairSums {
  float pe;     // sum of phi over the 12 points
  float g[3];   // sum of (c/|r|) r
  float gs;     // sum over site-i points of (c/|r|) r . xi1 / sqrt(3)
};

template <int KIND>
__device__ __forceinline__ PairSums pair_grad(const float4* tab, const float d[3], float ai, float aj, float Fi,
                                              float Fj) {
  const float D2 = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
  const float bi = D2 + ai * ai, bj = D2 + aj * aj;
  float su = 0.0f, ui = 0.0f, pe = 0.0f, ti[3], tj[3];
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    const float t1 = 2.0f * ai * d[c], t2 = 2.0f * aj * d[c];
    // site-i points: r = d +- ai e_c ; site-j points: r = d -+ aj e_c
    const float r2[4] = {bi + t1, bi - t1, bj - t2, bj + t2};
    float u[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const float x = fmaxf(r2[q], 1e-12f);
      const float rinv = rsq(x);
      float phi, cc;
      bool far = false;
      point_coef<KIND>(tab, x * rinv, Fi, Fj, phi, cc, far);
      pe += phi;
      u[q] = cc * rinv;
    }
    ti[c] = u[0] - u[1];
    tj[c] = u[2] - u[3];
    ui += u[0] + u[1];
    su += (u[0] + u[1]) + (u[2] + u[3]);
  }
  PairSums p;
  p.pe = pe;
#pragma unroll
  for (int c = 0; c < 3; ++c) p.g[c] = INVH * fmaf(d[c], su, ai * ti[c] - aj * tj[c]);
  p.gs = INVH * fmaf(d[0], ti[0], fmaf(d[1], ti[1], fmaf(d[2], ti[2], ai * ui)));
  return p;
}

__global__ __launch_bounds__(TB) void k_grad32(const LatDev* __restrict__ lats, const int32_t* __restrict__ b2l,
                                               const float4* __restrict__ gtab) {
  __shared__ float4 tab[NGTAB * NT];
  stage_grad_tables(tab, gtab);
  const int blk = xcd_block();
  const LatDev& S = lats[b2l[blk]];
  const int lb = blk - S.blk0;
  const int s = lb * TB + threadIdx.x;
  const int N = S.npd + S.nh;
  const double* __restrict__ X = S.X;
  const double* __restrict__ xo = S.x;
  const float2* __restrict__ aux = S.aux;
  double fs = 0.0;
  if (s < N) {
    const bool pd = s < S.npd;
    const int row = pd ? s : s - S.npd;
    const Ell LP = pd ? S.pdpd : S.hpd;
    const Ell LH = pd ? S.pdh : S.hh;
    const Site a = load_site(X, s);
    const float ai = SQ6F * (float)a.s;
    const float Fi = aux[s].x;
    const double xi = pd ? 1.0 : xo[row];
    double e = 0.0, g0 = 0.0, g1 = 0.0, g2 = 0.0, gs = 0.0;
    const int np = LP.cnt[row];
    for (int t = 0; t < np; ++t) {
      const int j = LP.idx[(size_t)t * LP.rows + row];
      const Site b = load_site(X, j);
      const float2 nb = aux[j];
      float d[3];
      pair_vec32(S, a, b, d);
      const float aj = SQ6F * (float)b.s;
      const PairSums p = pd ? pair_grad<0>(tab, d, ai, aj, Fi, nb.x) : pair_grad<2>(tab, d, ai, aj, Fi, nb.x);
      e += p.pe; g0 += p.g[0]; g1 += p.g[1]; g2 += p.g[2]; gs += p.gs;
    }
    const int nhn = LH.cnt[row];
    for (int t = 0; t < nhn; ++t) {
      const int h = LH.idx[(size_t)t * LH.rows + row];
      const Site b = load_site(X, S.npd + h);
      const float2 nb = aux[S.npd + h];
      float d[3];
      pair_vec32(S, a, b, d);
      const float aj = SQ6F * (float)b.s;
      const PairSums p = pd ? pair_grad<1>(tab, d, ai, aj, Fi, nb.x) : pair_grad<3>(tab, d, ai, aj, Fi, nb.x);
      const double w = xo[h];  // occupancy of the H neighbor
      e += w * p.pe; g0 += w * p.g[0]; g1 += w * p.g[1]; g2 += w * p.g[2]; gs += w * p.gs;
    }
    const double T = S.T;
    const double V = S.emb[s] + 0.5 * QW * xi * e;
    const double sc = QW * xi / T;
    double* __restrict__ G = S.G;
    G[4 * (size_t)s + 0] = sc * g0;
    G[4 * (size_t)s + 1] = sc * g1;
    G[4 * (size_t)s + 2] = sc * g2;
    G[4 * (size_t)s + 3] = SQ6 * sc * gs - 3.0 * KB * xi / a.s;
    fs = V / T - 3.0 * KB * xi * log(a.s);
  }
  // block reduction in the (now free) table memory
  __syncthreads();
  double* red = reinterpret_cast<double*>(tab);
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
// Each unordered pair {s, j} is evaluated once, by its owner:
//   own(s, j) = (s < j) XOR ((s ^ j) & 1),
// which gives every site about half of each of its lists.  Every row is split
// at neighbor-build time into a compact owned half and a compact received half
// (same order as the row), so each lane loops over its own pairs only.  The
// owner adds its own share and writes the partner's share into the partner's
// received slot; a gather kernel adds the received shares in FP64.  Each
// buffer entry is written by exactly one lane, so results are deterministic.

// 12-point sums of two densities at the same points: the density of the
// neighbor's species (for this site) and of this site's species (for the
// neighbor); one sum when the species are the same.  NB_H / ME_H: neighbor /
// this site is an H site.  FAST: no per-point branch (see tab_vf).
template <bool NB_H, bool ME_H, bool FAST>
__device__ __forceinline__ void pair_density2(const float4* tab, const float d[3], float ai, float aj, float& sa,
                                              float& sb, bool& far) {
  const float D2 = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
  const float bi = D2 + ai * ai, bj = D2 + aj * aj;
  sa = 0.0f;
  sb = 0.0f;
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    const float ti = 2.0f * ai * d[c], tj = 2.0f * aj * d[c];
    const float r2[4] = {bi + ti, bi - ti, bj - tj, bj + tj};
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const float x = fmaxf(r2[q], 1e-12f);
      const float r = x * rsq(x);
      sa += dens<NB_H, FAST>(tab, r, far);
      if (NB_H != ME_H) sb += dens<ME_H, FAST>(tab, r, far);
    }
  }
  if (NB_H == ME_H) sb = sa;
}

// Weighted Gauss-Hermite sums of the two densities at the 27 nodes of a pair
// (s = sqrt(sigma_i^2 + sigma_j^2)): a rolled loop over the z offset a2, the
// 9 nodes of each plane unrolled (a fully unrolled 27 spills at 64 VGPRs).
template <bool NB_H, bool ME_H, bool FAST>
__device__ __forceinline__ void pair_density2_gh(const float4* tab, const float d[3], float sd, float& sa,
                                                 float& sb, bool& far) {
  const float D2 = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
  const float s3 = 3.0f * sd * sd, tt = 2.0f * GHZF * sd;
  const float t0 = tt * d[0], t1 = tt * d[1], t2 = tt * d[2];
  sa = 0.0f;
  sb = 0.0f;
#pragma unroll 1
  for (int a2 = -1; a2 <= 1; ++a2) {
    const float base = D2 + s3 * (float)(a2 * a2) + (float)a2 * t2;
    float pa = 0.0f, pb = 0.0f;
#pragma unroll
    for (int a1 = -1; a1 <= 1; ++a1) {
#pragma unroll
      for (int a0 = -1; a0 <= 1; ++a0) {
        const float w = ghwf(a0) * ghwf(a1);
        const float x = fmaxf(base + s3 * (float)(a0 * a0 + a1 * a1) + (a0 * t0 + a1 * t1), 1e-12f);
        const float r = x * rsq(x);
        pa += w * dens<NB_H, FAST>(tab, r, far);
        if (NB_H != ME_H) pb += w * dens<ME_H, FAST>(tab, r, far);
      }
    }
    const float w2 = a2 ? 1.0f / 6.0f : 2.0f / 3.0f;
    sa += w2 * pa;
    sb += w2 * pb;
  }
  if (NB_H == ME_H) sb = sa;
}

template <int RULE, bool NB_H, bool ME_H, bool FAST>
__device__ __forceinline__ void pair_dens_rule(const float4* tab, const float d[3], float as, float bs, float& mine,
                                               float& theirs, bool& far) {
  if (RULE == GH3) pair_density2_gh<NB_H, ME_H, FAST>(tab, d, sqrtf(as * as + bs * bs), mine, theirs, far);
  else pair_density2<NB_H, ME_H, FAST>(tab, d, SQ6F * as, SQ6F * bs, mine, theirs, far);
}
// all points without a branch; the pair again with the tails if a point fell
// outside the table
template <int RULE, bool NB_H, bool ME_H>
__device__ __attribute__((noinline)) float2 pair_dens_tails(const float4* tab, float d0, float d1, float d2, float as,
                                                            float bs) {
  const float d[3] = {d0, d1, d2};
  bool far = false;
  float mine, theirs;
  pair_dens_rule<RULE, NB_H, ME_H, false>(tab, d, as, bs, mine, theirs, far);
  return make_float2(mine, theirs);
}
template <int RULE, bool NB_H, bool ME_H>
__device__ __forceinline__ void pair_dens(const float4* tab, const float d[3], float as, float bs, float& mine,
                                          float& theirs) {
  bool far = false;
  pair_dens_rule<RULE, NB_H, ME_H, true>(tab, d, as, bs, mine, theirs, far);
  if (far) {
    const float2 v = pair_dens_tails<RULE, NB_H, ME_H>(tab, d[0], d[1], d[2], as, bs);
    mine = v.x;
    theirs = v.y;
  }
}

// Partner list of each list: PdPd<->PdPd, PdH<->HPd, HH<->HH.
__device__ __forceinline__ const Ell& partner_of(const LatDev& S, int list) {
  return list == 0 ? S.pdpd : (list == 1 ? S.hpd : (list == 2 ? S.pdh : S.hh));
}

// At least 6 waves per SIMD (at most 80 VGPRs): the 12-point pass takes 71
// without spills (7 waves), the Gauss-Hermite planes 80
template <int RULE>
__global__ __launch_bounds__(TB) __attribute__((amdgpu_waves_per_eu(6))) void k_dens_own(
    const LatDev* __restrict__ lats, const int32_t* __restrict__ ob2l, const float4* __restrict__ gtab) {
  __shared__ float4 tab[NT];  // f_Pd (f_H is evaluated directly)
  const int w = xcd_block();
  const LatDev& S = lats[ob2l[w]];
  if (S.rule != RULE) return;  // whole workgroup: the other rule's instance serves this lattice
  stage_tables(tab, gtab, NT);
  const int K = S.kl32;  // lanes per site: lane ph takes owned slots ph, ph + K, ...
  const int s = (w - S.owg0_32) * (TB / K) + (int)threadIdx.x / K;
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
  const double wme = pd ? 1.0 : xo[row];  // shares leave weighted by the sender's occupancy
  double rho = 0.0;
#pragma unroll
  for (int side = 0; side < 2; ++side) {  // Pd neighbors, then H neighbors
    const int list = pd ? side : 2 + side;
    const Ell& L = list == 0 ? S.pdpd : (list == 1 ? S.pdh : (list == 2 ? S.hpd : S.hh));
    const Ell& P = partner_of(S, list);
    const A2C_GLOBAL int32_t* __restrict__ oj = gmem(L.oj);
    const A2C_GLOBAL int32_t* __restrict__ orv = gmem(L.orv);
    A2C_GLOBAL float* __restrict__ prb = gmem(reinterpret_cast<float*>(P.rbuf));
    const size_t rows = L.rows, prows = P.rows;
    const int n = gmem(L.ocnt)[row];
    for (int t = ph; t < n; t += K) {
      const size_t e = (size_t)t * rows + row;
      const int jl = oj[e];
      const int rv = orv[e];
      const Site b = load_site(X, side == 0 ? jl : npd + jl);
      const double wj = side == 0 ? 1.0 : xo[jl];
      float d[3];
      pair_vec32(img, a, b, d);
      float mine, theirs;  // neighbor-type density (for me), my-type density (for the neighbor)
      const float as = (float)a.s, bs = (float)b.s;
      if (list == 0) pair_dens<RULE, false, false>(tab, d, as, bs, mine, theirs);
      else if (list == 1) pair_dens<RULE, true, false>(tab, d, as, bs, mine, theirs);
      else if (list == 2) pair_dens<RULE, false, true>(tab, d, as, bs, mine, theirs);
      else pair_dens<RULE, true, true>(tab, d, as, bs, mine, theirs);
      rho += wj * (double)mine;
      if (rv >= 0) prb[(size_t)rv * prows + jl] = (float)(wme * (double)theirs);
    }
  }
  rho = site_sum(rho, K);
  if (ph == 0) gmem(S.rho)[s] = rho;
}

__global__ __launch_bounds__(TB) void k_dens_recv(const LatDev* __restrict__ lats, const int32_t* __restrict__ b2l) {
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
    const int list = pd ? side : 2 + side;
    const Ell& L = list == 0 ? S.pdpd : (list == 1 ? S.pdh : (list == 2 ? S.hpd : S.hh));
    const int n = L.rcnt[row];
    // shares arrive weighted by their senders
    for_row_shares<RECV_RU_SCALAR>(gmem(reinterpret_cast<const float*>(L.rbuf)), L.rows, row, n, [&](float v) { rho += (double)v; });
  }
  rho *= S.rule == GH3 ? 1.0 : QW;
  S.rho[s] = rho;
  const double xs = pd ? 1.0 : xo[row];
  const eam::VD F = pd ? eam::F_Pd_vd(rho) : eam::F_H_vd(rho);
  S.aux[s] = make_float2((float)F.d, (float)xs);
  S.emb[s] = xs * F.v;
}

struct PairSums2 {
  float pe;    // sum of phi
  float g[3];  // sum of (c/|r|) r           (q-gradient of site s; the partner gets -g)
  float gsi;   // site-s sigma term (x sqrt(6) QW x_s / T later)
  float gsj;   // partner sigma term
};

template <int KIND, bool FAST>
__device__ __forceinline__ PairSums2 pair_grad2(const float4* tab, const float d[3], float ai, float aj, float Fi,
                                                float Fj, bool& far) {
  const float D2 = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
  const float bi = D2 + ai * ai, bj = D2 + aj * aj;
  float su = 0.0f, ui = 0.0f, uj = 0.0f, pe = 0.0f, ti[3], tj[3];
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    const float t1 = 2.0f * ai * d[c], t2 = 2.0f * aj * d[c];
    const float r2[4] = {bi + t1, bi - t1, bj - t2, bj + t2};
    float u[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const float x = fmaxf(r2[q], 1e-12f);
      const float rinv = rsq(x);
      float phi, cc;
      point_coef<KIND, FAST>(tab, x * rinv, Fi, Fj, phi, cc, far);
      pe += phi;
      u[q] = cc * rinv;
    }
    ti[c] = u[0] - u[1];
    tj[c] = u[2] - u[3];
    ui += u[0] + u[1];
    uj += u[2] + u[3];
  }
  su = ui + uj;
  PairSums2 p;
  p.pe = pe;
#pragma unroll
  for (int c = 0; c < 3; ++c) p.g[c] = INVH * fmaf(d[c], su, ai * ti[c] - aj * tj[c]);
  p.gsi = INVH * fmaf(d[0], ti[0], fmaf(d[1], ti[1], fmaf(d[2], ti[2], ai * ui)));
  p.gsj = INVH * fmaf(aj, uj, -fmaf(d[0], tj[0], fmaf(d[1], tj[1], d[2] * tj[2])));
  return p;
}

// pair_grad2 for the points inside the table, with the table rows of one
// axis's four points read before any of them is used (their LDS latencies
// overlap); far is set when a point lies outside [1, 6) A.
template <int KIND>
__device__ __forceinline__ PairSums2 pair_grad2_rows(const float4* tab, const float d[3], float ai, float aj,
                                                     float Fi, float Fj, bool& far) {
  constexpr int WP = KIND == 0 ? G_PHPD : (KIND == 3 ? G_PHH : G_PHPDH);
  constexpr bool NEED_F = KIND != 3;  // f_Pd' row
  const float D2 = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
  const float bi = D2 + ai * ai, bj = D2 + aj * aj;
  float ui = 0.0f, uj = 0.0f, pe = 0.0f, ti[3], tj[3];
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    const float t1 = 2.0f * ai * d[c], t2 = 2.0f * aj * d[c];
    const float r2[4] = {bi + t1, bi - t1, bj - t2, bj + t2};
    float rinv[4], r[4], u[4];
    Pos p[4];
    float4 rp[4], rf[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const float x = fmaxf(r2[q], 1e-12f);
      rinv[q] = rsq(x);
      r[q] = x * rinv[q];
      p[q] = tab_pos(r[q]);
      far |= p[q].far;
      rp[q] = tab[WP * NT + p[q].k];
      if (NEED_F) rf[q] = tab[G_FPD * NT + p[q].k];
    }
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      float phi, dphi, cu;
      cubic_vd(rp[q], p[q].u, phi, dphi);
      if (KIND == 0) {
        cu = fmaf(Fi + Fj, quad(rf[q], p[q].u), dphi);
      } else {
        const float hdfh = (FH_KF / INVH) * fH_exp(r[q]);
        if (KIND == 3) {
          cu = fmaf(Fi + Fj, hdfh, dphi);
        } else {
          const float Fpd = KIND == 1 ? Fi : Fj;
          const float Fh = KIND == 1 ? Fj : Fi;
          cu = fmaf(Fpd, hdfh, fmaf(Fh, quad(rf[q], p[q].u), dphi));
        }
      }
      pe += phi;
      u[q] = cu * rinv[q];
    }
    ti[c] = u[0] - u[1];
    tj[c] = u[2] - u[3];
    ui += u[0] + u[1];
    uj += u[2] + u[3];
  }
  const float su = ui + uj;
  PairSums2 pr;
  pr.pe = pe;
#pragma unroll
  for (int c = 0; c < 3; ++c) pr.g[c] = INVH * fmaf(d[c], su, ai * ti[c] - aj * tj[c]);
  pr.gsi = INVH * fmaf(d[0], ti[0], fmaf(d[1], ti[1], fmaf(d[2], ti[2], ai * ui)));
  pr.gsj = INVH * fmaf(aj, uj, -fmaf(d[0], tj[0], fmaf(d[1], tj[1], d[2] * tj[2])));
  return pr;
}

// Gauss-Hermite pair sums.  With u_k = w_k (c/|r_k|) and the node offsets
// a_k, n_k = |a_k|^2:  U = sum u, Va = sum u a, Wn = sum u n;
//   q-gradient  sum u r = d U + sqrt(3) s Va,
//   Sz = sum u (r . z) = sqrt(3) d . Va + 3 s Wn,  d/dsigma_k = Sz sigma_k / s.
template <int KIND, bool FAST>
__device__ __forceinline__ PairSums2 pair_grad2_gh(const float4* tab, const float d[3], float si, float sj, float Fi,
                                                   float Fj, bool& far) {
  const float sd = sqrtf(si * si + sj * sj);
  const float D2 = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
  const float s3 = 3.0f * sd * sd, tt = 2.0f * GHZF * sd;
  const float t0 = tt * d[0], t1 = tt * d[1], t2 = tt * d[2];
  float pe = 0.0f, U = 0.0f, V0 = 0.0f, V1 = 0.0f, V2 = 0.0f, Wn = 0.0f;
#pragma unroll 1  // one z plane (9 nodes, unrolled) per trip, as in pair_density2_gh
  for (int a2 = -1; a2 <= 1; ++a2) {
    const float base = D2 + s3 * (float)(a2 * a2) + (float)a2 * t2;
    float p9 = 0.0f, u9 = 0.0f, v0 = 0.0f, v1 = 0.0f, n9 = 0.0f;
#pragma unroll
    for (int a1 = -1; a1 <= 1; ++a1) {
#pragma unroll
      for (int a0 = -1; a0 <= 1; ++a0) {
        const int n = a0 * a0 + a1 * a1;
        const float w = ghwf(a0) * ghwf(a1);
        const float x = fmaxf(base + s3 * (float)n + (a0 * t0 + a1 * t1), 1e-12f);
        const float rinv = rsq(x);
        float phi, cc;
        point_coef<KIND, FAST>(tab, x * rinv, Fi, Fj, phi, cc, far);
        p9 += w * phi;
        const float u = w * cc * rinv;
        u9 += u;
        if (a0) v0 += a0 * u;
        if (a1) v1 += a1 * u;
        if (n) n9 += n * u;
      }
    }
    const float w2 = a2 ? 1.0f / 6.0f : 2.0f / 3.0f;
    pe += w2 * p9;
    U += w2 * u9;
    V0 += w2 * v0;
    V1 += w2 * v1;
    V2 += (float)a2 * w2 * u9;
    Wn += w2 * fmaf((float)(a2 * a2), u9, n9);
  }
  PairSums2 p;
  p.pe = pe;
  const float sv = GHZF * sd;
  p.g[0] = INVH * fmaf(d[0], U, sv * V0);
  p.g[1] = INVH * fmaf(d[1], U, sv * V1);
  p.g[2] = INVH * fmaf(d[2], U, sv * V2);
  const float sz = INVH * fmaf(GHZF, fmaf(d[0], V0, fmaf(d[1], V1, d[2] * V2)), 3.0f * sd * Wn);
  p.gsi = sz * (si / sd);
  p.gsj = sz * (sj / sd);
  return p;
}

// all points without a branch; the pair again with the tails if a point fell
// outside the table
template <int RULE, int KIND>
__device__ __attribute__((noinline)) PairSums2 pair_grad_tails(const float4* tab, float d0, float d1, float d2,
                                                               float as, float bs, float Fi, float Fj) {
  const float d[3] = {d0, d1, d2};
  bool far = false;
  return RULE == GH3 ? pair_grad2_gh<KIND, false>(tab, d, as, bs, Fi, Fj, far)
                     : pair_grad2<KIND, false>(tab, d, SQ6F * as, SQ6F * bs, Fi, Fj, far);
}
template <int RULE, int KIND>
__device__ __forceinline__ PairSums2 pair_grad(const float4* tab, const float d[3], float as, float bs, float Fi,
                                               float Fj) {
  bool far = false;
  PairSums2 p = RULE == GH3 ? pair_grad2_gh<KIND, true>(tab, d, as, bs, Fi, Fj, far)
                            : pair_grad2_rows<KIND>(tab, d, SQ6F * as, SQ6F * bs, Fi, Fj, far);
  if (far) p = pair_grad_tails<RULE, KIND>(tab, d[0], d[1], d[2], as, bs, Fi, Fj);
  return p;
}

template <int RULE>
__global__ __launch_bounds__(TB) void k_grad_own(const LatDev* __restrict__ lats, const int32_t* __restrict__ ob2l,
                                                 const float4* __restrict__ gtab) {
  __shared__ float4 tab[NGTAB * NT];
  const int w = xcd_block();
  const LatDev& S = lats[ob2l[w]];
  if (S.rule != RULE) return;  // whole workgroup: the other rule's instance serves this lattice
  stage_grad_tables(tab, gtab);
  const int K = S.kl32;  // lanes per site: lane ph takes owned slots ph, ph + K, ...
  const int s = (w - S.owg0_32) * (TB / K) + (int)threadIdx.x / K;
  const int ph = (int)threadIdx.x % K;
  const int npd = S.npd;
  const int N = npd + S.nh;
  if (s >= N) return;  // whole lane groups
  // everything the pair loops use, out of the descriptor and into registers
  const A2C_GLOBAL double4* __restrict__ X = gmem(reinterpret_cast<const double4*>(S.X));
  const A2C_GLOBAL double* __restrict__ xo = gmem(S.x);
  const A2C_GLOBAL float2* __restrict__ aux = gmem(S.aux);
  const Image img = image_of(S);
  const bool pd = s < npd;
  const int row = pd ? s : s - npd;
  const Site a = load_site(X, s);
  const float Fi = aux[s].x;
  const double wme = pd ? 1.0 : xo[row];  // shares leave weighted by the sender's occupancy
  double e = 0.0, g0 = 0.0, g1 = 0.0, g2 = 0.0, gs = 0.0;
#pragma unroll
  for (int side = 0; side < 2; ++side) {
    const int list = pd ? side : 2 + side;
    const Ell& L = list == 0 ? S.pdpd : (list == 1 ? S.pdh : (list == 2 ? S.hpd : S.hh));
    const Ell& P = partner_of(S, list);
    const A2C_GLOBAL int32_t* __restrict__ oj = gmem(L.oj);
    const A2C_GLOBAL int32_t* __restrict__ orv = gmem(L.orv);
    A2C_GLOBAL float4* __restrict__ prb = gmem(reinterpret_cast<float4*>(P.rbuf));
    const size_t rows = L.rows, prows = P.rows;
    const int n = gmem(L.ocnt)[row];
    const int off = side == 0 ? 0 : npd;  // site index of neighbor 0 of this list
    // the next pair's neighbor, slot, state, F' and weight (loaded one pair ahead)
    int jl_n = 0, rv_n = -1;
    Site b_n{};
    float Fj_n = 0.0f;
    double w_n = 1.0;
    if (ph < n) {
      const size_t e0 = (size_t)ph * rows + row;
      jl_n = oj[e0];
      rv_n = orv[e0];
      b_n = load_site(X, off + jl_n);
      Fj_n = aux[off + jl_n].x;
      if (side) w_n = xo[jl_n];
    }
    for (int t = ph; t < n; t += K) {
      const int jl = jl_n, rv = rv_n;
      const Site b = b_n;
      const float Fj = Fj_n;
      const double wj = w_n;
      if (t + K < n) {  // the next pair's gathers overlap this pair's points
        const size_t en = (size_t)(t + K) * rows + row;
        jl_n = oj[en];
        rv_n = orv[en];
        b_n = load_site(X, off + jl_n);
        Fj_n = aux[off + jl_n].x;
        if (side) w_n = xo[jl_n];
      }
      float d[3];
      pair_vec32(img, a, b, d);
      const float as = (float)a.s, bs = (float)b.s;
      PairSums2 p;
      if (list == 0) p = pair_grad<RULE, 0>(tab, d, as, bs, Fi, Fj);
      else if (list == 1) p = pair_grad<RULE, 1>(tab, d, as, bs, Fi, Fj);
      else if (list == 2) p = pair_grad<RULE, 2>(tab, d, as, bs, Fi, Fj);
      else p = pair_grad<RULE, 3>(tab, d, as, bs, Fi, Fj);
      e += wj * p.pe;
      g0 += wj * p.g[0]; g1 += wj * p.g[1]; g2 += wj * p.g[2];
      gs += wj * p.gsi;
      if (rv >= 0)
        prb[(size_t)rv * prows + jl] =
            make_float4((float)(-wme * p.g[0]), (float)(-wme * p.g[1]), (float)(-wme * p.g[2]), (float)(wme * p.gsj));
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

__global__ __launch_bounds__(TB) void k_grad_recv(const LatDev* __restrict__ lats, const int32_t* __restrict__ b2l) {
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
      const int list = pd ? side : 2 + side;
      const Ell& L = list == 0 ? S.pdpd : (list == 1 ? S.pdh : (list == 2 ? S.hpd : S.hh));
      const int n = L.rcnt[row];
      // shares arrive weighted by their senders
      for_row_shares<RECV_RU_WIDE>(gmem(reinterpret_cast<const float4*>(L.rbuf)), L.rows, row, n, [&](const float4& v) {
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

namespace {
// ---------------------------------------------------------------- tables
template <class Fn>
void fill_table(float4* out, Fn g) {
  const long double h = H64, eps = 1e-13L;
  for (int i = 0; i < NT; ++i) {
    const long double ra = (long double)R0 + i * h + eps, rb = (long double)R0 + (i + 1) * h - eps;
    const eam::VDT<long double> A = g(ra), B = g(rb);
    const long double c0 = A.v, c1 = h * A.d;
    const long double c2 = 3 * (B.v - A.v) - h * (2 * A.d + B.d);
    const long double c3 = 2 * (A.v - B.v) + h * (A.d + B.d);
    out[i] = make_float4((float)c0, (float)c1, (float)c2, (float)c3);
  }
}

}  // namespace

void host_tables(std::vector<float4>& t) {
  static const std::vector<float4> built = [] {  // once per process (~0.4 ms of long-double work)
    using LD = long double;
    std::vector<float4> u((size_t)NTAB * NT, make_float4(0.f, 0.f, 0.f, 0.f));
    fill_table(&u[T_FPD * NT], [](LD r) { return eam::f_Pd_vd<LD>(r); });
    fill_table(&u[T_FH * NT], [](LD r) { return eam::f_H_vd<LD>(r); });
    fill_table(&u[T_PHPD * NT], [](LD r) { return eam::phi_Pd_vd<LD>(r, eam::f_Pd_vd<LD>(r)); });
    fill_table(&u[T_PHPDH * NT], [](LD r) { return eam::phi_morse_vd<0, LD>(r); });
    fill_table(&u[T_PHH * NT], [](LD r) { return eam::phi_morse_vd<1, LD>(r); });
    return u;
  }();
  t = built;
}

int build_tables(a2c_fe_ctx* c) {
  DeviceTables t;
  A2C_HIP(c, device_tables(c->desc.device, &t));
  c->dtab = t.t32;
  if (!c->daux) A2C_HIP(c, dmalloc(&c->daux, sizeof(float2) * std::max(c->N, 1)));
  return A2C_OK;
}

int run_eval32(const Launch& L, Timing* tm) {
  if (L.mixed_full && L.rules == (1 << A2C_QUAD_R12)) {  // every pair from both ends (12-point rule only)
    if (tm) tm->mark(L.stream, "density32");
    if (L.nblk > 0) {
      hipLaunchKernelGGL(k_density32, dim3(L.nblk), dim3(TB), 0, L.stream, L.lats, L.b2l, L.tab);
      if (hipGetLastError() != hipSuccess) return A2C_EHIP;
    }
    if (tm) tm->mark(L.stream, "grad32");
    if (L.nblk > 0) {
      hipLaunchKernelGGL(k_grad32, dim3(L.nblk), dim3(TB), 0, L.stream, L.lats, L.b2l, L.tab);
      if (hipGetLastError() != hipSuccess) return A2C_EHIP;
    }
  } else if (L.nblk > 0) {
    if (tm) tm->mark(L.stream, "dens_own");
    if (L.rules & (1 << A2C_QUAD_R12))
      hipLaunchKernelGGL(k_dens_own<A2C_QUAD_R12>, dim3(L.onblk), dim3(TB), 0, L.stream, L.lats, L.ob2l, L.tab);
    if (L.rules & (1 << GH3))
      hipLaunchKernelGGL(k_dens_own<GH3>, dim3(L.onblk), dim3(TB), 0, L.stream, L.lats, L.ob2l, L.tab);
    if (hipGetLastError() != hipSuccess) return A2C_EHIP;
    if (tm) tm->mark(L.stream, "dens_recv");
    hipLaunchKernelGGL(k_dens_recv, dim3(L.nblk), dim3(TB), 0, L.stream, L.lats, L.b2l);
    if (hipGetLastError() != hipSuccess) return A2C_EHIP;
    if (tm) tm->mark(L.stream, "grad_own");
    if (L.rules & (1 << A2C_QUAD_R12))
      hipLaunchKernelGGL(k_grad_own<A2C_QUAD_R12>, dim3(L.onblk), dim3(TB), 0, L.stream, L.lats, L.ob2l, L.tab);
    if (L.rules & (1 << GH3))
      hipLaunchKernelGGL(k_grad_own<GH3>, dim3(L.onblk), dim3(TB), 0, L.stream, L.lats, L.ob2l, L.tab);
    if (hipGetLastError() != hipSuccess) return A2C_EHIP;
    if (tm) tm->mark(L.stream, "grad_recv");
    hipLaunchKernelGGL(k_grad_recv, dim3(L.nblk), dim3(TB), 0, L.stream, L.lats, L.b2l);
    if (hipGetLastError() != hipSuccess) return A2C_EHIP;
  }
  if (tm) tm->mark(L.stream, "final_sum");
  if (int rc = run_final_sum(L)) return rc;
  if (tm) tm->mark(L.stream, nullptr);
  return A2C_OK;
}

}  // namespace a2c

// Host-only: the FP32 tables of the mixed path (no GPU needed).  Layout: 5
// tables (f_Pd, f_H, phi_Pd, phi_PdH, phi_H) x 500 intervals x {c0,c1,c2,c3},
// interval k covers r in [1 + 0.01 k, 1 + 0.01 (k+1)), value
// c0 + u(c1 + u(c2 + u c3)) with u = (r - 1)/0.01 - k.
extern "C" int a2c_eam_tables(float* out, int32_t cap) {
  std::vector<float4> t;
  a2c::host_tables(t);
  const int n = (int)t.size() * 4;
  if (out && cap >= n) std::memcpy(out, t.data(), sizeof(float) * n);
  return n;
}
