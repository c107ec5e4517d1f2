// Objective, gradient, density and chemical potential of a context with a
// general model: a tabulated potential (a2c_fe_set_potential_*) and/or a
// node table (a2c_fe_set_quadrature).  The reference's loop nests
// (minimizer.cpp:140-157 and :162-351 for the objective and gradient,
// :588-802 for the chemical potential, electrondensity.cpp:12-56 for the
// density) with the EAM functions and the rule as data:
//
//   k_gen_density  one lane per site over its full lists: rho, F(rho),
//                  F'(rho), the embedding energy
//   k_gen_grad     one lane per site over its full lists: every pair's energy
//                  and coefficient c = F'_i f_j' + F'_j f_i' + phi' at each
//                  node, q- and sigma-gradient, the site's share of f, and a
//                  fixed-shape block sum (k_final_sum completes f)
//   k_gen_chem     per H site: dV/dx_a + the correction terms, gamma, cap
//
// A node k of a 6-D table puts the pair vector at r_k = d + a_i xi1_k -
// a_j xi2_k with a = sqrt(2) sigma, so dr_k/dsigma_i = sqrt(2) xi1_k; a
// relative-coordinate rule (Gauss-Hermite) at r_k = d + s z_k with
// s^2 = sigma_i^2 + sigma_j^2, dr_k/dsigma_i = (sigma_i / s) z_k.  Both ends of
// a pair see the same average when the table is swap-symmetric (enforced
// when it is set), which the reference's pair terms assume.
//
// Precision: FP64 throughout, or (MIXED) the pair terms of each node in
// FP32 from FP32 pieces of the r-functions, summed per pair in FP32 and per
// site in FP64.  F(rho) and F'(rho) are FP64 in both.
#include <cmath>

#include "eam64.hpp"
#include "fe_internal.hpp"

namespace a2c {
namespace {

constexpr int TB = SITE_BLOCK;
constexpr double KB = 8.6173324e-5;  // minimizer.cpp:92
constexpr double SQ2 = 1.4142135623730950488;

struct Site {
  double q[3];
  double s;
};
__device__ __forceinline__ Site load_site(const double* __restrict__ X, int s) {
  const double4 v = *reinterpret_cast<const double4*>(X + 4 * (size_t)s);
  return {{v.x, v.y, v.z}, v.w};
}
__device__ __forceinline__ void pair_vec(const LatDev& S, const Site& a, const Site& b, double d[3]) {
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    double qj = b.q[c];
    if (S.periodic[c]) qj = qj - S.box[c] * rint((qj - a.q[c]) * S.ibox[c]);
    d[c] = a.q[c] - qj;
  }
}

// ------------------------------------------------------------ pieces
// Value and slope of a piecewise cubic at x (coefficients c0..c3 in the
// local coordinate u of piece k).  Past the last node: 0 for the r-functions,
// the end tangent for the embedding functions (emb).
template <class R, class C>
__device__ __forceinline__ eam::VDT<R> piece_vd(const GenFn& F, const C* __restrict__ tab, R x, bool emb) {
  const R p = x * (R)F.ih;
  int k = (int)p;
  if (k < 0) k = 0;
  if (k >= F.n - 1) {
    if (emb) return {(R)F.yend + (R)F.dend * (x - (R)(F.n - 1) / (R)F.ih), (R)F.dend};
    return {R(0), R(0)};
  }
  const R u = p - (R)k;
  const C c = tab[F.off + k];
  const R v = (R)c.x + u * ((R)c.y + u * ((R)c.z + u * (R)c.w));
  const R d = (R)c.y + u * (R(2) * (R)c.z + u * R(3) * (R)c.w);
  return {v, d * (R)F.ih};
}

// Embedding functions in FP64: closed forms or pieces.
template <int POT>
__device__ __forceinline__ double F_val(const GenDev& g, bool pd, double rho) {
  if (POT == 0) return pd ? eam::F_Pd(rho) : eam::F_H(rho);
  return piece_vd<double>(g.fn[pd ? GEN_F_PD : GEN_F_H], g.tab, rho, true).v;
}
template <int POT, int DM>
__device__ __forceinline__ double F_der(const GenDev& g, bool pd, double rho, double eps) {
  if (DM == A2C_DERIV_FD) return (F_val<POT>(g, pd, rho + eps) - F_val<POT>(g, pd, rho - eps)) / (2.0 * eps);
  if (POT == 0) return pd ? eam::F_Pd_vd(rho).d : eam::F_H_vd(rho).d;
  return piece_vd<double>(g.fn[pd ? GEN_F_PD : GEN_F_H], g.tab, rho, true).d;
}

// r-functions: density f (fn GEN_RHO_*) and pair phi (fn GEN_RPHI_*, stored
// as r phi).  POT 0 in FP64: the closed forms; otherwise the pieces (FP64, or
// FP32 for R = float).
template <int POT, class R>
struct RFun {
  const GenDev& g;
  __device__ __forceinline__ R val(int fn, R r) const {
    if (POT == 0 && sizeof(R) == 8) {
      const double x = (double)r;
      switch (fn) {
        case GEN_RHO_PD: return (R)eam::f_Pd(x);
        case GEN_RHO_H: return (R)eam::f_H(x);
        case GEN_RPHI_PD: return (R)eam::phi_Pd(x);
        case GEN_RPHI_PDH: return (R)eam::phi_morse<0>(x);
        default: return (R)eam::phi_morse<1>(x);
      }
    }
    const eam::VDT<R> z = look(fn, r);
    return fn >= GEN_RPHI_PD ? z.v / r : z.v;
  }
  __device__ __forceinline__ eam::VDT<R> vd(int fn, R r) const {
    if (POT == 0 && sizeof(R) == 8) {
      const double x = (double)r;
      eam::VD o;
      switch (fn) {
        case GEN_RHO_PD: o = eam::f_Pd_vd(x); break;
        case GEN_RHO_H: o = eam::f_H_vd(x); break;
        case GEN_RPHI_PD: o = eam::phi_Pd_vd(x, eam::f_Pd_vd(x)); break;
        case GEN_RPHI_PDH: o = eam::phi_morse_vd<0>(x); break;
        default: o = eam::phi_morse_vd<1>(x); break;
      }
      return {(R)o.v, (R)o.d};
    }
    const eam::VDT<R> z = look(fn, r);
    if (fn < GEN_RPHI_PD) return z;
    const R phi = z.v / r;
    return {phi, (z.d - phi) / r};  // (r phi)' = phi + r phi'
  }
  __device__ __forceinline__ eam::VDT<R> look(int fn, R r) const {
    if (sizeof(R) == 8) return piece_vd<R>(g.fn[fn], g.tab, r, false);
    return piece_vd<R>(g.fn32[fn], g.tab32, r, false);
  }
  // slope by the reference's central difference (DM = FD) or analytically
  template <int DM>
  __device__ __forceinline__ eam::VDT<R> vd_dm(int fn, R r, R eps) const {
    if (DM == A2C_DERIV_FD) return {val(fn, r), (val(fn, r + eps) - val(fn, r - eps)) / (R(2) * eps)};
    return vd(fn, r);
  }
};

// The rule's nodes staged in LDS: xi1 (or z), xi2, w as R.
template <class R>
__device__ __forceinline__ void stage_nodes(R* nd, const GenDev& g) {
  for (int i = threadIdx.x; i < g.K * 7; i += blockDim.x) nd[i] = (R)g.nodes[(i / 7) * 8 + i % 7];
  __syncthreads();
}

// Node k of a pair: r_k (in R) and the factor that turns r . xi1 into
// r . dr/dsigma_i.
template <class R>
struct PairFrame {
  R d[3], a, b;  // r_k = d + a xi1_k - b xi2_k
  R fac;         // dr/dsigma_i = fac * xi1_k
};
template <class R>
__device__ __forceinline__ PairFrame<R> frame(const GenDev& g, const double d[3], double si, double sj) {
  PairFrame<R> f;
  f.d[0] = (R)d[0];
  f.d[1] = (R)d[1];
  f.d[2] = (R)d[2];
  if (g.rel) {
    const double s = sqrt(si * si + sj * sj);
    f.a = (R)s;
    f.b = R(0);
    f.fac = (R)(si / s);
  } else {
    f.a = (R)(SQ2 * si);
    f.b = (R)(SQ2 * sj);
    f.fac = (R)SQ2;
  }
  return f;
}
template <class R>
__device__ __forceinline__ void node_r(const PairFrame<R>& f, const R* nd, int k, R r[3]) {
  const R* n = nd + 7 * k;
#pragma unroll
  for (int c = 0; c < 3; ++c) r[c] = f.d[c] + f.a * n[c] - f.b * n[3 + c];
}

// ------------------------------------------------------------ density
template <int POT, int DM, class R>
__global__ __launch_bounds__(TB) void k_gen_density(const LatDev* __restrict__ lat, GenDev g, double eps) {
  __shared__ R nd[A2C_QUAD_MAX_NODES * 7];
  stage_nodes(nd, g);
  const LatDev& S = *lat;
  const int s = xcd_block() * TB + threadIdx.x;
  const int N = S.npd + S.nh;
  if (s >= N) return;
  const RFun<POT, R> fun{g};
  const double* __restrict__ X = S.X;
  const double* __restrict__ xo = S.x;
  const bool pd = s < S.npd;
  const int row = pd ? s : s - S.npd;
  const Site a = load_site(X, s);
  const int K = g.K;
  double rho = 0.0;
  for (int side = 0; side < 2; ++side) {  // Pd neighbors, then H neighbors (electrondensity.cpp:20-48)
    const Ell L = side == 0 ? (pd ? S.pdpd : S.hpd) : (pd ? S.pdh : S.hh);
    const int fn = side == 0 ? GEN_RHO_PD : GEN_RHO_H;
    const int n = L.cnt[row];
    for (int t = 0; t < n; ++t) {
      const int jl = L.idx[(size_t)t * L.rows + row];
      const int j = side == 0 ? jl : S.npd + jl;
      const Site b = load_site(X, j);
      double d[3];
      pair_vec(S, a, b, d);
      const PairFrame<R> f = frame<R>(g, d, a.s, b.s);
      R acc = R(0);
      for (int k = 0; k < K; ++k) {
        R r[3];
        node_r(f, nd, k, r);
        acc += nd[7 * k + 6] * fun.val(fn, sqrt(r[0] * r[0] + r[1] * r[1] + r[2] * r[2]));
      }
      rho += (side == 0 ? 1.0 : xo[jl]) * (double)acc;
    }
  }
  S.rho[s] = rho;
  const double F = F_val<POT>(g, pd, rho);
  S.Fp[s] = F_der<POT, DM>(g, pd, rho, eps);
  S.emb[s] = pd ? F : xo[row] * F;
}

// ------------------------------------------------------------ gradient
// Kinds: 0 Pd-Pd, 1 Pd site / H neighbor, 2 H site / Pd neighbor, 3 H-H.
template <int POT, int DM, class R>
__device__ __forceinline__ void node_terms(const RFun<POT, R>& fun, int kind, R rn, R Fi, R Fj, R eps, R& phi,
                                           R& c) {
  if (kind == 0) {
    const eam::VDT<R> f = fun.template vd_dm<DM>(GEN_RHO_PD, rn, eps);
    const eam::VDT<R> p = fun.template vd_dm<DM>(GEN_RPHI_PD, rn, eps);
    phi = p.v;
    c = (Fi + Fj) * f.d + p.d;
  } else if (kind == 3) {
    const eam::VDT<R> h = fun.template vd_dm<DM>(GEN_RHO_H, rn, eps);
    const eam::VDT<R> p = fun.template vd_dm<DM>(GEN_RPHI_H, rn, eps);
    phi = p.v;
    c = (Fi + Fj) * h.d + p.d;
  } else {
    const eam::VDT<R> f = fun.template vd_dm<DM>(GEN_RHO_PD, rn, eps);
    const eam::VDT<R> h = fun.template vd_dm<DM>(GEN_RHO_H, rn, eps);
    const eam::VDT<R> p = fun.template vd_dm<DM>(GEN_RPHI_PDH, rn, eps);
    phi = p.v;
    // site i's density takes the neighbor's f, the neighbor's takes i's
    c = kind == 1 ? Fi * h.d + Fj * f.d + p.d : Fi * f.d + Fj * h.d + p.d;
  }
}

template <int POT, int DM, class R>
__global__ __launch_bounds__(TB) void k_gen_grad(const LatDev* __restrict__ lat, GenDev g, double eps) {
  __shared__ R nd[A2C_QUAD_MAX_NODES * 7];
  __shared__ double red[TB];
  stage_nodes(nd, g);
  const LatDev& S = *lat;
  const int lb = xcd_block();
  const int s = lb * TB + threadIdx.x;
  const int N = S.npd + S.nh;
  double fs = 0.0;
  if (s < N) {
    const RFun<POT, R> fun{g};
    const double* __restrict__ X = S.X;
    const double* __restrict__ xo = S.x;
    const double* __restrict__ Fp = S.Fp;
    const bool pd = s < S.npd;
    const int row = pd ? s : s - S.npd;
    const Site a = load_site(X, s);
    const double Fi = Fp[s];
    const double xi = pd ? 1.0 : xo[row];
    const int K = g.K;
    double e = 0.0, gq[3] = {0.0, 0.0, 0.0}, gs = 0.0;
    for (int side = 0; side < 2; ++side) {
      const Ell L = side == 0 ? (pd ? S.pdpd : S.hpd) : (pd ? S.pdh : S.hh);
      const int kind = side == 0 ? (pd ? 0 : 2) : (pd ? 1 : 3);
      const int n = L.cnt[row];
      for (int t = 0; t < n; ++t) {
        const int jl = L.idx[(size_t)t * L.rows + row];
        const int j = side == 0 ? jl : S.npd + jl;
        const Site b = load_site(X, j);
        double d[3];
        pair_vec(S, a, b, d);
        const PairFrame<R> f = frame<R>(g, d, a.s, b.s);
        R pe = R(0), px = R(0), py = R(0), pz = R(0), ps = R(0);
        for (int k = 0; k < K; ++k) {
          R r[3];
          node_r(f, nd, k, r);
          const R rn = sqrt(r[0] * r[0] + r[1] * r[1] + r[2] * r[2]);
          R phi, c;
          node_terms<POT, DM, R>(fun, kind, rn, (R)Fi, (R)Fp[j], (R)eps, phi, c);
          const R w = nd[7 * k + 6];
          pe += w * phi;
          const R u = w * c / rn;
          px += u * r[0];
          py += u * r[1];
          pz += u * r[2];
          ps += u * (r[0] * nd[7 * k] + r[1] * nd[7 * k + 1] + r[2] * nd[7 * k + 2]);
        }
        const double wj = side == 0 ? 1.0 : xo[jl];  // H neighbors weigh by their occupancy
        e += wj * (double)pe;
        gq[0] += wj * (double)px;
        gq[1] += wj * (double)py;
        gq[2] += wj * (double)pz;
        gs += wj * (double)(ps * f.fac);
      }
    }
    // V share of site s: embedding + half of every pair (minimizer.cpp:169-286)
    const double V = S.emb[s] + 0.5 * xi * e;
    const double T = S.T;
    const double sc = xi / T;
    double* __restrict__ G = S.G;
    G[4 * (size_t)s + 0] = sc * gq[0];
    G[4 * (size_t)s + 1] = sc * gq[1];
    G[4 * (size_t)s + 2] = sc * gq[2];
    G[4 * (size_t)s + 3] = sc * gs - 3.0 * KB * xi / a.s;
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

// ------------------------------------------------------------ chemical potential
// minimizer.cpp:588-802 per H site a (FP64): F_H(rho_a) + the H-Pd and H-H
// pair sums with the neighbor's F' + the correction terms (corr from
// k_chem_pre), gamma = ln(x/(1-x)) + sum/(kB T) + 1.5 for 0 < x < 1, capped.
template <int POT>
__global__ __launch_bounds__(TB) void k_gen_chem(const LatDev* __restrict__ lat, GenDev g,
                                                 const double* __restrict__ corr, double ch0, double ch1, double ch2,
                                                 double ch3, double gamma_bd, double* __restrict__ gamma,
                                                 int32_t* __restrict__ full) {
  __shared__ double nd[A2C_QUAD_MAX_NODES * 7];
  stage_nodes(nd, g);
  const LatDev& S = *lat;
  const int a = blockIdx.x * TB + threadIdx.x;
  if (a >= S.nh || full[a] == 1) return;  // fully occupied reservoir: gamma untouched
  const RFun<POT, double> fun{g};
  const double* __restrict__ X = S.X;
  const double* __restrict__ xo = S.x;
  const double* __restrict__ Fp = S.Fp;
  const int s = S.npd + a;
  const Site si = load_site(X, s);
  const double T = S.T;
  const int K = g.K;
  double sum = F_val<POT>(g, false, S.rho[s]);
  const double n = (double)S.hpd.cnt[a];
  double Xa = 0.0;
  const int nhh = S.hh.cnt[a];
  for (int t = 0; t < nhh; ++t) Xa += xo[S.hh.idx[(size_t)t * S.hh.rows + a]];
  sum += 0.5 * T * ch0 * (ch1 * n * n + 2.0 * ch2 * n * Xa + ch3 * Xa * Xa);
  for (int side = 0; side < 2; ++side) {  // H-Pd (:679-698), then H-H (:701-719)
    const Ell L = side == 0 ? S.hpd : S.hh;
    const int m = L.cnt[a];
    for (int t = 0; t < m; ++t) {
      const int jl = L.idx[(size_t)t * L.rows + a];
      const int j = side == 0 ? jl : S.npd + jl;
      const Site b = load_site(X, j);
      double d[3];
      pair_vec(S, si, b, d);
      const PairFrame<double> f = frame<double>(g, d, si.s, b.s);
      double sf = 0.0, sp = 0.0;
      for (int k = 0; k < K; ++k) {
        double r[3];
        node_r(f, nd, k, r);
        const double rn = sqrt(r[0] * r[0] + r[1] * r[1] + r[2] * r[2]);
        const double w = nd[7 * k + 6];
        sf += w * fun.val(GEN_RHO_H, rn);
        sp += w * fun.val(side == 0 ? GEN_RPHI_PDH : GEN_RPHI_H, rn);
      }
      const double term = Fp[j] * sf + sp + corr[j];
      sum += side == 0 ? term : xo[jl] * term;
    }
  }
  const double x = xo[a];
  const double raw = (x > 0.0 && x < 1.0) ? log(x / (1.0 - x)) + sum / (KB * T) : 0.0;
  double gm = raw + 1.5;
  if (gm > gamma_bd) {
    gm = gamma_bd;
    if (x > 0.9999) full[a] = 1;
  }
  gamma[a] = gm;
}

// The kernel instance for (potential, derivative mode, precision).  The
// FP32 instances read the r-functions from FP32 pieces whatever the
// potential (the closed forms have pieces of their own, general.cpp) and F
// like the FP64 ones.  Central differences of step ~1e-5 are meaningless in
// FP32, so FD mode runs the FP64 instances in either precision.
template <template <int, int, class> class Pick>
auto pick(const GenDev& g, int deriv_mode, bool fp32) {
  const bool fd = deriv_mode == A2C_DERIV_FD, tab = g.pot == 1;
  if (fd) return tab ? Pick<1, A2C_DERIV_FD, double>::fn : Pick<0, A2C_DERIV_FD, double>::fn;
  if (fp32) return tab ? Pick<1, A2C_DERIV_ANALYTIC, float>::fn : Pick<0, A2C_DERIV_ANALYTIC, float>::fn;
  return tab ? Pick<1, A2C_DERIV_ANALYTIC, double>::fn : Pick<0, A2C_DERIV_ANALYTIC, double>::fn;
}
template <int P, int D, class R> struct DensK { static constexpr auto fn = k_gen_density<P, D, R>; };
template <int P, int D, class R> struct GradK { static constexpr auto fn = k_gen_grad<P, D, R>; };

}  // namespace

int run_density_gen(const LatDev* dlat, int N, const GenDev& g, int deriv_mode, double eps, hipStream_t st) {
  if (N <= 0) return A2C_OK;
  // the density feeding F' is FP64 in both precisions
  hipLaunchKernelGGL(pick<DensK>(g, deriv_mode, false), dim3(site_blocks(N)), dim3(TB), 0, st, dlat, g, eps);
  return hipGetLastError() == hipSuccess ? A2C_OK : A2C_EHIP;
}

int run_eval_gen(const LatDev* dlat, int N, const GenDev& g, int deriv_mode, double eps, bool fp32, hipStream_t st,
                 Timing* tm) {
  const int nblk = site_blocks(N);
  if (nblk > 0) {
    if (tm) tm->mark(st, "gen_density");
    hipLaunchKernelGGL(pick<DensK>(g, deriv_mode, fp32), dim3(nblk), dim3(TB), 0, st, dlat, g, eps);
    if (tm) tm->mark(st, "gen_grad");
    hipLaunchKernelGGL(pick<GradK>(g, deriv_mode, fp32), dim3(nblk), dim3(TB), 0, st, dlat, g, eps);
    if (hipGetLastError() != hipSuccess) return A2C_EHIP;
  }
  if (tm) tm->mark(st, "final_sum");
  Launch L{};
  L.lats = dlat;
  L.nlat = 1;
  L.stream = st;
  if (int rc = run_final_sum(L)) return rc;
  if (tm) tm->mark(st, nullptr);
  return A2C_OK;
}

int launch_density_gen(a2c_fe_ctx* c, const double* dX) {
  if (int rc = sync_lattice(c, dX, c->dG, c->dF)) return rc;
  if (run_density_gen(c->dlat, c->N, c->gen, c->desc.deriv_mode, c->desc.fd_eps, c->stream))
    return set_err(c, A2C_EHIP, "general density launch failed");
  return A2C_OK;
}

int launch_eval_gen(a2c_fe_ctx* c, const double* dX, double* dG, double* dF, bool timed, Timing* tm) {
  if (int rc = sync_lattice(c, dX, dG, dF)) return rc;
  if (run_eval_gen(c->dlat, c->N, c->gen, c->desc.deriv_mode, c->desc.fd_eps,
                   c->desc.precision == A2C_PREC_MIXED, c->stream, timed ? tm : nullptr))
    return set_err(c, A2C_EHIP, "general eval launch failed");
  return A2C_OK;
}

int launch_chem_gen(a2c_fe_ctx* c, const double* dX, const double* coe_h, double gamma_bd, double* dgam,
                    int32_t* dfull, const double* dcorr) {
  (void)dX;  // the descriptor already points at dX (launch_density_gen)
  const int nh = c->desc.n_h;
  if (nh <= 0) return A2C_OK;
  hipLaunchKernelGGL(c->gen.pot == 1 ? k_gen_chem<1> : k_gen_chem<0>, dim3((nh + TB - 1) / TB), dim3(TB), 0,
                     c->stream, (const LatDev*)c->dlat, c->gen, dcorr, coe_h[0], coe_h[1], coe_h[2], coe_h[3],
                     gamma_bd, dgam, dfull);
  A2C_HIP(c, hipGetLastError());
  return A2C_OK;
}

}  // namespace a2c
