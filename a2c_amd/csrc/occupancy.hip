// Occupancy kernels (FP64):
//   chemical potential  Minimizer::calculateChemicalPotential
//                       (minimizer.cpp:588-802)
//   diffusion step      integrateDiffusion (integrate.cpp:9-87)
//
// Chemical potential of H site a (A.5 of the reference's algebra):
//   S_a = F_H(rho_a) + 1/2 T cH0 (cH1 n^2 + 2 cH2 n X_a + cH3 X_a^2)
//       + sum_{j in HPd(a)} [ sum_k w (F'_Pd(rho_j) f_H + phi_PdH) + P_j ]
//       + sum_{b in HH(a)} x_b [ sum_k w (F'_H(rho_b) f_H + phi_H) + Q_b ]
//   n = |HPd(a)|, X_a = sum_{b in HH(a)} x_b,
//   P_j = T cP0 (cP2 |PdPd(j)| + cP3 sum_{l in PdH(j)} x_l),
//   Q_b = T cH0 (cH2 |HPd(b)| + cH3 sum_{c in HH(b)} x_c).
// The reference's double loop over HH(a) for the cH3 term (:671-675) is the
// square X_a^2; P_j and Q_b are computed once per site (k_chem_pre) instead of
// once per (site, neighbor).  gamma = ln(x/(1-x)) + S_a/(kB T) + 1.5 for
// 0 < x < 1 (else 1.5), capped at gamma_bd with full_H = 1 when x > 0.9999
// (:721-722, :793-800); full_H sites keep gamma.  The pair sums sum_k w use
// the context's quadrature rule (12-point, w = 1/12, or Gauss-Hermite).
//
// Diffusion: f_a = gamma_a - ln(x_a/(1-x_a)) (k_diff_f, once per site), then
// x_a += (rate x_deriv) dt with x_deriv = sum over HH1(a) of
// x_b(1-x_a)e^{-(f_a-f_b)/2} - x_a(1-x_b)e^{(f_a-f_b)/2} (k_diff_step).  With
// h = e^{f/2}, U = x h and V = (1-x)/h per site (one exp per site in
// k_diff_f) the edge term is V_a U_b - U_a V_b: no exp per edge.  A site with
// |f| >= 700 (h near the end of the double range) has U = V = NaN, and its
// edges take the reference's form.  Either form is evaluated from both ends
// with the same operands and without contraction, so the edge term is exactly
// antisymmetric; the total changes only by the rounding of the site sums (plus
// clamps and full_H reservoirs, as in the reference).
#include <cmath>
#include <mutex>
#include <string>

#include "eam64.hpp"
#include "fe_internal.hpp"

namespace a2c {
namespace {

constexpr int TB = 256;
constexpr double QW = 1.0 / 12.0;
constexpr double SQ6 = 2.449489742783178098;

struct Site {
  double q[3];
  double s;
};
__device__ __forceinline__ Site load_site(const double* __restrict__ X, int s) {
  const double4 v = *reinterpret_cast<const double4*>(X + 4 * (size_t)s);
  return {{v.x, v.y, v.z}, v.w};
}
__device__ __forceinline__ void pair_vec(const DevSys& S, const Site& a, const Site& b, double d[3]) {
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    double qj = b.q[c];
    if (S.periodic[c]) qj = qj - S.box[c] * rint((qj - a.q[c]) * S.ibox[c]);
    d[c] = a.q[c] - qj;
  }
}

// Sums of f_H and of phi (phi_PdH or phi_H) over the nodes of one pair:
// unweighted over the 12 points (the caller applies 1/12), weighted over the
// 27 Gauss-Hermite nodes.
template <int HH, int RULE>
__device__ __forceinline__ void pair_fh_phi(const double d[3], double si, double sj, double& sf, double& sp) {
  sf = 0.0;
  sp = 0.0;
  if (RULE == A2C_QUAD_GH3) {
    const double sd = sqrt(si * si + sj * sj);
#pragma unroll 1
    for (int k = 0; k < 27; ++k) {
      double z[3], w;
      gh_node(k, z, w);
      const double r0 = d[0] + sd * z[0], r1 = d[1] + sd * z[1], r2 = d[2] + sd * z[2];
      const double rn = sqrt(r0 * r0 + r1 * r1 + r2 * r2);
      sf += w * eam::f_H(rn);
      sp += w * eam::phi_morse<HH>(rn);
    }
    return;
  }
  const double ai = SQ6 * si, aj = SQ6 * sj;
#pragma unroll
  for (int p = 0; p < 12; ++p) {
    const int c = (p >> 1) % 3;
    const double sg = (p & 1) ? -1.0 : 1.0;
    double r[3] = {d[0], d[1], d[2]};
    r[c] += (p < 6) ? sg * ai : -sg * aj;
    const double rn = sqrt(r[0] * r[0] + r[1] * r[1] + r[2] * r[2]);
    sf += eam::f_H(rn);
    sp += eam::phi_morse<HH>(rn);
  }
}

// Per-site correction sums: P_j for Pd sites, Q_b for H sites.
__global__ void k_chem_pre(DevSys S, const double* __restrict__ xo, double T, double cp0, double cp2, double cp3,
                           double ch0, double ch2, double ch3, double* __restrict__ corr) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  const int N = S.npd + S.nh;
  if (s >= N) return;
  const bool pd = s < S.npd;
  const int row = pd ? s : s - S.npd;
  const Ell LP = pd ? S.pdpd : S.hpd;
  const Ell LH = pd ? S.pdh : S.hh;
  double X = 0.0;
  const int nh = LH.cnt[row];
  for (int t = 0; t < nh; ++t) X += xo[LH.idx[(size_t)t * LH.rows + row]];
  const double n = (double)LP.cnt[row];
  corr[s] = pd ? T * cp0 * (cp2 * n + cp3 * X) : T * ch0 * (ch2 * n + ch3 * X);
}

template <int RULE>
__global__ __launch_bounds__(TB) void k_chem(DevSys S, const double* __restrict__ X, const double* __restrict__ xo,
                                             const double* __restrict__ rho, const double* __restrict__ Fp,
                                             const double* __restrict__ corr, double ch0, double ch1, double ch2,
                                             double ch3, double gamma_bd, double* __restrict__ gamma,
                                             int32_t* __restrict__ full) {
  const int a = blockIdx.x * blockDim.x + threadIdx.x;
  if (a >= S.nh) return;
  if (full[a] == 1) return;  // fully occupied reservoir: gamma untouched
  const int s = S.npd + a;
  const Site si = load_site(X, s);
  constexpr double qw = RULE == A2C_QUAD_GH3 ? 1.0 : QW;
  const double T = S.T;
  double sum = eam::F_H(rho[s]);
  // correction of the site itself
  const double n = (double)S.hpd.cnt[a];
  double Xa = 0.0;
  const int nhh = S.hh.cnt[a];
  for (int t = 0; t < nhh; ++t) Xa += xo[S.hh.idx[(size_t)t * S.hh.rows + a]];
  sum += 0.5 * T * ch0 * (ch1 * n * n + 2.0 * ch2 * n * Xa + ch3 * Xa * Xa);
  const int npn = S.hpd.cnt[a];
  for (int t = 0; t < npn; ++t) {  // H-Pd pairs (minimizer.cpp:679-698)
    const int j = S.hpd.idx[(size_t)t * S.hpd.rows + a];
    const Site b = load_site(X, j);
    double d[3];
    pair_vec(S, si, b, d);
    double sf, sp;
    pair_fh_phi<0, RULE>(d, si.s, b.s, sf, sp);
    sum += qw * (Fp[j] * sf + sp) + corr[j];
  }
  for (int t = 0; t < nhh; ++t) {  // H-H pairs (:701-719)
    const int h = S.hh.idx[(size_t)t * S.hh.rows + a];
    const Site b = load_site(X, S.npd + h);
    double d[3];
    pair_vec(S, si, b, d);
    double sf, sp;
    pair_fh_phi<1, RULE>(d, si.s, b.s, sf, sp);
    sum += xo[h] * (qw * (Fp[S.npd + h] * sf + sp) + corr[S.npd + h]);
  }
  const double x = xo[a];
  const double raw = (x > 0.0 && x < 1.0) ? log(x / (1.0 - x)) + sum / (KB * T) : 0.0;
  double g = raw + 1.5;
  if (g > gamma_bd) {
    g = gamma_bd;
    if (x > 0.9999) full[a] = 1;
  }
  gamma[a] = g;
}

constexpr double F_FACTOR_MAX = 700.0;  // |f| below which e^{+-f/2} and the edge products stay finite

__device__ __forceinline__ double diff_f(double gamma, double x) { return gamma - log(x / (1.0 - x)); }

__global__ void k_diff_f(int nh, const double* __restrict__ gamma, const double* __restrict__ xo,
                         double2* __restrict__ uv) {
  const int a = blockIdx.x * blockDim.x + threadIdx.x;
  if (a >= nh) return;
  const double x = xo[a];
  const double fa = diff_f(gamma[a], x);
  double2 w = make_double2(NAN, NAN);
  if (fabs(fa) < F_FACTOR_MAX) {  // false for NaN
    const double h = exp(0.5 * fa);
    w = make_double2(x * h, (1.0 - x) / h);
  }
  uv[a] = w;
}

// The reference's edge term (integrate.cpp:40-47) of (a, b), seen from a:
// for the edges of sites whose factors are NaN (out of line: rare).
__device__ __attribute__((noinline)) double edge_ref(const double* __restrict__ gamma,
                                                    const double* __restrict__ xo, int a, int b) {
#pragma clang fp contract(off)
  const double xa = xo[a], xb = xo[b];
  const double df = diff_f(gamma[a], xa) - diff_f(gamma[b], xb);
  return xb * (1.0 - xa) * exp(-df / 2.0) - xa * (1.0 - xb) * exp(df / 2.0);
}

// One explicit step per H site, and per block the four balance sums of the
// step (fixed-order block trees -> part[4 * block]): the increments applied
// (non-reservoir sites), the absolute edge terms of every site (the scale of
// the rounding), the clamp lifts (clamped minus unclamped occupancy) and the
// H the reservoirs supplied (minus the increments a full_H site would have
// received; it keeps x).  The edge terms are exactly antisymmetric, so the
// applied increments equal the reservoirs' supply up to the rounding of the
// site sums.
__global__ __launch_bounds__(TB) void k_diff_step(Ell hh1, const double* __restrict__ xo,
                                                  const double* __restrict__ gamma, const double2* __restrict__ uv,
                                                  const int32_t* __restrict__ full, double pref, double dt,
                                                  double xLb, double xUb, double* __restrict__ xn,
                                                  double* __restrict__ part) {
#pragma clang fp contract(off)
  __shared__ double red[4][TB];
  const int a = blockIdx.x * blockDim.x + threadIdx.x;
  double inc = 0.0, gross = 0.0, lift = 0.0, supply = 0.0;
  if (a < hh1.rows) {
    const double xa = xo[a];
    const double2 wa = uv[a];
    double xd = 0.0, xg = 0.0;
    const int n = hh1.cnt[a];
    const size_t rows = hh1.rows;
    constexpr int B = 4;  // edges whose gathers are in flight together (8 or 12: slower at 64^3 cells)
    for (int t0 = 0; t0 < n; t0 += B) {
      int b[B];
      double2 wb[B];
#pragma unroll
      for (int k = 0; k < B; ++k) b[k] = t0 + k < n ? hh1.idx[(size_t)(t0 + k) * rows + a] : -1;
#pragma unroll
      for (int k = 0; k < B; ++k) wb[k] = b[k] >= 0 ? uv[b[k]] : make_double2(0.0, 0.0);
#pragma unroll
      for (int k = 0; k < B; ++k) {
        if (b[k] < 0) continue;
        const double e = (isnan(wa.x) || isnan(wb[k].x)) ? edge_ref(gamma, xo, a, b[k])
                                                          : wa.y * wb[k].x - wa.x * wb[k].y;
        xd += e;
        xg += fabs(e);
      }
    }
    gross = pref * xg * dt;
    if (full[a] == 1) {  // reservoir: keeps x (integrate.cpp:50, :62, :77)
      supply = -(pref * xd * dt);
      xn[a] = xa;
    } else {
      inc = pref * xd * dt;
      const double v0 = xa + inc;
      double v = v0;
      if (v < xLb) v = xLb;
      else if (v > xUb) v = xUb;
      lift = v - v0;
      xn[a] = v;
    }
  }
  red[0][threadIdx.x] = inc;
  red[1][threadIdx.x] = gross;
  red[2][threadIdx.x] = lift;
  red[3][threadIdx.x] = supply;
  __syncthreads();
  for (int w = TB / 2; w > 0; w >>= 1) {
    if (threadIdx.x < w)
#pragma unroll
      for (int k = 0; k < 4; ++k) red[k][threadIdx.x] += red[k][threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x < 4) part[4 * (size_t)blockIdx.x + threadIdx.x] = red[threadIdx.x][0];
}

// The block sums of k_diff_step (one workgroup of BT lanes, each summing the
// blocks b = lane + BT i in order, four blocks' loads in flight; then a fixed
// tree): bal[0..3].
constexpr int BT = 1024;
__global__ __launch_bounds__(BT) void k_diff_balance(const double4* __restrict__ part, int nblk,
                                                     double* __restrict__ bal) {
  __shared__ double red[4][BT];
  double v[4] = {0.0, 0.0, 0.0, 0.0};
  for (int b0 = threadIdx.x; b0 < nblk; b0 += 4 * BT) {
    double4 p[4];
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const int b = b0 + i * BT;
      p[i] = b < nblk ? part[b] : make_double4(0.0, 0.0, 0.0, 0.0);
    }
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      v[0] += p[i].x;
      v[1] += p[i].y;
      v[2] += p[i].z;
      v[3] += p[i].w;
    }
  }
#pragma unroll
  for (int k = 0; k < 4; ++k) red[k][threadIdx.x] = v[k];
  __syncthreads();
  for (int w = BT / 2; w > 0; w >>= 1) {
    if (threadIdx.x < w)
#pragma unroll
      for (int k = 0; k < 4; ++k) red[k][threadIdx.x] += red[k][threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x < 4) bal[threadIdx.x] = red[threadIdx.x][0];
}

}  // namespace

int launch_chem(a2c_fe_ctx* c, const double* dX, const double* coe_pd, const double* coe_h, double gamma_bd,
                double* dgam, int32_t* dfull, double* dcorr) {
  const DevSys S = c->sys();
  const int N = c->N;
  if (int rc = launch_density64(c, dX)) return rc;  // rho, F'(rho) at X
  if (N > 0) {
    hipLaunchKernelGGL(k_chem_pre, dim3((N + TB - 1) / TB), dim3(TB), 0, c->stream, S, c->dx, S.T, coe_pd[0],
                       coe_pd[2], coe_pd[3], coe_h[0], coe_h[2], coe_h[3], dcorr);
    A2C_HIP(c, hipGetLastError());
  }
  if (S.nh > 0 && c->gen_on) return launch_chem_gen(c, dX, coe_h, gamma_bd, dgam, dfull, dcorr);
  if (S.nh > 0) {
    hipLaunchKernelGGL(S.rule == A2C_QUAD_GH3 ? k_chem<A2C_QUAD_GH3> : k_chem<A2C_QUAD_R12>,
                       dim3((S.nh + TB - 1) / TB), dim3(TB), 0, c->stream, S, dX, c->dx, c->drho, c->dFp, dcorr,
                       coe_h[0], coe_h[1], coe_h[2], coe_h[3], gamma_bd, dgam, dfull);
    A2C_HIP(c, hipGetLastError());
  }
  return A2C_OK;
}

int diffuse_blocks(int nh) { return (nh + TB - 1) / TB; }

int launch_diffuse(a2c_fe_ctx* c, double dt, double prefactor, double xLb, double xUb, const double* dgam,
                   const int32_t* dfull, const DiffuseScratch& w, double* dxnew, Timing* tm) {
  const int nh = c->desc.n_h;
  if (nh == 0) return A2C_OK;
  const int nblk = diffuse_blocks(nh);
  if (tm) tm->mark(c->stream, "diff_f");
  hipLaunchKernelGGL(k_diff_f, dim3(nblk), dim3(TB), 0, c->stream, nh, dgam, c->dx,
                     reinterpret_cast<double2*>(w.uv));
  A2C_HIP(c, hipGetLastError());
  if (tm) tm->mark(c->stream, "diff_step");
  hipLaunchKernelGGL(k_diff_step, dim3(nblk), dim3(TB), 0, c->stream, c->ell[4], c->dx, dgam,
                     reinterpret_cast<const double2*>(w.uv), dfull, prefactor, dt, xLb, xUb, dxnew, w.part);
  A2C_HIP(c, hipGetLastError());
  if (tm) tm->mark(c->stream, "diff_balance");
  hipLaunchKernelGGL(k_diff_balance, dim3(1), dim3(BT), 0, c->stream, reinterpret_cast<const double4*>(w.part),
                     nblk, w.bal);
  A2C_HIP(c, hipGetLastError());
  if (tm) tm->mark(c->stream, nullptr);
  return A2C_OK;
}

}  // namespace a2c

// ------------------------------------------------------------------ no-alloc
// a2c_fe_diffuse_device_noalloc: the diffusion step on device arrays with its
// scratch in this file-scope array -- the two edge-term factors for up to
// A2C_DIFFUSE_STATIC_MAX_H H sites, then 4 balance partials per block of 256
// sites, then the balance.
// Each GPU holds its own copy for the life of the process.  Every kernel
// overwrites what it reads later (f per site, one partial triple per block,
// the balance), so nothing needs clearing between calls; calls that overlap
// on one GPU share the array and race.
namespace a2c {
namespace {
constexpr size_t STATIC_BLK = (A2C_DIFFUSE_STATIC_MAX_H + TB - 1) / TB;
constexpr size_t STATIC_DOUBLES = 2 * (size_t)A2C_DIFFUSE_STATIC_MAX_H + 4 * STATIC_BLK + 4;
static __device__ __align__(256) double scratch[STATIC_DOUBLES];

constexpr int MAX_DEVICES = 64;
std::mutex g_scratch_mu;
double* g_scratch_ptr[MAX_DEVICES] = {};

// Device address of this device's copy of scratch (host code never uses the
// array's own name: on the host it is only a stand-in for the symbol).
int static_scratch(a2c_fe_ctx* c, double** out) {
  const int dev = c->desc.device;
  if (dev < 0 || dev >= MAX_DEVICES) return set_err(c, A2C_EINVAL, "device ordinal out of range");
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    int cur = 0;
    A2C_HIP(c, hipGetDevice(&cur));
    if (cur != dev) A2C_HIP(c, hipSetDevice(dev));
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (cur != dev) (void)hipSetDevice(cur);
    if (e != hipSuccess || !p)
      return set_err(c, A2C_EHIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
    g_scratch_ptr[dev] = static_cast<double*>(p);
  }
  *out = g_scratch_ptr[dev];
  return A2C_OK;
}
}  // namespace
}  // namespace a2c

extern "C" int a2c_fe_diffuse_device_noalloc(a2c_fe_ctx* c, double dt, double nu, double Qm, double xLb, double xUb,
                                             const double* d_gamma, const int32_t* d_full_h, double* d_x,
                                             double* d_balance) {
  using namespace a2c;
  if (!c) return A2C_EINVAL;
  c->err.clear();
  const int nh = c->desc.n_h;
  if (nh > A2C_DIFFUSE_STATIC_MAX_H)
    return set_err(c, A2C_EINVAL,
                   "a2c_fe_diffuse_device_noalloc: n_h = " + std::to_string(nh) + " exceeds its static scratch (" +
                       std::to_string(A2C_DIFFUSE_STATIC_MAX_H) + " H sites); use a2c_fe_diffuse");
  if (!c->have_nbr || !c->have_x) return set_err(c, A2C_ESTATE, "neighbors/occupancy not set");
  if (nh && (!d_gamma || !d_full_h || !d_x)) return set_err(c, A2C_EINVAL, "null argument");
  if (d_x == c->dx) return set_err(c, A2C_EINVAL, "d_x must not be the context occupancy");
  if (!(xLb <= xUb)) return set_err(c, A2C_EINVAL, "xLb > xUb");
  if (nh == 0) return A2C_OK;
  double* base = nullptr;
  if (int rc = static_scratch(c, &base)) return rc;
  const size_t MH = A2C_DIFFUSE_STATIC_MAX_H;
  const DiffuseScratch w{base, base + 2 * MH, d_balance ? d_balance : base + 2 * MH + 4 * STATIC_BLK};
  const double pref = nu * std::exp(-Qm / (KB * c->desc.temperature));  // integrate.cpp:26
  if (int rc = launch_diffuse(c, dt, pref, xLb, xUb, d_gamma, d_full_h, w, d_x)) return rc;
  A2C_HIP(c, hipMemcpyAsync(c->dx, d_x, 8 * (size_t)nh, hipMemcpyDeviceToDevice, c->stream));
  return A2C_OK;
}

