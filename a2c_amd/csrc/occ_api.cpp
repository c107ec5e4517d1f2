// C-ABI of the occupancy kernels: chemical potential and diffusion step.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "fe_internal.hpp"

using namespace a2c;

namespace {

// Context-owned scratch of the occupancy calls, from the library's device
// cache at the first call and kept until the context is destroyed.
int occ_scratch(a2c_fe_ctx* c) {
  const size_t nh = std::max(c->desc.n_h, 1), N = std::max(c->N, 1);
  if (c->occ_g) return A2C_OK;
  const size_t nblk = (size_t)diffuse_blocks((int)nh);
  A2C_HIP(c, dmalloc(&c->occ_g, 8 * nh));
  A2C_HIP(c, dmalloc(&c->occ_full, 4 * nh));
  A2C_HIP(c, dmalloc(&c->occ_corr, 8 * N));
  A2C_HIP(c, dmalloc(&c->occ_uv, 16 * nh));
  A2C_HIP(c, dmalloc(&c->occ_xn, 8 * nh));
  A2C_HIP(c, dmalloc(&c->occ_part, 8 * 4 * nblk));
  A2C_HIP(c, dmalloc(&c->occ_bal, 8 * 4));
  return A2C_OK;
}

}  // namespace

extern "C" int a2c_fe_chem_potential(a2c_fe_ctx* c, const double* X, const double* coe_pd, const double* coe_h,
                                     double gamma_bd, double* gamma, int32_t* full_h) {
  if (!c) return A2C_EINVAL;
  c->err.clear();
  if (!c->have_nbr || !c->have_x) return set_err(c, A2C_ESTATE, "neighbors/occupancy not set");
  if ((!X && c->N) || !coe_pd || !coe_h || (c->desc.n_h && (!gamma || !full_h)))
    return set_err(c, A2C_EINVAL, "null argument");
  const size_t nh = c->desc.n_h, N = c->N;
  if (int rc = occ_scratch(c)) return rc;
  double *dg = c->occ_g, *dcorr = c->occ_corr;
  int32_t* dfull = c->occ_full;
  if (N) {
    if (int rc = host_staging(c)) return rc;
    std::copy(X, X + 4 * N, c->hX);
    A2C_HIP(c, hipMemcpyAsync(c->dX, c->hX, 32 * N, hipMemcpyHostToDevice, c->stream));
  }
  if (nh) {
    A2C_HIP(c, hipMemcpyAsync(dg, gamma, 8 * nh, hipMemcpyHostToDevice, c->stream));
    A2C_HIP(c, hipMemcpyAsync(dfull, full_h, 4 * nh, hipMemcpyHostToDevice, c->stream));
  }
  if (int rc = launch_chem(c, c->dX, coe_pd, coe_h, gamma_bd, dg, dfull, dcorr)) return rc;
  if (nh) {
    A2C_HIP(c, hipMemcpyAsync(gamma, dg, 8 * nh, hipMemcpyDeviceToHost, c->stream));
    A2C_HIP(c, hipMemcpyAsync(full_h, dfull, 4 * nh, hipMemcpyDeviceToHost, c->stream));
  }
  A2C_HIP(c, hipStreamSynchronize(c->stream));
  return A2C_OK;
}

extern "C" int a2c_fe_diffuse(a2c_fe_ctx* c, double dt, double nu, double Qm, double xLb, double xUb,
                              const double* gamma, const int32_t* full_h, double* x) {
  if (!c) return A2C_EINVAL;
  c->err.clear();
  if (!c->have_nbr || !c->have_x) return set_err(c, A2C_ESTATE, "neighbors/occupancy not set");
  const size_t nh = c->desc.n_h;
  if (nh && (!gamma || !full_h || !x)) return set_err(c, A2C_EINVAL, "null argument");
  if (!(xLb <= xUb)) return set_err(c, A2C_EINVAL, "xLb > xUb");
  if (nh == 0) return A2C_OK;
  const double pref = nu * std::exp(-Qm / (KB * c->desc.temperature));  // integrate.cpp:26
  if (int rc = occ_scratch(c)) return rc;
  double *dg = c->occ_g, *dxn = c->occ_xn;
  int32_t* dfull = c->occ_full;
  A2C_HIP(c, hipMemcpyAsync(dg, gamma, 8 * nh, hipMemcpyHostToDevice, c->stream));
  A2C_HIP(c, hipMemcpyAsync(dfull, full_h, 4 * nh, hipMemcpyHostToDevice, c->stream));
  const DiffuseScratch w{c->occ_uv, c->occ_part, c->occ_bal};
  if (int rc = launch_diffuse(c, dt, pref, xLb, xUb, dg, dfull, w, dxn)) return rc;
  A2C_HIP(c, hipMemcpyAsync(c->dx, dxn, 8 * nh, hipMemcpyDeviceToDevice, c->stream));
  A2C_HIP(c, hipMemcpyAsync(x, dxn, 8 * nh, hipMemcpyDeviceToHost, c->stream));
  double bal[4];
  A2C_HIP(c, hipMemcpyAsync(bal, c->occ_bal, sizeof(bal), hipMemcpyDeviceToHost, c->stream));
  A2C_HIP(c, hipStreamSynchronize(c->stream));
  c->last_balance = a2c_diffusion_balance{bal[0], bal[1], bal[2], bal[3]};
  return A2C_OK;
}

extern "C" int a2c_fe_diffusion_balance(const a2c_fe_ctx* c, a2c_diffusion_balance* out) {
  if (!c || !out) return A2C_EINVAL;
  *out = c->last_balance;
  return A2C_OK;
}

extern "C" int a2c_fe_time_diffuse(a2c_fe_ctx* c, double dt, double nu, double Qm, double xLb, double xUb,
                                   const double* d_gamma, const int32_t* d_full_h, double* d_x, int32_t iters,
                                   a2c_timing* out) {
  if (!c) return A2C_EINVAL;
  c->err.clear();
  if (!c->have_nbr || !c->have_x) return set_err(c, A2C_ESTATE, "neighbors/occupancy not set");
  const int nh = c->desc.n_h;
  if (!out || iters <= 0 || iters > 100000 || (nh && (!d_gamma || !d_full_h || !d_x)))
    return set_err(c, A2C_EINVAL, "bad timing args");
  if (d_x == c->dx) return set_err(c, A2C_EINVAL, "d_x must not be the context occupancy");
  if (!(xLb <= xUb)) return set_err(c, A2C_EINVAL, "xLb > xUb");
  std::memset(out, 0, sizeof(*out));
  if (nh == 0) return A2C_OK;
  if (int rc = occ_scratch(c)) return rc;
  const double pref = nu * std::exp(-Qm / (KB * c->desc.temperature));  // integrate.cpp:26
  const DiffuseScratch w{c->occ_uv, c->occ_part, c->occ_bal};
  constexpr int per = 4;  // marks per step: three kernels and the end
  std::vector<hipEvent_t> ev((size_t)per * iters + 2);
  for (auto& e : ev) A2C_HIP(c, hipEventCreate(&e));
  struct Destroy {
    std::vector<hipEvent_t>& v;
    ~Destroy() {
      for (auto e : v) hipEventDestroy(e);
    }
  } ds{ev};
  Timing tm;
  const char* names[8] = {};
  A2C_HIP(c, hipEventRecord(ev[0], c->stream));
  for (int it = 0; it < iters; ++it) {
    tm.ev = ev.data() + 1 + (size_t)per * it;
    tm.cap = per;
    tm.n = 0;
    tm.on = true;
    if (int rc = launch_diffuse(c, dt, pref, xLb, xUb, d_gamma, d_full_h, w, d_x, &tm)) return rc;
    if (tm.n != per) return set_err(c, A2C_EHIP, "diffusion timing marks");
    if (it == 0)
      for (int k = 0; k < per - 1; ++k) names[k] = tm.name[k];
  }
  hipEvent_t tend = ev[1 + (size_t)per * iters];
  A2C_HIP(c, hipEventRecord(tend, c->stream));
  A2C_HIP(c, hipEventSynchronize(tend));
  float tot = 0.f;
  A2C_HIP(c, hipEventElapsedTime(&tot, ev[0], tend));
  double acc[per - 1] = {0};
  for (int it = 0; it < iters; ++it) {
    hipEvent_t* e = ev.data() + 1 + (size_t)per * it;
    for (int k = 0; k < per - 1; ++k) {
      float ms = 0.f;
      A2C_HIP(c, hipEventElapsedTime(&ms, e[k], e[k + 1]));
      acc[k] += ms;
    }
  }
  out->n_kernels = per - 1;
  for (int k = 0; k < per - 1; ++k) {
    std::strncpy(out->names[k], names[k] ? names[k] : "?", 31);
    out->avg_ms[k] = acc[k] / iters;
  }
  out->total_ms = tot;
  return A2C_OK;
}
