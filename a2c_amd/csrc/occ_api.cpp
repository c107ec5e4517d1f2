// C-ABI of the occupancy kernels: chemical potential and diffusion step.
#include <hip/hip_runtime.h>

#include <cmath>
#include <vector>

#include "fe_internal.hpp"

using namespace a2c;

namespace {

// Device scratch owned by one call; freed on every return path.
struct Scratch {
  std::vector<void*> p;
  ~Scratch() {
    for (void* q : p) (void)hipFree(q);
  }
  template <class T>
  hipError_t alloc(T** out, size_t n) {
    void* q = nullptr;
    hipError_t e = hipMalloc(&q, sizeof(T) * (n ? n : 1));
    if (e == hipSuccess) {
      p.push_back(q);
      *out = static_cast<T*>(q);
    }
    return e;
  }
};

}  // namespace

extern "C" int a2c_fe_chem_potential(a2c_fe_ctx* c, const double* X, const double* coe_pd, const double* coe_h,
                                     double gamma_bd, double* gamma, int32_t* full_h) {
  if (!c) return A2C_EINVAL;
  c->err.clear();
  if (!c->have_nbr || !c->have_x) return set_err(c, A2C_ESTATE, "neighbors/occupancy not set");
  if ((!X && c->N) || !coe_pd || !coe_h || (c->desc.n_h && (!gamma || !full_h)))
    return set_err(c, A2C_EINVAL, "null argument");
  const size_t nh = c->desc.n_h, N = c->N;
  Scratch sc;
  double *dg = nullptr, *dcorr = nullptr;
  int32_t* dfull = nullptr;
  A2C_HIP(c, sc.alloc(&dg, nh));
  A2C_HIP(c, sc.alloc(&dfull, nh));
  A2C_HIP(c, sc.alloc(&dcorr, N));
  if (N) {
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
  Scratch sc;
  double *dg = nullptr, *df = nullptr, *dxn = nullptr;
  int32_t* dfull = nullptr;
  A2C_HIP(c, sc.alloc(&dg, nh));
  A2C_HIP(c, sc.alloc(&df, nh));
  A2C_HIP(c, sc.alloc(&dxn, nh));
  A2C_HIP(c, sc.alloc(&dfull, nh));
  A2C_HIP(c, hipMemcpyAsync(dg, gamma, 8 * nh, hipMemcpyHostToDevice, c->stream));
  A2C_HIP(c, hipMemcpyAsync(dfull, full_h, 4 * nh, hipMemcpyHostToDevice, c->stream));
  if (int rc = launch_diffuse(c, dt, pref, xLb, xUb, dg, dfull, df, dxn)) return rc;
  A2C_HIP(c, hipMemcpyAsync(c->dx, dxn, 8 * nh, hipMemcpyDeviceToDevice, c->stream));
  A2C_HIP(c, hipMemcpyAsync(x, dxn, 8 * nh, hipMemcpyDeviceToHost, c->stream));
  A2C_HIP(c, hipStreamSynchronize(c->stream));
  return A2C_OK;
}
