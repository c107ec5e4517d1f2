// a2c_minimizer.hpp -- C++ host mirror of the reference's legacy DMD
// interface over the C-ABI of liba2c_fe.so (header-only).
//
// Same names, argument lists, argument meaning and container types as the
// reference (a caller written against minimizer.h compiles against these
// signatures with a2c:: types):
//   a2c::Minimizer::Minimizer(argc, argv) minimizer.h:82
//   a2c::Minimizer::initialize            minimizer.h:84-88   (minimizer.cpp:484-546)
//   a2c::Minimizer::minimizeFreeEntropy   minimizer.h:89-92   (minimizer.cpp:548-586)
//   a2c::Minimizer::formFunctionAndGradient                   (minimizer.cpp:83-369)
//   a2c::Minimizer::calculateChemicalPotential minimizer.h:93 (minimizer.cpp:588-802)
//   a2c::Minimizer::findLocalNeighbors    minimizer.h:94-98   (minimizer.cpp:833-1144)
//   a2c::findNeighbors                    neighbors.cpp:8-113
//   a2c::integrateDiffusion               integrate.cpp:9-87
// Differences: errors are returned as A2C_* status codes (the reference exits
// or aborts); the lists live on the GPU and are exported into the
// vector<vector<int>> containers on request; the occupancy x0 is copied to
// the device at initialize/minimize time instead of being held by pointer.
#pragma once
#include <cmath>
#include <stdexcept>
#include <vector>

#include "a2c_fe.h"

namespace a2c {

struct Vec3D {  // the subset of Vector3D.h:47-88 the interface needs
  double v[3];
  Vec3D() : v{0.0, 0.0, 0.0} {}
  Vec3D(double x, double y, double z) : v{x, y, z} {}
  double& operator[](int i) { return v[i]; }
  double operator[](int i) const { return v[i]; }
};

struct InputFile {  // the fields of the reference's missing input.h this path reads
  double eps = 1e-5, rc = 5.35, T0 = 300.0, xUb = 1.0 - 1e-16, xLb = 1e-3, v = 0.0, Qm = 0.0, a_Pd = 3.89;
  double box[3] = {0.0, 0.0, 0.0};
  int periodic[3] = {0, 0, 0};
  int mixed_precision = 0;
};
struct Input {
  InputFile file;
};

using Lists = std::vector<std::vector<int>>;

namespace detail {
inline void pack(const std::vector<Vec3D>& qP, const std::vector<Vec3D>& qH, const std::vector<double>& sP,
                 const std::vector<double>& sH, std::vector<double>& X) {
  X.resize(4 * (qP.size() + qH.size()));
  size_t k = 0;
  for (size_t i = 0; i < qP.size(); ++i) { X[k++] = qP[i][0]; X[k++] = qP[i][1]; X[k++] = qP[i][2]; X[k++] = sP[i]; }
  for (size_t i = 0; i < qH.size(); ++i) { X[k++] = qH[i][0]; X[k++] = qH[i][1]; X[k++] = qH[i][2]; X[k++] = sH[i]; }
}
inline void unpack(const std::vector<double>& X, std::vector<Vec3D>& qP, std::vector<Vec3D>& qH,
                   std::vector<double>& sP, std::vector<double>& sH) {
  size_t k = 0;
  for (size_t i = 0; i < qP.size(); ++i) { qP[i] = Vec3D(X[k], X[k + 1], X[k + 2]); sP[i] = X[k + 3]; k += 4; }
  for (size_t i = 0; i < qH.size(); ++i) { qH[i] = Vec3D(X[k], X[k + 1], X[k + 2]); sH[i] = X[k + 3]; k += 4; }
}
inline void export_list(a2c_fe_ctx* c, int which, size_t rows, Lists& out) {
  std::vector<int32_t> off(rows + 1);
  if (a2c_fe_get_neighbors(c, which, off.data(), nullptr)) throw std::runtime_error(a2c_fe_last_error(c));
  std::vector<int32_t> idx(off[rows] > 0 ? off[rows] : 1);
  if (a2c_fe_get_neighbors(c, which, off.data(), idx.data())) throw std::runtime_error(a2c_fe_last_error(c));
  out.assign(rows, {});
  for (size_t r = 0; r < rows; ++r) out[r].assign(idx.begin() + off[r], idx.begin() + off[r + 1]);
}
inline a2c_fe_desc describe(const Input& in, size_t nPd, size_t nH) {
  a2c_fe_desc d;
  a2c_fe_desc_default(&d);
  d.n_pd = (int32_t)nPd;
  d.n_h = (int32_t)nH;
  for (int c = 0; c < 3; ++c) { d.box[c] = in.file.box[c]; d.periodic[c] = in.file.periodic[c]; }
  d.rc = in.file.rc;
  d.a_pd = in.file.a_Pd;
  d.temperature = in.file.T0;
  d.deriv_mode = A2C_DERIV_FD;  // the reference differentiates by central differences
  d.fd_eps = in.file.eps;
  d.precision = in.file.mixed_precision ? A2C_PREC_MIXED : A2C_PREC_FP64;
  return d;
}
}  // namespace detail

// findNeighbors (neighbors.cpp:8-113): lists at rc, self excluded; returns the
// largest neighbor count (maxNeib).  Builds on a temporary context.
inline int findNeighbors(double a_Pd, double rc, std::vector<Vec3D>& q_Pd0, std::vector<Vec3D>& q_H0, Lists& PdPd,
                         Lists& PdH, Lists& HPd, Lists& HH, const Input* in = nullptr) {
  Input def;
  const Input& I = in ? *in : def;
  a2c_fe_desc d = detail::describe(I, q_Pd0.size(), q_H0.size());
  d.rc = rc;
  d.a_pd = a_Pd;
  a2c_fe_ctx* c = nullptr;
  if (a2c_fe_create(&d, &c)) return -1;
  std::vector<double> X, sP(q_Pd0.size(), 0.1), sH(q_H0.size(), 0.1);
  detail::pack(q_Pd0, q_H0, sP, sH, X);
  a2c_nbr_stats st;
  int rc_ = a2c_fe_build_neighbors(c, X.data(), &st);
  if (!rc_) {
    detail::export_list(c, 0, q_Pd0.size(), PdPd);
    detail::export_list(c, 1, q_Pd0.size(), PdH);
    detail::export_list(c, 2, q_H0.size(), HPd);
    detail::export_list(c, 3, q_H0.size(), HH);
  }
  a2c_fe_destroy(c);
  return rc_ ? rc_ : st.max_neib;
}

// The reference's 12-point rule as gaussianQuadrature(6, QPp) returns it
// (quadrature.cpp:43-75): point (m, +-) displaces coordinate m of the 6-D pair
// (xi1 = coordinates 0-2 of site i, xi2 = 3-5 of site j) by +-sqrt(3);
// weight QWp = 1/12 (quadrature.cpp:39).  The device evaluates this rule in
// its structured form; initialize() checks that the caller's QPp/QWp are it.
inline bool is_reference_rule(const std::vector<std::vector<Vec3D>>& QPp, double QWp) {
  if (QPp.size() != 12 || std::fabs(QWp - 1.0 / 12.0) > 1e-15) return false;
  const double h = std::sqrt(3.0);
  std::vector<int> seen(12, 0);
  for (const auto& p : QPp) {
    if (p.size() != 2) return false;
    int nz = 0, m = -1, sg = 0;
    for (int c = 0; c < 6; ++c) {
      const double v = c < 3 ? p[0][c] : p[1][c - 3];
      if (std::fabs(v) < 1e-12) continue;
      if (std::fabs(std::fabs(v) - h) > 1e-12) return false;
      ++nz;
      m = c;
      sg = v > 0 ? 0 : 1;
    }
    if (nz != 1) return false;
    seen[2 * m + sg]++;
  }
  for (int k : seen)
    if (k != 1) return false;
  return true;
}

class Minimizer {
 public:
  Minimizer() = default;
  // minimizer.h:82 (PETSc initialisation in the reference; nothing to set up here)
  Minimizer(int* /*argc*/, char*** /*argv*/) {}
  Minimizer(const Minimizer&) = delete;
  Minimizer& operator=(const Minimizer&) = delete;
  ~Minimizer() { a2c_fe_destroy(ctx_); }

  // minimizer.h:84-88, the reference's argument list.  The neighbor lists
  // PdPd..HH are rebuilt on the device from q at rc (the same sets the
  // reference's findLocalNeighbors produces; the pointers may be null);
  // QPp/QWp must be the 12-point rule of quadrature.cpp (A2C_EINVAL
  // otherwise); coe_Pd/coe_H (4 values each, may be null: zeros) feed
  // calculateChemicalPotential.  x0 is copied, not held.
  int initialize(Input* input, Lists* /*PdPd*/, Lists* /*PdH*/, Lists* /*HPd*/, Lists* /*HH*/,
                 std::vector<Vec3D>& q_Pd0, std::vector<Vec3D>& q_H0, std::vector<double>& sigma_Pd0,
                 std::vector<double>& sigma_H0, std::vector<double>& x0, std::vector<std::vector<Vec3D>>* QPp,
                 double& QWp, std::vector<double>* coe_Pd, std::vector<double>* coe_H) {
    if (!input || (QPp && !is_reference_rule(*QPp, QWp))) return A2C_EINVAL;
    return setup(*input, q_Pd0, q_H0, sigma_Pd0, sigma_H0, x0, coe_Pd, coe_H);
  }
  // the same without the lists and the rule (they are implied)
  int initialize(Input* input, std::vector<Vec3D>& q_Pd0, std::vector<Vec3D>& q_H0, std::vector<double>& sigma_Pd0,
                 std::vector<double>& sigma_H0, std::vector<double>& x0, std::vector<double>* coe_Pd = nullptr,
                 std::vector<double>* coe_H = nullptr) {
    if (!input) return A2C_EINVAL;
    return setup(*input, q_Pd0, q_H0, sigma_Pd0, sigma_H0, x0, coe_Pd, coe_H);
  }

  // The TAO callback (minimizer.cpp:83-369) on host arrays in TAO layout.
  int formFunctionAndGradient(const double* X, double* f, double* G) { return a2c_fe_eval(ctx_, X, f, G); }

  // minimizer.h:89-92 (dt is unused by the reference's solve as well)
  int minimizeFreeEntropy(double /*dt*/, std::vector<Vec3D>& q_Pd0, std::vector<Vec3D>& q_H0,
                          std::vector<double>& sigma_Pd0, std::vector<double>& sigma_H0, std::vector<double>& x0,
                          std::vector<Vec3D>& q_Pd, std::vector<Vec3D>& q_H, std::vector<double>& sigma_Pd,
                          std::vector<double>& sigma_H, double& free_entropy, a2c_lmvm_report* report = nullptr) {
    if (int rc = a2c_fe_set_occupancy(ctx_, x0.data())) return rc;
    detail::pack(q_Pd0, q_H0, sigma_Pd0, sigma_H0, X_);
    a2c_lmvm_opts o;
    a2c_lmvm_opts_default(&o);
    a2c_lmvm_report rep;
    if (int rc = a2c_fe_minimize(ctx_, X_.data(), &o, X_.data(), &rep)) return rc;
    q_Pd.resize(nPd_); q_H.resize(nH_); sigma_Pd.resize(nPd_); sigma_H.resize(nH_);
    detail::unpack(X_, q_Pd, q_H, sigma_Pd, sigma_H);
    free_entropy = rep.f;
    if (report) *report = rep;
    return A2C_OK;
  }

  // minimizer.h:93, evaluated at the last minimized state
  int calculateChemicalPotential(std::vector<double>& x0, double gammabd, std::vector<double>& gamma,
                                 std::vector<int>& full_H) {
    if (int rc = a2c_fe_set_occupancy(ctx_, x0.data())) return rc;
    std::vector<int32_t> fh(full_H.begin(), full_H.end());
    gamma.resize(nH_);
    int rc = a2c_fe_chem_potential(ctx_, X_.data(), coe_Pd_.data(), coe_H_.data(), gammabd, gamma.data(), fh.data());
    full_H.assign(fh.begin(), fh.end());
    return rc;
  }

  // minimizer.h:94-98, the reference's argument list: the refined lists at rc
  // from the current state, PdPdnum and HH1.  The reference refines the
  // candidate lists *_ext; the device finds the neighbors directly, so they
  // are not read.
  int findLocalNeighbors(Lists& /*PdPd_ext*/, Lists& /*PdH_ext*/, Lists& /*HPd_ext*/, Lists& /*HH_ext*/, Lists& PdPd,
                         Lists& PdH, Lists& HPd, Lists& HH, Lists& PdPdnum, Lists& HH1, int& maxNeib) {
    return findLocalNeighbors(PdPd, PdH, HPd, HH, PdPdnum, HH1, maxNeib);
  }
  // the same without the candidate lists
  int findLocalNeighbors(Lists& PdPd, Lists& PdH, Lists& HPd, Lists& HH, Lists& PdPdnum, Lists& HH1, int& maxNeib) {
    if (int rc = a2c_fe_build_neighbors(ctx_, X_.data(), &stats_)) return rc;
    detail::export_list(ctx_, 0, nPd_, PdPd);
    detail::export_list(ctx_, 1, nPd_, PdH);
    detail::export_list(ctx_, 2, nH_, HPd);
    detail::export_list(ctx_, 3, nH_, HH);
    detail::export_list(ctx_, 4, nH_, HH1);
    std::vector<int32_t> num(3 * (nPd_ ? nPd_ : 1));
    if (int rc = a2c_fe_get_pdpdnum(ctx_, num.data())) return rc;
    PdPdnum.assign(nPd_, std::vector<int>(3));
    for (size_t i = 0; i < nPd_; ++i)
      for (int k = 0; k < 3; ++k) PdPdnum[i][k] = num[3 * i + k];
    maxNeib = stats_.max_neib;
    return A2C_OK;
  }

  a2c_fe_ctx* context() { return ctx_; }

 private:
  int setup(const Input& input, std::vector<Vec3D>& q_Pd0, std::vector<Vec3D>& q_H0, std::vector<double>& sigma_Pd0,
            std::vector<double>& sigma_H0, std::vector<double>& x0, std::vector<double>* coe_Pd,
            std::vector<double>* coe_H) {
    input_ = input;
    nPd_ = q_Pd0.size();
    nH_ = q_H0.size();
    if (sigma_Pd0.size() != nPd_ || sigma_H0.size() != nH_ || x0.size() != nH_) return A2C_EINVAL;
    if ((coe_Pd && coe_Pd->size() < 4) || (coe_H && coe_H->size() < 4)) return A2C_EINVAL;
    coe_Pd_ = coe_Pd ? *coe_Pd : std::vector<double>(4, 0.0);
    coe_H_ = coe_H ? *coe_H : std::vector<double>(4, 0.0);
    a2c_fe_desc d = detail::describe(input_, nPd_, nH_);
    a2c_fe_destroy(ctx_);
    ctx_ = nullptr;
    if (int rc = a2c_fe_create(&d, &ctx_)) return rc;
    if (int rc = a2c_fe_set_occupancy(ctx_, x0.data())) return rc;
    detail::pack(q_Pd0, q_H0, sigma_Pd0, sigma_H0, X_);
    return a2c_fe_build_neighbors(ctx_, X_.data(), &stats_);
  }

  Input input_;
  a2c_fe_ctx* ctx_ = nullptr;
  size_t nPd_ = 0, nH_ = 0;
  std::vector<double> X_;
  std::vector<double> coe_Pd_, coe_H_;
  a2c_nbr_stats stats_{};
};

// integrateDiffusion (integrate.cpp:9-87) on the minimizer's lattice: x0 ->
// x; full_H sites keep x.
inline int integrateDiffusion(Minimizer& m, Input& input, double dt, std::vector<double>& gamma,
                              std::vector<double>& x0, std::vector<double>& x, std::vector<int>& full_H) {
  if (int rc = a2c_fe_set_occupancy(m.context(), x0.data())) return rc;
  std::vector<int32_t> fh(full_H.begin(), full_H.end());
  x = x0;
  return a2c_fe_diffuse(m.context(), dt, input.file.v, input.file.Qm, input.file.xLb, input.file.xUb, gamma.data(),
                        fh.data(), x.data());
}

}  // namespace a2c
