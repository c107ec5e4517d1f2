// Diffusive-MD time step and its host helpers (include/a2c_dmd.h).
#include "a2c_dmd.h"

#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <limits>
#include <string>
#include <vector>

#include "fe_internal.hpp"

extern "C" {

int a2c_enforce_boundary_conditions(int32_t n_sub, const int32_t* sub, double gamma_bd, double xUb, double* x0,
                                    double* gamma, int32_t* full_h) {
  if (n_sub < 0 || (n_sub > 0 && (!sub || !x0 || !gamma || !full_h))) return A2C_EINVAL;
  for (int32_t i = 0; i < n_sub; ++i) {  // boundaryconditions.cpp:47-51
    gamma[sub[i]] = gamma_bd;
    full_h[sub[i]] = 1;
    x0[sub[i]] = xUb;
  }
  return A2C_OK;
}

int a2c_macro_results(int32_t n_h, const double* x, int32_t n_surf, const int32_t* surf, double* xH,
                      double* xH_interior) {
  if (n_h <= 0 || !x || !xH || !xH_interior || n_surf < 0 || (n_surf > 0 && !surf)) return A2C_EINVAL;
  std::vector<char> is_surf(n_h, 0);
  for (int32_t i = 0; i < n_surf; ++i)
    if (surf[i] >= 0 && surf[i] < n_h) is_surf[surf[i]] = 1;
  double all = 0.0, inner = 0.0;
  for (int32_t i = 0; i < n_h; ++i) {  // calculateresults.cpp:44-70
    all += x[i];
    if (!is_surf[i]) inner += x[i];
  }
  *xH = all / n_h;
  *xH_interior = inner / (n_h - n_surf);
  return A2C_OK;
}

int a2c_read_results(const char* path, int32_t n_pd, int32_t n_h, double xLb, double xUb, double* X, double* x,
                     double* gamma, int32_t* full_h) {
  if (!path || !X || n_pd < 0 || n_h < 0 || (n_h > 0 && (!x || !gamma || !full_h))) return A2C_EINVAL;
  std::ifstream in(path);
  if (!in) return A2C_EINVAL;  // readresults.cpp:21-24 exits instead
  for (int k = 0; k < 9; ++k) in.ignore(std::numeric_limits<std::streamsize>::max(), '\n');
  double t;
  for (int32_t i = 0; i < n_pd; ++i) {
    double* s = X + 4 * (size_t)i;
    if (!(in >> t >> t >> s[0] >> s[1] >> s[2] >> s[3] >> t >> t >> t)) return A2C_EINVAL;
  }
  for (int32_t i = 0; i < n_h; ++i) {
    double* s = X + 4 * ((size_t)n_pd + i);
    double fh;
    if (!(in >> t >> t >> s[0] >> s[1] >> s[2] >> s[3] >> x[i] >> gamma[i] >> fh)) return A2C_EINVAL;
    full_h[i] = (int32_t)fh;
  }
  for (int32_t i = 0; i < n_h; ++i) {  // readresults.cpp:48-51
    if (x[i] < xLb) x[i] = xLb;
    else if (x[i] > xUb) x[i] = xUb;
  }
  return A2C_OK;
}

int a2c_write_results(const char* path, int32_t n_pd, int32_t n_h, const double* X, const double* x,
                      const double* gamma, const int32_t* full_h, double time, int32_t step) {
  if (!path || !X || n_pd < 0 || n_h < 0 || (n_h > 0 && (!x || !gamma || !full_h))) return A2C_EINVAL;
  FILE* f = std::fopen(path, "w");
  if (!f) return A2C_EINVAL;
  std::fprintf(f, "# a2c DMD restart file\n# step %d\n# time %.17g\n# n_Pd %d\n# n_H %d\n", step, time, n_pd, n_h);
  std::fprintf(f, "# columns: index species qx qy qz sigma x gamma full_H\n# species: 0 Pd, 1 H\n#\n#\n");
  for (int32_t i = 0; i < n_pd; ++i) {
    const double* s = X + 4 * (size_t)i;
    std::fprintf(f, "%d 0 %.17g %.17g %.17g %.17g 1 0 0\n", i, s[0], s[1], s[2], s[3]);
  }
  for (int32_t i = 0; i < n_h; ++i) {
    const double* s = X + 4 * ((size_t)n_pd + i);
    std::fprintf(f, "%d 1 %.17g %.17g %.17g %.17g %.17g %.17g %d\n", i, s[0], s[1], s[2], s[3], x[i], gamma[i],
                 full_h[i]);
  }
  return std::fclose(f) == 0 ? A2C_OK : A2C_EINVAL;
}

int a2c_write_xyz(const char* path, int32_t n_pd, int32_t n_h, const double* X, const double* X0, const double* x,
                  const double* gamma, const double* box, const int32_t* periodic, double time) {
  if (!path || !X || !X0 || !box || !periodic || n_pd < 0 || n_h < 0 || (n_h > 0 && (!x || !gamma)))
    return A2C_EINVAL;
  FILE* f = std::fopen(path, "w");
  if (!f) return A2C_EINVAL;
  const int32_t n = n_pd + n_h;
  std::fprintf(f, "%d\n", n);
  std::fprintf(f, "Lattice=\"%e %e %e %e %e %e %e %e %e\" ", box[0], 0.0, 0.0, 0.0, box[1], 0.0, 0.0, 0.0, box[2]);
  std::fprintf(f, "Origin=\"%e %e %e\" ", 0.0, 0.0, 0.0);
  std::fprintf(f, "pbc=\"%s %s %s\" ", periodic[0] ? "T" : "F", periodic[1] ? "T" : "F", periodic[2] ? "T" : "F");
  std::fprintf(f, "Properties=\"species:S:1:pos:R:3:disp:R:3:type:I:1:atomic_freq:R:1:tag:I:1:molar_frac:R:1:"
                  "chem_potent:R:1\" Time=%e\n", time);
  for (int32_t s = 0; s < n; ++s) {
    const double* q = X + 4 * (size_t)s;
    const double* q0 = X0 + 4 * (size_t)s;
    const bool h = s >= n_pd;
    std::fprintf(f, "%16s  %16.10e  %16.10e  %16.10e  %16.10e  %16.10e  %16.10e  %4d  %16.10e  %4d  %16.10e  %16.10e\n",
                 h ? "H" : "Pd", q[0], q[1], q[2], q[0] - q0[0], q[1] - q0[1], q[2] - q0[2], h ? 1 : 0, q[3], 0,
                 h ? x[s - n_pd] : 1.0, h ? gamma[s - n_pd] : 0.0);
  }
  return std::fclose(f) == 0 ? A2C_OK : A2C_EINVAL;
}

void a2c_dmd_params_default(a2c_dmd_params* p) {
  std::memset(p, 0, sizeof(*p));
  p->dt = 1e-13;
  p->nu = 1e13;
  p->Qm = 0.23;
  p->xLb = 1e-3;
  p->xUb = 1.0 - 1e-16;
  p->gamma_bd = 1e300;
  p->rebuild_neighbors = 1;
  a2c_lmvm_opts_default(&p->lmvm);
}

}  // extern "C"

namespace {
// Total H, compensated (Neumaier): a running sum of ~10^6 occupancies, most
// of them the same clamp value, rounds with a systematic bias (~3e-7 of a
// total of 8e4 at 64^3 cells), far above the balance it is compared with.
double total_of(const double* x, int32_t n) {
  double s = 0.0, comp = 0.0;
  for (int32_t i = 0; i < n; ++i) {
    const double t = s + x[i];
    comp += std::fabs(s) >= std::fabs(x[i]) ? (s - t) + x[i] : (x[i] - t) + s;
    s = t;
  }
  return s + comp;
}
}  // namespace

extern "C" {

int a2c_dmd_step(a2c_fe_ctx* c, const a2c_dmd_params* p, int32_t n_sub, const int32_t* sub, int32_t n_surf,
                 const int32_t* surf, double* X, double* x, double* gamma, int32_t* full_h, a2c_dmd_report* rep) {
  if (!c || !p || !X || !rep || (c->desc.n_h > 0 && (!x || !gamma || !full_h))) return A2C_EINVAL;
  const int32_t nh = c->desc.n_h;
  std::memset(rep, 0, sizeof(*rep));
  if (int rc = a2c_enforce_boundary_conditions(n_sub, sub, p->gamma_bd, p->xUb, x, gamma, full_h)) return rc;
  if (int rc = a2c_fe_set_occupancy(c, x)) return rc;
  if (p->rebuild_neighbors || !c->have_nbr)
    if (int rc = a2c_fe_build_neighbors(c, X, nullptr)) return rc;
  a2c_lmvm_report lr;
  if (int rc = a2c_fe_minimize(c, X, &p->lmvm, X, &lr)) return rc;
  rep->f = lr.f;
  rep->lmvm_reason = lr.reason;
  rep->lmvm_iterations = lr.iterations;
  rep->lmvm_evals = lr.evals;
  if (int rc = a2c_fe_chem_potential(c, X, p->coe_pd, p->coe_h, p->gamma_bd, gamma, full_h)) return rc;
  const double before = total_of(x, nh);
  if (int rc = a2c_fe_diffuse(c, p->dt, p->nu, p->Qm, p->xLb, p->xUb, gamma, full_h, x)) return rc;
  const double after = total_of(x, nh);
  rep->total_h_before = before;
  rep->total_h_after = after;
  a2c_diffusion_balance bal;
  a2c_fe_diffusion_balance(c, &bal);
  rep->flux_sum = bal.flux_sum;
  rep->flux_gross = bal.flux_gross;
  rep->clamp_lift = bal.clamp_lift;
  rep->reservoir_supply = bal.reservoir_supply;
  if (nh > 0) a2c_macro_results(nh, x, n_surf, surf, &rep->xH, &rep->xH_interior);
  return A2C_OK;
}

}  // extern "C"
