// General model of a context: a tabulated potential (uniform grids, or a
// DYNAMO setfl file) and a node table for the phase-packet average.  The
// reference reads InteratomicPotential.File (IoData.cpp:523-534) without
// using it and hard-codes its rule (quadrature.cpp:43-75); here both are
// data, evaluated by the kernels of evalgen.hip.
//
// Interpolation: on each grid interval the cubic Hermite piece through the
// two nodes with node slopes from central differences -- fourth order
// (f[m-2] - 8 f[m-1] + 8 f[m+1] - f[m+2]) / 12 inside, second order at the
// second and second-to-last node, one-sided at the ends.  This is the
// interpolation the setfl format is written for; it reproduces cubics
// exactly on every interval not touching the two end intervals at either side.
#include <cmath>
#include <cstring>
#include <fstream>
#include <sstream>
#include <string>
#include <vector>

#include "a2c_fe.h"
#include "eam64.hpp"
#include "fe_internal.hpp"

namespace a2c {
namespace {

struct Pieces {
  std::vector<double> c;  // 4 per interval
  int n = 0;              // nodes
  double h = 0.0, yend = 0.0, dend = 0.0;
};

bool make_pieces(const double* y, int n, double h, Pieces& P) {
  if (n < 5 || !(h > 0.0) || !std::isfinite(h)) return false;
  for (int k = 0; k < n; ++k)
    if (!std::isfinite(y[k])) return false;
  std::vector<double> s(n);
  s[0] = y[1] - y[0];
  s[1] = 0.5 * (y[2] - y[0]);
  s[n - 2] = 0.5 * (y[n - 1] - y[n - 3]);
  s[n - 1] = y[n - 1] - y[n - 2];
  for (int m = 2; m < n - 2; ++m) s[m] = ((y[m - 2] - y[m + 2]) + 8.0 * (y[m + 1] - y[m - 1])) / 12.0;
  P.n = n;
  P.h = h;
  P.c.resize(4 * (size_t)(n - 1));
  for (int m = 0; m < n - 1; ++m) {
    const double dy = y[m + 1] - y[m];
    double* c = &P.c[4 * (size_t)m];
    c[0] = y[m];
    c[1] = s[m];
    c[2] = 3.0 * dy - 2.0 * s[m] - s[m + 1];
    c[3] = s[m] + s[m + 1] - 2.0 * dy;
  }
  P.yend = y[n - 1];
  P.dend = s[n - 1] / h;
  return true;
}

// The r-functions of the compiled-in potential sampled for the FP32 pieces:
// nodes every 1e-3 A up to 8 A (the quadrature points of a listed pair lie
// within rc + a few sigma; the tails are ~0 there).
constexpr int CF_N = 8001;
constexpr double CF_H = 1e-3;
double closed_r_function(int fn, double r) {
  switch (fn) {
    case GEN_RHO_PD: return eam::f_Pd(r);
    case GEN_RHO_H: return eam::f_H(r);
    case GEN_RPHI_PD: return r * eam::phi_Pd(r);
    case GEN_RPHI_PDH: return r * eam::phi_morse<0>(r);
    default: return r * eam::phi_morse<1>(r);
  }
}

void set_fn(GenFn& F, const Pieces& P, int off) {
  F.n = P.n;
  F.off = off;
  F.ih = 1.0 / P.h;
  F.yend = P.yend;
  F.dend = P.dend;
}

template <class T>
int upload(a2c_fe_ctx* c, T** dst, size_t& cap, const std::vector<T>& v) {
  if (v.size() > cap) {
    dfree(*dst);
    *dst = nullptr;
    cap = 0;
    if (dmalloc(dst, sizeof(T) * v.size()) != hipSuccess) return set_err(c, A2C_ENOMEM, "general model tables");
    cap = v.size();
  }
  A2C_HIP(c, hipMemcpy(*dst, v.data(), sizeof(T) * v.size(), hipMemcpyHostToDevice));
  return A2C_OK;
}

// Pieces of the seven functions (FP64) and of the five r-functions (FP32);
// fn[] null for the closed forms: FP64 kernels evaluate those directly.
int install_potential(a2c_fe_ctx* c, const Pieces* fn) {
  std::vector<double4> t64;
  std::vector<float4> t32;
  GenDev& g = c->gen;
  Pieces cf[GEN_NFN];
  if (fn) {
    for (int f = 0; f < GEN_NFN; ++f) {
      set_fn(g.fn[f], fn[f], (int)t64.size());
      for (int m = 0; m + 1 < fn[f].n; ++m) {
        const double* q = &fn[f].c[4 * (size_t)m];
        t64.push_back(make_double4(q[0], q[1], q[2], q[3]));
      }
    }
  } else {
    std::vector<double> y(CF_N);
    for (int f = GEN_RHO_PD; f < GEN_NFN; ++f) {
      for (int k = 0; k < CF_N; ++k) y[k] = closed_r_function(f, k * CF_H);
      make_pieces(y.data(), CF_N, CF_H, cf[f]);
    }
    fn = cf;
    t64.push_back(make_double4(0, 0, 0, 0));  // unused by the closed-form FP64 kernels
  }
  for (int f = GEN_RHO_PD; f < GEN_NFN; ++f) {
    set_fn(g.fn32[f], fn[f], (int)t32.size());
    for (int m = 0; m + 1 < fn[f].n; ++m) {
      const double* q = &fn[f].c[4 * (size_t)m];
      t32.push_back(make_float4((float)q[0], (float)q[1], (float)q[2], (float)q[3]));
    }
  }
  if (int rc = upload(c, &c->gen_tab, c->gen_tab_n, t64)) return rc;
  if (int rc = upload(c, &c->gen_tab32, c->gen_tab32_n, t32)) return rc;
  g.tab = c->gen_tab;
  g.tab32 = c->gen_tab32;
  return A2C_OK;
}

// the descriptor's rule as nodes: the 12-point rule of quadrature.cpp:43-75
// (xi = +-sqrt(3) e_m in 6-D, w = 1/12, m outer and + first), or the 27
// Gauss-Hermite nodes on the relative coordinate
void rule_nodes(int rule, std::vector<double>& nd, int& rel) {
  nd.clear();
  if (rule == A2C_QUAD_GH3) {
    rel = 1;
    for (int k = 0; k < 27; ++k) {
      const int a[3] = {k % 3 - 1, (k / 3) % 3 - 1, k / 9 - 1};
      double w = 1.0;
      for (int c = 0; c < 3; ++c) {
        nd.push_back(GH_Z * a[c]);
        w *= gh_w(a[c]);
      }
      for (int c = 0; c < 3; ++c) nd.push_back(0.0);
      nd.push_back(w);
      nd.push_back(0.0);
    }
    return;
  }
  rel = 0;
  const double v = std::sqrt(3.0);
  for (int m = 0; m < 6; ++m)
    for (int sg = 0; sg < 2; ++sg) {
      for (int c = 0; c < 6; ++c) nd.push_back(c == m ? (sg ? -v : v) : 0.0);
      nd.push_back(1.0 / 12.0);
      nd.push_back(0.0);
    }
}

int install_nodes(a2c_fe_ctx* c, const std::vector<double>& nd, int rel) {
  // one buffer of the largest table, kept for the context's lifetime
  if (!c->gen_nodes && dmalloc(&c->gen_nodes, sizeof(double) * A2C_QUAD_MAX_NODES * 8) != hipSuccess)
    return set_err(c, A2C_ENOMEM, "quadrature nodes");
  A2C_HIP(c, hipMemcpy(c->gen_nodes, nd.data(), sizeof(double) * nd.size(), hipMemcpyHostToDevice));
  c->gen.nodes = c->gen_nodes;
  c->gen.K = (int)(nd.size() / 8);
  c->gen.rel = rel;
  return A2C_OK;
}

// after a change: nodes of the descriptor's rule when no table is set; the
// general kernels are on while either part is set
int refresh(a2c_fe_ctx* c) {
  c->gen_on = c->gen_table_rule || c->gen.pot == 1;
  if (c->gen_on && !c->gen_table_rule) {
    std::vector<double> nd;
    int rel = 0;
    rule_nodes(c->desc.quad_rule, nd, rel);
    if (int rc = install_nodes(c, nd, rel)) return rc;
  }
  if (c->gen_on && !c->gen_tab32)  // the closed forms' FP32 pieces
    if (int rc = install_potential(c, nullptr)) return rc;
  c->lat_dirty = true;
  return A2C_OK;
}

// ---------------------------------------------------------------- setfl
struct Setfl {
  std::vector<std::string> el;
  int nrho = 0, nr = 0;
  double drho = 0, dr = 0, cut = 0;
  std::vector<std::vector<double>> F, f, z;  // per element; z per pair i >= j
};

bool read_setfl(const char* path, Setfl& s, std::string& err) {
  std::ifstream in(path);
  if (!in) { err = std::string("cannot open ") + path; return false; }
  std::string line;
  for (int i = 0; i < 3; ++i)
    if (!std::getline(in, line)) { err = "short header"; return false; }
  if (!std::getline(in, line)) { err = "missing element line"; return false; }
  {
    std::istringstream ls(line);
    int n = 0;
    if (!(ls >> n) || n < 1) { err = "bad element count"; return false; }
    for (int i = 0; i < n; ++i) {
      std::string e;
      if (!(ls >> e)) { err = "fewer element names than the count"; return false; }
      s.el.push_back(e);
    }
  }
  if (!std::getline(in, line)) { err = "missing grid line"; return false; }
  {
    std::istringstream ls(line);
    if (!(ls >> s.nrho >> s.drho >> s.nr >> s.dr >> s.cut) || s.nrho < 5 || s.nr < 5 || !(s.drho > 0) ||
        !(s.dr > 0)) {
      err = "bad grid line (n_rho d_rho n_r d_r cutoff; at least 5 nodes)";
      return false;
    }
  }
  auto values = [&](int n, std::vector<double>& v) {
    v.resize(n);
    for (int k = 0; k < n; ++k)
      if (!(in >> v[k])) return false;
    std::getline(in, line);  // rest of the last value line
    return true;
  };
  const int ne = (int)s.el.size();
  s.F.resize(ne);
  s.f.resize(ne);
  for (int i = 0; i < ne; ++i) {
    if (!std::getline(in, line)) { err = "missing element header"; return false; }
    if (!values(s.nrho, s.F[i]) || !values(s.nr, s.f[i])) { err = "short element tables"; return false; }
  }
  s.z.resize((size_t)ne * (ne + 1) / 2);
  for (size_t p = 0; p < s.z.size(); ++p)
    if (!values(s.nr, s.z[p])) { err = "short pair tables"; return false; }
  return true;
}

int element_index(const Setfl& s, const char* name) {
  for (size_t i = 0; i < s.el.size(); ++i)
    if (name && s.el[i] == name) return (int)i;
  return -1;
}
const std::vector<double>& pair_table(const Setfl& s, int a, int b) {
  const int i = a > b ? a : b, j = a > b ? b : a;
  return s.z[(size_t)i * (i + 1) / 2 + j];
}

}  // namespace

void gen_free(a2c_fe_ctx* c) {
  dfree(c->gen_nodes);
  dfree(c->gen_tab);
  dfree(c->gen_tab32);
  c->gen_nodes = nullptr;
  c->gen_tab = nullptr;
  c->gen_tab32 = nullptr;
  c->gen_tab_n = c->gen_tab32_n = 0;
}

}  // namespace a2c

using namespace a2c;

extern "C" {

int a2c_fe_set_potential_tables(a2c_fe_ctx* c, const a2c_eam_tabulated* t) {
  if (!c) return A2C_EINVAL;
  if (!t || !t->F_pd || !t->F_h || !t->f_pd || !t->f_h || !t->rphi_pdpd || !t->rphi_pdh || !t->rphi_hh)
    return set_err(c, A2C_EINVAL, "set_potential_tables: null table");
  const double* src[GEN_NFN] = {t->F_pd, t->F_h, t->f_pd, t->f_h, t->rphi_pdpd, t->rphi_pdh, t->rphi_hh};
  Pieces P[GEN_NFN];
  for (int f = 0; f < GEN_NFN; ++f) {
    const bool emb = f <= GEN_F_H;
    if (!make_pieces(src[f], emb ? t->n_rho : t->n_r, emb ? t->d_rho : t->d_r, P[f]))
      return set_err(c, A2C_EINVAL, "set_potential_tables: need >= 5 finite nodes and a positive step");
  }
  if (hipSetDevice(c->desc.device) != hipSuccess) return set_err(c, A2C_EHIP, "hipSetDevice");
  A2C_HIP(c, hipStreamSynchronize(c->stream));  // kernels in flight may read the old tables
  if (int rc = install_potential(c, P)) return rc;
  c->gen.pot = 1;
  return refresh(c);
}

int a2c_fe_set_potential_file(a2c_fe_ctx* c, const char* path, const char* host, const char* solute) {
  if (!c) return A2C_EINVAL;
  Setfl s;
  std::string err;
  if (!path || !read_setfl(path, s, err)) return set_err(c, A2C_EINVAL, "setfl: " + err);
  const int a = element_index(s, host), b = element_index(s, solute);
  if (a < 0 || b < 0) return set_err(c, A2C_EINVAL, "setfl: element not in the file");
  a2c_eam_tabulated t;
  t.n_rho = s.nrho;
  t.d_rho = s.drho;
  t.n_r = s.nr;
  t.d_r = s.dr;
  t.F_pd = s.F[a].data();
  t.F_h = s.F[b].data();
  t.f_pd = s.f[a].data();
  t.f_h = s.f[b].data();
  t.rphi_pdpd = pair_table(s, a, a).data();
  t.rphi_pdh = pair_table(s, a, b).data();
  t.rphi_hh = pair_table(s, b, b).data();
  return a2c_fe_set_potential_tables(c, &t);
}

// Host-only reader of a setfl file into caller arrays (sizes from the first
// call with null arrays): the tables a2c_fe_set_potential_file would install.
int a2c_eam_setfl_read(const char* path, const char* host, const char* solute, int32_t* n_rho, double* d_rho,
                       int32_t* n_r, double* d_r, double* cutoff, double* F_pd, double* F_h, double* f_pd,
                       double* f_h, double* rphi_pdpd, double* rphi_pdh, double* rphi_hh) {
  Setfl s;
  std::string err;
  if (!path || !read_setfl(path, s, err)) return A2C_EINVAL;
  const int a = element_index(s, host), b = element_index(s, solute);
  if (a < 0 || b < 0) return A2C_EINVAL;
  if (n_rho) *n_rho = s.nrho;
  if (d_rho) *d_rho = s.drho;
  if (n_r) *n_r = s.nr;
  if (d_r) *d_r = s.dr;
  if (cutoff) *cutoff = s.cut;
  auto put = [](double* dst, const std::vector<double>& v) {
    if (dst) std::memcpy(dst, v.data(), sizeof(double) * v.size());
  };
  put(F_pd, s.F[a]);
  put(F_h, s.F[b]);
  put(f_pd, s.f[a]);
  put(f_h, s.f[b]);
  put(rphi_pdpd, pair_table(s, a, a));
  put(rphi_pdh, pair_table(s, a, b));
  put(rphi_hh, pair_table(s, b, b));
  return A2C_OK;
}

int a2c_fe_clear_potential(a2c_fe_ctx* c) {
  if (!c) return A2C_EINVAL;
  if (hipSetDevice(c->desc.device) != hipSuccess) return set_err(c, A2C_EHIP, "hipSetDevice");
  A2C_HIP(c, hipStreamSynchronize(c->stream));
  c->gen.pot = 0;
  // the closed forms' FP32 pieces replace the tabulated ones
  if (int rc = install_potential(c, nullptr)) return rc;
  return refresh(c);
}

int a2c_fe_set_quadrature(a2c_fe_ctx* c, int32_t K, const double* xi, const double* w) {
  if (!c) return A2C_EINVAL;
  if (K < 0 || K > A2C_QUAD_MAX_NODES || (K > 0 && (!xi || !w)))
    return set_err(c, A2C_EINVAL, "set_quadrature: K out of range or null table");
  if (hipSetDevice(c->desc.device) != hipSuccess) return set_err(c, A2C_EHIP, "hipSetDevice");
  A2C_HIP(c, hipStreamSynchronize(c->stream));
  if (K == 0) {
    c->gen_table_rule = false;
    return refresh(c);
  }
  // swap symmetry: node k's partner carries (xi2, xi1) and the same weight
  for (int k = 0; k < K; ++k) {
    if (!std::isfinite(w[k])) return set_err(c, A2C_EINVAL, "set_quadrature: weight not finite");
    const double* a = xi + 6 * (size_t)k;
    double scale = 1.0;
    for (int e = 0; e < 6; ++e) {
      if (!std::isfinite(a[e])) return set_err(c, A2C_EINVAL, "set_quadrature: node not finite");
      scale = std::fmax(scale, std::fabs(a[e]));
    }
    bool found = false;
    for (int m = 0; m < K && !found; ++m) {
      const double* b = xi + 6 * (size_t)m;
      bool same = std::fabs(w[m] - w[k]) <= 1e-12 * std::fmax(1.0, std::fabs(w[k]));
      for (int e = 0; e < 3 && same; ++e)
        same = std::fabs(b[e] - a[3 + e]) <= 1e-12 * scale && std::fabs(b[3 + e] - a[e]) <= 1e-12 * scale;
      found = same;
    }
    if (!found)
      return set_err(c, A2C_EINVAL, "set_quadrature: the table is not swap-symmetric (node " + std::to_string(k) +
                                        " has no partner with xi1 and xi2 exchanged)");
  }
  std::vector<double> nd;
  for (int k = 0; k < K; ++k) {
    for (int e = 0; e < 6; ++e) nd.push_back(xi[6 * (size_t)k + e]);
    nd.push_back(w[k]);
    nd.push_back(0.0);
  }
  if (int rc = install_nodes(c, nd, 0)) return rc;
  c->gen_table_rule = true;
  return refresh(c);
}

}  // extern "C"
