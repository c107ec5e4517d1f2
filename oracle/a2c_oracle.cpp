/*
 * a2c_oracle.cpp -- FP64 CPU restatement of the a2c free-entropy hot path.
 *
 * TEST INFRASTRUCTURE ONLY (see a2c_oracle.h): the checker for the HIP path,
 * never shipped and never called by the product library.
 *
 * Built with g++ -O2 -ffp-contract=off (no -march=native, no fast-math): the
 * reference is plain x86-64 C++ built without FMA, and this file keeps the
 * reference's expression order so that FD-mode results follow its rounding.
 */
#include "a2c_oracle.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>
#ifdef _OPENMP
#include <omp.h>
#endif

#define NOIPA __attribute__((noipa))

namespace {

/* ---------------------------------------------------------------------------
 * EAM Pd-H closed forms: potential.cpp:16-205.  Each function keeps the
 * reference's expression order; constants are listed once in the K_* tables
 * (order of appearance in potential.cpp) and the bodies read the tables.
 * ------------------------------------------------------------------------- */
const double K_FPD[] = {  /* potential.cpp:18-31,35 */
  295878.9003038662, 0.20581955357385892, 0.081228755904399, 0.05298811034615951,
  2.4242616904962846, 1.4791899886249564, 2.1376274623740064, 1.2169215689822592,
  1.6486007989726832, 0.8159825255339774, 1.0749204110338482, 0.42007491336688396,
  0.5128056047933808, 0.12468685331167456, 0.251546};
const double K_FH[] = {   /* potential.cpp:47-51,54 */
  0.000197047, 1.18860, 9.99780, 60.0155, 0.0540638, -0.0296604};
const double K_FPDR[] = { /* potential.cpp:66-67,69-82,86-90 */
  0.001510569378297, -0.077976566126292,
  2.7267629107325706, 1.8716766113599643, 2.50290548851635, 1.668549182690922,
  2.0924467509943674, 1.3150372774478005, 1.564328475106985, 0.8987511149780485,
  1.009780903403673, 0.5124363774128722, 0.5304054524800665, 0.2169886022464641,
  0.1356566408715063, 0.035852347523891395,
  -0.02972698211669922, 0.6676807403564453, -255.8965835571289, 14673.409149169922,
  -2.597301181336601e7};
const double K_FHR[] = {  /* potential.cpp:107-108,110-111 */
  0.009986990526578, -0.013075667086408, 11.0025, 1.30927};
const double K_PHPD[] = { /* potential.cpp:128-129,131-143,148 */
  -0.016958780164980, 0.089019679316328,
  -79415.24035137112, 1.0699996145674568, 1.06015072612581, 0.42433991011376526,
  0.06169160085238687, 2.0586473420376348, 1.0683922574015199, 1.6696359816422877,
  0.7337878627470482, 1.1690370066230809, 0.3909805777737639, 0.2635598721249787,
  0.033551116514910245, 0.251546};
const double K_PHPDH[] = { /* potential.cpp:164-165,167-170 */
  -3.353001113951400e-04, 7.147062196813336e-04, 0.2494540, 4.82613, 2.13158, 1.50964};
const double K_PHH[] = {  /* potential.cpp:187-188,190-193 */
  -0.003702722457147, 0.005465945471933, 0.0661496, 3.67263, 1.47797, 2.51980};

/* Literal scale divisors and switch radii inside the bodies. */
const double S_FPD = 50.0, S_FPDR = 5.0, S_PHPD = 5.0;
const double R_FPDR = 5.0, R_FHR = 5.35, R_PHPD = 5.05, R_PHPDH = 5.35, R_PHH = 5.35;

/* value + derivative pair used for the analytic derivatives */
struct D2 { double v, d; };
inline D2 dmul(D2 a, D2 b) { return {a.v * b.v, a.d * b.v + a.v * b.d}; }

/* F_Pd: potential.cpp:16-38 */
NOIPA double F_Pd(double rho) {
  const double* k = K_FPD;
  rho /= S_FPD;
  double y = k[0] * (rho - k[1]) * (rho - k[2]) * rho * (rho + k[3]) *
             (rho * (rho - k[4]) + k[5]) * (rho * (rho - k[6]) + k[7]) *
             (rho * (rho - k[8]) + k[9]) * (rho * (rho - k[10]) + k[11]) *
             (rho * (rho - k[12]) + k[13]);
  y -= S_FPD * k[14] * rho;
  return y;
}
D2 F_Pd_a(double rho) {
  const double* k = K_FPD;
  double r = rho / S_FPD;
  D2 p = {k[0], 0.0};
  p = dmul(p, {r - k[1], 1.0});
  p = dmul(p, {r - k[2], 1.0});
  p = dmul(p, {r, 1.0});
  p = dmul(p, {r + k[3], 1.0});
  for (int q = 0; q < 5; ++q) {
    double b = k[4 + 2 * q], c = k[5 + 2 * q];
    p = dmul(p, {r * (r - b) + c, 2.0 * r - b});
  }
  return {p.v - S_FPD * k[14] * r, p.d / S_FPD - k[14]};
}

/* F_H: potential.cpp:45-57 */
NOIPA double F_H(double rho) {
  const double* k = K_FH;
  double c = k[0], d = k[1], a = k[2], b = k[3], e1 = k[4];
  double y = -c * (1.0 / (2.0 + d) * std::pow(rho + e1, 2.0 + d) -
                   (a + b) / (1.0 + d) * std::pow(rho + e1, 1.0 + d) +
                   a * b / d * std::pow(rho + e1, d));
  y -= k[5] * rho;
  return y;
}
D2 F_H_a(double rho) {
  const double* k = K_FH;
  double c = k[0], d = k[1], a = k[2], b = k[3], u = rho + k[4];
  double p2 = std::pow(u, 2.0 + d), p1 = std::pow(u, 1.0 + d), p0 = std::pow(u, d);
  double v = -c * (1.0 / (2.0 + d) * p2 - (a + b) / (1.0 + d) * p1 + a * b / d * p0) - k[5] * rho;
  double dv = -c * (p1 - (a + b) * p0 + a * b * std::pow(u, d - 1.0)) - k[5];
  return {v, dv};
}

/* f_Pd: potential.cpp:64-98 (tail compare on the rescaled value, :93-96) */
NOIPA double f_Pd(double r) {
  const double* k = K_FPDR;
  r /= S_FPDR;
  double rho_a = r * (r * (r - k[2]) + k[3]) * (r * (r - k[4]) + k[5]) *
                 (r * (r - k[6]) + k[7]) * (r * (r - k[8]) + k[9]) *
                 (r * (r - k[10]) + k[11]) * (r * (r - k[12]) + k[13]) *
                 (r * (r - k[14]) + k[15]);
  double y = k[16] + r * (k[17] + r * (k[18] + r * (k[19] + k[20] * rho_a)));
  if (r * S_FPDR <= 0.0)
    y = 0.0;
  else if (r * S_FPDR > R_FPDR)
    y = k[0] * std::exp(k[1] / k[0] * (r * S_FPDR - R_FPDR));
  return y;
}
D2 f_Pd_a(double r) {
  const double* k = K_FPDR;
  double t = r / S_FPDR, rr = t * S_FPDR;
  if (rr <= 0.0) return {0.0, 0.0};
  if (rr > R_FPDR) {
    double y = k[0] * std::exp(k[1] / k[0] * (rr - R_FPDR));
    return {y, k[1] / k[0] * y};
  }
  D2 p = {t, 1.0};
  for (int q = 0; q < 7; ++q) {
    double b = k[2 + 2 * q], c = k[3 + 2 * q];
    p = dmul(p, {t * (t - b) + c, 2.0 * t - b});
  }
  /* y = b1 + t(b2 + t(b3 + t(b4 + b5 p))) */
  D2 h = {k[19] + k[20] * p.v, k[20] * p.d};
  h = {k[18] + t * h.v, h.v + t * h.d};
  h = {k[17] + t * h.v, h.v + t * h.d};
  h = {k[16] + t * h.v, h.v + t * h.d};
  return {h.v, h.d / S_FPDR};
}

/* f_H: potential.cpp:105-119 */
NOIPA double f_H(double r) {
  const double* k = K_FHR;
  double y = k[2] * std::exp(-k[3] * r);
  if (r <= 0.0)
    y = 0.0;
  else if (r > R_FHR)
    y = k[0] * std::exp(k[1] / k[0] * (r - R_FHR));
  return y;
}
D2 f_H_a(double r) {
  const double* k = K_FHR;
  if (r <= 0.0) return {0.0, 0.0};
  if (r > R_FHR) {
    double y = k[0] * std::exp(k[1] / k[0] * (r - R_FHR));
    return {y, k[1] / k[0] * y};
  }
  double y = k[2] * std::exp(-k[3] * r);
  return {y, -k[3] * y};
}

/* phi_Pd: potential.cpp:126-155 */
NOIPA double phi_Pd(double r) {
  const double* k = K_PHPD;
  r /= S_PHPD;
  double y = k[2] * (r - k[3]) * (r - k[4]) * (r - k[5]) * (r + k[6]) *
             (r * (r - k[7]) + k[8]) * (r * (r - k[9]) + k[10]) *
             (r * (r - k[11]) + k[12]) * (r * (r - k[13]) + k[14]);
  r *= S_PHPD;
  y += 2 * k[15] * f_Pd(r);
  if (r <= 0.0)
    y = 0.0;
  else if (r > R_PHPD)
    y = k[0] * std::exp(k[1] / k[0] * (r - R_PHPD));
  return y;
}
D2 phi_Pd_a(double r) {
  const double* k = K_PHPD;
  double t = r / S_PHPD, rr = t * S_PHPD;
  if (rr <= 0.0) return {0.0, 0.0};
  if (rr > R_PHPD) {
    double y = k[0] * std::exp(k[1] / k[0] * (rr - R_PHPD));
    return {y, k[1] / k[0] * y};
  }
  D2 p = {k[2], 0.0};
  p = dmul(p, {t - k[3], 1.0});
  p = dmul(p, {t - k[4], 1.0});
  p = dmul(p, {t - k[5], 1.0});
  p = dmul(p, {t + k[6], 1.0});
  for (int q = 0; q < 4; ++q) {
    double b = k[7 + 2 * q], c = k[8 + 2 * q];
    p = dmul(p, {t * (t - b) + c, 2.0 * t - b});
  }
  D2 fp = f_Pd_a(rr);
  return {p.v + 2 * k[15] * fp.v, p.d / S_PHPD + 2 * k[15] * fp.d};
}

/* phi_PdH / phi_H: potential.cpp:162-200 (same Morse-like form) */
inline double morse(const double* k, double rsw, double r) {
  double D = k[2], al = k[3], be = k[4], r0 = k[5];
  double y = D * (be * std::exp(-al * (r - r0)) - al * std::exp(-be * (r - r0)));
  if (r <= 0.0)
    y = 0.0;
  else if (r > rsw)
    y = k[0] * std::exp(k[1] / k[0] * (r - rsw));
  return y;
}
inline D2 morse_a(const double* k, double rsw, double r) {
  if (r <= 0.0) return {0.0, 0.0};
  if (r > rsw) {
    double y = k[0] * std::exp(k[1] / k[0] * (r - rsw));
    return {y, k[1] / k[0] * y};
  }
  double D = k[2], al = k[3], be = k[4], r0 = k[5];
  double ea = std::exp(-al * (r - r0)), eb = std::exp(-be * (r - r0));
  return {D * (be * ea - al * eb), D * al * be * (eb - ea)};
}
NOIPA double phi_PdH(double r) { return morse(K_PHPDH, R_PHPDH, r); }
NOIPA double phi_H(double r) { return morse(K_PHH, R_PHH, r); }

/* OR_POT_TEST: g(r) = r^2 for every pair/density function, F(rho) = rho. */
int g_potential = OR_POT_EAM;

/* OR_POT_TAB: the seven functions as uniform-grid node values (pair
 * functions as r*phi), interpolated piecewise by cubic Hermite polynomials
 * in the Hermite basis h00 = 2u^3 - 3u^2 + 1, h10 = u^3 - 2u^2 + u,
 * h01 = 3u^2 - 2u^3, h11 = u^3 - u^2 with node slopes (per grid step) from
 * central differences: (y[m-2] - 8y[m-1] + 8y[m+1] - y[m+2])/12 inside,
 * (y[m+1] - y[m-1])/2 at the second and second-to-last node, one-sided at
 * the ends -- the interpolation DYNAMO setfl tables are written for.  Past
 * the last node r-functions are 0 and embedding functions continue along the
 * end slope.  Function order as OR_F_PD .. OR_PHI_H. */
struct Tab {
  std::vector<double> y, s;
  double h = 0.0;
};
Tab g_tab[OR_NFUNC];

void tab_slopes(Tab& t) {
  const std::vector<double>& y = t.y;
  const int n = (int)y.size();
  t.s.assign(n, 0.0);
  t.s[0] = y[1] - y[0];
  t.s[n - 1] = y[n - 1] - y[n - 2];
  t.s[1] = (y[2] - y[0]) / 2.0;
  t.s[n - 2] = (y[n - 1] - y[n - 3]) / 2.0;
  for (int m = 2; m <= n - 3; ++m) t.s[m] = (y[m - 2] - 8.0 * y[m - 1] + 8.0 * y[m + 1] - y[m + 2]) / 12.0;
}

/* value (d = 0) or slope (d = 1) of table t at x */
double tab_eval(const Tab& t, double x, bool emb, int d) {
  const int n = (int)t.y.size();
  const double p = x / t.h;
  int m = (int)std::floor(p);
  if (m < 0) m = 0;
  if (m >= n - 1) {
    if (!emb) return 0.0;
    const double slope = t.s[n - 1] / t.h;
    return d ? slope : t.y[n - 1] + slope * (x - (n - 1) * t.h);
  }
  const double u = p - m, u2 = u * u, u3 = u2 * u;
  if (!d)
    return (2 * u3 - 3 * u2 + 1) * t.y[m] + (u3 - 2 * u2 + u) * t.s[m] + (3 * u2 - 2 * u3) * t.y[m + 1] +
           (u3 - u2) * t.s[m + 1];
  return ((6 * u2 - 6 * u) * t.y[m] + (3 * u2 - 4 * u + 1) * t.s[m] + (6 * u - 6 * u2) * t.y[m + 1] +
          (3 * u2 - 2 * u) * t.s[m + 1]) / t.h;
}

double tab_val(int fn, double x) {
  const bool emb = fn == OR_F_PD || fn == OR_F_H;
  const double v = tab_eval(g_tab[fn], x, emb, 0);
  return fn >= OR_PHI_PD ? v / x : v;
}
double tab_der(int fn, double x) {
  const bool emb = fn == OR_F_PD || fn == OR_F_H;
  const double d = tab_eval(g_tab[fn], x, emb, 1);
  if (fn < OR_PHI_PD) return d;
  return (d * x - tab_eval(g_tab[fn], x, emb, 0)) / (x * x);  /* (z/x)' */
}

double eam_val(int fn, double x) {
  if (g_potential == OR_POT_TEST) return (fn == OR_F_PD || fn == OR_F_H) ? x : x * x;
  if (g_potential == OR_POT_TEST4) return (fn == OR_F_PD || fn == OR_F_H) ? x : x * x * x * x;
  if (g_potential == OR_POT_TAB) return tab_val(fn, x);
  switch (fn) {
    case OR_F_PD: return F_Pd(x);
    case OR_F_H: return F_H(x);
    case OR_RHO_PD: return f_Pd(x);
    case OR_RHO_H: return f_H(x);
    case OR_PHI_PD: return phi_Pd(x);
    case OR_PHI_PDH: return phi_PdH(x);
    case OR_PHI_H: return phi_H(x);
  }
  return NAN;
}

/* Derivatives.  FD mode is the reference's central difference with step
 * eps (potential.cpp:40-43,59-62,100-103,121-124,157-160,179-182,202-205). */
double eam_der(int fn, double x, int mode, double eps) {
  if (mode == OR_DERIV_FD) return (eam_val(fn, x + eps) - eam_val(fn, x - eps)) / (2.0 * eps);
  if (g_potential == OR_POT_TEST) return (fn == OR_F_PD || fn == OR_F_H) ? 1.0 : 2.0 * x;
  if (g_potential == OR_POT_TEST4) return (fn == OR_F_PD || fn == OR_F_H) ? 1.0 : 4.0 * x * x * x;
  if (g_potential == OR_POT_TAB) return tab_der(fn, x);
  switch (fn) {
    case OR_F_PD: return F_Pd_a(x).d;
    case OR_F_H: return F_H_a(x).d;
    case OR_RHO_PD: return f_Pd_a(x).d;
    case OR_RHO_H: return f_H_a(x).d;
    case OR_PHI_PD: return phi_Pd_a(x).d;
    case OR_PHI_PDH: return morse_a(K_PHPDH, R_PHPDH, x).d;
    case OR_PHI_H: return morse_a(K_PHH, R_PHH, x).d;
  }
  return NAN;
}

/* ---------------------------------------------------------------------------
 * Vec3D semantics (Vector3D.h:47-88): component-wise ops, norm =
 * sqrt(x*x+y*y+z*z) evaluated left to right (:83), dot left to right (:281).
 * ------------------------------------------------------------------------- */
struct V3 { double v[3]; };
inline V3 vadd(const V3& a, const V3& b) { return {{a.v[0] + b.v[0], a.v[1] + b.v[1], a.v[2] + b.v[2]}}; }
inline V3 vsub(const V3& a, const V3& b) { return {{a.v[0] - b.v[0], a.v[1] - b.v[1], a.v[2] - b.v[2]}}; }
inline V3 vscale(double c, const V3& a) { return {{c * a.v[0], c * a.v[1], c * a.v[2]}}; }
inline double vdot(const V3& a, const V3& b) { return a.v[0] * b.v[0] + a.v[1] * b.v[1] + a.v[2] * b.v[2]; }
inline double vnorm(const V3& a) { return std::sqrt(a.v[0] * a.v[0] + a.v[1] * a.v[1] + a.v[2] * a.v[2]); }

/* Quadrature: quadrature.cpp:43-75, node value sqrt(dim/2), weight 0.5/dim
 * (the weight of the commented twin rule, quadrature.cpp:39). */
const int QDIM = 6, QK = 2 * QDIM;
struct Quad { V3 xi1[QK], xi2[QK]; double w; };
Quad make_quad() {
  Quad q;
  int k = 0;
  for (int i = 0; i < QDIM; ++i) {
    for (int sgn = 0; sgn < 2; ++sgn) {
      double t[QDIM];
      for (int j = 0; j < QDIM; ++j) t[j] = (i == j) ? (sgn == 0 ? std::sqrt(QDIM / 2.0) : -std::sqrt(QDIM / 2.0)) : 0.0;
      q.xi1[k] = {{t[0], t[1], t[2]}};
      q.xi2[k] = {{t[3], t[4], t[5]}};
      ++k;
    }
  }
  q.w = 0.5 / QDIM;
  return q;
}
const Quad QUAD = make_quad();

/* Nodes of one pair under either rule: relative vector r_k, weight w_k and
 * dsi_k, the derivative of r_k w.r.t. sigma_i up to the factor sigfac (the
 * 12-point rule keeps the reference's sqrt(2) as a final factor on the
 * summed sigma gradient, minimizer.cpp:221,283). */
const int MAXK = 512;
struct Nodes {
  int K;
  V3 r[MAXK];
  double w[MAXK];
  V3 dsi[MAXK];
};
/* OR_QUAD_TABLE: a node table (xi1, xi2, w) used like the 12-point rule. */
struct QTable { std::vector<V3> xi1, xi2; std::vector<double> w; };
QTable g_qtab;
const double SQ3 = 1.7320508075688772935;
inline void pair_nodes(int rule, const V3& qi, double si, const V3& qj, double sj, Nodes& n) {
  const double sq2 = std::sqrt(2.0);
  if (rule == OR_QUAD_TABLE) {
    n.K = (int)g_qtab.w.size();
    for (int k = 0; k < n.K; ++k) {
      n.r[k] = vsub(vsub(vadd(qi, vscale(sq2 * si, g_qtab.xi1[k])), qj), vscale(sq2 * sj, g_qtab.xi2[k]));
      n.w[k] = g_qtab.w[k];
      n.dsi[k] = g_qtab.xi1[k];
    }
    return;
  }
  if (rule != OR_QUAD_GH3) {
    n.K = QK;
    for (int k = 0; k < QK; ++k) {
      n.r[k] = vsub(vsub(vadd(qi, vscale(sq2 * si, QUAD.xi1[k])), qj), vscale(sq2 * sj, QUAD.xi2[k]));
      n.w[k] = QUAD.w;
      n.dsi[k] = QUAD.xi1[k];
    }
    return;
  }
  static const double z1[3] = {-SQ3, 0.0, SQ3}, w1[3] = {1.0 / 6.0, 2.0 / 3.0, 1.0 / 6.0};
  const V3 d = vsub(qi, qj);
  const double s = std::sqrt(si * si + sj * sj);
  const double ds = s > 0.0 ? si / s : 0.0;
  n.K = 27;
  int k = 0;
  for (int a = 0; a < 3; ++a)
    for (int b = 0; b < 3; ++b)
      for (int c = 0; c < 3; ++c, ++k) {
        const V3 z = {{z1[a], z1[b], z1[c]}};
        n.r[k] = vadd(d, vscale(s, z));
        n.w[k] = w1[a] * w1[b] * w1[c];
        n.dsi[k] = vscale(ds, z);
      }
}
inline double sigma_factor(int rule) { return rule == OR_QUAD_GH3 ? 1.0 : std::sqrt(2.0); }

const double KB = 8.6173324e-5; /* minimizer.cpp:92 */

/* ---------------------------------------------------------------------------
 * System view over the 4N host layout (minimizer.cpp:398-418,464-473).
 * ------------------------------------------------------------------------- */
struct Sys {
  const or_params* p;
  const double* X;
  int npd, nh;
  V3 q(int s) const { return {{X[4 * s], X[4 * s + 1], X[4 * s + 2]}}; }
  double sig(int s) const { return X[4 * s + 3]; }
};

/* Periodic image of q_j closest to q_i: Delta -= L*nearbyint(Delta/L) on the
 * periodic axes.  Extension (the reference has no minimum image,
 * neighbors.cpp:78-82); identity on non-periodic axes. */
inline V3 image_of(const or_params* p, const V3& qi, const V3& qj) {
  V3 r = qj;
  for (int c = 0; c < 3; ++c)
    if (p->periodic[c]) {
      double d = qj.v[c] - qi.v[c];
      double n = std::nearbyint(d / p->box[c]);
      r.v[c] = qj.v[c] - p->box[c] * n;
    }
  return r;
}

int nthreads(const or_params* p) { return p->threads > 0 ? p->threads : 1; }

}  // namespace

/* ===========================================================================
 * Neighbor lists.  Sets follow neighbors.cpp:74-96 and minimizer.cpp:923-1005:
 * j != i and ((dx*dx + dy*dy) + dz*dz) <= rc*rc with d = q_j - q_i (minimum
 * image on periodic axes); PdPdnum shells d2 < 0.75a^2, 0.75a^2 < d2 <
 * 1.25a^2, else (:935-947); HH1 = H neighbors with d2 < 0.75a^2 (:999-1004).
 * Lists are sorted by index (the reference's order is KD-tree order).
 * ========================================================================= */
struct or_nbr {
  std::vector<int32_t> off[5], idx[5];
  std::vector<int32_t> pdpdnum;
  int maxneib = 0;
};

extern "C" or_nbr* or_nbr_build(const or_params* p, const double* X) {
  const int npd = p->n_pd, nh = p->n_h, N = npd + nh;
  for (int c = 0; c < 3; ++c)
    if (p->periodic[c] && !(p->box[c] >= 2.0 * p->rc)) return nullptr;
  /* cell grid */
  double lo[3], hi[3];
  int nc[3];
  for (int c = 0; c < 3; ++c) {
    if (p->periodic[c]) { lo[c] = 0.0; hi[c] = p->box[c]; }
    else {
      lo[c] = 1e300; hi[c] = -1e300;
      for (int s = 0; s < N; ++s) { lo[c] = std::min(lo[c], X[4 * s + c]); hi[c] = std::max(hi[c], X[4 * s + c]); }
      if (N == 0) { lo[c] = 0; hi[c] = 1; }
    }
    nc[c] = std::max(1, (int)std::floor((hi[c] - lo[c]) / p->rc));
    if (p->periodic[c] && nc[c] < 3) nc[c] = 1;
  }
  auto cell_of = [&](int s, int c) {
    double x = X[4 * s + c];
    if (p->periodic[c]) x -= p->box[c] * std::floor(x / p->box[c]);
    int k = (int)std::floor((x - lo[c]) / (hi[c] - lo[c]) * nc[c]);
    return std::min(std::max(k, 0), nc[c] - 1);
  };
  const int ncell = nc[0] * nc[1] * nc[2];
  std::vector<int> cstart(ncell + 1, 0), cid(N), order(N);
  for (int s = 0; s < N; ++s) {
    cid[s] = (cell_of(s, 2) * nc[1] + cell_of(s, 1)) * nc[0] + cell_of(s, 0);
    cstart[cid[s] + 1]++;
  }
  for (int c = 0; c < ncell; ++c) cstart[c + 1] += cstart[c];
  {
    std::vector<int> fill(cstart.begin(), cstart.end() - 1);
    for (int s = 0; s < N; ++s) order[fill[cid[s]]++] = s;
  }
  const double rc2 = p->rc * p->rc;
  const double sh1 = p->a_pd * p->a_pd * 0.75, sh2 = p->a_pd * p->a_pd * 1.25;

  std::vector<std::vector<int32_t>> nb(N);
  std::vector<std::vector<double>> nbd2(N);
#pragma omp parallel for schedule(dynamic, 64) num_threads(nthreads(p))
  for (int s = 0; s < N; ++s) {
    int ci[3] = {cid[s] % nc[0], (cid[s] / nc[0]) % nc[1], cid[s] / (nc[0] * nc[1])};
    int lst[3][3], cnt[3];
    for (int c = 0; c < 3; ++c) {
      cnt[c] = 0;
      for (int o = -1; o <= 1; ++o) {
        int k = ci[c] + o;
        if (p->periodic[c]) k = ((k % nc[c]) + nc[c]) % nc[c];
        else if (k < 0 || k >= nc[c]) continue;
        bool dup = false;
        for (int t = 0; t < cnt[c]; ++t) dup |= (lst[c][t] == k);
        if (!dup) lst[c][cnt[c]++] = k;
      }
    }
    std::vector<std::pair<int32_t, double>> found;
    const double qi[3] = {X[4 * s], X[4 * s + 1], X[4 * s + 2]};
    for (int a = 0; a < cnt[2]; ++a)
      for (int b = 0; b < cnt[1]; ++b)
        for (int e = 0; e < cnt[0]; ++e) {
          int c = (lst[2][a] * nc[1] + lst[1][b]) * nc[0] + lst[0][e];
          for (int t = cstart[c]; t < cstart[c + 1]; ++t) {
            int j = order[t];
            if (j == s) continue;
            double d[3];
            for (int k = 0; k < 3; ++k) {
              d[k] = X[4 * j + k] - qi[k];
              if (p->periodic[k]) d[k] = d[k] - p->box[k] * std::nearbyint(d[k] / p->box[k]);
            }
            double d2 = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
            if (d2 > rc2) continue;
            found.push_back({j, d2});
          }
        }
    std::sort(found.begin(), found.end());
    for (auto& f : found) { nb[s].push_back(f.first); nbd2[s].push_back(f.second); }
  }

  or_nbr* out = new or_nbr;
  out->pdpdnum.assign((size_t)npd * 3, 0);
  for (int w = 0; w < 5; ++w) out->off[w].push_back(0);
  for (int s = 0; s < N; ++s) {
    out->maxneib = std::max(out->maxneib, (int)nb[s].size());
    bool is_pd = s < npd;
    for (size_t t = 0; t < nb[s].size(); ++t) {
      int j = nb[s][t];
      double d2 = nbd2[s][t];
      if (is_pd) {
        if (j < npd) {
          out->idx[0].push_back(j);
          if (d2 < sh1) out->pdpdnum[3 * s]++;
          else if (d2 > sh1 && d2 < sh2) out->pdpdnum[3 * s + 1]++;
          else out->pdpdnum[3 * s + 2]++;
        } else out->idx[1].push_back(j - npd);
      } else {
        if (j < npd) out->idx[2].push_back(j);
        else {
          out->idx[3].push_back(j - npd);
          if (d2 < sh1) out->idx[4].push_back(j - npd);
        }
      }
    }
    if (is_pd) {
      out->off[0].push_back((int32_t)out->idx[0].size());
      out->off[1].push_back((int32_t)out->idx[1].size());
    } else {
      out->off[2].push_back((int32_t)out->idx[2].size());
      out->off[3].push_back((int32_t)out->idx[3].size());
      out->off[4].push_back((int32_t)out->idx[4].size());
    }
  }
  return out;
}

extern "C" int or_nbr_get(const or_nbr* nb, int which, const int32_t** off, const int32_t** idx) {
  if (!nb || which < 0 || which > 4) return -1;
  *off = nb->off[which].data();
  *idx = nb->idx[which].data();
  return (int)nb->off[which].size() - 1;
}
extern "C" const int32_t* or_nbr_pdpdnum(const or_nbr* nb) { return nb->pdpdnum.data(); }
extern "C" int or_nbr_max(const or_nbr* nb) { return nb->maxneib; }
extern "C" void or_nbr_free(or_nbr* nb) { delete nb; }

/* ===========================================================================
 * EAM / quadrature exports
 * ========================================================================= */
extern "C" double or_eam(int fn, double x) { return eam_val(fn, x); }
extern "C" double or_eam_d(int fn, double x, int mode, double eps) { return eam_der(fn, x, mode, eps); }
extern "C" int or_eam_constants(int fn, double* out, int cap) {
  const double* t = nullptr;
  int n = 0;
  switch (fn) {
    case OR_F_PD: t = K_FPD; n = sizeof(K_FPD) / 8; break;
    case OR_F_H: t = K_FH; n = sizeof(K_FH) / 8; break;
    case OR_RHO_PD: t = K_FPDR; n = sizeof(K_FPDR) / 8; break;
    case OR_RHO_H: t = K_FHR; n = sizeof(K_FHR) / 8; break;
    case OR_PHI_PD: t = K_PHPD; n = sizeof(K_PHPD) / 8; break;
    case OR_PHI_PDH: t = K_PHPDH; n = sizeof(K_PHPDH) / 8; break;
    case OR_PHI_H: t = K_PHH; n = sizeof(K_PHH) / 8; break;
    default: return -1;
  }
  for (int i = 0; i < n && i < cap; ++i) out[i] = t[i];
  return n;
}
extern "C" int or_quadrature(double* xi, double* w) {
  for (int k = 0; k < QK; ++k) {
    for (int c = 0; c < 3; ++c) { xi[6 * k + c] = QUAD.xi1[k].v[c]; xi[6 * k + 3 + c] = QUAD.xi2[k].v[c]; }
    w[k] = QUAD.w;
  }
  return QK;
}

extern "C" int or_set_tabulated(int n_rho, double d_rho, int n_r, double d_r, const double* const* tables) {
  if (n_rho < 5 || n_r < 5 || !(d_rho > 0) || !(d_r > 0) || !tables) return -1;
  for (int f = 0; f < OR_NFUNC; ++f) {
    const bool emb = f == OR_F_PD || f == OR_F_H;
    const int n = emb ? n_rho : n_r;
    g_tab[f].y.assign(tables[f], tables[f] + n);
    g_tab[f].h = emb ? d_rho : d_r;
    tab_slopes(g_tab[f]);
  }
  return 0;
}

extern "C" int or_set_quad_table(int K, const double* xi, const double* w) {
  if (K < 1 || K > MAXK || !xi || !w) return -1;
  g_qtab = QTable();
  for (int k = 0; k < K; ++k) {
    g_qtab.xi1.push_back({{xi[6 * k], xi[6 * k + 1], xi[6 * k + 2]}});
    g_qtab.xi2.push_back({{xi[6 * k + 3], xi[6 * k + 4], xi[6 * k + 5]}});
    g_qtab.w.push_back(w[k]);
  }
  return 0;
}

/* One table evaluated on its own (test access): value and slope at x. */
extern "C" int or_tab_eval(int n, double h, const double* y, int emb, double x, double* v, double* d) {
  if (n < 5 || !(h > 0) || !y) return -1;
  Tab t;
  t.y.assign(y, y + n);
  t.h = h;
  tab_slopes(t);
  if (v) *v = tab_eval(t, x, emb != 0, 0);
  if (d) *d = tab_eval(t, x, emb != 0, 1);
  return 0;
}

/* ===========================================================================
 * Densities: electrondensity.cpp:12-56 (Pd branch :20-33, H branch :35-48).
 * ========================================================================= */
namespace {
struct Lists {
  const int32_t *off[5], *idx[5];
  int n(int w, int i) const { return off[w][i + 1] - off[w][i]; }
  int at(int w, int i, int t) const { return idx[w][off[w][i] + t]; }
};
Lists lists_of(const or_nbr* nb) {
  Lists L;
  for (int w = 0; w < 5; ++w) { L.off[w] = nb->off[w].data(); L.idx[w] = nb->idx[w].data(); }
  return L;
}

double site_density(const Sys& S, const Lists& L, const double* x, int s) {
  const int rule = S.p->quad_rule;
  double rho = 0.0;
  const bool pd = s < S.npd;
  const int i = pd ? s : s - S.npd;
  const int wl0 = pd ? 0 : 2, wl1 = pd ? 1 : 3;
  const V3 qi = S.q(s);
  const double si = S.sig(s);
  Nodes n;
  for (int t = 0; t < L.n(wl0, i); ++t) { /* Pd neighbors */
    int j = L.at(wl0, i, t);
    V3 qj = image_of(S.p, qi, S.q(j));
    pair_nodes(rule, qi, si, qj, S.sig(j), n);
    for (int k = 0; k < n.K; ++k) rho += eam_val(OR_RHO_PD, vnorm(n.r[k])) * n.w[k];
  }
  for (int t = 0; t < L.n(wl1, i); ++t) { /* H neighbors, weighted by occupancy */
    int a = L.at(wl1, i, t), j = S.npd + a;
    V3 qj = image_of(S.p, qi, S.q(j));
    pair_nodes(rule, qi, si, qj, S.sig(j), n);
    for (int k = 0; k < n.K; ++k) rho += x[a] * eam_val(OR_RHO_H, vnorm(n.r[k])) * n.w[k];
  }
  return rho;
}
}  // namespace

extern "C" int or_density(const or_params* p, const or_nbr* nb, const double* X,
                          const double* x, double* rho) {
  if (!p || !nb || !X) return -1;
  g_potential = p->potential;
  Sys S{p, X, p->n_pd, p->n_h};
  Lists L = lists_of(nb);
  const int N = p->n_pd + p->n_h;
#pragma omp parallel for schedule(static) num_threads(nthreads(p))
  for (int s = 0; s < N; ++s) rho[s] = site_density(S, L, x, s);
  return 0;
}

/* ===========================================================================
 * Objective + gradient: minimizer.cpp:83-369.  Pd sites :162-224 (PdPd
 * :172-194, PdH :197-219, scaling :220-223), H sites :225-286 (HPd :234-256,
 * HH :259-281, scaling :282-285), f = sum_V/T - 3 kB sum_sigma (:353).
 * Each per-site term is recomputed inside the k loop as the reference does.
 * ========================================================================= */
extern "C" int or_eval(const or_params* p, const or_nbr* nb, const double* X,
                       const double* x0, double* fun, double* G) {
  if (!p || !nb || !X || !fun || !G) return -1;
  g_potential = p->potential;
  const int npd = p->n_pd, nh = p->n_h, N = npd + nh;
  const int md = p->deriv_mode;
  const double eps = p->eps, T = p->T, sq2 = sigma_factor(p->quad_rule);
  Sys S{p, X, npd, nh};
  Lists L = lists_of(nb);
  std::vector<double> rho(N), sv(N), ss(N);
  /* Phase I (:140-157) */
#pragma omp parallel for schedule(static) num_threads(nthreads(p))
  for (int s = 0; s < N; ++s) rho[s] = site_density(S, L, x0, s);
  const double* rPd = rho.data();
  const double* rH = rho.data() + npd;

  /* Phase II (:162-351) */
#pragma omp parallel for schedule(static) num_threads(nthreads(p))
  for (int s = 0; s < N; ++s) {
    const bool pd = s < npd;
    const int i = pd ? s : s - npd;
    const V3 qi = S.q(s);
    const double sigmai = S.sig(s);
    double sum_V = 0.0;
    V3 Fq = {{0, 0, 0}};
    double Fs = 0.0;
    Nodes n;
    auto acc = [&](double c, const V3& r, int k) { /* Fq += c*r; Fsigma += c*r . dr/dsigma_i */
      V3 cr = vscale(c, r);
      Fq = vadd(Fq, cr);
      Fs += vdot(cr, n.dsi[k]);
    };
    if (pd) {
      sum_V += eam_val(OR_F_PD, rPd[i]);
      for (int t = 0; t < L.n(0, i); ++t) { /* Pd-Pd pairs :172-194 */
        int jj = L.at(0, i, t);
        V3 qj = image_of(p, qi, S.q(jj));
        pair_nodes(p->quad_rule, qi, sigmai, qj, S.sig(jj), n);
        for (int k = 0; k < n.K; ++k) {
          const V3& r = n.r[k];
          const double QWp = n.w[k];
          acc(QWp * eam_der(OR_F_PD, rPd[i], md, eps) * eam_der(OR_RHO_PD, vnorm(r), md, eps) / vnorm(r), r, k);
          acc(QWp * eam_der(OR_F_PD, rPd[jj], md, eps) * eam_der(OR_RHO_PD, vnorm(r), md, eps) / vnorm(r), r, k);
          sum_V += 0.5 * QWp * eam_val(OR_PHI_PD, vnorm(r));
          acc(QWp * eam_der(OR_PHI_PD, vnorm(r), md, eps) / vnorm(r), r, k);
        }
      }
      for (int t = 0; t < L.n(1, i); ++t) { /* Pd-H pairs :197-219 */
        int jj = L.at(1, i, t);
        V3 qj = image_of(p, qi, S.q(npd + jj));
        pair_nodes(p->quad_rule, qi, sigmai, qj, S.sig(npd + jj), n);
        for (int k = 0; k < n.K; ++k) {
          const V3& r = n.r[k];
          const double QWp = n.w[k];
          acc(x0[jj] * QWp * eam_der(OR_F_PD, rPd[i], md, eps) * eam_der(OR_RHO_H, vnorm(r), md, eps) / vnorm(r), r, k);
          acc(x0[jj] * QWp * eam_der(OR_F_H, rH[jj], md, eps) * eam_der(OR_RHO_PD, vnorm(r), md, eps) / vnorm(r), r, k);
          sum_V += 0.5 * x0[jj] * QWp * eam_val(OR_PHI_PDH, vnorm(r));
          acc(x0[jj] * QWp * eam_der(OR_PHI_PDH, vnorm(r), md, eps) / vnorm(r), r, k);
        }
      }
      Fq = vscale(1.0 / T, Fq);
      Fs *= 1.0 / T * sq2;
      Fs -= 3.0 * KB / sigmai;
      ss[s] = std::log(sigmai);
    } else {
      sum_V += x0[i] * eam_val(OR_F_H, rH[i]);
      for (int t = 0; t < L.n(2, i); ++t) { /* H-Pd pairs :234-256 */
        int jj = L.at(2, i, t);
        V3 qj = image_of(p, qi, S.q(jj));
        pair_nodes(p->quad_rule, qi, sigmai, qj, S.sig(jj), n);
        for (int k = 0; k < n.K; ++k) {
          const V3& r = n.r[k];
          const double QWp = n.w[k];
          acc(QWp * eam_der(OR_F_H, rH[i], md, eps) * eam_der(OR_RHO_PD, vnorm(r), md, eps) / vnorm(r), r, k);
          acc(QWp * eam_der(OR_F_PD, rPd[jj], md, eps) * eam_der(OR_RHO_H, vnorm(r), md, eps) / vnorm(r), r, k);
          sum_V += 0.5 * x0[i] * QWp * eam_val(OR_PHI_PDH, vnorm(r));
          acc(QWp * eam_der(OR_PHI_PDH, vnorm(r), md, eps) / vnorm(r), r, k);
        }
      }
      for (int t = 0; t < L.n(3, i); ++t) { /* H-H pairs :259-281 */
        int jj = L.at(3, i, t);
        V3 qj = image_of(p, qi, S.q(npd + jj));
        pair_nodes(p->quad_rule, qi, sigmai, qj, S.sig(npd + jj), n);
        for (int k = 0; k < n.K; ++k) {
          const V3& r = n.r[k];
          const double QWp = n.w[k];
          acc(x0[jj] * QWp * eam_der(OR_F_H, rH[i], md, eps) * eam_der(OR_RHO_H, vnorm(r), md, eps) / vnorm(r), r, k);
          acc(x0[jj] * QWp * eam_der(OR_F_H, rH[jj], md, eps) * eam_der(OR_RHO_H, vnorm(r), md, eps) / vnorm(r), r, k);
          sum_V += 0.5 * x0[i] * x0[jj] * QWp * eam_val(OR_PHI_H, vnorm(r));
          acc(x0[jj] * QWp * eam_der(OR_PHI_H, vnorm(r), md, eps) / vnorm(r), r, k);
        }
      }
      Fq = vscale(1.0 / T * x0[i], Fq);
      Fs *= 1.0 / T * sq2 * x0[i];
      Fs -= 3.0 * KB / sigmai * x0[i];
      ss[s] = std::log(sigmai) * x0[i];
    }
    sv[s] = sum_V;
    G[4 * s] = Fq.v[0]; G[4 * s + 1] = Fq.v[1]; G[4 * s + 2] = Fq.v[2]; G[4 * s + 3] = Fs;
  }
  double sum_V = 0.0, sum_sigma = 0.0;
  for (int s = 0; s < N; ++s) { sum_V += sv[s]; sum_sigma += ss[s]; }
  *fun = 1.0 / T * sum_V - 3.0 * KB * sum_sigma;
  return 0;
}

/* ===========================================================================
 * Chemical potential: minimizer.cpp:588-802 (per H site :658-723, correction
 * terms :670-676,:694-696,:716-718, gamma update and cap :793-800).
 * ========================================================================= */
extern "C" int or_chem_potential(const or_params* p, const or_nbr* nb, const double* X,
                                 const double* x0, const double* coe_Pd, const double* coe_H,
                                 double gammabd, double* gamma, int32_t* full_H) {
  if (!p || !nb || !X || !gamma || !full_H) return -1;
  g_potential = p->potential;
  const int npd = p->n_pd, nh = p->n_h, N = npd + nh;
  const int md = p->deriv_mode;
  const double eps = p->eps, T = p->T;
  Sys S{p, X, npd, nh};
  Lists L = lists_of(nb);
  std::vector<double> rho(N), out(nh, 0.0);
#pragma omp parallel for schedule(static) num_threads(nthreads(p))
  for (int s = 0; s < N; ++s) rho[s] = site_density(S, L, x0, s);
  const double* rPd = rho.data();
  const double* rH = rho.data() + npd;
#pragma omp parallel for schedule(static) num_threads(nthreads(p))
  for (int i = 0; i < nh; ++i) {
    if (full_H[i] == 1) continue;
    const int s = npd + i;
    const V3 qi = S.q(s);
    const double sigmai = S.sig(s);
    Nodes n;
    double sum_Vx = 0.0;
    sum_Vx += eam_val(OR_F_H, rH[i]);
    const double nHPd = (double)L.n(2, i);
    double tmp = coe_H[1] * nHPd * nHPd;
    for (int j = 0; j < L.n(3, i); ++j) {
      tmp += 2.0 * coe_H[2] * nHPd * x0[L.at(3, i, j)];
      for (int l = 0; l < L.n(3, i); ++l) tmp += coe_H[3] * x0[L.at(3, i, j)] * x0[L.at(3, i, l)];
    }
    sum_Vx += 0.5 * T * coe_H[0] * tmp;
    for (int j = 0; j < L.n(2, i); ++j) { /* H-Pd :679-698 */
      int jj = L.at(2, i, j);
      V3 qj = image_of(p, qi, S.q(jj));
      pair_nodes(p->quad_rule, qi, sigmai, qj, S.sig(jj), n);
      for (int k = 0; k < n.K; ++k) {
        const V3& r = n.r[k];
        const double QWp = n.w[k];
        sum_Vx += QWp * eam_der(OR_F_PD, rPd[jj], md, eps) * eam_val(OR_RHO_H, vnorm(r));
        sum_Vx += QWp * eam_val(OR_PHI_PDH, vnorm(r));
      }
      sum_Vx += T * coe_Pd[0] * coe_Pd[2] * (double)L.n(0, jj);
      for (int l = 0; l < L.n(1, jj); ++l) sum_Vx += T * coe_Pd[0] * coe_Pd[3] * x0[L.at(1, jj, l)];
    }
    for (int j = 0; j < L.n(3, i); ++j) { /* H-H :701-719 */
      int jj = L.at(3, i, j);
      V3 qj = image_of(p, qi, S.q(npd + jj));
      pair_nodes(p->quad_rule, qi, sigmai, qj, S.sig(npd + jj), n);
      for (int k = 0; k < n.K; ++k) {
        const V3& r = n.r[k];
        const double QWp = n.w[k];
        sum_Vx += x0[jj] * QWp * eam_der(OR_F_H, rH[jj], md, eps) * eam_val(OR_RHO_H, vnorm(r));
        sum_Vx += x0[jj] * QWp * eam_val(OR_PHI_H, vnorm(r));
      }
      sum_Vx += x0[jj] * T * coe_H[0] * coe_H[2] * (double)L.n(2, jj);
      for (int l = 0; l < L.n(3, jj); ++l) sum_Vx += x0[jj] * T * coe_H[0] * coe_H[3] * x0[L.at(3, jj, l)];
    }
    if (x0[i] > 0.0 && x0[i] < 1.0) out[i] = std::log(x0[i] / (1.0 - x0[i])) + sum_Vx / (KB * T);
  }
  for (int i = 0; i < nh; ++i) {
    if (full_H[i] == 1) continue;
    gamma[i] = out[i] + 1.5;
    if (gamma[i] > gammabd) {
      gamma[i] = gammabd;
      if (x0[i] > 0.9999) full_H[i] = 1;
    }
  }
  return 0;
}

/* ===========================================================================
 * Diffusion step: integrate.cpp:9-87 (prefactor :26, flux :52-57, update
 * :57, clamp :76-81).  full_H sites keep their value.
 * ========================================================================= */
extern "C" int or_diffuse(const or_params* p, const or_nbr* nb, double dt, double v, double Qm,
                          double xLb, double xUb, const double* gamma, const int32_t* full_H,
                          const double* x0, double* x) {
  if (!p || !nb) return -1;
  const int nh = p->n_h;
  const double T = p->T;
  const double prefactor = v * std::exp(-Qm / (KB * T));
  Lists L = lists_of(nb);
  std::vector<double> nx(nh, 0.0);
#pragma omp parallel for schedule(static) num_threads(nthreads(p))
  for (int i = 0; i < nh; ++i) {
    if (full_H[i] == 1) continue;
    double x_deriv = 0.0;
    double fi = gamma[i] - std::log(x0[i] / (1.0 - x0[i]));
    for (int t = 0; t < L.n(4, i); ++t) {
      int j = L.at(4, i, t);
      double fj = gamma[j] - std::log(x0[j] / (1.0 - x0[j]));
      x_deriv += x0[j] * (1.0 - x0[i]) * std::exp(-(fi - fj) / 2.0) - x0[i] * (1.0 - x0[j]) * std::exp((fi - fj) / 2.0);
    }
    nx[i] = x0[i] + prefactor * x_deriv * dt;
  }
  for (int i = 0; i < nh; ++i) {
    if (full_H[i] == 1) continue;
    x[i] = nx[i];
    if (x[i] < xLb) x[i] = xLb;
    else if (x[i] > xUb) x[i] = xUb;
  }
  return 0;
}
