// A caller written against the reference's legacy interface (minimizer.h:82-98,
// neighbors.cpp:8, integrate.cpp:9), in the order its DMD step implies:
// findNeighbors -> Minimizer(argc, argv) -> initialize(input, lists, q, sigma,
// x0, QPp, QWp, coe) -> findLocalNeighbors(ext lists ...) ->
// minimizeFreeEntropy -> calculateChemicalPotential -> integrateDiffusion.
// It only renames the reference's types into a2c::.  Prints one JSON object.
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "a2c_lattice.h"
#include "a2c_minimizer.hpp"

using namespace a2c;
using std::vector;

// gaussianQuadrature(6, QP) of quadrature.cpp:43-75 as the caller builds it
static void reference_rule(vector<vector<Vec3D>>& QP) {
  const int dim = 6;
  for (int i = 0; i < dim; ++i) {
    double p[6] = {0, 0, 0, 0, 0, 0}, m[6] = {0, 0, 0, 0, 0, 0};
    p[i] = std::sqrt(dim / 2.0);
    m[i] = -std::sqrt(dim / 2.0);
    QP.push_back({Vec3D(p[0], p[1], p[2]), Vec3D(p[3], p[4], p[5])});
    QP.push_back({Vec3D(m[0], m[1], m[2]), Vec3D(m[3], m[4], m[5])});
  }
}

static void print_list(const char* name, const vector<vector<int>>& L) {
  std::printf("\"%s\": [", name);
  for (size_t r = 0; r < L.size(); ++r) {
    std::printf("%s[", r ? "," : "");
    for (size_t k = 0; k < L[r].size(); ++k) std::printf("%s%d", k ? "," : "", L[r][k]);
    std::printf("]");
  }
  std::printf("], ");
}

int main(int argc, char** argv) {
  const int cells = argc > 1 ? std::atoi(argv[1]) : 3;
  a2c_lattice_desc ld;
  a2c_lattice_desc_default(&ld);
  ld.nx = ld.ny = ld.nz = cells;
  ld.seed = 21;
  ld.disp_std = 0.02;
  ld.sigma_lo = ld.sigma_hi = 0.10;
  ld.x_lo = 0.05;
  ld.x_hi = 0.5;
  const int n = 4 * cells * cells * cells;
  vector<double> X(8 * (size_t)n), xs(n), box(3);
  if (a2c_lattice_fcc_pdh(&ld, X.data(), xs.data(), box.data())) return 1;

  Input input;
  input.file.rc = 5.35;
  input.file.a_Pd = 3.89;
  input.file.T0 = 400.0;
  input.file.eps = 1e-5;
  input.file.v = 1e13;
  input.file.Qm = 0.23;
  input.file.xLb = 1e-3;
  input.file.xUb = 1.0 - 1e-16;
  for (int c = 0; c < 3; ++c) { input.file.box[c] = box[c]; input.file.periodic[c] = 1; }

  vector<Vec3D> q_Pd0, q_H0, q_Pd, q_H;
  vector<double> sigma_Pd0, sigma_H0, sigma_Pd, sigma_H, x0(xs);
  for (int s = 0; s < 2 * n; ++s) {
    const double* v = &X[4 * (size_t)s];
    (s < n ? q_Pd0 : q_H0).push_back(Vec3D(v[0], v[1], v[2]));
    (s < n ? sigma_Pd0 : sigma_H0).push_back(v[3]);
  }

  vector<vector<int>> PdPd_ext, PdH_ext, HPd_ext, HH_ext, PdPd, PdH, HPd, HH, PdPdnum, HH1;
  const int maxNeib0 = findNeighbors(input.file.a_Pd, input.file.rc, q_Pd0, q_H0, PdPd_ext, PdH_ext, HPd_ext, HH_ext,
                                     &input);
  vector<vector<Vec3D>> QPp;
  reference_rule(QPp);
  double QWp = 0.5 / 6;  // quadrature.cpp:39
  vector<double> coe_Pd(4, 0.0), coe_H(4, 0.0);

  Minimizer minimizer(&argc, &argv);
  int rc = minimizer.initialize(&input, &PdPd_ext, &PdH_ext, &HPd_ext, &HH_ext, q_Pd0, q_H0, sigma_Pd0, sigma_H0, x0,
                                &QPp, QWp, &coe_Pd, &coe_H);
  if (rc) { std::printf("{\"error\": \"initialize %d\"}\n", rc); return 2; }
  int maxNeib = 0;
  rc = minimizer.findLocalNeighbors(PdPd_ext, PdH_ext, HPd_ext, HH_ext, PdPd, PdH, HPd, HH, PdPdnum, HH1, maxNeib);
  if (rc) { std::printf("{\"error\": \"findLocalNeighbors %d\"}\n", rc); return 2; }
  double free_entropy = 0.0;
  rc = minimizer.minimizeFreeEntropy(1e-13, q_Pd0, q_H0, sigma_Pd0, sigma_H0, x0, q_Pd, q_H, sigma_Pd, sigma_H,
                                     free_entropy);
  if (rc) { std::printf("{\"error\": \"minimizeFreeEntropy %d\"}\n", rc); return 2; }
  vector<double> gamma(q_H0.size(), 0.0);
  vector<int> full_H(q_H0.size(), 0);
  rc = minimizer.calculateChemicalPotential(x0, 1e300, gamma, full_H);
  if (rc) { std::printf("{\"error\": \"calculateChemicalPotential %d\"}\n", rc); return 2; }
  vector<double> x;
  rc = integrateDiffusion(minimizer, input, 1e-13, gamma, x0, x, full_H);
  if (rc) { std::printf("{\"error\": \"integrateDiffusion %d\"}\n", rc); return 2; }
  // a rule other than the reference's is refused
  vector<vector<Vec3D>> bad(QPp);
  bad[0][0][0] *= 2.0;
  Minimizer other;
  const int rc_bad = other.initialize(&input, &PdPd_ext, &PdH_ext, &HPd_ext, &HH_ext, q_Pd0, q_H0, sigma_Pd0,
                                      sigma_H0, x0, &bad, QWp, &coe_Pd, &coe_H);

  std::printf("{\"maxNeib0\": %d, \"maxNeib\": %d, \"free_entropy\": %.17g, \"rc_bad_rule\": %d, ", maxNeib0, maxNeib,
              free_entropy, rc_bad);
  print_list("PdPd", PdPd);
  print_list("HH1", HH1);
  print_list("PdPdnum", PdPdnum);
  std::printf("\"X\": [");
  for (int s = 0; s < 2 * n; ++s) {
    const Vec3D& q = s < n ? q_Pd[s] : q_H[s - n];
    const double sg = s < n ? sigma_Pd[s] : sigma_H[s - n];
    std::printf("%s%.17g,%.17g,%.17g,%.17g", s ? "," : "", q[0], q[1], q[2], sg);
  }
  std::printf("], \"gamma\": [");
  for (size_t a = 0; a < gamma.size(); ++a) std::printf("%s%.17g", a ? "," : "", gamma[a]);
  std::printf("], \"x\": [");
  for (size_t a = 0; a < x.size(); ++a) std::printf("%s%.17g", a ? "," : "", x[a]);
  std::printf("]}\n");
  return 0;
}
