// a2c_dmd -- diffusive-MD host program over liba2c_fe.so (C++).
//
//   a2c_dmd <input.st>
//
// Reads an input deck in the reference's language (deck.hpp) and runs the
// DMD time loop: each step rebuilds the neighbor lists, enforces the
// subsurface boundary condition, relaxes the phase packets (device BLMVM),
// computes the chemical potential and advances the H occupancy
// (include/a2c_dmd.h).  Restart files use the format readresults.cpp reads.
//
// The specimen comes from one of two places:
//  - the reference's objects (specimen.hpp): Species, Material, Lattice,
//    RegionalInitialCondition, as SpaceInitializer builds them; H sites
//    with Tag = 1 are the subsurface reservoir (LatticeVariables.h:33);
//    snapshots in the reference's VTP / PVD / XYZ formats (output.hpp),
//    frame 0 forced before the loop as Main.cpp:118 does;
//  - or, with no Lattice object, a generated FCC Pd + octahedral-H block:
//    Specimen.Cells (or Cells_x/_y/_z), .LatticeConstant, .Seed,
//    .Displacement, .Sigma, .HFraction, .SlabCells, .SlabHFraction,
//    .Periodic (True/False), reservoir = H sites in the first
//    Diffusion.SubsurfaceLayers layers of cells along x; XYZ snapshots.
//
// Keys (reference names where the reference has them, IoData.cpp):
//   InteratomicPotential.File (a DYNAMO setfl file, elements Pd and H),
//   .Tolerance, .CutOffDistance                               (:523-534)
//   Diffusion.Temperature, .AttemptFrequency, .ActivationEnergy,
//   .SoluteChemicalPotential (the reservoir potential)        (:639-651)
//   Time.TimeStep, Time.MaxIts                                 (:696-699)
//   Output.Prefix, .Solution, .Frequency, .TimeInterval, .Format,
//   .Precision                                                 (:719-746)
// and, where the reference has no key:
//   Restart.File (start from a restart file instead)
//   Diffusion.GammaBoundary (same as SoluteChemicalPotential),
//   Diffusion.xLb, .xUb (occupancy clamp)
//   Solver.Precision (Mixed/FP64), Solver.Derivative (Analytic/FD),
//   Solver.GradientTolerance, Solver.MaxIts, Solver.SigmaUpperBound
#include <cstdio>
#include <memory>
#include <string>
#include <vector>

#include "a2c_dmd.h"
#include "a2c_fe.h"
#include "a2c_lattice.h"
#include "deck.hpp"
#include "output.hpp"
#include "specimen.hpp"

namespace {

// The specimen in the engine's layout, with where it came from.
struct Run {
  int npd = 0, nh = 0;
  std::vector<double> X, x;
  double box[3] = {0, 0, 0}, a = 3.89;
  int32_t periodic[3] = {0, 0, 0};
  std::vector<int32_t> sub;  // reservoir H sites
  // set when the specimen came from the reference's objects
  std::unique_ptr<a2c::SpecimenInput> in;
  std::vector<a2c::LatticeVariables> LVS;
  std::unique_ptr<a2c::PackedSpecimen> packed;
};

bool has_lattice_objects(const a2c::Deck& deck) {
  const auto& v = deck.values();
  auto it = v.lower_bound("Lattice[");
  return it != v.end() && it->first.compare(0, 8, "Lattice[") == 0;
}

void from_objects(const a2c::Deck& deck, double xLb, double xUb, Run& r) {
  r.in.reset(new a2c::SpecimenInput(a2c::read_specimen(deck)));
  r.LVS = a2c::setup_lattice_variables(*r.in);
  r.packed.reset(new a2c::PackedSpecimen(a2c::pack_pd_h(*r.in, r.LVS)));
  const a2c::PackedSpecimen& P = *r.packed;
  r.npd = P.n_pd;
  r.nh = P.n_h;
  r.X = P.X;
  r.x = P.x;
  for (double& v : r.x) v = v < xLb ? xLb : v > xUb ? xUb : v;
  for (int c = 0; c < 3; ++c) { r.box[c] = P.box[c]; r.periodic[c] = P.periodic[c]; }
  if (P.a > 0) r.a = P.a;
  for (int a = 0; a < r.nh; ++a)
    if (P.tag_h[a] == 1) r.sub.push_back(a);
}

bool generated(const a2c::Deck& deck, double xLb, double xUb, Run& r) {
  a2c_lattice_desc L;
  a2c_lattice_desc_default(&L);
  const long long cells = deck.integer("Specimen.Cells", 6);
  L.nx = (int)deck.integer("Specimen.Cells_x", cells);
  L.ny = (int)deck.integer("Specimen.Cells_y", cells);
  L.nz = (int)deck.integer("Specimen.Cells_z", cells);
  L.a = deck.num("Specimen.LatticeConstant", 3.89);
  L.seed = (uint64_t)deck.integer("Specimen.Seed", 1);
  L.disp_std = deck.num("Specimen.Displacement", 0.02);
  L.sigma_lo = L.sigma_hi = deck.num("Specimen.Sigma", 0.10);
  L.x_lo = L.x_hi = deck.num("Specimen.HFraction", 0.10);
  L.slab_cells = (int)deck.integer("Specimen.SlabCells", 0);
  L.slab_x = deck.num("Specimen.SlabHFraction", 0.6);
  L.x_clip_lo = xLb;
  L.x_clip_hi = xUb;
  r.npd = r.nh = 4 * L.nx * L.ny * L.nz;
  r.X.resize(4 * (size_t)(r.npd + r.nh));
  r.x.resize(r.nh);
  r.a = L.a;
  if (a2c_lattice_fcc_pdh(&L, r.X.data(), r.x.data(), r.box)) return false;
  const bool per = deck.str("Specimen.Periodic", "True") != "False";
  for (int c = 0; c < 3; ++c) r.periodic[c] = per;
  const int layers = (int)deck.integer("Diffusion.SubsurfaceLayers", 0);
  for (int a = 0; a < r.nh; ++a)
    if ((a / 4) % L.nx < layers) r.sub.push_back(a);
  return true;
}

}  // namespace

int main(int argc, char** argv) {
  if (argc < 2) {
    std::fprintf(stderr, "usage: %s <input file>\n", argv[0]);
    return 2;
  }
  a2c::Deck deck;
  Run run;
  const double xLb_default = 1e-3, xUb_default = 1.0 - 1e-16;
  double xLb = xLb_default, xUb = xUb_default;
  std::unique_ptr<a2c::SolutionOutput> snapshots;
  try {
    deck = a2c::Deck::parse_file(argv[1]);
    xLb = deck.num("Diffusion.xLb", xLb_default);
    xUb = deck.num("Diffusion.xUb", xUb_default);
    if (has_lattice_objects(deck)) {
      from_objects(deck, xLb, xUb, run);
      snapshots.reset(new a2c::SolutionOutput(deck, *run.in));
    } else if (!generated(deck, xLb, xUb, run)) {
      std::fprintf(stderr, "ERROR: bad specimen\n");
      return 1;
    }
  } catch (const std::exception& e) {
    std::fprintf(stderr, "ERROR: %s\n", e.what());
    return 1;
  }
  const int npd = run.npd, nh = run.nh, N = npd + nh;
  std::vector<double>& X = run.X;
  std::vector<double>& x = run.x;
  std::vector<double> gamma(nh, 0.0);
  std::vector<int32_t> full(nh, 0);
  if (N == 0) {
    std::fprintf(stderr, "ERROR: the specimen has no sites\n");
    return 1;
  }
  if (deck.has("Restart.File") &&
      a2c_read_results(deck.str("Restart.File", "").c_str(), npd, nh, xLb, xUb, X.data(), x.data(), gamma.data(),
                       full.data())) {
    std::fprintf(stderr, "ERROR: cannot read the restart file\n");
    return 1;
  }
  const std::vector<double> X0 = X;  // reference positions for the displacement column
  const std::vector<int32_t>& sub = run.sub;
  // engine
  a2c_fe_desc d;
  a2c_fe_desc_default(&d);
  d.n_pd = npd;
  d.n_h = nh;
  for (int c = 0; c < 3; ++c) { d.box[c] = run.box[c]; d.periodic[c] = run.periodic[c]; }
  d.rc = deck.num("InteratomicPotential.CutOffDistance", 5.35);
  d.fd_eps = deck.num("InteratomicPotential.Tolerance", 1e-5);
  d.a_pd = run.a;
  d.temperature = deck.num("Diffusion.Temperature", 300.0);
  d.precision = deck.str("Solver.Precision", "Mixed") == "FP64" ? A2C_PREC_FP64 : A2C_PREC_MIXED;
  d.deriv_mode = deck.str("Solver.Derivative", "Analytic") == "FD" ? A2C_DERIV_FD : A2C_DERIV_ANALYTIC;
  a2c_fe_ctx* ctx = nullptr;
  if (int rc = a2c_fe_create(&d, &ctx)) {
    std::fprintf(stderr, "ERROR: cannot create the engine (%d)\n", rc);
    return 1;
  }
  // InteratomicPotential.File: a tabulated potential (setfl) for the Pd and H sites
  const std::string pot_file = deck.str("InteratomicPotential.File", "");
  if (!pot_file.empty() && a2c_fe_set_potential_file(ctx, pot_file.c_str(), "Pd", "H")) {
    std::fprintf(stderr, "ERROR: %s\n", a2c_fe_last_error(ctx));
    a2c_fe_destroy(ctx);
    return 1;
  }
  a2c_dmd_params p;
  a2c_dmd_params_default(&p);
  p.dt = deck.num("Time.TimeStep", 1e-13);
  p.nu = deck.num("Diffusion.AttemptFrequency", 1e13);
  p.Qm = deck.num("Diffusion.ActivationEnergy", 0.23);
  p.xLb = xLb;
  p.xUb = xUb;
  p.gamma_bd = deck.num("Diffusion.SoluteChemicalPotential", deck.num("Diffusion.GammaBoundary", 1e300));
  p.lmvm.gatol = deck.num("Solver.GradientTolerance", 1e-8);
  p.lmvm.grtol = 0.0;
  p.lmvm.max_it = (int)deck.integer("Solver.MaxIts", 2000);
  p.lmvm.sigma_ub = deck.num("Solver.SigmaUpperBound", p.lmvm.sigma_ub);
  const int steps = (int)deck.integer("Time.MaxIts", 10);
  const std::string prefix = deck.str("Output.Prefix", "");
  const std::string base = deck.str("Output.Solution", "solution");
  const int freq = (int)deck.integer("Output.Frequency", 0);
  std::printf("a2c_dmd: %d Pd + %d H sites (%zu reservoir), T = %g K, dt = %g s, %d steps\n", npd, nh, sub.size(),
              d.temperature, p.dt, steps);
  int status = 0;
  auto snapshot = [&](double t, int st, bool force) {
    if (!snapshots) return;
    a2c::unpack_pd_h(*run.packed, X.data(), x.data(), gamma.data(), run.LVS);
    try {
      snapshots->write(t, p.dt, st, run.LVS, force);
    } catch (const std::exception& e) {
      std::fprintf(stderr, "WARNING: %s\n", e.what());
    }
  };
  snapshot(0.0, 0, true);
  for (int st = 1; st <= steps; ++st) {
    a2c_dmd_report r;
    if (int rc = a2c_dmd_step(ctx, &p, (int32_t)sub.size(), sub.data(), (int32_t)sub.size(), sub.data(), X.data(),
                              x.data(), gamma.data(), full.data(), &r)) {
      std::fprintf(stderr, "ERROR: step %d failed (%d): %s\n", st, rc, a2c_fe_last_error(ctx));
      status = 1;
      break;
    }
    std::printf("step %5d  f = %.12e  lmvm %d it (reason %d)  sum x = %.15e  dH = %+.3e  xH = %.6f\n", st, r.f,
                r.lmvm_iterations, r.lmvm_reason, r.total_h_after, r.total_h_after - r.total_h_before, r.xH);
    snapshot(st * p.dt, st, false);
    if (freq > 0 && st % freq == 0) {
      const std::string path = prefix + base + "." + std::to_string(st) + ".dat";
      if (a2c_write_results(path.c_str(), npd, nh, X.data(), x.data(), gamma.data(), full.data(), st * p.dt, st))
        std::fprintf(stderr, "WARNING: cannot write %s\n", path.c_str());
      if (snapshots) continue;
      char frame[32];
      std::snprintf(frame, sizeof frame, "_L0_%04d.xyz", st / freq);
      const std::string xyz = prefix + base + frame;
      if (a2c_write_xyz(xyz.c_str(), npd, nh, X.data(), X0.data(), x.data(), gamma.data(), run.box, run.periodic,
                        st * p.dt))
        std::fprintf(stderr, "WARNING: cannot write %s\n", xyz.c_str());
    }
  }
  a2c_fe_destroy(ctx);
  return status;
}
