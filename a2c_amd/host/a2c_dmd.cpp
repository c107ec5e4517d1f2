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
// Keys (reference names where the reference has them, IoData.cpp):
//   InteratomicPotential.Tolerance, .CutOffDistance          (:523-534)
//   Diffusion.Temperature, .AttemptFrequency, .ActivationEnergy (:628-646)
//   Time.TimeStep, Time.MaxIts                                 (:696-699)
//   Output.Prefix, Output.Solution, Output.Frequency           (:728-734)
// and, for what the reference's specimen setup would provide:
//   Specimen.Cells (or Cells_x/_y/_z), .LatticeConstant, .Seed, .Displacement,
//   .Sigma, .HFraction, .SlabCells, .SlabHFraction, .Periodic (True/False)
//   Restart.File (start from a restart file instead of a generated specimen)
//   Diffusion.GammaBoundary, Diffusion.SubsurfaceLayers (reservoir sites:
//   H sites in the first layers of cells along x), Diffusion.xLb, .xUb
//   Solver.Precision (Mixed/FP64), Solver.Derivative (Analytic/FD),
//   Solver.GradientTolerance, Solver.MaxIts, Solver.SigmaUpperBound
#include <cstdio>
#include <string>
#include <vector>

#include "a2c_dmd.h"
#include "a2c_fe.h"
#include "a2c_lattice.h"
#include "deck.hpp"

int main(int argc, char** argv) {
  if (argc < 2) {
    std::fprintf(stderr, "usage: %s <input file>\n", argv[0]);
    return 2;
  }
  a2c::Deck deck;
  try {
    deck = a2c::Deck::parse_file(argv[1]);
  } catch (const std::exception& e) {
    std::fprintf(stderr, "ERROR: %s\n", e.what());
    return 1;
  }
  // specimen
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
  const double xLb = deck.num("Diffusion.xLb", 1e-3), xUb = deck.num("Diffusion.xUb", 1.0 - 1e-16);
  L.x_clip_lo = xLb;
  L.x_clip_hi = xUb;
  const int npd = 4 * L.nx * L.ny * L.nz, nh = npd, N = npd + nh;
  std::vector<double> X(4 * (size_t)N), x(nh), gamma(nh, 0.0), box(3);
  std::vector<int32_t> full(nh, 0);
  if (a2c_lattice_fcc_pdh(&L, X.data(), x.data(), box.data())) {
    std::fprintf(stderr, "ERROR: bad specimen\n");
    return 1;
  }
  if (deck.has("Restart.File") &&
      a2c_read_results(deck.str("Restart.File", "").c_str(), npd, nh, xLb, xUb, X.data(), x.data(), gamma.data(),
                       full.data())) {
    std::fprintf(stderr, "ERROR: cannot read the restart file\n");
    return 1;
  }
  const std::vector<double> X0 = X;  // reference positions for the displacement column
  // reservoir (subsurface) sites: H sites of the first layers of cells along x
  std::vector<int32_t> sub;
  const int layers = (int)deck.integer("Diffusion.SubsurfaceLayers", 0);
  for (int a = 0; a < nh; ++a)
    if ((a / 4) % L.nx < layers) sub.push_back(a);
  // engine
  a2c_fe_desc d;
  a2c_fe_desc_default(&d);
  d.n_pd = npd;
  d.n_h = nh;
  const bool per = deck.str("Specimen.Periodic", "True") != "False";
  for (int c = 0; c < 3; ++c) { d.box[c] = box[c]; d.periodic[c] = per; }
  d.rc = deck.num("InteratomicPotential.CutOffDistance", 5.35);
  d.fd_eps = deck.num("InteratomicPotential.Tolerance", 1e-5);
  d.a_pd = L.a;
  d.temperature = deck.num("Diffusion.Temperature", 300.0);
  d.precision = deck.str("Solver.Precision", "Mixed") == "FP64" ? A2C_PREC_FP64 : A2C_PREC_MIXED;
  d.deriv_mode = deck.str("Solver.Derivative", "Analytic") == "FD" ? A2C_DERIV_FD : A2C_DERIV_ANALYTIC;
  a2c_fe_ctx* ctx = nullptr;
  if (int rc = a2c_fe_create(&d, &ctx)) {
    std::fprintf(stderr, "ERROR: cannot create the engine (%d)\n", rc);
    return 1;
  }
  a2c_dmd_params p;
  a2c_dmd_params_default(&p);
  p.dt = deck.num("Time.TimeStep", 1e-13);
  p.nu = deck.num("Diffusion.AttemptFrequency", 1e13);
  p.Qm = deck.num("Diffusion.ActivationEnergy", 0.23);
  p.xLb = xLb;
  p.xUb = xUb;
  p.gamma_bd = deck.num("Diffusion.GammaBoundary", 1e300);
  p.lmvm.gatol = deck.num("Solver.GradientTolerance", 1e-8);
  p.lmvm.grtol = 0.0;
  p.lmvm.max_it = (int)deck.integer("Solver.MaxIts", 2000);
  p.lmvm.sigma_ub = deck.num("Solver.SigmaUpperBound", p.lmvm.sigma_ub);
  const int steps = (int)deck.integer("Time.MaxIts", 10);
  const std::string prefix = deck.str("Output.Prefix", "");
  const std::string base = deck.str("Output.Solution", "solution");
  const int freq = (int)deck.integer("Output.Frequency", 0);
  std::printf("a2c_dmd: %d Pd + %d H sites, T = %g K, dt = %g s, %d steps\n", npd, nh, d.temperature, p.dt, steps);
  int status = 0;
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
    if (freq > 0 && st % freq == 0) {
      const std::string path = prefix + base + "." + std::to_string(st) + ".dat";
      if (a2c_write_results(path.c_str(), npd, nh, X.data(), x.data(), gamma.data(), full.data(), st * p.dt, st))
        std::fprintf(stderr, "WARNING: cannot write %s\n", path.c_str());
      char frame[32];
      std::snprintf(frame, sizeof frame, "_L0_%04d.xyz", st / freq);
      const std::string xyz = prefix + base + frame;
      const int32_t perio[3] = {per, per, per};
      if (a2c_write_xyz(xyz.c_str(), npd, nh, X.data(), X0.data(), x.data(), gamma.data(), box.data(), perio,
                        st * p.dt))
        std::fprintf(stderr, "WARNING: cannot write %s\n", xyz.c_str());
    }
  }
  a2c_fe_destroy(ctx);
  return status;
}
