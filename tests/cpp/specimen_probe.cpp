// Test driver for a2c_amd/host/specimen.hpp and output.hpp:
//   specimen_probe <deck> <out.json>
// builds the LatticeVariables from the deck, packs them into the engine's
// layout, writes snapshot 0 (forced, as Main.cpp does before its loop), sets
// gamma = -x at H sites and x = 0.5 on the first H site, unpacks, writes
// snapshot 1 at t = dt, and dumps the packed arrays as JSON.
#include <cstdio>
#include <fstream>

#include "output.hpp"
#include "specimen.hpp"

int main(int argc, char** argv) {
  if (argc < 3) return 2;
  try {
    a2c::Deck d = a2c::Deck::parse_file(argv[1]);
    a2c::SpecimenInput in = a2c::read_specimen(d);
    std::vector<a2c::LatticeVariables> LVS = a2c::setup_lattice_variables(in);
    a2c::PackedSpecimen P = a2c::pack_pd_h(in, LVS);
    a2c::SolutionOutput out(d, in);
    const double dt = d.num("Time.TimeStep", 1e-13);
    out.write(0.0, dt, 0, LVS, true);
    std::vector<double> gamma(P.n_h), x(P.x);
    for (int a = 0; a < P.n_h; ++a) gamma[a] = -x[a];
    if (P.n_h) x[0] = 0.5;
    a2c::unpack_pd_h(P, P.X.data(), x.data(), gamma.data(), LVS);
    out.write(dt, dt, 1, LVS, false);
    std::ofstream o(argv[2]);
    o.precision(17);
    o << "{\"n_pd\": " << P.n_pd << ", \"n_h\": " << P.n_h << ", \"a\": " << P.a << ", \"frames\": " << out.frame();
    o << ", \"box\": [" << P.box[0] << ", " << P.box[1] << ", " << P.box[2] << "], \"periodic\": [" << P.periodic[0]
      << ", " << P.periodic[1] << ", " << P.periodic[2] << "], \"sizes\": [";
    for (size_t l = 0; l < LVS.size(); ++l) o << (l ? ", " : "") << LVS[l].size;
    o << "], \"X\": [";
    for (size_t k = 0; k < P.X.size(); ++k) o << (k ? ", " : "") << P.X[k];
    o << "], \"x\": [";
    for (size_t k = 0; k < P.x.size(); ++k) o << (k ? ", " : "") << P.x[k];
    o << "], \"tag_h\": [";
    for (size_t k = 0; k < P.tag_h.size(); ++k) o << (k ? ", " : "") << P.tag_h[k];
    o << "]}\n";
  } catch (const std::exception& e) {
    std::printf("ERROR %s\n", e.what());
    return 1;
  }
  return 0;
}
