// output.hpp -- solution snapshots of the LatticeVariables in the reference's
// on-disk formats (Output.cpp:8-96, :104-318, :322-494; Utils.cpp:291-311):
//   <prefix><base>_L<l>_<frame>.vtp   VTK PolyData, ASCII: points q; point
//       data SiteID, MaterialID, Tag (Int8), AtomicFrequency, MolarFraction
//       and ChemicalPotential (nSpecies_max components, padded with 0)
//   <prefix><base>_L<l>.pvd           collection, one DataSet per snapshot
//   <prefix><base>_L<l>_<frame>.xyz   extended XYZ: material name, q,
//       q - q0, site id, sigma, tag, x and gamma per species
// Frames count from 0 with at least four digits.  Output.Format (VTP, XYZ,
// VTP_and_XYZ), Output.Precision (Low %7.4e, Medium %10.6e, High %16.10e),
// Output.Frequency / Output.TimeInterval, Output.Prefix, Output.Solution.
#pragma once
#include <cmath>
#include <cstdio>
#include <stdexcept>
#include <string>
#include <vector>

#include "deck.hpp"
#include "specimen.hpp"

namespace a2c {

class SolutionOutput {
 public:
  SolutionOutput(const Deck& d, const SpecimenInput& in) {
    const std::string f = d.str("Output.Format", "VTP");
    if (f == "VTP") format_ = VTP;
    else if (f == "XYZ") format_ = XYZ;
    else if (f == "VTP_and_XYZ" || f == "XYZ_and_VTP") format_ = BOTH;
    else throw std::runtime_error("unknown value '" + f + "' for Output.Format");
    const std::string p = d.str("Output.Precision", "Medium");
    precision_ = p == "Low" ? 0 : p == "Medium" ? 1 : p == "High" ? 2 : -1;
    if (precision_ < 0) throw std::runtime_error("unknown value '" + p + "' for Output.Precision");
    prefix_ = d.str("Output.Prefix", "");
    base_ = d.str("Output.Solution", "solution");
    frequency_ = (int)d.integer("Output.Frequency", 0);
    frequency_dt_ = d.num("Output.TimeInterval", -1.0);
    for (size_t m = 0; m < in.materials.size(); ++m)
      material_.push_back(in.materials[m].name.empty() ? "Mat" + std::to_string(m) : in.materials[m].name);
    lattices_ = in.lattices;
    if (format_ != XYZ)
      for (size_t l = 0; l < lattices_.size(); ++l) {
        FILE* f2 = std::fopen(pvd_name((int)l).c_str(), "w");
        if (!f2) throw std::runtime_error("Cannot open file '" + pvd_name((int)l) + "' for output.");
        std::fputs("<?xml version=\"1.0\"?>\n<VTKFile type=\"Collection\" version=\"0.1\"\n"
                   "byte_order=\"LittleEndian\">\n  <Collection>\n",
                   f2);
        std::fputs(kPvdTail, f2);
        std::fclose(f2);
      }
  }

  // Output::OutputSolution: writes when the step or time interval says so
  // (isTimeToWrite), or once more at this time when forced.
  bool write(double time, double dt, int step, const std::vector<LatticeVariables>& LVS, bool force) {
    bool due;
    if (force) due = time > last_ + 0.1 * dt;
    else if (frequency_dt_ > 0) due = std::floor(time / frequency_dt_) != std::floor((time - dt) / frequency_dt_);
    else due = frequency_ > 0 && step % frequency_ == 0;
    if (!due) return false;
    for (size_t l = 0; l < LVS.size(); ++l) {
      if (format_ != XYZ) write_vtp(time, (int)l, LVS[l]);
      if (format_ != VTP) write_xyz(time, (int)l, LVS[l]);
    }
    ++frame_;
    last_ = time;
    return true;
  }
  int frame() const { return frame_; }

 private:
  enum { VTP = 0, XYZ = 1, BOTH = 2 } format_ = VTP;
  int precision_ = 1, frequency_ = 0, frame_ = 0;
  double frequency_dt_ = -1.0, last_ = -1.0;
  std::string prefix_, base_;
  std::vector<std::string> material_;
  std::vector<LatticeInfo> lattices_;
  // the closing lines every snapshot rewrites: 27 bytes
  static constexpr const char* kPvdTail = "  </Collection>\n</VTKFile>\n";

  std::string pvd_name(int l) const { return prefix_ + base_ + "_L" + std::to_string(l) + ".pvd"; }
  std::string frame_name(int l, const char* ext) const {
    char b[64];
    std::snprintf(b, sizeof b, "_L%d_%04d.%s", l, frame_, ext);
    return base_ + b;
  }
  const char* num() const { return precision_ == 2 ? "%16.10e" : precision_ == 1 ? "%10.6e" : "%7.4e"; }

  // value then a space; species arrays padded to ns with "0 "
  void per_species(FILE* f, const std::vector<double>& v, int ns) const {
    int i = 0;
    for (; i < (int)v.size(); ++i) {
      std::fprintf(f, num(), v[i]);
      std::fputc(' ', f);
    }
    for (; i < ns; ++i) std::fputs("0 ", f);
  }

  void write_vtp(double time, int l, const LatticeVariables& LV) {
    const std::string name = frame_name(l, "vtp"), full = prefix_ + name;
    FILE* f = std::fopen(full.c_str(), "w");
    if (!f) throw std::runtime_error("Cannot open file '" + full + "' for output.");
    std::fprintf(f, "<?xml version=\"1.0\"?>\n<VTKFile type=\"PolyData\" version=\"0.1\"\n"
                    "byte_order=\"LittleEndian\">\n  <PolyData>\n    <Piece NumberOfPoints=\"%zu\">\n",
                 LV.q.size());
    std::fputs("      <Points>\n        <DataArray type=\"Float32\" NumberOfComponents=\"3\" format=\"ascii\">\n", f);
    for (const V3& q : LV.q) {
      std::fputs("          ", f);
      for (int c = 0; c < 3; ++c) {
        if (c) std::fputc(' ', f);
        std::fprintf(f, num(), q[c]);
      }
      std::fputc('\n', f);
    }
    std::fputs("        </DataArray>\n      </Points>\n", f);
    std::fputs("      <PointData Scalars=\"ScalarResults\" Vectors=\"VectorResults\">\n", f);
    const std::vector<int>* ints[3] = {&LV.siteid, &LV.matid, &LV.tag};
    const char* int_names[3] = {"SiteID", "MaterialID", "Tag"};
    for (int a = 0; a < 3; ++a) {
      std::fprintf(f, "        <DataArray type=\"Int8\" NumberOfComponents=\"1\" Name=\"%s\" format=\"ascii\">\n"
                      "          ",
                   int_names[a]);
      for (int v : *ints[a]) std::fprintf(f, "%d ", v);
      std::fputs("\n        </DataArray>\n", f);
    }
    std::fputs("        <DataArray type=\"Float32\" NumberOfComponents=\"1\" Name=\"AtomicFrequency\" "
               "format=\"ascii\">\n          ",
               f);
    for (double s : LV.sigma) {
      std::fprintf(f, num(), s);
      std::fputc(' ', f);
    }
    std::fputs("\n        </DataArray>\n", f);
    const int ns = LV.nSpecies_max;
    const std::vector<std::vector<double>>* arrs[2] = {&LV.x, &LV.gamma};
    const char* arr_names[2] = {"MolarFraction", "ChemicalPotential"};
    for (int a = 0; a < 2; ++a) {
      std::fprintf(f, "        <DataArray type=\"Float32\" NumberOfComponents=\"%d\" Name=\"%s\" format=\"ascii\">\n"
                      "          ",
                   ns, arr_names[a]);
      for (const auto& v : *arrs[a]) per_species(f, v, ns);
      std::fputs("\n        </DataArray>\n", f);
    }
    std::fputs("      </PointData>\n    </Piece>\n  </PolyData>\n</VTKFile>\n", f);
    std::fclose(f);
    // the collection: overwrite its closing lines with the new data set
    FILE* p = std::fopen(pvd_name(l).c_str(), "r+");
    if (!p) throw std::runtime_error("Cannot open file '" + pvd_name(l) + "' for output.");
    std::fseek(p, -27, SEEK_END);
    std::fprintf(p, "  <DataSet timestep=\"%e\" file=\"%s\"/>\n", time, name.c_str());
    std::fputs(kPvdTail, p);
    std::fclose(p);
  }

  void write_xyz(double time, int l, const LatticeVariables& LV) {
    const std::string full = prefix_ + frame_name(l, "xyz");
    FILE* f = std::fopen(full.c_str(), "w");
    if (!f) throw std::runtime_error("Cannot open file '" + full + "' for output.");
    const LatticeInfo& L = lattices_[l];
    const int ns = LV.nSpecies_max;
    std::fprintf(f, "%d\n", (int)LV.q.size());
    std::fprintf(f, "Lattice=\"%e %e %e %e %e %e %e %e %e\" ", L.abc[0][0], L.abc[0][1], L.abc[0][2], L.abc[1][0],
                 L.abc[1][1], L.abc[1][2], L.abc[2][0], L.abc[2][1], L.abc[2][2]);
    std::fprintf(f, "Origin=\"%e %e %e\" ", L.o[0], L.o[1], L.o[2]);
    std::fprintf(f, "pbc=\"%s %s %s\" ", L.periodic[0] ? "T" : "F", L.periodic[1] ? "T" : "F",
                 L.periodic[2] ? "T" : "F");
    std::fprintf(f, "Properties=\"species:S:1:pos:R:3:disp:R:3:type:I:1:atomic_freq:R:1:tag:I:1:"
                    "molar_frac:R:%d:chem_potent:R:%d\" Time=%e\n",
                 ns, ns, time);
    const char* sep = precision_ == 2 ? "  " : " ";
    auto real = [&](double v) {
      std::fprintf(f, num(), v);
      std::fputs(sep, f);
    };
    for (int i = 0; i < (int)LV.q.size(); ++i) {
      std::fprintf(f, "%16s%s", material_[LV.matid[i]].c_str(), sep);
      for (int c = 0; c < 3; ++c) real(LV.q[i][c]);
      for (int c = 0; c < 3; ++c) real(LV.q[i][c] - LV.q0[i][c]);
      std::fprintf(f, "%4d%s", LV.siteid[i], sep);
      real(LV.sigma[i]);
      std::fprintf(f, "%4d%s", LV.tag[i], sep);
      for (const auto* arr : {&LV.x[i], &LV.gamma[i]}) {
        int s = 0;
        for (; s < (int)arr->size(); ++s) real((*arr)[s]);
        for (; s < ns; ++s) real(0.0);
      }
      std::fputc('\n', f);
    }
    std::fclose(f);
  }
};

}  // namespace a2c
