// specimen.hpp -- the specimen as the reference's set-up builds it from the
// deck's Species, Material, Lattice and RegionalInitialCondition objects
// (Main.cpp:42-109, IoData.cpp:845-880, MaterialOperator.cpp:19-49,
// LatticeStructure.cpp:24-150, SpaceInitializer.cpp:41-298, :604-877,
// :883-1061, LatticeVariables.cpp:60-165): one LatticeVariables per lattice
// (LatticeVariables.h:14-50), then packed into the engine's site layout --
// Pd block then H block, [qx, qy, qz, sigma] per site, occupancy x per H
// site -- and unpacked back after a run.
//
// Deck objects and keys are the reference's (IoData.cpp:28-621):
//   Species[i].SpeciesName
//   Material[m] { MaterialName; Species0..Species11; MinimumMolarFraction }
//   Lattice[l] { OAx..OCz (lattice vectors); Ox, Oy, Oz; PeriodicOA/OB/OC;
//                MinDistance; Site[s] { La, Lb, Lc; MaterialID; Tag;
//                AtomicFrequency; MeanDisplacementX/Y/Z;
//                LocalSpecies[k] { SpeciesID; Diffusive; MolarFraction;
//                MinimumMolarFraction } }; Plane[g], Sphere[g],
//                Parallelepiped[g] (the domain) }
//   RegionalInitialCondition[r] { LatticeID; MaterialID; SiteID; Tag;
//                AtomicFrequency; MeanDisplacementX/Y/Z; LocalSpecies[k];
//                Plane[g], Sphere[g], Parallelepiped[g] (the region) }
// Spheroid, CylinderAndCone, CylinderWithSphericalCaps and CustomGeometry
// (a plugin the reference dlopens) are refused with an error.
//
// The reference's conventions are kept, including the ones a deck must work
// around: a site without AtomicFrequency has sigma = -1 (refused when
// packed); a regional condition sets sigma and Tag of every site it covers
// whether or not it names them; a site exactly on a parallelepiped face or
// sphere surface is outside; overlapping lattices drop the sites within
// MinDistance (a box test, max |dq_c| <= dmin) of another lattice's sites.
#pragma once
#include <algorithm>
#include <array>
#include <cfloat>
#include <cmath>
#include <cstdint>
#include <map>
#include <set>
#include <stdexcept>
#include <string>
#include <unordered_map>
#include <vector>

#include "deck.hpp"

namespace a2c {

using V3 = std::array<double, 3>;

inline V3 v3_add(const V3& a, const V3& b) { return {a[0] + b[0], a[1] + b[1], a[2] + b[2]}; }
inline V3 v3_sub(const V3& a, const V3& b) { return {a[0] - b[0], a[1] - b[1], a[2] - b[2]}; }
inline V3 v3_scale(double c, const V3& a) { return {c * a[0], c * a[1], c * a[2]}; }
inline double v3_dot(const V3& a, const V3& b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }
inline V3 v3_cross(const V3& a, const V3& b) {
  return {a[1] * b[2] - a[2] * b[1], a[2] * b[0] - a[0] * b[2], a[0] * b[1] - a[1] * b[0]};
}
inline double v3_norm(const V3& a) { return std::sqrt(a[0] * a[0] + a[1] * a[1] + a[2] * a[2]); }

// LatticeVariables.h:14-50
struct LatticeVariables {
  int lattice_id = -1;
  V3 lmin{0, 0, 0}, lmax{0, 0, 0};
  int size = 0;
  std::vector<int> siteid, matid;
  std::vector<V3> l, q, q0;
  std::vector<double> sigma;
  std::vector<int> tag;
  int nSpecies_max = 0;
  std::vector<std::vector<int>> diffusive, species_id;
  std::vector<std::vector<double>> x, gamma;

  // LatticeVariables.cpp:60-83
  void find_bounding_box_and_size() {
    lmin = {DBL_MAX, DBL_MAX, DBL_MAX};
    lmax = {-DBL_MAX, -DBL_MAX, -DBL_MAX};
    for (const V3& p : l)
      for (int c = 0; c < 3; ++c) {
        lmin[c] = std::min(lmin[c], p[c]);
        lmax[c] = std::max(lmax[c], p[c]);
      }
    size = (int)l.size();
  }
};

// One user geometry (IoData.h PlaneData, SphereData, ParallelepipedData) with
// the reference's membership test (SpaceInitializer.cpp:625-689).
struct Geometry {
  enum Kind { PLANE = 0, SPHERE = 3, PARALLELEPIPED = 4 } kind = PLANE;
  int order = 0;
  bool intersect = false;  // Inclusion = Intersection (else Union)
  bool exterior = false;   // Side = Exterior
  V3 p{0, 0, 0}, n{0, 0, 0};  // plane: point and unit normal; sphere: center
  double radius = -1.0;
  V3 O{0, 0, 0}, OA{0, 0, 0}, OB{0, 0, 0}, OC{0, 0, 0};  // parallelepiped corner and edges
  V3 face_o[6], face_n[6];

  // DistancePointToParallelepiped.h:38-73: six faces, unit outward normals
  void setup_faces() {
    const V3 A = v3_add(O, OA), B = v3_add(O, OB), C = v3_add(O, OC);
    const V3 D = v3_add(A, OB), E = v3_add(A, OC), F = v3_add(B, OC);
    const V3 o[6] = {O, O, A, B, O, C};
    const V3 e1[6] = {OB, OA, v3_sub(D, A), v3_sub(F, B), OC, v3_sub(E, C)};
    const V3 e2[6] = {OA, OC, v3_sub(E, A), v3_sub(D, B), OB, v3_sub(F, C)};
    for (int i = 0; i < 6; ++i) {
      face_o[i] = o[i];
      const V3 nn = v3_cross(e1[i], e2[i]);
      face_n[i] = v3_scale(1.0 / v3_norm(nn), nn);
    }
  }
  // Calculate(q) >= 0 of the sphere and parallelepiped distance functions:
  // a point on the surface is outside.  For the parallelepiped the signed
  // distance is negative only when the point is strictly on the inner side
  // of all six face planes (DistancePointToParallelepiped.h:83-109).
  bool outside(const V3& q) const {
    if (kind == SPHERE) return v3_norm(v3_sub(q, p)) - radius >= 0.0;
    for (int i = 0; i < 6; ++i)
      if (v3_dot(v3_sub(q, face_o[i]), face_n[i]) >= 0.0) return true;
    return false;
  }
  // SpaceInitializer.cpp:630-635, :666-673: the test is applied only when it
  // can change the answer.
  bool apply(bool in, const V3& q) const {
    if (!((in && intersect) || (!in && !intersect))) return in;
    if (kind == PLANE) return v3_dot(v3_sub(q, p), n) > 0.0;
    const bool out = outside(q);
    return (out && exterior) || (!out && !exterior);
  }
};

struct LocalSpeciesData {  // IoData.cpp:380-404
  int species_id = -1;
  bool diffusive = false;
  double molar_fraction = -1.0e-3, min_molar_fraction = -1.0e-4;
};

struct MaterialInfo {  // IoData.cpp:540-572, :845-880; MaterialOperator.cpp:19-49
  std::string name;
  std::vector<int> species;  // local -> global species id
  double xmin = 0.0;
  int local_of(int global) const {
    for (size_t i = 0; i < species.size(); ++i)
      if (species[i] == global) return (int)i;
    return -1;
  }
};

struct SiteGroup {  // IoData.cpp:412-445; LatticeStructure.cpp:78-146
  V3 l{-1, -1, -1};
  int matid = -1, tag = 0;
  double sigma = -1.0;
  V3 disp{0, 0, 0};
  std::vector<int> species_id, diffusive;
  std::vector<double> x0, xmin;
};

struct LatticeInfo {  // IoData.cpp:450-508; LatticeStructure.cpp:24-76
  V3 abc[3], o{0, 0, 0};
  bool periodic[3] = {false, false, false};
  double dmin = 0.0;
  std::vector<SiteGroup> sites;
  std::vector<Geometry> geometry;  // sorted by OperationOrder
  int nSpecies_max = 0;
  V3 xyz(const V3& l) const {  // LatticeStructure.h:71
    return v3_add(v3_add(v3_add(o, v3_scale(l[0], abc[0])), v3_scale(l[1], abc[1])), v3_scale(l[2], abc[2]));
  }
};

struct RegionInfo {  // IoData.cpp:577-622
  int lattice = -1, material = -1, site = -1, tag = 0;
  double sigma = -1.0;
  V3 disp{0, 0, 0};
  std::vector<LocalSpeciesData> species;
  std::vector<Geometry> geometry;
};

struct SpecimenInput {
  std::vector<std::string> species;  // global species names
  std::vector<MaterialInfo> materials;
  std::vector<LatticeInfo> lattices;
  std::vector<RegionInfo> regions;
};

namespace specimen_detail {

[[noreturn]] inline void fail(const std::string& m) { throw std::runtime_error(m); }

// the ids i of the deck's objects "name[i]", checked to be 0..n-1 as the
// reference's object maps are (Main.cpp:46-60)
inline std::vector<int> object_ids(const Deck& d, const std::string& name) {
  std::set<int> ids;
  const std::string pre = name + "[";
  const auto& v = d.values();
  for (auto it = v.lower_bound(pre); it != v.end() && it->first.compare(0, pre.size(), pre) == 0; ++it) {
    const size_t e = it->first.find(']', pre.size());
    ids.insert(std::atoi(it->first.substr(pre.size(), e - pre.size()).c_str()));
  }
  std::vector<int> out(ids.begin(), ids.end());
  for (size_t k = 0; k < out.size(); ++k)
    if (out[k] != (int)k) fail("Detected error in " + name + " indices (id = " + std::to_string(out[k]) + ")");
  return out;
}

inline std::string at(const std::string& pre, int i) { return pre + "[" + std::to_string(i) + "]"; }

inline int token(const Deck& d, const std::string& key, const std::vector<std::string>& names, int def) {
  if (!d.has(key)) return def;
  const std::string s = d.str(key, "");
  for (size_t k = 0; k < names.size(); ++k)
    if (s == names[k]) return (int)k;
  fail("unknown value '" + s + "' for " + key);
}

inline V3 vec(const Deck& d, const std::string& pre, const char* x, const char* y, const char* z, double def) {
  return {d.num(pre + x, def), d.num(pre + y, def), d.num(pre + z, def)};
}

inline std::vector<LocalSpeciesData> local_species(const Deck& d, const std::string& pre) {
  std::vector<LocalSpeciesData> out;
  for (int k : object_ids(d, pre + ".LocalSpecies")) {
    const std::string s = at(pre + ".LocalSpecies", k) + ".";
    LocalSpeciesData ls;
    ls.species_id = (int)d.integer(s + "SpeciesID", -1);
    ls.diffusive = token(d, s + "Diffusive", {"False", "True"}, 0) == 1;
    ls.molar_fraction = d.num(s + "MolarFraction", -1.0e-3);
    ls.min_molar_fraction = d.num(s + "MinimumMolarFraction", -1.0e-4);
    out.push_back(ls);
  }
  return out;
}

// SpaceInitializer.cpp:326-595 for the geometries this build supports
inline std::vector<Geometry> geometries(const Deck& d, const std::string& pre) {
  for (const char* g : {"Spheroid", "CylinderAndCone", "CylinderWithSphericalCaps"})
    if (!object_ids(d, pre + "." + g).empty()) fail(pre + ": geometry " + g + " is not supported");
  if (d.str(pre + ".CustomGeometry.GeometrySpecifier", "") != "")
    fail(pre + ": CustomGeometry (a loaded plugin) is not supported");
  const std::vector<std::string> incl = {"Override", "Intersection", "Union"}, side = {"Interior", "Exterior"};
  std::vector<Geometry> g;
  for (int k : object_ids(d, pre + ".Plane")) {
    const std::string s = at(pre + ".Plane", k) + ".";
    Geometry G;
    G.kind = Geometry::PLANE;
    G.p = vec(d, s, "Point_x", "Point_y", "Point_z", 0.0);
    const V3 nn = vec(d, s, "Normal_x", "Normal_y", "Normal_z", 0.0);
    const double nrm = v3_norm(nn);
    if (nrm == 0.0) fail(s + "Normal is zero");
    G.n = v3_scale(1.0 / nrm, nn);  // DistancePointToPlane.h:24-31
    G.intersect = token(d, s + "Inclusion", incl, 0) == 1;
    G.order = (int)d.integer(s + "OperationOrder", 0);
    if (token(d, s + "Inclusion", incl, 0) == 0) fail(s + "Inclusion must be Intersection or Union");
    g.push_back(G);
  }
  for (int k : object_ids(d, pre + ".Sphere")) {
    const std::string s = at(pre + ".Sphere", k) + ".";
    Geometry G;
    G.kind = Geometry::SPHERE;
    G.p = vec(d, s, "Center_x", "Center_y", "Center_z", 0.0);
    G.radius = d.num(s + "Radius", -1.0);
    G.exterior = token(d, s + "Side", side, 0) == 1;
    G.intersect = token(d, s + "Inclusion", incl, 0) == 1;
    G.order = (int)d.integer(s + "OperationOrder", 0);
    if (token(d, s + "Inclusion", incl, 0) == 0) fail(s + "Inclusion must be Intersection or Union");
    g.push_back(G);
  }
  for (int k : object_ids(d, pre + ".Parallelepiped")) {
    const std::string s = at(pre + ".Parallelepiped", k) + ".";
    Geometry G;
    G.kind = Geometry::PARALLELEPIPED;
    G.O = vec(d, s, "Px", "Py", "Pz", 0.0);
    G.OA = v3_sub(vec(d, s, "Ax", "Ay", "Az", 0.0), G.O);
    G.OB = v3_sub(vec(d, s, "Bx", "By", "Bz", 0.0), G.O);
    G.OC = v3_sub(vec(d, s, "Cx", "Cy", "Cz", 0.0), G.O);
    // SpaceInitializer.cpp:827-832
    if (v3_norm(G.OA) == 0 || v3_norm(G.OB) == 0 || v3_norm(G.OC) == 0 ||
        v3_dot(v3_cross(G.OA, G.OB), G.OC) <= 0.0)
      fail(s + " overlapping vertices or violation of the right-hand rule");
    G.setup_faces();
    G.exterior = token(d, s + "Side", side, 0) == 1;
    G.intersect = token(d, s + "Inclusion", incl, 0) == 1;
    G.order = (int)d.integer(s + "OperationOrder", 0);
    if (token(d, s + "Inclusion", incl, 0) == 0) fail(s + "Inclusion must be Intersection or Union");
    g.push_back(G);
  }
  std::vector<Geometry> sorted(g.size());
  std::vector<bool> used(g.size(), false);
  for (const Geometry& G : g) {
    if (G.order < 0 || G.order >= (int)g.size() || used[G.order])
      fail(pre + ": a geometry has an invalid or conflicting OperationOrder (" + std::to_string(G.order) + ")");
    used[G.order] = true;
    sorted[G.order] = G;
  }
  return sorted;
}

// GeoTools/BoundingBoxes.cpp:86-109, :114-192 (scaling factor 1): the extent
// along each lattice direction, divided by that vector's length
inline void bounding_box(const LatticeInfo& L, V3& lmin, V3& lmax) {
  lmin = {DBL_MAX, DBL_MAX, DBL_MAX};
  lmax = {-DBL_MAX, -DBL_MAX, -DBL_MAX};
  double dn[3];
  V3 du[3];
  for (int i = 0; i < 3; ++i) {
    dn[i] = v3_norm(L.abc[i]);
    du[i] = v3_scale(1.0 / dn[i], L.abc[i]);
  }
  for (const Geometry& G : L.geometry) {
    if (G.kind == Geometry::PLANE || G.exterior) continue;
    V3 lo, hi;
    if (G.kind == Geometry::SPHERE) {
      const V3 Op = v3_sub(G.p, L.o);
      for (int i = 0; i < 3; ++i) {
        const double pr = v3_dot(Op, du[i]);
        lo[i] = (pr - G.radius) / dn[i];
        hi[i] = (pr + G.radius) / dn[i];
      }
    } else {
      const V3 A = v3_add(G.O, G.OA);
      const V3 v[8] = {G.O, A, v3_add(G.O, G.OB), v3_add(G.O, G.OC), v3_add(A, G.OB), v3_add(A, G.OC),
                       v3_add(v3_add(G.O, G.OB), G.OC), v3_add(v3_add(A, G.OB), G.OC)};
      lo = {DBL_MAX, DBL_MAX, DBL_MAX};
      hi = {-DBL_MAX, -DBL_MAX, -DBL_MAX};
      for (const V3& p : v) {
        const V3 Ov = v3_sub(p, L.o);
        for (int j = 0; j < 3; ++j) {
          const double c = v3_dot(Ov, du[j]);
          lo[j] = std::min(lo[j], c);
          hi[j] = std::max(hi[j], c);
        }
      }
      for (int i = 0; i < 3; ++i) {
        lo[i] /= dn[i];
        hi[i] /= dn[i];
      }
    }
    for (int i = 0; i < 3; ++i) {
      lmin[i] = std::min(lmin[i], lo[i]);
      lmax[i] = std::max(lmax[i], hi[i]);
    }
  }
  for (int i = 0; i < 3; ++i)
    if (!std::isfinite(lmin[i]) || !std::isfinite(lmax[i]) || lmin[i] >= lmax[i])
      fail("Unable to find a bounding box for a lattice domain. Domain is unbounded.");
}

}  // namespace specimen_detail

// The deck's objects, checked as the reference checks them.
inline SpecimenInput read_specimen(const Deck& d) {
  using namespace specimen_detail;
  SpecimenInput in;
  for (int i : object_ids(d, "Species")) in.species.push_back(d.str(at("Species", i) + ".SpeciesName", ""));
  for (int m : object_ids(d, "Material")) {
    const std::string s = at("Material", m) + ".";
    MaterialInfo M;
    M.name = d.str(s + "MaterialName", "");
    int sp[12];
    for (int k = 0; k < 12; ++k) sp[k] = (int)d.integer(s + "Species" + std::to_string(k), -1);
    int n = 0;
    while (n < 12 && sp[n] >= 0) ++n;  // IoData.cpp:860-877
    if (n == 0) fail("Material[" + std::to_string(m) + "] does not have any species.");
    for (int k = n; k < 12; ++k)
      if (sp[k] >= 0) fail("Found gap(s) in Material[" + std::to_string(m) + "] species indices.");
    std::set<int> seen;
    for (int k = 0; k < n; ++k) {
      if (sp[k] >= (int)in.species.size())
        fail("Material[" + std::to_string(m) + "] has undefined species (id=" + std::to_string(sp[k]) + ").");
      seen.insert(sp[k]);
      M.species.push_back(sp[k]);
    }
    if ((int)seen.size() != n) fail("Detected duplicate species IDs in Material[" + std::to_string(m) + "].");
    M.xmin = d.num(s + "MinimumMolarFraction", 0.0);
    if (M.xmin < 0.0) fail("Detected negative molar fraction in Material[" + std::to_string(m) + "].");
    in.materials.push_back(M);
  }
  for (int li : object_ids(d, "Lattice")) {
    const std::string pre = at("Lattice", li), s = pre + ".";
    LatticeInfo L;
    L.abc[0] = {d.num(s + "OAx", 1.0), d.num(s + "OAy", 0.0), d.num(s + "OAz", 0.0)};
    L.abc[1] = {d.num(s + "OBx", 0.0), d.num(s + "OBy", 1.0), d.num(s + "OBz", 0.0)};
    L.abc[2] = {d.num(s + "OCx", 0.0), d.num(s + "OCy", 0.0), d.num(s + "OCz", 1.0)};
    L.o = vec(d, s, "Ox", "Oy", "Oz", 0.0);
    const char* per[3] = {"PeriodicOA", "PeriodicOB", "PeriodicOC"};
    for (int c = 0; c < 3; ++c) {
      L.periodic[c] = token(d, s + per[c], {"False", "True"}, 0) == 1;
      if (v3_norm(L.abc[c]) <= 0.0) fail(pre + ": a lattice vector is zero");
    }
    // LatticeStructure.cpp:59-61: dmin defaults to the shortest lattice vector
    L.dmin = d.num(s + "MinDistance", 0.0);
    if (L.dmin <= 0.0) L.dmin = std::min(v3_norm(L.abc[0]), std::min(v3_norm(L.abc[1]), v3_norm(L.abc[2])));
    const std::vector<int> sites = object_ids(d, pre + ".Site");
    if (sites.empty()) fail("No lattice sites specified for " + pre + ".");
    for (int si : sites) {
      const std::string t = at(pre + ".Site", si), u = t + ".";
      SiteGroup S;
      S.l = {d.num(u + "La", -1.0), d.num(u + "Lb", -1.0), d.num(u + "Lc", -1.0)};
      for (int c = 0; c < 3; ++c)
        if (S.l[c] < 0.0 || S.l[c] >= 1.0) fail(t + ": site coords must be in [0, 1).");
      S.matid = (int)d.integer(u + "MaterialID", -1);
      if (S.matid < 0 || S.matid >= (int)in.materials.size()) fail(t + ": non-existent material id.");
      S.tag = (int)d.integer(u + "Tag", 0);
      S.sigma = d.num(u + "AtomicFrequency", -1.0);
      S.disp = vec(d, u, "MeanDisplacementX", "MeanDisplacementY", "MeanDisplacementZ", 0.0);
      const MaterialInfo& M = in.materials[S.matid];
      const int ns = (int)M.species.size();
      L.nSpecies_max = std::max(L.nSpecies_max, ns);
      S.species_id = M.species;
      S.diffusive.assign(ns, 0);
      S.x0.assign(ns, M.xmin);
      S.xmin.assign(ns, M.xmin);
      std::set<int> seen;
      for (const LocalSpeciesData& ls : local_species(d, t)) {
        const int loc = M.local_of(ls.species_id);
        if (loc < 0) fail(t + " has species " + std::to_string(ls.species_id) + ", which is not in its material.");
        if (!seen.insert(ls.species_id).second) fail(t + " has duplicate (global) species.");
        S.diffusive[loc] = ls.diffusive ? 1 : 0;
        if (ls.molar_fraction >= 0.0) S.x0[loc] = ls.molar_fraction;
        if (ls.min_molar_fraction >= 0.0) S.xmin[loc] = ls.min_molar_fraction;
      }
      L.sites.push_back(S);
    }
    L.geometry = geometries(d, pre);
    in.lattices.push_back(L);
  }
  for (int r : object_ids(d, "RegionalInitialCondition")) {
    const std::string pre = at("RegionalInitialCondition", r), s = pre + ".";
    RegionInfo R;
    R.lattice = (int)d.integer(s + "LatticeID", -1);
    R.material = (int)d.integer(s + "MaterialID", -1);
    R.site = (int)d.integer(s + "SiteID", -1);
    R.tag = (int)d.integer(s + "Tag", 0);
    R.sigma = d.num(s + "AtomicFrequency", -1.0);
    R.disp = vec(d, s, "MeanDisplacementX", "MeanDisplacementY", "MeanDisplacementZ", 0.0);
    R.species = local_species(d, pre);
    R.geometry = geometries(d, pre);
    if (R.lattice >= (int)in.lattices.size()) fail(pre + ": unknown lattice ID.");
    if (R.material >= (int)in.materials.size()) fail(pre + ": unknown MaterialID.");
    in.regions.push_back(R);
  }
  return in;
}

// SpaceInitializer::CreateOneLattice + CheckAndAddSitesInCell
// (SpaceInitializer.cpp:120-298, :604-719), on one process
inline void create_lattice(const LatticeInfo& L, int id, LatticeVariables& LV) {
  LV = LatticeVariables();
  LV.lattice_id = id;
  V3 lmin, lmax;
  specimen_detail::bounding_box(L, lmin, lmax);
  int lo[3], hi[3];
  for (int i = 0; i < 3; ++i) {
    lo[i] = (int)std::floor(lmin[i]);
    hi[i] = (int)std::ceil(lmax[i]) + 1;
  }
  for (int lk = lo[2]; lk < hi[2]; ++lk)
    for (int lj = lo[1]; lj < hi[1]; ++lj)
      for (int li = lo[0]; li < hi[0]; ++li)
        for (int sid = 0; sid < (int)L.sites.size(); ++sid) {
          const SiteGroup& S = L.sites[sid];
          const V3 l = v3_add(S.l, V3{(double)li, (double)lj, (double)lk});
          const V3 q = L.xyz(l);
          bool in = true;
          for (const Geometry& G : L.geometry) in = G.apply(in, q);
          if (!in) continue;
          LV.l.push_back(l);
          LV.q0.push_back(q);
          LV.siteid.push_back(sid);
          LV.matid.push_back(S.matid);
          LV.sigma.push_back(S.sigma);
          LV.q.push_back(v3_add(q, S.disp));
          LV.tag.push_back(S.tag);
          LV.species_id.push_back(S.species_id);
          LV.diffusive.push_back(S.diffusive);
          LV.x.push_back(S.x0);
          LV.gamma.push_back(std::vector<double>(S.species_id.size(), 0.0));
        }
  LV.find_bounding_box_and_size();
  LV.nSpecies_max = L.nSpecies_max;
}

// SpaceInitializer::ApplyRegionalICOnOneLattice (SpaceInitializer.cpp:883-1061)
inline void apply_region(const RegionInfo& R, LatticeVariables& LV) {
  for (int i = 0; i < LV.size; ++i) {
    if (LV.matid[i] != R.material || LV.siteid[i] != R.site) continue;
    V3& q = LV.q[i];
    bool in = true;
    for (const Geometry& G : R.geometry) in = G.apply(in, q);
    if (!in) continue;
    q = v3_add(q, R.disp);
    LV.sigma[i] = R.sigma;
    LV.tag[i] = R.tag;
    for (const LocalSpeciesData& ls : R.species) {
      if (ls.molar_fraction < 0.0) continue;
      for (size_t s = 0; s < LV.species_id[i].size(); ++s)
        if (LV.species_id[i][s] == ls.species_id) LV.x[i][s] = ls.molar_fraction;
    }
  }
}

// LatticeVariables::RemoveConflictingSites (LatticeVariables.cpp:87-165): the
// reference's KD-tree query is a box test, max_c |q_c - q2_c| <= dmin.
inline int remove_conflicting_sites(LatticeVariables& LV, const std::vector<V3>& q2, double dmin) {
  if (q2.empty() || LV.size == 0) return 0;
  auto key = [dmin](const V3& p, int dx, int dy, int dz) {
    const int64_t i = (int64_t)std::floor(p[0] / dmin) + dx, j = (int64_t)std::floor(p[1] / dmin) + dy,
                  k = (int64_t)std::floor(p[2] / dmin) + dz;
    return (uint64_t)((i & 0x1FFFFF) | ((j & 0x1FFFFF) << 21) | ((k & 0x1FFFFF) << 42));
  };
  std::unordered_map<uint64_t, std::vector<int>> grid;
  for (int j = 0; j < (int)q2.size(); ++j) grid[key(q2[j], 0, 0, 0)].push_back(j);
  std::vector<char> stay(LV.size, 1);
  int erased = 0;
  for (int i = 0; i < LV.size; ++i) {
    bool hit = false;
    for (int dz = -1; dz <= 1 && !hit; ++dz)
      for (int dy = -1; dy <= 1 && !hit; ++dy)
        for (int dx = -1; dx <= 1 && !hit; ++dx) {
          auto it = grid.find(key(LV.q[i], dx, dy, dz));
          if (it == grid.end()) continue;
          for (int j : it->second) {
            const V3& p = q2[j];
            if (std::max(std::fabs(p[0] - LV.q[i][0]), std::max(std::fabs(p[1] - LV.q[i][1]),
                                                                 std::fabs(p[2] - LV.q[i][2]))) <= dmin) {
              hit = true;
              break;
            }
          }
        }
    if (hit) {
      stay[i] = 0;
      ++erased;
    }
  }
  if (!erased) return 0;
  int c = 0;
  for (int i = 0; i < LV.size; ++i) {
    if (!stay[i]) continue;
    LV.siteid[c] = LV.siteid[i];
    LV.matid[c] = LV.matid[i];
    LV.l[c] = LV.l[i];
    LV.q[c] = LV.q[i];
    LV.q0[c] = LV.q0[i];
    LV.sigma[c] = LV.sigma[i];
    LV.tag[c] = LV.tag[i];
    LV.diffusive[c] = LV.diffusive[i];
    LV.species_id[c] = LV.species_id[i];
    LV.x[c] = LV.x[i];
    LV.gamma[c] = LV.gamma[i];
    ++c;
  }
  for (auto* v : {&LV.siteid, &LV.matid, &LV.tag}) v->resize(c);
  for (auto* v : {&LV.l, &LV.q, &LV.q0}) v->resize(c);
  LV.sigma.resize(c);
  LV.diffusive.resize(c);
  LV.species_id.resize(c);
  LV.x.resize(c);
  LV.gamma.resize(c);
  bool found = false;
  for (const auto& xi : LV.x) found = found || (int)xi.size() == LV.nSpecies_max;
  if (!found) specimen_detail::fail("After removing overlapping sites on a lattice, nSpecies_max has reduced.");
  LV.find_bounding_box_and_size();
  return erased;
}

// SpaceInitializer::SetupLatticeVariables (SpaceInitializer.cpp:41-115)
inline std::vector<LatticeVariables> setup_lattice_variables(const SpecimenInput& in) {
  const int nl = (int)in.lattices.size();
  std::vector<LatticeVariables> LVS(nl);
  for (int l = 0; l < nl; ++l) create_lattice(in.lattices[l], l, LVS[l]);
  for (const RegionInfo& R : in.regions) {
    if (R.lattice < 0)
      for (auto& LV : LVS) apply_region(R, LV);
    else
      apply_region(R, LVS[R.lattice]);
  }
  for (int i = 0; i < nl; ++i)
    for (int j = 0; j < nl; ++j) {
      if (j == i) continue;
      if (j < i && in.lattices[j].dmin <= in.lattices[i].dmin) continue;  // already resolved
      remove_conflicting_sites(LVS[i], LVS[j].q, in.lattices[i].dmin);
    }
  return LVS;
}

// The engine's layout.  A site whose material holds the solute species
// (default "H") is an H site, with occupancy = the solute's molar fraction;
// a site whose material holds the host species (default "Pd") and not the
// solute is a Pd site.  Pd sites of lattice 0, 1, ... in LatticeVariables
// order, then the H sites likewise.  Periodic axes need lattice vectors
// along x, y, z and the same cell count on every lattice that has sites;
// the box edge is |lattice vector| x cells.
struct PackedSpecimen {
  int32_t n_pd = 0, n_h = 0;
  std::vector<double> X, x;      // 4 (n_pd + n_h), n_h
  std::vector<int32_t> tag_h;    // LatticeVariables tag of each H site
  std::vector<int32_t> lattice, index, slot;  // where each packed site lives; solute slot (-1 for Pd)
  double box[3] = {0, 0, 0};
  int32_t periodic[3] = {0, 0, 0};
  double a = 0.0;  // length of lattice vector OA of the first lattice with Pd sites
};

inline PackedSpecimen pack_pd_h(const SpecimenInput& in, const std::vector<LatticeVariables>& LVS,
                                const std::string& host = "Pd", const std::string& solute = "H") {
  using specimen_detail::fail;
  int g_host = -1, g_sol = -1;
  for (size_t k = 0; k < in.species.size(); ++k) {
    if (in.species[k] == host) g_host = (int)k;
    if (in.species[k] == solute) g_sol = (int)k;
  }
  if (g_host < 0) fail("no species named " + host);
  PackedSpecimen P;
  std::vector<int32_t> pd_l, pd_i, h_l, h_i, h_s;
  for (const LatticeVariables& LV : LVS)
    for (int i = 0; i < LV.size; ++i) {
      int slot = -1, hs = -1;
      for (size_t s = 0; s < LV.species_id[i].size(); ++s) {
        if (LV.species_id[i][s] == g_sol) slot = (int)s;
        if (LV.species_id[i][s] == g_host) hs = (int)s;
      }
      if (!(LV.sigma[i] > 0.0))
        fail("Lattice[" + std::to_string(LV.lattice_id) + "] site " + std::to_string(i) +
             " has no positive AtomicFrequency");
      if (slot >= 0) {
        h_l.push_back(LV.lattice_id);
        h_i.push_back(i);
        h_s.push_back(slot);
      } else if (hs >= 0) {
        pd_l.push_back(LV.lattice_id);
        pd_i.push_back(i);
      } else {
        fail("Lattice[" + std::to_string(LV.lattice_id) + "] site " + std::to_string(i) + " holds neither " + host +
             " nor " + solute);
      }
    }
  P.n_pd = (int32_t)pd_i.size();
  P.n_h = (int32_t)h_i.size();
  const int32_t n = P.n_pd + P.n_h;
  P.X.resize(4 * (size_t)n);
  P.x.resize(P.n_h);
  P.tag_h.resize(P.n_h);
  for (int32_t s = 0; s < n; ++s) {
    const bool h = s >= P.n_pd;
    const int32_t l = h ? h_l[s - P.n_pd] : pd_l[s], i = h ? h_i[s - P.n_pd] : pd_i[s];
    const LatticeVariables& LV = LVS[l];
    for (int c = 0; c < 3; ++c) P.X[4 * (size_t)s + c] = LV.q[i][c];
    P.X[4 * (size_t)s + 3] = LV.sigma[i];
    P.lattice.push_back(l);
    P.index.push_back(i);
    P.slot.push_back(h ? h_s[s - P.n_pd] : -1);
    if (h) {
      P.x[s - P.n_pd] = LV.x[i][h_s[s - P.n_pd]];
      P.tag_h[s - P.n_pd] = LV.tag[i];
    }
  }
  for (size_t l = 0; l < LVS.size(); ++l)
    if (P.a == 0.0 && std::count(pd_l.begin(), pd_l.end(), (int32_t)l)) P.a = v3_norm(in.lattices[l].abc[0]);
  for (int c = 0; c < 3; ++c) {
    double lo = DBL_MAX, hi = -DBL_MAX;
    for (int32_t s = 0; s < n; ++s) {
      lo = std::min(lo, P.X[4 * (size_t)s + c]);
      hi = std::max(hi, P.X[4 * (size_t)s + c]);
    }
    P.box[c] = n ? hi - lo : 0.0;
    bool any = false, first = true;
    double edge = 0.0;
    for (size_t l = 0; l < LVS.size(); ++l) {
      const LatticeInfo& L = in.lattices[l];
      if (!L.periodic[c]) continue;
      any = true;
      for (int k = 0; k < 3; ++k)
        if (k != c && L.abc[c][k] != 0.0) fail("a periodic axis needs its lattice vector along x, y or z");
      if (LVS[l].size == 0) continue;
      int cmin = INT32_MAX, cmax = INT32_MIN;
      for (int i = 0; i < LVS[l].size; ++i) {
        const int cell = (int)std::lround(LVS[l].l[i][c] - L.sites[LVS[l].siteid[i]].l[c]);
        cmin = std::min(cmin, cell);
        cmax = std::max(cmax, cell);
      }
      const double e = std::fabs(L.abc[c][c]) * (cmax - cmin + 1);
      if (!first && e != edge) fail("lattices disagree on the periodic box");
      edge = e;
      first = false;
    }
    if (any) {
      for (size_t l = 0; l < LVS.size(); ++l)
        if (LVS[l].size && !in.lattices[l].periodic[c]) fail("lattices disagree on periodic flags");
      P.periodic[c] = 1;
      P.box[c] = edge;
    }
  }
  return P;
}

// Results back into the LatticeVariables: q, sigma of every site; the
// solute's x and gamma at H sites.
inline void unpack_pd_h(const PackedSpecimen& P, const double* X, const double* x, const double* gamma,
                        std::vector<LatticeVariables>& LVS) {
  const int32_t n = P.n_pd + P.n_h;
  for (int32_t s = 0; s < n; ++s) {
    LatticeVariables& LV = LVS[P.lattice[s]];
    const int32_t i = P.index[s];
    for (int c = 0; c < 3; ++c) LV.q[i][c] = X[4 * (size_t)s + c];
    LV.sigma[i] = X[4 * (size_t)s + 3];
    if (s >= P.n_pd) {
      LV.x[i][P.slot[s]] = x[s - P.n_pd];
      if (gamma) LV.gamma[i][P.slot[s]] = gamma[s - P.n_pd];
    }
  }
}

}  // namespace a2c
