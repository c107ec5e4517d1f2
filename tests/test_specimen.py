"""Specimen set-up from the deck's Species / Material / Lattice /
RegionalInitialCondition objects (SpaceInitializer.cpp:41-115, :604-719,
:883-1061; LatticeStructure.cpp:24-150), the pack into the engine's Pd/H
layout, and the VTP / PVD / XYZ snapshots (Output.cpp:8-96, :133-318,
:322-494).  Pinned by the seeded FCC generator (include/a2c_lattice.h): an
FCC Pd + octahedral-H deck cut by a parallelepiped, with a regional slab,
packs to exactly the generator's positions, sigma, occupancies and box."""
import json
import os
import subprocess
import xml.etree.ElementTree as ET

import numpy as np
import pytest

import a2c_amd as A

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PD = [(0, 0, 0), (0.5, 0.5, 0), (0.5, 0, 0.5), (0, 0.5, 0.5)]
H = [(0.5, 0, 0), (0, 0.5, 0), (0, 0, 0.5), (0.5, 0.5, 0.5)]


@pytest.fixture(scope="module")
def probe(tmp_path_factory):
    exe = str(tmp_path_factory.mktemp("spec") / "specimen_probe")
    subprocess.run(["g++", "-O1", "-std=c++17", "-Wall", "-Werror", "-I" + os.path.join(ROOT, "a2c_amd", "host"),
                    os.path.join(ROOT, "tests", "cpp", "specimen_probe.cpp"), "-o", exe], check=True)
    return exe


def fcc_deck(n, a=3.89, slab=2, fmt="VTP_and_XYZ", precision="High", extra=""):
    e = -0.01  # box corners just below the lattice planes: sites on a face are outside
    L = n * a + e
    sites = []
    for k, (u, v, w) in enumerate(PD + H):
        h = k >= 4
        sp = ("under LocalSpecies[0] { SpeciesID = 1; Diffusive = True; MolarFraction = 0.1; }" if h else
              "under LocalSpecies[0] { SpeciesID = 0; MolarFraction = 1.0; }")
        sites.append(f"under Site[{k}] {{ La = {u}; Lb = {v}; Lc = {w}; MaterialID = {1 if h else 0}; "
                     f"AtomicFrequency = 0.10; {sp} }}")
    regions = []
    for r in range(4):
        regions.append(f"""under RegionalInitialCondition[{r}] {{
          LatticeID = 0; MaterialID = 1; SiteID = {4 + r}; Tag = 1; AtomicFrequency = 0.10;
          under LocalSpecies[0] {{ SpeciesID = 1; MolarFraction = 0.6; }}
          under Parallelepiped[0] {{ Px = {e}; Py = {e}; Pz = {e}; Ax = {slab * a + e}; Ay = {e}; Az = {e};
            Bx = {e}; By = {L}; Bz = {e}; Cx = {e}; Cy = {e}; Cz = {L}; Inclusion = Intersection; }}
        }}""")
    return f"""
    under Species[0] {{ SpeciesName = "Pd"; }}
    under Species[1] {{ SpeciesName = "H"; }}
    under Material[0] {{ MaterialName = "Pd"; Species0 = 0; }}
    under Material[1] {{ MaterialName = "Interstitial"; Species0 = 1; MinimumMolarFraction = 1.0e-3; }}
    under Lattice[0] {{
      OAx = {a}; OAy = 0; OAz = 0; OBx = 0; OBy = {a}; OBz = 0; OCx = 0; OCy = 0; OCz = {a};
      PeriodicOA = True; PeriodicOB = True; PeriodicOC = True;
      {' '.join(sites)}
      under Parallelepiped[0] {{ Px = {e}; Py = {e}; Pz = {e}; Ax = {L}; Ay = {e}; Az = {e};
        Bx = {e}; By = {L}; Bz = {e}; Cx = {e}; Cy = {e}; Cz = {L}; Inclusion = Intersection; }}
    }}
    {' '.join(regions)}
    Time.TimeStep = 2.0e-13;
    under Output {{ Format = {fmt}; Precision = {precision}; Prefix = "out/"; Solution = "spec"; Frequency = 1; }}
    {extra}
    """


def run(probe, tmp_path, text):
    (tmp_path / "out").mkdir(exist_ok=True)
    (tmp_path / "in.st").write_text(text)
    r = subprocess.run([probe, "in.st", "packed.json"], cwd=tmp_path, capture_output=True, text=True)
    if r.returncode:
        return r.stdout + r.stderr
    return json.loads((tmp_path / "packed.json").read_text())


@pytest.mark.parametrize("n", [2, 3])
def test_fcc_deck_packs_to_the_generator(probe, tmp_path, n):
    out = run(probe, tmp_path, fcc_deck(n))
    assert isinstance(out, dict), out
    X, x, box, npd, nh = A.lattice(n, disp_std=0.0, sigma=(0.10, 0.10), x=(0.1, 0.1), slab_cells=2, slab_x=0.6)
    assert out["n_pd"] == npd == 4 * n ** 3 and out["n_h"] == nh and out["sizes"] == [8 * n ** 3]
    assert np.array_equal(np.array(out["X"]), X)  # positions and order, bit for bit
    assert np.array_equal(np.array(out["x"]), x)
    assert out["box"] == list(box) and out["periodic"] == [1, 1, 1] and out["a"] == 3.89
    cell_x = (np.arange(nh) // 4) % n
    assert np.array_equal(np.array(out["tag_h"]), (cell_x < 2).astype(int))


def test_snapshots_in_the_reference_layout(probe, tmp_path):
    n = 2
    out = run(probe, tmp_path, fcc_deck(n))
    assert out["frames"] == 2
    d = tmp_path / "out"
    for f in ("spec_L0_0000.vtp", "spec_L0_0001.vtp", "spec_L0_0000.xyz", "spec_L0_0001.xyz", "spec_L0.pvd"):
        assert (d / f).exists(), f
    pvd = ET.parse(d / "spec_L0.pvd").getroot()
    sets = pvd.findall("./Collection/DataSet")
    assert [s.get("file") for s in sets] == ["spec_L0_0000.vtp", "spec_L0_0001.vtp"]
    assert float(sets[1].get("timestep")) == 2e-13
    root = ET.parse(d / "spec_L0_0001.vtp").getroot()
    piece = root.find("./PolyData/Piece")
    N = 8 * n ** 3
    assert int(piece.get("NumberOfPoints")) == N
    pts = np.array(piece.find("./Points/DataArray").text.split(), float).reshape(N, 3)
    arrays = {a.get("Name"): a for a in piece.findall("./PointData/DataArray")}
    assert list(arrays) == ["SiteID", "MaterialID", "Tag", "AtomicFrequency", "MolarFraction", "ChemicalPotential"]
    sid = np.array(arrays["SiteID"].text.split(), int)
    mat = np.array(arrays["MaterialID"].text.split(), int)
    assert np.array_equal(sid, np.tile(np.arange(8), n ** 3)) and np.array_equal(mat, (sid >= 4).astype(int))
    # the reference's site order: cells with la fastest, site groups within a cell
    X, x, box, npd, nh = A.lattice(n, disp_std=0.0, slab_cells=2, slab_x=0.6)
    q = X.reshape(-1, 4)[:, :3]
    assert np.allclose(pts[sid < 4], q[:npd], rtol=1e-10, atol=0)
    xf = np.array(arrays["MolarFraction"].text.split(), float)
    gf = np.array(arrays["ChemicalPotential"].text.split(), float)
    assert arrays["MolarFraction"].get("NumberOfComponents") == "1"
    x1 = x.copy()
    x1[0] = 0.5
    assert np.allclose(xf[sid >= 4], x1, rtol=1e-10) and np.all(xf[sid < 4] == 1.0)
    assert np.allclose(gf[sid >= 4], -x, rtol=1e-10) and np.all(gf[sid < 4] == 0.0)
    lines = (d / "spec_L0_0001.xyz").read_text().splitlines()
    assert int(lines[0]) == N and len(lines) == N + 2
    assert lines[1].startswith('Lattice="3.890000e+00 0.000000e+00') and 'pbc="T T T"' in lines[1]
    assert "type:I:1:atomic_freq:R:1:tag:I:1:molar_frac:R:1:chem_potent:R:1\" Time=2.000000e-13" in lines[1]
    row = lines[2].split()
    assert row[0] == "Pd" and len(row) == 12 and float(row[8]) == 0.1


def test_output_precision_and_format(probe, tmp_path):
    out = run(probe, tmp_path, fcc_deck(2, fmt="XYZ", precision="Low"))
    assert isinstance(out, dict), out
    d = tmp_path / "out"
    assert not (d / "spec_L0.pvd").exists() and not (d / "spec_L0_0000.vtp").exists()
    row = (d / "spec_L0_0000.xyz").read_text().splitlines()[2]
    assert " 0.0000e+00 0.0000e+00 0.0000e+00 0.0000e+00 0.0000e+00 0.0000e+00    0 1.0000e-01    0 " in row


def test_deck_errors(probe, tmp_path):
    bad_site = fcc_deck(2).replace("La = 0.5; Lb = 0.5; Lc = 0.5;", "La = 1.5; Lb = 0.5; Lc = 0.5;")
    assert "site coords must be in [0, 1)" in run(probe, tmp_path, bad_site)
    override = fcc_deck(2).replace("Inclusion = Intersection; }\n    }", "}\n    }")
    assert "Inclusion must be Intersection or Union" in run(probe, tmp_path, override)
    assert "not supported" in run(probe, tmp_path, fcc_deck(2, extra="under Lattice[0] { under Spheroid[0] { Radius = 1; } }"))
    no_sigma = fcc_deck(2).replace("AtomicFrequency = 0.10; under LocalSpecies[0] { SpeciesID = 0;",
                                   "under LocalSpecies[0] { SpeciesID = 0;")
    assert "AtomicFrequency" in run(probe, tmp_path, no_sigma)


def test_sphere_domain_and_two_lattices(probe, tmp_path):
    # Pd cut by a sphere on lattice 0, H on lattice 1 offset by a/2 along x.
    # Lattice 0 drops its sites with an H site within a box of half-width
    # MinDistance; lattice 1 is not checked against lattice 0 when lattice
    # 0's MinDistance is not larger (SpaceInitializer.cpp:89-104).
    a = 3.89
    base = f"""
    under Species[0] {{ SpeciesName = "Pd"; }}
    under Species[1] {{ SpeciesName = "H"; }}
    under Material[0] {{ MaterialName = "Pd"; Species0 = 0; }}
    under Material[1] {{ MaterialName = "Hsite"; Species0 = 1; }}
    under Lattice[0] {{ OAx = {a}; OBy = {a}; OCz = {a}; MinDistance = D0;
      under Site[0] {{ La = 0; Lb = 0; Lc = 0; MaterialID = 0; AtomicFrequency = 0.1; }}
      under Sphere[0] {{ Center_x = 0; Center_y = 0; Center_z = 0; Radius = {2.5 * a}; Inclusion = Intersection; }} }}
    under Lattice[1] {{ OAx = {a}; OBy = {a}; OCz = {a}; Ox = {a / 2}; MinDistance = 3.0;
      under Site[0] {{ La = 0; Lb = 0; Lc = 0; MaterialID = 1; AtomicFrequency = 0.1;
        under LocalSpecies[0] {{ SpeciesID = 1; MolarFraction = 0.2; }} }}
      under Sphere[0] {{ Center_x = 0; Center_y = 0; Center_z = 0; Radius = {1.5 * a}; Inclusion = Intersection; }} }}
    Output.Format = XYZ; Output.Prefix = "out/";
    """
    g = np.arange(-3, 4)
    grid = np.array([(i, j, k) for k in g for j in g for i in g], float) * a
    pd = grid[np.linalg.norm(grid, axis=1) < 2.5 * a]
    h = grid + [a / 2, 0, 0]
    h = h[np.linalg.norm(h, axis=1) < 1.5 * a]
    for d0 in (1.0, 2.0):
        cheb = np.max(np.abs(pd[:, None, :] - h[None, :, :]), axis=2).min(axis=1)
        out = run(probe, tmp_path, base.replace("D0", str(d0)))
        assert isinstance(out, dict), out
        assert out["n_pd"] == int(np.sum(cheb > d0)) and out["n_h"] == len(h)
        assert out["periodic"] == [0, 0, 0] and np.all(np.array(out["x"]) == 0.2)
    assert int(np.sum(cheb > 2.0)) < len(pd)  # the 2.0 case removed sites
