"""DMD steps on the GPU: each step's chemical potential and diffusion update
against the oracle from the same relaxed state; total H conserved to round-off
when no reservoir or clamp acts; slab initial condition of the coupled
workload (occupancy 0.6 in the first cell layers, xLb elsewhere)."""
import numpy as np
import pytest

import a2c_amd as A
import oracle as O

pytestmark = pytest.mark.gpu


def test_dmd_steps_conserve_and_match_oracle():
    X, x, box, npd, nh = A.lattice((8, 4, 4), seed=5, disp_std=0.02, x=(1e-3, 1e-3), slab_cells=2, slab_x=0.6)
    T = 400.0
    fe = A.FreeEntropy(npd, nh, box, (1, 1, 1), T=T, precision="mixed")
    fe.set_occupancy(x)
    fe.findNeighbors(X)
    prm = A.dmd_params(dt=1e-13, nu=1e13, Qm=0.23, xLb=1e-3, xUb=1 - 1e-16, gamma_bd=1e300,
                       lmvm={"gatol": 1e-8, "grtol": 0.0})
    g = np.zeros(nh)
    fh = np.zeros(nh, np.int32)
    orc = O.Oracle(npd, nh, box, (1, 1, 1), T=T, threads=8)
    for step in range(3):
        x_before = x.copy()
        X_start = X.copy()
        rep = A.dmd_step(fe, prm, X, x, g, fh)
        assert rep["lmvm_reason"] > 0
        # oracle at the relaxed state of this step, with the lists of the
        # step's start (the step builds its lists before relaxing, as
        # findLocalNeighbors precedes minimizeFreeEntropy in the reference)
        orc.build_neighbors(X_start)
        go, fo = orc.chem_potential(X, x_before, np.zeros(4), np.zeros(4), 1e300, np.zeros(nh), np.zeros(nh, np.int32))
        assert np.max(np.abs(g - go) / np.maximum(np.abs(go), 1.0)) <= 1e-10
        assert np.array_equal(fh, fo)
        # the diffusion step from the GPU's own gamma (the gamma bar above
        # allows ~1e-8 absolute, which the flux would carry into x)
        xo = orc.diffuse(1e-13, 1e13, 0.23, 1e-3, 1 - 1e-16, g, fh, x_before)
        assert np.max(np.abs(x - xo)) <= 1e-13
        assert rep["total_h_before"] == pytest.approx(x_before.sum(), rel=1e-15)
        # balance of the step: the net flux is zero to round-off of its scale,
        # and the change of total H is the clamp lift (the oracle's step with
        # the clamp opened gives the unclamped occupancy)
        assert abs(rep["flux_sum"]) <= 1e-12 * rep["flux_gross"] and rep["flux_gross"] > 0
        xu = orc.diffuse(1e-13, 1e13, 0.23, -1e300, 1e300, g, fh, x_before)
        assert rep["clamp_lift"] == pytest.approx((xo - xu).sum(), rel=1e-9, abs=1e-14)
        dH = rep["total_h_after"] - rep["total_h_before"]
        assert abs(dH - rep["clamp_lift"]) <= 1e-12 * rep["total_h_before"]
        clamped = np.any(x <= 1e-3) and np.any(x_before > 1e-3)
        if not clamped:
            assert abs(rep["total_h_after"] - rep["total_h_before"]) <= 1e-12 * rep["total_h_before"]
    # H moved out of the slab
    ix = (np.arange(nh) // 4) % 8
    assert x[ix >= 2].sum() > (ix >= 2).sum() * 1e-3


def test_host_program_runs_a_deck(tmp_path):
    """The C++ host program: deck -> specimen -> DMD steps -> restart files
    that the reader (readresults.cpp format) accepts."""
    import os
    import subprocess
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    deck = tmp_path / "slab.st"
    deck.write_text("""
    under Specimen { Cells_x = 8; Cells_y = 3; Cells_z = 3; Seed = 5; Displacement = 0.02;
                     HFraction = 0.001; SlabCells = 2; SlabHFraction = 0.6; }
    under InteratomicPotential { CutOffDistance = 5.35; Tolerance = 1.0e-5; }
    under Diffusion { Temperature = 400; AttemptFrequency = 1.0e13; ActivationEnergy = 0.23; }
    under Time { TimeStep = 1.0e-13; MaxIts = 2; }
    under Output { Prefix = "%s/"; Solution = "slab"; Frequency = 1; }
    under Solver { Precision = Mixed; GradientTolerance = 1.0e-8; }
    """ % tmp_path)
    r = subprocess.run([os.path.join(root, "a2c_amd", "a2c_dmd"), str(deck)], capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 0, r.stderr
    lines = [l for l in r.stdout.splitlines() if l.startswith("step")]
    assert len(lines) == 2 and all("reason 1" in l for l in lines)
    npd = nh = 4 * 8 * 3 * 3
    X, x, g, fh = A.readResults(str(tmp_path / "slab.2.dat"), npd, nh)
    assert X.size == 4 * (npd + nh) and np.all(x >= 1e-3)


def test_host_program_runs_the_reference_objects(tmp_path):
    """The same specimen from the reference's Lattice / Material / Species /
    RegionalInitialCondition objects and from the generated block gives the
    same DMD steps, bit for bit; the object path writes the reference's
    VTP / PVD / XYZ snapshots (frame 0 before the loop)."""
    import os
    import subprocess
    import xml.etree.ElementTree as ET
    from test_specimen import fcc_deck
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    exe = os.path.join(root, "a2c_amd", "a2c_dmd")
    common = """
    under InteratomicPotential { CutOffDistance = 5.35; Tolerance = 1.0e-5; }
    under Diffusion { Temperature = 400; AttemptFrequency = 1.0e13; ActivationEnergy = 0.23;
                      SoluteChemicalPotential = 0.5; }
    under Time { TimeStep = 1.0e-13; MaxIts = 2; }
    under Solver { Precision = Mixed; GradientTolerance = 1.0e-8; }
    """
    (tmp_path / "out").mkdir()
    objects = tmp_path / "objects.st"
    objects.write_text(fcc_deck(4, fmt="VTP_and_XYZ", precision="Medium").replace('Prefix = "out/"',
                       'Prefix = "%s/out/"' % tmp_path) + common)
    block = tmp_path / "block.st"
    block.write_text("""
    under Specimen { Cells = 4; Displacement = 0.0; Sigma = 0.10; HFraction = 0.1; SlabCells = 2;
                     SlabHFraction = 0.6; Periodic = True; }
    Diffusion.SubsurfaceLayers = 2;
    """ + common)
    print("host program: 4^3 cells from objects and from the block", flush=True)
    runs = []
    for deck in (objects, block):
        r = subprocess.run([exe, str(deck)], capture_output=True, text=True, timeout=120)
        assert r.returncode == 0, r.stderr
        runs.append([l for l in r.stdout.splitlines() if l.startswith(("a2c_dmd", "step"))])
    assert runs[0] == runs[1] and len(runs[0]) == 3
    assert "(128 reservoir)" in runs[0][0]
    sets = ET.parse(tmp_path / "out" / "spec_L0.pvd").getroot().findall("./Collection/DataSet")
    assert [float(s.get("timestep")) for s in sets] == [0.0, 1e-13, 2e-13]
    piece = ET.parse(tmp_path / "out" / "spec_L0_0002.vtp").getroot().find("./PolyData/Piece")
    assert int(piece.get("NumberOfPoints")) == 8 * 4 ** 3
    assert (tmp_path / "out" / "spec_L0_0002.xyz").exists()
