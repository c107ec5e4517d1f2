"""Host helpers of the DMD step: boundary conditions (boundaryconditions.cpp:
30-55), macro results (calculateresults.cpp:8-78), restart reader
(readresults.cpp:10-55) and the matching writer."""
import numpy as np
import pytest

import a2c_amd as A


def test_enforce_boundary_conditions():
    x = np.full(10, 0.3)
    g = np.zeros(10)
    fh = np.zeros(10, np.int32)
    A.enforceBoundaryConditions([1, 7], -2.0, 1 - 1e-16, x, g, fh)
    assert list(np.nonzero(fh)[0]) == [1, 7]
    assert g[1] == g[7] == -2.0 and x[1] == x[7] == 1 - 1e-16
    assert np.all(x[[0, 2, 3, 4, 5, 6, 8, 9]] == 0.3)


def test_macro_results():
    x = np.arange(10, dtype=float) / 10
    a, b = A.calculateMacroResults(x, [0, 9])
    assert a == pytest.approx(x.mean(), rel=1e-15)
    assert b == pytest.approx(x[1:9].sum() / 8, rel=1e-15)


def test_restart_round_trip(tmp_path):
    X, x, box, npd, nh = A.lattice(2, seed=3, disp_std=0.05, sigma=(0.08, 0.15), x=(0.0, 0.6), x_clip=(0, 1))
    rng = np.random.default_rng(0)
    g = rng.normal(size=nh)
    fh = (rng.random(nh) < 0.2).astype(np.int32)
    x[0], x[1] = 0.0, 1.0  # clamped on read, as the reference does
    p = str(tmp_path / "run.7.dat")
    A.writeResults(p, npd, nh, X, x, g, fh, time=1e-12, step=7)
    X2, x2, g2, fh2 = A.readResults(p, npd, nh, xLb=1e-3, xUb=1 - 1e-16)
    assert np.array_equal(X2, X) and np.array_equal(g2, g) and np.array_equal(fh2, fh)
    assert x2[0] == 1e-3 and x2[1] == 1 - 1e-16 and np.array_equal(x2[2:], x[2:])
    lines = open(p).read().splitlines()
    assert len(lines) == 9 + npd + nh and all(len(l.split()) == 9 for l in lines[9:])


def test_missing_restart_file_is_an_error(tmp_path):
    try:
        A.readResults(str(tmp_path / "nope.0.dat"), 4, 4)
    except A.A2CError as e:
        assert e.code == A.A2C_EINVAL
    else:
        raise AssertionError("expected an error")


def test_xyz_snapshot_layout(tmp_path):
    X, x, box, npd, nh = A.lattice(2, seed=3, disp_std=0.05)
    X0 = X.copy()
    X0[:3] -= 0.01
    g = np.linspace(-1, 1, nh)
    p = str(tmp_path / "s_L0_0001.xyz")
    A.writeXYZ(p, npd, nh, X, X0, x, g, box, (1, 1, 0), time=2e-13)
    lines = open(p).read().splitlines()
    assert int(lines[0]) == npd + nh and len(lines) == 2 + npd + nh
    assert 'pbc="T T F"' in lines[1] and "Properties=\"species:S:1:pos:R:3:disp:R:3" in lines[1]
    first = lines[2].split()
    assert first[0] == "Pd" and abs(float(first[4]) - 0.01) < 1e-12 and len(first) == 12
    last = lines[-1].split()
    assert last[0] == "H" and float(last[-1]) == pytest.approx(1.0)
