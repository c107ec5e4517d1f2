"""The C++ mirror of the reference's interface (include/a2c_minimizer.hpp),
driven by a caller written against the reference's call shapes
(tests/cpp/minimizer_caller.cpp: findNeighbors, Minimizer(argc, argv),
initialize with the prebuilt lists and the quadrature QPp/QWp,
findLocalNeighbors with the candidate lists, minimizeFreeEntropy,
calculateChemicalPotential, integrateDiffusion).  Its results are those of
the library's own entry points and of the oracle (FD derivatives, as the
reference and the mirror use)."""
import json
import os
import subprocess

import numpy as np
import pytest

import a2c_amd as A
import oracle as O
from helpers import csr_rows

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.fixture(scope="module")
def caller(tmp_path_factory):
    exe = str(tmp_path_factory.mktemp("mirror") / "minimizer_caller")
    lib = os.path.join(ROOT, "a2c_amd")
    subprocess.run(["g++", "-O1", "-std=c++17", "-I" + os.path.join(ROOT, "include"),
                    os.path.join(ROOT, "tests", "cpp", "minimizer_caller.cpp"), "-L" + lib, "-la2c_fe",
                    "-Wl,-rpath," + lib, "-o", exe], check=True)
    return exe


def test_reference_call_shapes_give_the_library_results(caller):
    print("mirror caller: 3^3 cells", flush=True)
    r = subprocess.run([caller, "3"], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    out = json.loads(r.stdout)
    assert out["rc_bad_rule"] == -1  # a rule other than quadrature.cpp's is refused
    # the same specimen through the library's Python binding
    X, x, box, npd, nh = A.lattice(3, seed=21, disp_std=0.02, sigma=(0.10, 0.10), x=(0.05, 0.5))
    per = (1, 1, 1)
    orc = O.Oracle(npd, nh, box, per, T=400.0, deriv="fd", threads=8)
    mx = orc.build_neighbors(X)
    assert out["maxNeib0"] == mx and out["maxNeib"] == mx
    oo, oi = orc.neighbors("PdPd")
    assert [list(v) for v in csr_rows(oo, oi)] == out["PdPd"]
    oo, oi = orc.neighbors("HH1")
    assert [list(v) for v in csr_rows(oo, oi)] == out["HH1"]
    assert np.array_equal(orc.pdpdnum(), np.array(out["PdPdnum"], dtype=np.int32))
    with A.FreeEntropy(npd, nh, box, per, T=400.0, deriv="fd") as fe:
        fe.set_occupancy(x)
        fe.findNeighbors(X)
        Xr, rep = fe.minimizeFreeEntropy(X)
    # with the reference's central differences the gradient carries ~1e-6 of
    # its scale in round-off, so TAO's default tolerances end on max_funcs;
    # the mirror must stop where the library stops
    assert rep["reason"] != 0
    assert out["free_entropy"] == rep["f"]
    assert np.array_equal(np.array(out["X"]), Xr)
    # chemical potential and diffusion at the relaxed state against the oracle
    go, fo = orc.chem_potential(Xr, x, np.zeros(4), np.zeros(4), 1e300, np.zeros(nh), np.zeros(nh, np.int32))
    g = np.array(out["gamma"])
    assert np.max(np.abs(g - go) / np.maximum(np.abs(go), 1.0)) <= 1e-8
    xo = orc.diffuse(1e-13, 1e13, 0.23, 1e-3, 1 - 1e-16, g, np.zeros(nh, np.int32), x)
    assert np.max(np.abs(np.array(out["x"]) - xo)) <= 1e-14
