"""GPU parity of the general model -- a tabulated potential (a2c_fe_set_potential_*,
here a setfl file of the Pd-H functions sampled on fine grids) and node
tables (a2c_fe_set_quadrature) -- against the oracle's restatement, which the
CPU suite pins (tests/test_oracle_general.py).

Bars as for the compiled-in model: FP64 objective and densities to 1e-10
relative, gradient to 1e-8 of its largest component; mixed objective to 1e-6
relative and gradient to 1e-4 in the 2-norm against the FP64 general path
(FP32 pieces of the r-functions); chemical potential to 1e-10."""
import numpy as np
import pytest

import a2c_amd as A
import oracle as O
from helpers import product_rule_2pt, sampled_tables, specimen, write_setfl

pytestmark = pytest.mark.gpu

FP64_REL = 1e-10
FP64_GRAD = 1e-8

CASES = [  # (cells, periodic, displacement, sigma, x, T, seed)
    (3, True, 0.02, (0.10, 0.10), (0.10, 0.10), 300.0, 1),
    (4, True, 0.15, (0.08, 0.15), (0.0, 0.6), 500.0, 2),
    (3, False, 0.05, (0.08, 0.15), (0.01, 0.6), 400.0, 3),
]


@pytest.fixture(scope="module")
def setfl(tmp_path_factory):
    t = sampled_tables()
    p = str(tmp_path_factory.mktemp("pot") / "PdH_sampled.eam.alloy")
    write_setfl(p, ["Pd", "H"], {"Pd": t["F_pd"], "H": t["F_h"]}, {"Pd": t["f_pd"], "H": t["f_h"]},
                {("Pd", "Pd"): t["rphi_pdpd"], ("H", "Pd"): t["rphi_pdh"], ("H", "H"): t["rphi_hh"]},
                t["n_rho"], t["d_rho"], t["n_r"], t["d_r"], 8.0)
    return p, t


def pair(case, pot, quad, setfl_file=None, tables=None, deriv="analytic", precision="fp64", rule="r12"):
    """Engine context and oracle on the same specimen and model."""
    n, per, disp, sig, xr, T, seed = case
    X, x, box, per3, npd, nh = specimen(n, seed=seed, disp=disp, sigma=sig, x=xr, periodic=per)
    fe = A.FreeEntropy(npd, nh, box, per3, T=T, deriv=deriv, precision=precision, quad=rule)
    if pot == "tab":
        fe.set_potential_file(setfl_file, "Pd", "H")
        O.set_tabulated(tables)
    if quad is not None:
        fe.set_quadrature(*quad)
        O.set_quad_table(*quad)
    fe.set_occupancy(x)
    fe.findNeighbors(X)
    orc = O.Oracle(npd, nh, box, per3, T=T, deriv=deriv, threads=8, potential="tab" if pot == "tab" else "eam",
                   quad="table" if quad is not None else rule)
    orc.build_neighbors(X)
    return fe, orc, X, x


def check64(fe, orc, X, x):
    f, G = fe.eval(X)
    fo, Go = orc.eval(X, x)
    assert abs(f - fo) <= FP64_REL * abs(fo)
    assert np.max(np.abs(G - Go)) <= FP64_GRAD * np.abs(Go).max()
    rho = fe.density(X)
    ro = orc.density(X, x)
    assert np.max(np.abs(rho - ro) / np.abs(ro)) <= FP64_REL
    return f, G


@pytest.mark.parametrize("case", CASES)
def test_tabulated_potential_fp64(setfl, case):
    path, t = setfl
    print("general: setfl potential, 12-point rule,", case, flush=True)
    fe, orc, X, x = pair(case, "tab", None, path, t)
    with fe:
        check64(fe, orc, X, x)


@pytest.mark.parametrize("case", CASES[:2])
def test_node_table_with_the_compiled_in_potential(case):
    quad = product_rule_2pt()
    print("general: closed forms, 64-node table,", case, flush=True)
    fe, orc, X, x = pair(case, "eam", quad)
    with fe:
        check64(fe, orc, X, x)


def test_tabulated_potential_with_a_node_table_and_the_gauss_hermite_rule(setfl):
    path, t = setfl
    case = CASES[1]
    print("general: setfl + 64-node table; setfl + Gauss-Hermite", flush=True)
    fe, orc, X, x = pair(case, "tab", product_rule_2pt(), path, t)
    with fe:
        check64(fe, orc, X, x)
    fe, orc, X, x = pair(case, "tab", None, path, t, rule="gh3")
    with fe:
        check64(fe, orc, X, x)


def test_reference_nodes_as_a_table_match_the_12_point_kernels():
    """The general kernels with the reference's 12 nodes as a table against
    the pair-once 12-point kernels of the same context."""
    case = CASES[1]
    print("general: the 12 reference nodes as a table", flush=True)
    n, per, disp, sig, xr, T, seed = case
    X, x, box, per3, npd, nh = specimen(n, seed=seed, disp=disp, sigma=sig, x=xr, periodic=per)
    with A.FreeEntropy(npd, nh, box, per3, T=T) as fe:
        fe.set_occupancy(x)
        fe.findNeighbors(X)
        f0, G0 = fe.eval(X)
        fe.set_quadrature(*O.quadrature())
        f1, G1 = fe.eval(X)
        fe.set_quadrature()  # back to the descriptor's rule and the pair-once kernels
        f2, G2 = fe.eval(X)
    assert abs(f1 - f0) <= FP64_REL * abs(f0) and np.max(np.abs(G1 - G0)) <= FP64_GRAD * np.abs(G0).max()
    assert f2 == f0 and np.array_equal(G2, G0)


def test_fd_mode_and_chemical_potential(setfl):
    path, t = setfl
    case = CASES[0]
    print("general: setfl, FD derivatives, chemical potential", flush=True)
    fe, orc, X, x = pair(case, "tab", None, path, t, deriv="fd")
    with fe:
        f, G = fe.eval(X)
        fo, Go = orc.eval(X, x)
        assert abs(f - fo) <= FP64_REL * abs(fo)
        assert np.max(np.abs(G - Go)) <= 1e-7 * np.abs(Go).max()
    fe, orc, X, x = pair(case, "tab", None, path, t)
    with fe:
        nh = fe.n_h
        coe = (np.array([0.3, 0.0, 0.2, -0.1]), np.array([0.5, 0.07, -0.03, 0.02]))
        f0 = np.zeros(nh, np.int32)
        f0[::13] = 1
        g0 = np.full(nh, 2.0)
        gg, fg = fe.calculateChemicalPotential(X, coe[0], coe[1], 1e9, g0.copy(), f0.copy())
        go, fo = orc.chem_potential(X, x, coe[0], coe[1], 1e9, g0, f0)
        assert np.array_equal(fg, fo)
        assert np.max(np.abs(gg - go) / np.maximum(np.abs(go), 1.0)) <= 1e-10


def test_mixed_precision_against_fp64(setfl):
    path, t = setfl
    print("general: mixed vs FP64 (setfl; closed forms with a table)", flush=True)
    for pot, quad in (("tab", None), ("eam", product_rule_2pt())):
        res = []
        for prec in ("fp64", "mixed"):
            fe, orc, X, x = pair(CASES[1], pot, quad, path, t, precision=prec)
            with fe:
                res.append(fe.eval(X))
        (f64, G64), (f32, G32) = res
        assert abs(f32 - f64) <= 1e-6 * abs(f64)
        assert np.linalg.norm(G32 - G64) <= 1e-4 * np.linalg.norm(G64)


def test_relaxation_with_a_tabulated_potential(setfl):
    """Mixed-then-FP64 LMVM on the general kernels: converges, and its final
    objective is the oracle's at the returned state."""
    path, t = setfl
    case = (3, True, 0.05, (0.10, 0.10), (0.1, 0.1), 400.0, 4)
    print("general: relaxation, setfl potential", flush=True)
    fe, orc, X, x = pair(case, "tab", None, path, t, precision="mixed")
    with fe:
        Xr, rep = fe.minimizeFreeEntropy(X, gatol=1e-8, grtol=0.0)
    assert rep["reason"] > 0, rep
    fo, Go = orc.eval(Xr, x)
    assert abs(rep["f"] - fo) <= FP64_REL * abs(fo)
    assert np.linalg.norm(Go) <= 1.5e-8


def test_bad_tables_are_refused(setfl):
    path, t = setfl
    case = CASES[0]
    n, per, disp, sig, xr, T, seed = case
    X, x, box, per3, npd, nh = specimen(n, seed=seed, disp=disp, sigma=sig, x=xr, periodic=per)
    with A.FreeEntropy(npd, nh, box, per3, T=T) as fe:
        xi, w = product_rule_2pt()
        w2 = w.copy()
        w2[1] *= 2.0  # breaks swap symmetry: node 1's partner (node 8) keeps the old weight
        with pytest.raises(A.A2CError):
            fe.set_quadrature(xi, w2)
        with pytest.raises(A.A2CError):
            fe.set_quadrature(np.zeros((600, 6)), np.ones(600))
        with pytest.raises(A.A2CError):
            fe.set_potential_file(path, "Pd", "He")
        bad = dict(t)
        bad["n_r"] = 3
        bad.update({k: t[k][:3] for k in ("f_pd", "f_h", "rphi_pdpd", "rphi_pdh", "rphi_hh")})
        with pytest.raises(A.A2CError):
            fe.set_potential_tables(bad)


def test_host_program_takes_the_potential_file(setfl, tmp_path):
    """InteratomicPotential.File in a deck: the host program relaxes with the
    tabulated potential; with the Pd-H functions sampled finely its step
    objectives are the compiled-in ones to the interpolation error."""
    import os
    import re
    import subprocess
    path, t = setfl
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    base = """
    under Specimen { Cells = 4; Seed = 5; Displacement = 0.02; HFraction = 0.1; }
    under Diffusion { Temperature = 400; AttemptFrequency = 1.0e13; ActivationEnergy = 0.23; }
    under Time { TimeStep = 1.0e-13; MaxIts = 1; }
    under Solver { Precision = FP64; GradientTolerance = 1.0e-8; }
    """
    out = []
    print("host program: compiled-in potential, then the setfl file", flush=True)
    for extra in ("", 'InteratomicPotential.File = "%s";' % path):
        deck = tmp_path / "d.st"
        deck.write_text(base + extra)
        r = subprocess.run([os.path.join(root, "a2c_amd", "a2c_dmd"), str(deck)], capture_output=True, text=True,
                           timeout=120)
        assert r.returncode == 0, r.stderr
        out.append(float(re.search(r"f = (\S+)", r.stdout).group(1)))
    assert abs(out[1] - out[0]) <= 1e-7 * abs(out[0])


def test_general_model_at_12_cubed(setfl):
    """12^3 cells (13,824 sites, displacements 0.03 A, widths 0.08-0.15 A):
    the setfl potential in FP64 and mixed precision and the 64-node table,
    against the oracle at the FP64 bars (mixed: 1e-7 objective)."""
    path, t = setfl
    case = (12, True, 0.03, (0.08, 0.15), (0.0, 0.6), 500.0, 12)
    print("general: 12^3 cells, setfl FP64 / mixed, 64-node table", flush=True)
    fe, orc, X, x = pair(case, "tab", None, path, t)
    with fe:
        f64, _ = check64(fe, orc, X, x)
        fe.set_precision("mixed")
        fm, Gm = fe.eval(X)
        fo, Go = orc.eval(X, x)
        assert abs(fm - fo) <= 1e-7 * abs(fo)
        assert np.linalg.norm(Gm - Go) <= 1e-4 * np.linalg.norm(Go)
    fe, orc, X, x = pair(case, "eam", product_rule_2pt())
    with fe:
        check64(fe, orc, X, x)
