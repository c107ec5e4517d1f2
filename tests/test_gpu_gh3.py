"""GPU parity of the 3-point Gauss-Hermite product rule (27 nodes on the
relative coordinate) against the oracle's restatement of it, which the CPU
suite pins by the rule's exact quartic moment and by finite differences.

Bars are those of the 12-point rule: FP64 objective and densities to 1e-10
relative, gradient to 1e-8 of its largest component; mixed objective to 1e-7
relative, gradient to 1e-5 in the 2-norm and 5e-5 of the largest component.
A batch may hold lattices of both rules; each gives the same bits alone and
in the mix."""
import numpy as np
import pytest

import a2c_amd as A
import oracle as O
from helpers import specimen

pytestmark = pytest.mark.gpu

FP64_REL = 1e-10
FP64_GRAD = 1e-8
MIXED_OBJ_TOL = 1e-7
MIXED_GRAD_TOL = 1e-5
MIXED_GRAD_MAX = 5e-5

CASES = [  # (cells, periodic, displacement, sigma, x, T, seed)
    (3, True, 0.02, (0.10, 0.10), (0.10, 0.10), 300.0, 1),
    (4, True, 0.15, (0.08, 0.15), (0.0, 0.6), 500.0, 2),
    (3, False, 0.05, (0.08, 0.15), (0.01, 0.6), 400.0, 3),
]


def pair(case, deriv="analytic", precision="fp64"):
    n, per, disp, sig, xr, T, seed = case
    X, x, box, per3, npd, nh = specimen(n, seed=seed, disp=disp, sigma=sig, x=xr, periodic=per)
    fe = A.FreeEntropy(npd, nh, box, per3, T=T, deriv=deriv, precision=precision, quad="gh3")
    fe.set_occupancy(x)
    fe.findNeighbors(X)
    orc = O.Oracle(npd, nh, box, per3, T=T, deriv=deriv, threads=8, quad="gh3")
    orc.build_neighbors(X)
    return fe, orc, X, x


@pytest.mark.parametrize("case", CASES)
def test_density_fp64_gh3(case):
    fe, orc, X, x = pair(case)
    rg = fe.density(X)
    ro = orc.density(X, x)
    assert np.max(np.abs(rg - ro) / np.abs(ro)) <= FP64_REL


@pytest.mark.parametrize("case", CASES)
def test_objective_gradient_fp64_gh3(case):
    fe, orc, X, x = pair(case)
    fg, Gg = fe.formFunctionAndGradient(X)
    fo, Go = orc.eval(X, x)
    assert abs(fg - fo) <= FP64_REL * abs(fo)
    assert np.max(np.abs(Gg - Go)) <= FP64_GRAD * np.max(np.abs(Go))
    # the rule matters: the 12-point objective differs well beyond the bar
    o12 = O.Oracle(fe.n_pd, fe.n_h, list(orc.p.box), list(orc.p.periodic), T=case[5], threads=8)
    o12.build_neighbors(X)
    f12, _ = o12.eval(X, x)
    assert abs(f12 - fo) > 1e3 * FP64_REL * abs(fo)


@pytest.mark.parametrize("case", CASES[:2])
def test_objective_gradient_fd_mode_gh3(case):
    fe, orc, X, x = pair(case, deriv="fd")
    fg, Gg = fe.formFunctionAndGradient(X)
    fo, Go = orc.eval(X, x)
    assert abs(fg - fo) <= FP64_REL * abs(fo)
    assert np.linalg.norm(Gg - Go) <= 1e-5 * np.linalg.norm(Go)


@pytest.mark.parametrize("case", CASES)
def test_objective_gradient_mixed_gh3(case):
    fe, orc, X, x = pair(case, precision="mixed")
    fg, Gg = fe.formFunctionAndGradient(X)
    fo, Go = orc.eval(X, x)
    assert abs(fg - fo) <= MIXED_OBJ_TOL * abs(fo)
    assert np.linalg.norm(Gg - Go) <= MIXED_GRAD_TOL * np.linalg.norm(Go)
    assert np.max(np.abs(Gg - Go)) <= MIXED_GRAD_MAX * np.max(np.abs(Go))
    fg2, Gg2 = fe.formFunctionAndGradient(X)
    assert fg2 == fg and np.array_equal(Gg2, Gg)  # deterministic


def test_mixed_matches_fp64_kernel_gh3_at_12_cubed():
    """13,824 sites: the FP64 kernel against the oracle, mixed against both."""
    X, x, box, per3, npd, nh = specimen(12, seed=2, disp=0.03, sigma=(0.08, 0.15), x=(0.0, 0.6))
    print("gh3: 12^3 cells, FP64 and mixed against the oracle", flush=True)
    fe = A.FreeEntropy(npd, nh, box, per3, T=500.0, quad="gh3")
    fe.set_occupancy(x)
    fe.findNeighbors(X)
    f64, G64 = fe.eval(X)
    orc = O.Oracle(npd, nh, box, per3, T=500.0, threads=8, quad="gh3")
    orc.build_neighbors(X)
    fo, Go = orc.eval(X, x)
    assert abs(f64 - fo) <= FP64_REL * abs(fo)
    assert np.max(np.abs(G64 - Go)) <= FP64_GRAD * np.abs(Go).max()
    fe.set_precision("mixed")
    f32, G32 = fe.eval(X)
    assert abs(f32 - f64) <= MIXED_OBJ_TOL * abs(f64)
    assert abs(f32 - fo) <= MIXED_OBJ_TOL * abs(fo)
    assert np.linalg.norm(G32 - G64) <= MIXED_GRAD_TOL * np.linalg.norm(G64)


@pytest.mark.parametrize("deriv", ["analytic", "fd"])
def test_chemical_potential_gh3(deriv):
    X, x, box, per, npd, nh = specimen(4, seed=5, disp=0.03, sigma=(0.08, 0.15), x=(0.01, 0.6))
    fe = A.FreeEntropy(npd, nh, box, per, T=400.0, deriv=deriv, quad="gh3")
    fe.set_occupancy(x)
    fe.findNeighbors(X)
    orc = O.Oracle(npd, nh, box, per, T=400.0, deriv=deriv, threads=8, quad="gh3")
    orc.build_neighbors(X)
    cp, ch = np.array([0.3, 0.0, 0.2, -0.1]), np.array([0.5, 0.07, -0.03, 0.02])
    g0 = np.full(nh, 3.0)
    f0 = np.zeros(nh, np.int32)
    f0[::17] = 1
    gg, fg = fe.calculateChemicalPotential(X, cp, ch, 1e9, g0.copy(), f0.copy())
    go, fo = orc.chem_potential(X, x, cp, ch, 1e9, g0, f0)
    assert np.array_equal(fg, fo)
    tol = 1e-10 if deriv == "analytic" else 1e-8
    assert np.max(np.abs(gg - go) / np.maximum(np.abs(go), 1.0)) <= tol


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_batch_with_both_rules_matches_single_lattices(precision):
    specs = []
    for i, (n, seed, quad) in enumerate([(3, 1, "gh3"), (4, 2, "r12"), (3, 3, "r12"), (4, 4, "gh3")]):
        X, x, box, per3, npd, nh = specimen(n, seed=seed, disp=0.03, sigma=(0.08, 0.15), x=(0.01, 0.6))
        specs.append(dict(X=X, x=x, box=box, periodic=per3, n_pd=npd, n_h=nh, T=300.0 + 50 * i, quad=quad))
    with A.Batch(precision=precision, max_lattices=8) as B:
        slots = [B.add(**s) for s in specs]
        f = B.eval(slots, precision=precision)
        for sl, s, fb in zip(slots, specs, f):
            with A.FreeEntropy(s["n_pd"], s["n_h"], s["box"], s["periodic"], T=s["T"], precision=precision,
                               quad=s["quad"]) as fe:
                fe.set_occupancy(s["x"])
                fe.findNeighbors(s["X"])
                f1, G1 = fe.eval(s["X"])
            _, Gb = B.get_state(sl)
            assert fb == f1 and np.array_equal(Gb, G1), s["quad"]


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_relaxation_gh3_is_stationary_for_the_oracle(precision):
    """The plain 6^3-cell lattice (1,728 sites, x = 0.10, sigma = 0.10 A,
    displacements N(0, 0.02 A), T = 300 K) under the Gauss-Hermite rule: one
    evaluation, then a relaxation whose end state the oracle sees as
    stationary."""
    X, x, box, per, npd, nh = specimen(6, seed=1, disp=0.02, sigma=(0.10, 0.10), x=(0.10, 0.10))
    fe = A.FreeEntropy(npd, nh, box, per, T=300.0, precision=precision, quad="gh3")
    fe.set_occupancy(x)
    fe.findNeighbors(X)
    Xr, rep = fe.minimizeFreeEntropy(X, gatol=1e-9, grtol=0.0)
    assert rep["reason"] > 0, rep
    orc = O.Oracle(npd, nh, box, per, T=300.0, threads=8, quad="gh3")
    orc.build_neighbors(X)
    f0, _ = orc.eval(X, x)
    fo, Go = orc.eval(Xr, x)
    assert fo < f0
    assert abs(rep["f"] - fo) <= FP64_REL * abs(fo)
    assert np.linalg.norm(Go) <= 1e-8, np.linalg.norm(Go)
