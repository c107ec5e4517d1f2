"""GPU parity of the HIP path against the FP64 oracle, through the C-ABI.

Bars: neighbor sets bit-exact.  FP64: densities and objective to 1e-10
relative, gradient to 1e-8 of its largest component.  These are the
conditioning of the reference's own product polynomials: f_Pd sums terms of
size ~1e4 to values of 1e-3..1, so two FP64 evaluation orders (FMA or not,
association of the pair vector) differ by up to ~1e-11 in rho and ~1e-9 in
f_Pd' near r = 5 A.  FD-derivative mode: to the FD round-off of that form
(1e-5 of ||G||, see the oracle tests)."""
FP64_REL = 1e-10
FP64_GRAD = 1e-8
import numpy as np
import pytest

import a2c_amd as A
import oracle as O
from helpers import csr_rows, specimen

pytestmark = pytest.mark.gpu

CASES = [  # (cells, periodic, displacement, sigma, x, T, seed)
    (3, True, 0.02, (0.10, 0.10), (0.10, 0.10), 300.0, 1),
    (4, True, 0.15, (0.08, 0.15), (0.0, 0.6), 500.0, 2),
    (3, False, 0.05, (0.08, 0.15), (0.01, 0.6), 400.0, 3),
    (6, True, 0.02, (0.10, 0.10), (0.10, 0.10), 300.0, 1),  # the plain 6^3 workload
]


def pair(case, deriv="analytic", precision="fp64"):
    n, per, disp, sig, xr, T, seed = case
    X, x, box, per3, npd, nh = specimen(n, seed=seed, disp=disp, sigma=sig, x=xr, periodic=per)
    fe = A.FreeEntropy(npd, nh, box, per3, T=T, deriv=deriv, precision=precision)
    fe.set_occupancy(x)
    st = fe.findNeighbors(X)
    orc = O.Oracle(npd, nh, box, per3, T=T, deriv=deriv, threads=8)
    mx = orc.build_neighbors(X)
    return fe, orc, X, x, st, mx


@pytest.mark.parametrize("case", CASES)
def test_neighbor_sets_bit_exact(case):
    fe, orc, X, x, st, mx = pair(case)
    assert st["max_neib"] == mx
    for k in ["PdPd", "PdH", "HPd", "HH", "HH1"]:
        go, gi = fe.neighbors(k)
        oo, oi = orc.neighbors(k)
        assert np.array_equal(go, oo), k
        assert np.array_equal(gi, oi), k
    assert np.array_equal(fe.pdpdnum(), orc.pdpdnum())


@pytest.mark.parametrize("case", CASES)
def test_density_fp64(case):
    fe, orc, X, x, *_ = pair(case)
    rg = fe.density(X)
    ro = orc.density(X, x)
    assert np.max(np.abs(rg - ro) / np.abs(ro)) <= FP64_REL


@pytest.mark.parametrize("case", CASES)
def test_objective_gradient_fp64(case):
    fe, orc, X, x, *_ = pair(case)
    fg, Gg = fe.formFunctionAndGradient(X)
    fo, Go = orc.eval(X, x)
    assert abs(fg - fo) <= FP64_REL * abs(fo)
    assert np.max(np.abs(Gg - Go)) <= FP64_GRAD * np.max(np.abs(Go))


@pytest.mark.parametrize("case", CASES[:2])
def test_objective_gradient_fd_mode(case):
    fe, orc, X, x, *_ = pair(case, deriv="fd")
    fg, Gg = fe.formFunctionAndGradient(X)
    fo, Go = orc.eval(X, x)
    assert abs(fg - fo) <= FP64_REL * abs(fo)
    assert np.linalg.norm(Gg - Go) <= 1e-5 * np.linalg.norm(Go)


def test_eval_is_deterministic():
    fe, orc, X, x, *_ = pair(CASES[1])
    f1, G1 = fe.eval(X)
    f2, G2 = fe.eval(X)
    assert f1 == f2 and np.array_equal(G1, G2)


def test_tao_adapter_and_device_entry_point():
    fe, orc, X, x, *_ = pair(CASES[0])
    f, G = fe.eval(X)
    import ctypes as C
    f2 = C.c_double()
    G2 = np.zeros_like(G)
    rc = A.lib.a2c_tao_objgrad(None, A._dptr(X), C.byref(f2), A._dptr(G2), fe._ctx)
    assert rc == 0 and f2.value == f and np.array_equal(G2, G)
    dX, dG, df = fe.dev_alloc(X.nbytes), fe.dev_alloc(X.nbytes), fe.dev_alloc(8)
    try:
        fe.upload(dX, X)
        fe.eval_device(dX, dG, df)
        fe.sync()
        G3 = fe.download(np.zeros_like(G), dG)
        f3 = fe.download(np.zeros(1), df)[0]
        assert f3 == f and np.array_equal(G3, G)
    finally:
        for p in (dX, dG, df):
            fe.dev_free(p)


def test_no_hydrogen_occupancy_and_pd_only():
    n, per, disp, sig, xr, T, seed = CASES[0]
    X, x, box, per3, npd, nh = specimen(n, seed=seed, disp=disp)
    x0 = np.zeros(nh)
    fe = A.FreeEntropy(npd, nh, box, per3)
    fe.set_occupancy(x0)
    fe.findNeighbors(X)
    orc = O.Oracle(npd, nh, box, per3)
    orc.build_neighbors(X)
    fg, Gg = fe.eval(X)
    fo, Go = orc.eval(X, x0)
    assert abs(fg - fo) <= FP64_REL * abs(fo)
    assert np.max(np.abs(Gg - Go)) <= FP64_GRAD * np.max(np.abs(Go))
    # Pd-only lattice (n_h = 0)
    Xp = X[: 4 * npd]
    fe2 = A.FreeEntropy(npd, 0, box, per3)
    fe2.set_occupancy(np.zeros(0))
    fe2.findNeighbors(Xp)
    orc2 = O.Oracle(npd, 0, box, per3)
    orc2.build_neighbors(Xp)
    f2, G2 = fe2.eval(Xp)
    fo2, Go2 = orc2.eval(Xp, np.zeros(0))
    assert abs(f2 - fo2) <= FP64_REL * abs(fo2)


def test_eval_before_neighbors_is_an_error():
    X, x, box, per3, npd, nh = specimen(3)
    fe = A.FreeEntropy(npd, nh, box, per3)
    fe.set_occupancy(x)
    with pytest.raises(A.A2CError) as e:
        fe.eval(X)
    assert e.value.code == A.A2C_ESTATE


def test_neighbor_overflow_is_an_error():
    X, x, box, per3, npd, nh = specimen(3)
    fe = A.FreeEntropy(npd, nh, box, per3, max_neighbors=40)
    fe.set_occupancy(x)
    with pytest.raises(A.A2CError) as e:
        fe.findNeighbors(X)
    assert e.value.code == A.A2C_ENBR_OVERFLOW


# ---------------------------------------------------------------- mixed path
# Bar: objective within 1e-7 relative of the FP64 oracle (the stated
# tolerance); gradient within 1e-5 relative in the 2-norm and 5e-5 of the
# largest component.  In near-perfect lattices a site's gradient is the small
# residual of pair terms ~30x larger, so FP32 rounding of the pair terms shows
# up at ~2e-5 of the largest component.
MIXED_OBJ_TOL = 1e-7
MIXED_GRAD_TOL = 1e-5
MIXED_GRAD_MAX = 5e-5


@pytest.mark.parametrize("case", CASES)
def test_objective_gradient_mixed(case):
    fe, orc, X, x, *_ = pair(case, precision="mixed")
    fg, Gg = fe.formFunctionAndGradient(X)
    fo, Go = orc.eval(X, x)
    assert abs(fg - fo) <= MIXED_OBJ_TOL * abs(fo)
    assert np.linalg.norm(Gg - Go) <= MIXED_GRAD_TOL * np.linalg.norm(Go)
    assert np.max(np.abs(Gg - Go)) <= MIXED_GRAD_MAX * np.max(np.abs(Go))


def test_mixed_matches_fp64_kernel_at_12_cubed():
    """13,824 sites: mixed vs the FP64 kernel (itself pinned to the oracle)."""
    case = (12, True, 0.03, (0.08, 0.15), (0.0, 0.6), 500.0, 2)
    n, per, disp, sig, xr, T, seed = case
    X, x, box, per3, npd, nh = specimen(n, seed=seed, disp=disp, sigma=sig, x=xr)
    fe = A.FreeEntropy(npd, nh, box, per3, T=T)
    fe.set_occupancy(x)
    fe.findNeighbors(X)
    f64, G64 = fe.eval(X)
    fe.set_precision("mixed")
    f32, G32 = fe.eval(X)
    assert abs(f32 - f64) <= MIXED_OBJ_TOL * abs(f64)
    assert np.linalg.norm(G32 - G64) <= MIXED_GRAD_TOL * np.linalg.norm(G64)
    f32b, G32b = fe.eval(X)
    assert f32b == f32 and np.array_equal(G32b, G32)  # deterministic


def test_mixed_path_below_its_tables_falls_back_to_fp64():
    """A pair closer than 1 A (outside the FP32 tables) makes the mixed
    objective non-finite instead of silently clamping; the relaxation then
    runs in FP64 from its first evaluation and converges."""
    X, x, box, per3, npd, nh = specimen(3)
    X = X.copy()
    X[4 * npd: 4 * npd + 3] = X[0:3] + np.array([0.9, 0.0, 0.0])  # H site 0 at 0.9 A from Pd site 0
    fe = A.FreeEntropy(npd, nh, box, per3, precision="mixed")
    fe.set_occupancy(x)
    fe.findNeighbors(X)
    with pytest.raises(A.A2CError) as e:
        fe.eval(X)
    assert e.value.code == A.A2C_ENONFINITE
    fe64 = A.FreeEntropy(npd, nh, box, per3, precision="fp64")
    fe64.set_occupancy(x)
    fe64.findNeighbors(X)
    f64, _ = fe64.eval(X)
    assert np.isfinite(f64)
    _, rep = fe.minimizeFreeEntropy(X, gatol=1e-6, grtol=0.0, max_it=400)
    assert rep["reason"] != -4  # not stopped on the non-finite mixed value
    assert rep["evals"] - rep["evals_fp64"] <= 1  # only the first evaluation was mixed


def test_contexts_and_batches_share_the_tables_and_stage_lazily():
    """The EAM tables live once per device: destroying a context or a batch
    leaves every other context's results bitwise unchanged.  A context first
    driven through device arrays allocates its pinned host staging at its
    first host-array call, with the same result."""
    n, per, disp, sig, xr, T, seed = CASES[1]
    X, x, box, per3, npd, nh = specimen(n, seed=seed, disp=disp, sigma=sig, x=xr, periodic=per)
    print("shared tables: contexts and a batch created and destroyed around one context", flush=True)

    def ctx(prec):
        fe = A.FreeEntropy(npd, nh, box, per3, T=T, precision=prec)
        fe.set_occupancy(x)
        fe.findNeighbors(X)
        return fe

    keep = {p: ctx(p) for p in ("mixed", "fp64")}
    ref = {p: fe.eval(X) for p, fe in keep.items()}
    for p in ("mixed", "fp64"):
        ctx(p).close()
    with A.Batch(precision="mixed", max_lattices=2) as B:
        sl = B.add(X, x, npd, nh, box, per3, T=T)
        fb = B.eval([sl])[0]
    assert abs(fb - ref["mixed"][0]) <= 1e-7 * abs(ref["mixed"][0])
    for p, fe in keep.items():
        f, G = fe.eval(X)
        assert f == ref[p][0] and np.array_equal(G, ref[p][1]), p
        fe.close()
    # device arrays first, host arrays afterwards
    fe = A.FreeEntropy(npd, nh, box, per3, T=T, precision="mixed")
    fe.set_occupancy(x)
    dX, dG, df = fe.dev_alloc(X.nbytes), fe.dev_alloc(X.nbytes), fe.dev_alloc(8)
    try:
        fe.upload(dX, X)
        fe.findNeighbors_device(dX)
        fe.eval_device(dX, dG, df)
        Gd = fe.download(np.zeros(X.size), dG)
        fd = fe.download(np.zeros(1), df)[0]
    finally:
        for p in (dX, dG, df):
            fe.dev_free(p)
    f, G = fe.eval(X)
    fe.close()
    assert f == fd and np.array_equal(G, Gd)
    assert f == ref["mixed"][0] and np.array_equal(G, ref["mixed"][1])
