"""GPU parity of the chemical potential (minimizer.cpp:588-802) and the
diffusion step (integrate.cpp:9-87) against the FP64 oracle; conservation of
total H at a larger size."""
import numpy as np
import pytest

import a2c_amd as A
import oracle as O
from helpers import specimen

pytestmark = pytest.mark.gpu


def setup(n=4, seed=5, T=400.0, deriv="analytic"):
    X, x, box, per, npd, nh = specimen(n, seed=seed, disp=0.03, sigma=(0.08, 0.15), x=(0.01, 0.6))
    fe = A.FreeEntropy(npd, nh, box, per, T=T, deriv=deriv)
    fe.set_occupancy(x)
    fe.findNeighbors(X)
    orc = O.Oracle(npd, nh, box, per, T=T, deriv=deriv, threads=8)
    orc.build_neighbors(X)
    return fe, orc, X, x


@pytest.mark.parametrize("deriv", ["analytic", "fd"])
@pytest.mark.parametrize("coe", [(np.zeros(4), np.zeros(4)),
                                 (np.array([0.3, 0.0, 0.2, -0.1]), np.array([0.5, 0.07, -0.03, 0.02]))])
def test_chemical_potential_parity(deriv, coe):
    fe, orc, X, x = setup(deriv=deriv)
    nh = fe.n_h
    g0 = np.full(nh, 3.0)
    f0 = np.zeros(nh, np.int32)
    f0[::17] = 1  # reservoir sites keep gamma
    gg, fg = fe.calculateChemicalPotential(X, coe[0], coe[1], 1e9, g0.copy(), f0.copy())
    go, fo = orc.chem_potential(X, x, coe[0], coe[1], 1e9, g0, f0)
    assert np.array_equal(fg, fo)
    tol = 1e-10 if deriv == "analytic" else 1e-8
    assert np.max(np.abs(gg - go) / np.maximum(np.abs(go), 1.0)) <= tol


def test_chemical_potential_cap():
    fe, orc, X, x = setup()
    x = x.copy()
    x[:5] = 0.99995
    fe.set_occupancy(x)
    nh = fe.n_h
    gbd = np.median(orc.chem_potential(X, x, np.zeros(4), np.zeros(4), 1e9, np.zeros(nh), np.zeros(nh, np.int32))[0])
    gg, fg = fe.calculateChemicalPotential(X, np.zeros(4), np.zeros(4), gbd, np.zeros(nh), np.zeros(nh, np.int32))
    go, fo = orc.chem_potential(X, x, np.zeros(4), np.zeros(4), gbd, np.zeros(nh), np.zeros(nh, np.int32))
    assert np.array_equal(fg, fo) and fg.sum() > 0
    assert np.max(np.abs(gg - go)) <= 1e-10 * np.max(np.abs(go))


def test_chemical_potential_at_empty_and_full_sites():
    """x = 0 and x = 1 exactly: the reference skips ln(x / (1 - x)) and the
    pair sum there, so gamma = 1.5 (minimizer.cpp:785-787, :792); below that
    cap gamma_bd takes over and only the x = 1 sites are marked full
    (:793-799)."""
    fe, orc, X, x = setup()
    x = x.copy()
    x[0:7:2] = 0.0
    x[1:8:2] = 1.0
    fe.set_occupancy(x)
    nh = fe.n_h
    for gbd in (1e9, 1.0):
        gg, fg = fe.calculateChemicalPotential(X, np.zeros(4), np.zeros(4), gbd, np.zeros(nh), np.zeros(nh, np.int32))
        go, fo = orc.chem_potential(X, x, np.zeros(4), np.zeros(4), gbd, np.zeros(nh), np.zeros(nh, np.int32))
        assert np.array_equal(fg, fo)
        edge = np.zeros(nh, bool)
        edge[:8] = True
        assert np.all(go[edge] == min(1.5, gbd)) and np.array_equal(gg[edge], go[edge])
        assert np.all(fg[1:8:2] == (1 if gbd < 1.5 else 0)) and np.all(fg[0:7:2] == 0)
        assert np.max(np.abs(gg - go) / np.maximum(np.abs(go), 1.0)) <= 1e-10


def test_diffusion_parity_and_context_update():
    fe, orc, X, x = setup()
    nh = fe.n_h
    rng = np.random.default_rng(1)
    gam = rng.normal(0.0, 0.8, nh)
    full = np.zeros(nh, np.int32)
    full[::11] = 1
    args = (1e-13, 1e13, 0.23, 1e-3, 1 - 1e-16)
    xg = fe.integrateDiffusion(*args, gam, full)
    xo = orc.diffuse(*args, gam, full, x)
    assert np.max(np.abs(xg - xo)) <= 1e-14
    # the context occupancy advanced: a second step starts from xg
    xg2 = fe.integrateDiffusion(*args, gam, full)
    xo2 = orc.diffuse(*args, gam, full, xo)
    assert np.max(np.abs(xg2 - xo2)) <= 1e-14


def test_diffusion_parity_with_sites_far_from_equilibrium():
    """Sites whose f = gamma - ln(x/(1-x)) is beyond +-700 take the reference
    form of the edge term (e^{f/2} would leave the double range); their
    neighbors' increments are huge and clamped, as in the oracle."""
    fe, orc, X, x = setup(seed=9)
    nh = fe.n_h
    gam = np.random.default_rng(4).normal(0.0, 0.8, nh)
    gam[::17] = 800.0
    gam[5::23] = -800.0
    full = np.zeros(nh, np.int32)
    args = (1e-13, 1e13, 0.23, 1e-3, 1 - 1e-16)
    print("diffusion: %d sites at gamma = +-800" % (len(gam[::17]) + len(gam[5::23])), flush=True)
    xg = fe.integrateDiffusion(*args, gam, full)
    xo = orc.diffuse(*args, gam, full, x)
    assert np.all(np.isfinite(xg)) and np.any(xg == 1e-3) and np.any(xg == 1 - 1e-16)
    assert np.max(np.abs(xg - xo)) <= 1e-14


def test_diffusion_conserves_total_h_at_size():
    """24^3 cells (55,296 H sites), no reservoir, no clamp: sum x changes
    only by round-off of the site sums."""
    X, x, box, per, npd, nh = specimen(24, seed=8, disp=0.02, x=(0.2, 0.5))
    fe = A.FreeEntropy(npd, nh, box, per, T=400.0)
    fe.set_occupancy(x)
    fe.findNeighbors(X)
    gam = np.random.default_rng(3).normal(0.0, 1.0, nh)
    x1 = fe.integrateDiffusion(1e-13, 1e13, 0.23, 1e-3, 1 - 1e-16, gam, np.zeros(nh, np.int32))
    assert np.all(x1 > 1e-3) and np.any(x1 != x)
    assert abs(x1.sum() - x.sum()) <= 1e-12 * x.sum()
    b = fe.diffusion_balance()
    assert b["clamp_lift"] == 0.0 and b["flux_gross"] > 0
    assert abs(b["flux_sum"]) <= 1e-12 * b["flux_gross"]


def test_timed_diffusion_equals_a_step_and_keeps_the_occupancy():
    """a2c_fe_time_diffuse: every timed step starts from the context
    occupancy, writes what a2c_fe_diffuse returns (bitwise) and leaves the
    context occupancy as it was."""
    fe, orc, X, x = setup()
    nh = fe.n_h
    rng = np.random.default_rng(2)
    gam = rng.normal(0.0, 0.8, nh)
    full = np.zeros(nh, np.int32)
    full[::7] = 1
    args = (1e-13, 1e13, 0.23, 1e-3, 1 - 1e-16)
    print("timed diffusion: %d H sites" % nh, flush=True)
    dg, df, dx = fe.dev_alloc(8 * nh), fe.dev_alloc(4 * nh), fe.dev_alloc(8 * nh)
    try:
        fe.upload(dg, gam)
        fe.upload(df, full)
        tm = fe.time_diffuse(*args, dg, df, dx, 4)
        xt = fe.download(np.zeros(nh), dx)
    finally:
        for p in (dg, df, dx):
            fe.dev_free(p)
    assert list(tm["kernels_ms"]) == ["diff_f", "diff_step", "diff_balance"]
    assert all(v > 0 for v in tm["kernels_ms"].values()) and tm["total_ms"] > 0
    xs = fe.integrateDiffusion(*args, gam, full)  # from the unchanged occupancy x
    assert np.array_equal(xt, xs)
    assert np.max(np.abs(xs - orc.diffuse(*args, gam, full, x))) <= 1e-14
