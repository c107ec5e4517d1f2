"""Pins of the oracle's 3-point-per-axis Gauss-Hermite rule on the relative
coordinate (27 nodes; a builder addition, not in the reference).  Every pair
function depends only on r = d + sqrt(2)(s_i xi1 - s_j xi2), which is
N(d, (s_i^2 + s_j^2) I), so E g(|r|) = sum_k w_k g(|d + s z_k|)."""
import numpy as np
import pytest

import oracle as O
from helpers import KB, csr_rows, specimen


def make(n=3, seed=11, disp=0.03, sigma=(0.08, 0.15), x=(0.05, 0.6), periodic=True, T=500.0, **kw):
    X, xo, box, per, npd, nh = specimen(n, seed=seed, disp=disp, sigma=sigma, x=x, periodic=periodic)
    orc = O.Oracle(npd, nh, box, per, T=T, **kw)
    orc.build_neighbors(X)
    return orc, X, xo


def moment4(orc, X):
    """Closed-form densities for f = r^4: E|d + s z|^4 = |d|^4 + 10 s^2 |d|^2 + 15 s^4."""
    npd = orc.n_pd
    q = X.reshape(-1, 4)[:, :3]
    s = X.reshape(-1, 4)[:, 3]
    box = np.array(orc.p.box[:])

    def E(i, j):
        d = q[i] - q[j]
        d = d - box * np.rint(d / box)
        d2, s2 = d @ d, s[i] ** 2 + s[j] ** 2
        return d2 * d2 + 10 * s2 * d2 + 15 * s2 * s2

    return E


def test_gh3_is_exact_for_the_quartic_moment_and_r12_is_not():
    orc, X, x = make(potential="test4", quad="gh3")
    orc12, _, _ = make(potential="test4", quad="r12")
    npd, nh = orc.n_pd, orc.n_h
    E = moment4(orc, X)
    PdPd, PdH = csr_rows(*orc.neighbors("PdPd")), csr_rows(*orc.neighbors("PdH"))
    HPd, HH = csr_rows(*orc.neighbors("HPd")), csr_rows(*orc.neighbors("HH"))
    want = np.zeros(npd + nh)
    for i in range(npd):
        want[i] = sum(E(i, j) for j in PdPd[i]) + sum(x[a] * E(i, npd + a) for a in PdH[i])
    for a in range(nh):
        want[npd + a] = sum(E(npd + a, j) for j in HPd[a]) + sum(x[b] * E(npd + a, npd + b) for b in HH[a])
    rho = orc.density(X, x)
    assert np.max(np.abs(rho - want) / np.abs(want)) <= 1e-12
    rho12 = orc12.density(X, x)
    assert np.max(np.abs(rho12 - want) / np.abs(want)) > 1e-6  # the 12-point rule is exact to degree 3 only


def test_gh3_quadratic_moment_matches_r12():
    orc, X, x = make(potential="test", quad="gh3")
    orc12, _, _ = make(potential="test", quad="r12")
    f, G = orc.eval(X, x)
    f12, G12 = orc12.eval(X, x)
    assert f == pytest.approx(f12, rel=1e-12)
    assert np.max(np.abs(G - G12)) <= 1e-12 * np.abs(G12).max()


@pytest.mark.parametrize("periodic", [True, False])
def test_gh3_gradient_vs_richardson_fd(periodic):
    orc, X, x = make(periodic=periodic, quad="gh3")
    f, G = orc.eval(X, x)
    rng = np.random.default_rng(7)
    scale = np.abs(G).max()
    for k in list(rng.choice(X.size, 12, replace=False)) + [3, 4 * orc.n_pd + 3]:  # include two widths
        vals = []
        for h in (2e-4, 1e-4):
            Xp, Xm = X.copy(), X.copy()
            Xp[k] += h
            Xm[k] -= h
            vals.append((orc.eval(Xp, x)[0] - orc.eval(Xm, x)[0]) / (2 * h))
        rich = (4 * vals[1] - vals[0]) / 3
        assert abs(G[k] - rich) <= 1e-6 * scale, (k, G[k], rich)


def test_gh3_translation_invariance_and_fd_mode():
    orc, X, x = make(periodic=False, quad="gh3")
    f, G = orc.eval(X, x)
    g = G.reshape(-1, 4)
    assert np.abs(g[:, :3].sum(axis=0)).max() <= 1e-10 * np.abs(g[:, :3]).max()
    orc_fd, _, _ = make(periodic=False, quad="gh3", deriv="fd")
    f2, G2 = orc_fd.eval(X, x)
    assert f2 == f
    assert np.linalg.norm(G2 - G) <= 1e-5 * np.linalg.norm(G)


def test_gh3_chemical_potential_is_occupancy_derivative():
    orc, X, x = make(n=3, quad="gh3")
    npd, nh, T = orc.n_pd, orc.n_h, orc.p.T
    sig = X.reshape(-1, 4)[:, 3]
    gam, full = orc.chem_potential(X, x, np.zeros(4), np.zeros(4), 1e9, np.zeros(nh), np.zeros(nh, np.int32))

    def V(xx):  # f*T + 3 kB T sum x ln sigma  (the objective without the entropy term)
        f, _ = orc.eval(X, xx)
        return f * T + 3 * KB * T * (np.log(sig[:npd]).sum() + (xx * np.log(sig[npd:])).sum())

    for a in (0, 5, nh - 1):
        vals = []
        for h in (2e-4, 1e-4):
            xp, xm = x.copy(), x.copy()
            xp[a] += h
            xm[a] -= h
            vals.append((V(xp) - V(xm)) / (2 * h))
        dV = (4 * vals[1] - vals[0]) / 3
        want = np.log(x[a] / (1 - x[a])) + dV / (KB * T) + 1.5
        assert gam[a] == pytest.approx(want, rel=1e-7, abs=1e-7)
