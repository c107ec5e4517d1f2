"""Pins of the oracle objective/gradient (minimizer.cpp:83-369), chemical
potential (:588-802) and diffusion step (integrate.cpp:9-87)."""
import numpy as np
import pytest

import oracle as O
from helpers import KB, csr_rows, specimen


def make(n=3, seed=11, disp=0.03, sigma=(0.08, 0.15), x=(0.05, 0.6), periodic=True, T=500.0, **kw):
    X, xo, box, per, npd, nh = specimen(n, seed=seed, disp=disp, sigma=sigma, x=x, periodic=periodic)
    orc = O.Oracle(npd, nh, box, per, T=T, **kw)
    orc.build_neighbors(X)
    return orc, X, xo


def test_test_potential_closed_form():
    """With f = phi = r^2 and F(rho) = rho, the Gaussian average of |r|^2 over
    r = d + sqrt2 (s_i xi1 - s_j xi2) is |d|^2 + 3(s_i^2 + s_j^2) exactly, and
    the 12-point rule is exact for it.  The objective then has a closed form in
    the neighbor pairs alone (numpy below), pinning the pair vector, the
    sqrt(2) factors, the weight and the V/T - 3kB sum x ln(sigma) assembly."""
    orc, X, x = make(potential="test")
    npd, nh, T = orc.n_pd, orc.n_h, orc.p.T
    q = X.reshape(-1, 4)[:, :3]
    s = X.reshape(-1, 4)[:, 3]
    box = np.array(orc.p.box[:])

    def E(i, j):
        d = q[i] - q[j]
        d = d - box * np.rint(d / box)
        return d @ d + 3 * (s[i] ** 2 + s[j] ** 2)

    PdPd, PdH = csr_rows(*orc.neighbors("PdPd")), csr_rows(*orc.neighbors("PdH"))
    HPd, HH = csr_rows(*orc.neighbors("HPd")), csr_rows(*orc.neighbors("HH"))
    V = 0.0
    for i in range(npd):
        rho = sum(E(i, j) for j in PdPd[i]) + sum(x[a] * E(i, npd + a) for a in PdH[i])
        V += rho + 0.5 * rho  # F(rho) = rho; pair sums equal the density sums
    for a in range(nh):
        rho = sum(E(npd + a, j) for j in HPd[a]) + sum(x[b] * E(npd + a, npd + b) for b in HH[a])
        V += x[a] * rho + 0.5 * x[a] * rho
    ent = np.log(s[:npd]).sum() + (x * np.log(s[npd:])).sum()
    want = V / T - 3 * KB * ent
    f, G = orc.eval(X, x)
    assert f == pytest.approx(want, rel=1e-12)


def test_sigma_to_zero_is_static_eam():
    """As sigma -> 0 every quadrature point collapses onto d, so V is the
    static EAM energy over the neighbor pairs."""
    orc, X, x = make(sigma=(1e-5, 1e-5))
    npd, nh, T = orc.n_pd, orc.n_h, orc.p.T
    f, G = orc.eval(X, x)
    Xs = X.reshape(-1, 4)
    V = T * (f + 3 * KB * (np.log(Xs[:npd, 3]).sum() + (x * np.log(Xs[npd:, 3])).sum()))
    q = Xs[:, :3]
    box = np.array(orc.p.box[:])

    def r(i, j):
        d = q[i] - q[j]
        return np.linalg.norm(d - box * np.rint(d / box))

    PdPd, PdH = csr_rows(*orc.neighbors("PdPd")), csr_rows(*orc.neighbors("PdH"))
    HPd, HH = csr_rows(*orc.neighbors("HPd")), csr_rows(*orc.neighbors("HH"))
    Vs = 0.0
    for i in range(npd):
        rp = np.array([r(i, j) for j in PdPd[i]])
        rh = np.array([r(i, npd + a) for a in PdH[i]])
        rho = O.eam(O.RHO_PD, rp).sum() + (x[PdH[i]] * O.eam(O.RHO_H, rh)).sum()
        Vs += O.eam(O.F_PD, [rho])[0] + 0.5 * O.eam(O.PHI_PD, rp).sum() + 0.5 * (x[PdH[i]] * O.eam(O.PHI_PDH, rh)).sum()
    for a in range(nh):
        rp = np.array([r(npd + a, j) for j in HPd[a]])
        rh = np.array([r(npd + a, npd + b) for b in HH[a]])
        rho = O.eam(O.RHO_PD, rp).sum() + (x[HH[a]] * O.eam(O.RHO_H, rh)).sum()
        Vs += x[a] * (O.eam(O.F_H, [rho])[0] + 0.5 * O.eam(O.PHI_PDH, rp).sum() + 0.5 * (x[HH[a]] * O.eam(O.PHI_H, rh)).sum())
    assert V == pytest.approx(Vs, rel=1e-8)


@pytest.mark.parametrize("periodic", [True, False])
def test_gradient_vs_richardson_fd(periodic):
    orc, X, x = make(periodic=periodic)
    f, G = orc.eval(X, x)
    rng = np.random.default_rng(5)
    scale = np.abs(G).max()
    for k in rng.choice(X.size, 16, replace=False):
        vals = []
        for h in (2e-4, 1e-4):
            Xp, Xm = X.copy(), X.copy()
            Xp[k] += h
            Xm[k] -= h
            vals.append((orc.eval(Xp, x)[0] - orc.eval(Xm, x)[0]) / (2 * h))
        rich = (4 * vals[1] - vals[0]) / 3
        assert abs(G[k] - rich) <= 1e-6 * scale, (k, G[k], rich)


def test_translation_invariance_and_zero_net_force():
    orc, X, x = make(periodic=False)
    f, G = orc.eval(X, x)
    Xt = X.reshape(-1, 4).copy()
    Xt[:, :3] += np.array([0.37, -1.1, 2.9])
    ft, Gt = orc.eval(Xt.ravel(), x)
    assert ft == pytest.approx(f, rel=1e-13)
    # positions round differently after the shift; the Pd polynomials cancel
    assert np.allclose(Gt, G, rtol=0, atol=1e-9 * np.abs(G).max())
    net = G.reshape(-1, 4)[:, :3].sum(axis=0)
    assert np.abs(net).max() <= 1e-11 * np.abs(G).max()


def test_sigma_gradient_entropy_dominates_at_small_sigma():
    orc, X, x = make(sigma=(1e-3, 1e-3))
    f, G = orc.eval(X, x)
    gs = G.reshape(-1, 4)[:, 3]
    assert (gs[: orc.n_pd] < 0).all()  # -3 kB / sigma wins: f decreases as sigma grows


def test_objective_is_unbounded_as_one_width_grows():
    """Spreading one Pd packet: past a barrier near 1 A its pair terms vanish
    and only the entropy -3 kB x ln(sigma) / the embedding at rho -> 0 remain,
    so df/dsigma -> -3 kB / sigma exactly and f decreases without bound (the
    reason for the optimizer's optional sigma_ub)."""
    orc, X, x = make(n=4, sigma=(0.10, 0.10), x=(1e-3, 1e-3), T=400.0)
    f0, _ = orc.eval(X, x)
    fs, gs = [], []
    for s in (1.0, 30.0, 100.0, 300.0):
        Y = X.copy()
        Y[3] = s
        f, G = orc.eval(Y, x)
        fs.append(f)
        gs.append(G[3])
    assert fs[0] > f0  # the barrier
    assert fs[1] > fs[2] > fs[3]  # then downhill for good
    for s, g in zip((30.0, 100.0, 300.0), gs[1:]):
        assert g == pytest.approx(-3.0 * KB / s, rel=1e-9)
    assert fs[2] - fs[3] == pytest.approx(3.0 * KB * np.log(3.0), rel=1e-9)


def test_fd_mode_matches_analytic():
    """Reference FD derivatives (eps 1e-5) agree with the analytic ones up to
    the FD round-off of the reference's f_Pd polynomial form (~1e-6 rel)."""
    orc, X, x = make()
    orc_fd, _, _ = make(deriv="fd")
    f, G = orc.eval(X, x)
    f2, G2 = orc_fd.eval(X, x)
    assert f2 == f  # the objective uses no derivative
    assert np.linalg.norm(G2 - G) <= 1e-5 * np.linalg.norm(G)


def test_chemical_potential_is_occupancy_derivative():
    """With coe = 0, gamma - 1.5 - ln(x/(1-x)) = (dV/dx_a)/(kB T), checked by
    Richardson differences of V in x_a (minimizer.cpp:664-722)."""
    orc, X, x = make(n=3)
    npd, nh, T = orc.n_pd, orc.n_h, orc.p.T
    sig = X.reshape(-1, 4)[:, 3]
    gam, full = orc.chem_potential(X, x, np.zeros(4), np.zeros(4), 1e9, np.zeros(nh), np.zeros(nh, np.int32))

    def V(xx):
        f, _ = orc.eval(X, xx)
        return T * (f + 3 * KB * (np.log(sig[:npd]).sum() + (xx * np.log(sig[npd:])).sum()))

    for a in [0, 5, 17, nh - 1]:
        vals = []
        for h in (2e-4, 1e-4):
            xp, xm = x.copy(), x.copy()
            xp[a] += h
            xm[a] -= h
            vals.append((V(xp) - V(xm)) / (2 * h))
        dVdx = (4 * vals[1] - vals[0]) / 3
        want = np.log(x[a] / (1 - x[a])) + dVdx / (KB * T) + 1.5
        assert gam[a] == pytest.approx(want, rel=1e-8, abs=1e-8)
    assert (full == 0).all()


def test_chemical_potential_cap_and_full_sites():
    orc, X, x = make(n=3)
    nh = orc.n_h
    x = x.copy()
    x[0], x[1] = 0.99995, 0.5
    g0 = np.full(nh, 7.0)
    full0 = np.zeros(nh, np.int32)
    full0[2] = 1
    gam, full = orc.chem_potential(X, x, np.zeros(4), np.zeros(4), -1e9, g0, full0)
    assert gam[2] == 7.0 and full[2] == 1  # untouched
    assert gam[0] == -1e9 and full[0] == 1  # capped and x > 0.9999
    assert gam[1] == -1e9 and full[1] == 0  # capped only


def test_chemical_potential_correction_terms():
    """A.5 correction terms, restated in vectorised form from the neighbor
    counts: 1/2 T c0H (c1H n^2 + 2 c2H n X + c3H X^2) + sum_j T c0P (c2P |PdPd(j)|
    + c3P sum x) + sum_b x_b T c0H (c2H |HPd(b)| + c3H sum x)."""
    orc, X, x = make(n=3)
    nh, T = orc.n_h, orc.p.T
    cp, ch = np.array([0.3, 0.0, 0.2, -0.1]), np.array([0.5, 0.07, -0.03, 0.02])
    g0, _ = orc.chem_potential(X, x, np.zeros(4), np.zeros(4), 1e9, np.zeros(nh), np.zeros(nh, np.int32))
    g1, _ = orc.chem_potential(X, x, cp, ch, 1e9, np.zeros(nh), np.zeros(nh, np.int32))
    PdPd, PdH = csr_rows(*orc.neighbors("PdPd")), csr_rows(*orc.neighbors("PdH"))
    HPd, HH = csr_rows(*orc.neighbors("HPd")), csr_rows(*orc.neighbors("HH"))
    for a in range(0, nh, 7):
        n = len(HPd[a])
        Xa = x[HH[a]].sum()
        corr = 0.5 * T * ch[0] * (ch[1] * n * n + 2 * ch[2] * n * Xa + ch[3] * Xa * Xa)
        corr += sum(T * cp[0] * (cp[2] * len(PdPd[j]) + cp[3] * x[PdH[j]].sum()) for j in HPd[a])
        corr += sum(x[b] * T * ch[0] * (ch[2] * len(HPd[b]) + ch[3] * x[HH[b]].sum()) for b in HH[a])
        assert g1[a] - g0[a] == pytest.approx(corr / (KB * T), rel=1e-9, abs=1e-9)


def test_diffusion_two_site_closed_form_and_conservation():
    orc, X, x = make(n=3)
    nh, T = orc.n_h, orc.p.T
    rng = np.random.default_rng(2)
    gam = rng.normal(0, 0.5, nh)
    full = np.zeros(nh, np.int32)
    nu, Qm, dt = 1e13, 0.23, 1e-13  # 12 * rate * dt ~ 0.06
    x1 = orc.diffuse(dt, nu, Qm, 1e-3, 1 - 1e-16, gam, full, x)
    pref = nu * np.exp(-Qm / (KB * T))
    # closed form for one site
    HH1 = csr_rows(*orc.neighbors("HH1"))
    fa = gam - np.log(x / (1 - x))
    a = 4
    flux = sum(x[b] * (1 - x[a]) * np.exp(-(fa[a] - fa[b]) / 2) - x[a] * (1 - x[b]) * np.exp((fa[a] - fa[b]) / 2) for b in HH1[a])
    assert x1[a] == pytest.approx(x[a] + pref * flux * dt, rel=1e-13)
    # no clamp fired, no reservoir: total H conserved to round-off
    assert np.all(x1 > 1e-3) and np.all(x1 < 1 - 1e-16)
    assert abs(x1.sum() - x.sum()) <= 64 * np.finfo(float).eps * x.sum()


def test_diffusion_fixed_point():
    orc, X, x = make(n=3)
    gam = np.full(orc.n_h, 0.7)  # uniform chemical potential -> zero net flux on every edge
    x1 = orc.diffuse(1e-13, 1e13, 0.23, 1e-3, 1 - 1e-16, gam, np.zeros(orc.n_h, np.int32), x)
    assert np.allclose(x1, x, rtol=1e-12, atol=0)
