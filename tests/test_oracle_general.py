"""Pins of the oracle's general model -- a tabulated potential ("tab") and a
node table ("table") -- and of the library's setfl reader, on the CPU.

Tabulated functions: cubic Hermite pieces with central-difference node
slopes reproduce cubic data exactly away from the end intervals; tables
sampled from the compiled-in Pd-H closed forms reproduce the closed-form
objective and gradient.  Node table: the reference's 12 nodes as a table
give the 12-point rule's results; the 6-D two-point product rule gives the
closed-form quadratic moment |d|^2 + 3 (s_i^2 + s_j^2); gradients agree with
central differences of the objective.  setfl: a file written here reads back
bit for bit, element and pair order included."""
import numpy as np
import pytest

import a2c_amd as A
import oracle as O
from helpers import csr_rows, product_rule_2pt, sampled_tables, specimen, write_setfl


def make(n=3, seed=13, disp=0.03, sigma=(0.08, 0.15), x=(0.05, 0.6), periodic=True, T=500.0, **kw):
    X, xo, box, per, npd, nh = specimen(n, seed=seed, disp=disp, sigma=sigma, x=x, periodic=periodic)
    orc = O.Oracle(npd, nh, box, per, T=T, threads=8, **kw)
    orc.build_neighbors(X)
    return orc, X, xo


def test_pieces_reproduce_cubics_inside_and_extend_at_the_ends():
    h = 0.1
    k = np.arange(30)
    poly = lambda t: 0.7 * t ** 3 - 2.0 * t ** 2 + 0.5 * t + 1.0
    dpoly = lambda t: 2.1 * t ** 2 - 4.0 * t + 0.5
    y = poly(k * h)
    for t in np.linspace(0.2 + 1e-9, 2.7 - 1e-9, 41):  # intervals 2 .. 26: slopes from 5-point differences
        v, d = O.tab_eval(y, h, t)
        assert v == pytest.approx(poly(t), abs=1e-13) and d == pytest.approx(dpoly(t), abs=1e-12)
    v, d = O.tab_eval(y, h, 0.05)  # first interval: one-sided slope, not exact
    assert abs(v - poly(0.05)) > 1e-6
    # past the last node: 0 for an r-function, the end tangent for an embedding function
    assert O.tab_eval(y, h, 3.5) == (0.0, 0.0)
    ve, de = O.tab_eval(y, h, 3.5, emb=True)
    s_end = (y[-1] - y[-2]) / h
    assert de == pytest.approx(s_end, rel=1e-14) and ve == pytest.approx(y[-1] + s_end * (3.5 - 2.9), rel=1e-14)
    # nodes are interpolated exactly
    for m in (0, 7, 28):
        assert O.tab_eval(y, h, m * h)[0] == pytest.approx(y[m], abs=1e-15)


def test_tables_of_the_closed_forms_reproduce_the_closed_form_objective():
    O.set_tabulated(sampled_tables())
    orc, X, x = make(potential="eam")
    tab, _, _ = make(potential="tab")
    f0, G0 = orc.eval(X, x)
    f1, G1 = tab.eval(X, x)
    assert abs(f1 - f0) <= 1e-8 * abs(f0)
    # f_Pd's slope jumps by ~1% at its 5.0 A switch (value continuous); the
    # pieces round that corner over a grid step, which moves the gradient of
    # the Pd sites with quadrature points there by ~1e-5 of its scale
    assert np.linalg.norm(G1 - G0) <= 5e-5 * np.linalg.norm(G0)
    assert np.max(np.abs(G1 - G0)) <= 1e-4 * np.abs(G0).max()
    g0, _ = orc.chem_potential(X, x, np.zeros(4), np.zeros(4), 1e9, np.zeros(orc.n_h), np.zeros(orc.n_h, np.int32))
    g1, _ = tab.chem_potential(X, x, np.zeros(4), np.zeros(4), 1e9, np.zeros(orc.n_h), np.zeros(orc.n_h, np.int32))
    assert np.max(np.abs(g1 - g0)) <= 1e-6 * np.max(np.abs(g0))


def test_the_reference_nodes_as_a_table_give_the_12_point_rule():
    xi, w = O.quadrature()
    O.set_quad_table(xi, w)
    r12, X, x = make(quad="r12")
    tbl, _, _ = make(quad="table")
    f0, G0 = r12.eval(X, x)
    f1, G1 = tbl.eval(X, x)
    assert f1 == pytest.approx(f0, rel=1e-14)
    assert np.max(np.abs(G1 - G0)) <= 1e-13 * np.abs(G0).max()


def test_two_point_product_rule_gives_the_quadratic_moment():
    xi, w = product_rule_2pt()
    O.set_quad_table(xi, w)
    orc, X, x = make(potential="test", quad="table")
    npd, nh = orc.n_pd, orc.n_h
    q = X.reshape(-1, 4)[:, :3]
    s = X.reshape(-1, 4)[:, 3]
    box = np.array(orc.p.box[:])

    def E(i, j):
        d = q[i] - q[j]
        d = d - box * np.rint(d / box)
        return d @ d + 3.0 * (s[i] ** 2 + s[j] ** 2)

    PdPd, PdH = csr_rows(*orc.neighbors("PdPd")), csr_rows(*orc.neighbors("PdH"))
    HPd, HH = csr_rows(*orc.neighbors("HPd")), csr_rows(*orc.neighbors("HH"))
    want = np.zeros(npd + nh)
    for i in range(npd):
        want[i] = sum(E(i, j) for j in PdPd[i]) + sum(x[a] * E(i, npd + a) for a in PdH[i])
    for a in range(nh):
        want[npd + a] = sum(E(npd + a, j) for j in HPd[a]) + sum(x[b] * E(npd + a, npd + b) for b in HH[a])
    rho = orc.density(X, x)
    assert np.max(np.abs(rho - want) / np.abs(want)) <= 1e-12


@pytest.mark.parametrize("quad", ["table", "r12"])
def test_tabulated_gradient_vs_central_differences(quad):
    O.set_tabulated(sampled_tables())
    xi, w = product_rule_2pt()
    O.set_quad_table(xi, w)
    orc, X, x = make(n=2, periodic=False, potential="tab", quad=quad)
    f, G = orc.eval(X, x)
    scale = np.abs(G).max()
    rng = np.random.default_rng(3)
    for k in list(rng.choice(X.size, 10, replace=False)) + [3, 4 * orc.n_pd + 3]:
        vals = []
        for h in (2e-4, 1e-4):
            Xp, Xm = X.copy(), X.copy()
            Xp[k] += h
            Xm[k] -= h
            vals.append((orc.eval(Xp, x)[0] - orc.eval(Xm, x)[0]) / (2 * h))
        rich = (4 * vals[1] - vals[0]) / 3
        assert abs(G[k] - rich) <= 1e-6 * scale, (k, G[k], rich)
    assert np.max(np.abs(G.reshape(-1, 4)[:, :3].sum(axis=0))) <= 1e-10 * scale  # translation invariance


def test_setfl_reads_back_bit_for_bit(tmp_path):
    rng = np.random.default_rng(5)
    els = ["Ni", "H", "Pd"]  # the host is not the first element: pair order matters
    n_rho, n_r = 23, 31
    F = {e: rng.normal(size=n_rho) for e in els}
    f = {e: rng.normal(size=n_r) for e in els}
    rphi = {(els[i], els[j]): rng.normal(size=n_r) for i in range(3) for j in range(i + 1)}
    p = str(tmp_path / "x.eam.alloy")
    write_setfl(p, els, F, f, rphi, n_rho, 0.05, n_r, 0.2, 6.0)
    t = A.eam_setfl_read(p, "Pd", "H")
    assert (t["n_rho"], t["d_rho"], t["n_r"], t["d_r"], t["cutoff"]) == (n_rho, 0.05, n_r, 0.2, 6.0)
    assert np.array_equal(t["F_pd"], F["Pd"]) and np.array_equal(t["F_h"], F["H"])
    assert np.array_equal(t["f_pd"], f["Pd"]) and np.array_equal(t["f_h"], f["H"])
    assert np.array_equal(t["rphi_pdpd"], rphi[("Pd", "Pd")])
    assert np.array_equal(t["rphi_pdh"], rphi[("Pd", "H")])  # stored as (i, j) with i >= j
    assert np.array_equal(t["rphi_hh"], rphi[("H", "H")])
    with pytest.raises(A.A2CError):
        A.eam_setfl_read(p, "Pd", "He")
    bad = str(tmp_path / "short.eam.alloy")
    open(bad, "w").write(open(p).read()[:-400])
    with pytest.raises(A.A2CError):
        A.eam_setfl_read(bad, "Pd", "H")
