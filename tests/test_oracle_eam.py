"""Pins of the oracle's EAM functions (potential.cpp:16-205)."""
import json
import os

import numpy as np
import pytest

import oracle as O

GOLD = os.path.join(os.path.dirname(__file__), "golden")
ORDER = ["F_Pd", "F_H", "f_Pd", "f_H", "phi_Pd", "phi_PdH", "phi_H"]


def test_constants_match_reference_text():
    """Every named constant equals the value read from potential.cpp, in order."""
    ref = json.load(open(os.path.join(GOLD, "eam_constants.json")))
    for fn, name in enumerate(ORDER):
        want = [v for _, v in ref[name]["constants"]]
        got = list(O.eam_constants(fn))
        assert got == want, name


def test_switch_radii_and_scales_match_reference_text():
    ref = json.load(open(os.path.join(GOLD, "eam_constants.json")))
    assert ref["F_Pd"]["scale"] == 50.0
    assert ref["f_Pd"]["scale"] == 5.0 and ref["phi_Pd"]["scale"] == 5.0
    # the oracle switches from polynomial to tail exactly at these radii
    for fn, name in [(2, "f_Pd"), (3, "f_H"), (4, "phi_Pd"), (5, "phi_PdH"), (6, "phi_H")]:
        rs = ref[name]["switch"]
        k = O.eam_constants(fn)
        fr, fd = k[0], k[1]
        above = O.eam(fn, [np.nextafter(rs, 10) + 1e-9])[0]
        tail = fr * np.exp(fd / fr * (np.nextafter(rs, 10) + 1e-9 - rs))
        assert above == pytest.approx(tail, rel=1e-12), name


@pytest.mark.parametrize("fn,rs,tol", [(2, 5.0, 1e-10), (3, 5.35, 1e-10), (4, 5.05, 5e-6),
                                       (5, 5.35, 1e-10), (6, 5.35, 1e-10)])
def test_value_continuity_at_switch(fn, rs, tol):
    """The authors' funRc is the closed form's value at the switch radius, so a
    transcription error in any polynomial/Morse constant breaks continuity.
    (phi_Pd's funRc was evaluated with the polynomial f_Pd beyond 5.0; the
    reference's own mismatch is 2.7e-6 relative.)"""
    below = O.eam(fn, [rs])[0]  # the closed form applies up to and including rs
    funRc = O.eam_constants(fn)[0]
    assert abs(below - funRc) <= tol * abs(funRc)


@pytest.mark.parametrize("fn", range(7))
def test_analytic_derivative_vs_richardson(fn):
    xs = np.linspace(1.0, 40.0, 157) if fn < 2 else np.linspace(1.2, 6.3, 157)
    if fn >= 2:  # stay off the switch radii (C1, not C2, there)
        xs = xs[np.min(np.abs(xs[:, None] - np.array([5.0, 5.05, 5.35])[None]), axis=1) > 0.01]
    h = 2e-3
    d1 = (O.eam(fn, xs + h) - O.eam(fn, xs - h)) / (2 * h)
    d2 = (O.eam(fn, xs + h / 2) - O.eam(fn, xs - h / 2)) / h
    rich = (4 * d2 - d1) / 3
    a = O.eam_d(fn, xs)
    scale = np.abs(rich).max()
    assert np.max(np.abs(a - rich)) <= 2e-6 * scale


def test_fd_mode_is_reference_central_difference():
    """FD mode = (g(x+eps) - g(x-eps)) / (2 eps) with eps = 1e-5
    (potential.cpp:40-43 and the six siblings)."""
    xs = np.array([1.9, 2.75, 3.37, 3.89, 4.35, 4.76])
    for fn in range(2, 7):
        fd = O.eam_d(fn, xs, "fd", 1e-5)
        want = (O.eam(fn, xs + 1e-5) - O.eam(fn, xs - 1e-5)) / (2e-5)
        assert np.array_equal(fd, want)


def fcc_shell_distances(a, rc=5.35):
    pts = []
    basis = [(0, 0, 0), (.5, .5, 0), (.5, 0, .5), (0, .5, .5)]
    for i in range(-3, 4):
        for j in range(-3, 4):
            for k in range(-3, 4):
                for b in basis:
                    d = np.linalg.norm((np.array([i, j, k]) + b) * a)
                    if 0 < d <= rc:
                        pts.append(d)
    return np.array(pts)


def test_fcc_pd_static_energy_minimum():
    """Physics pin of F_Pd, f_Pd and phi_Pd together: static FCC Pd with the
    5.35 A neighbor cutoff has its energy minimum at a ~ 3.88-3.89 A (the
    lattice constant the reference's inputs use, 3.89) and a cohesive energy of
    about -3.96 eV/atom (Zhou et al.'s Pd fit)."""
    a_grid = np.linspace(3.80, 3.98, 37)
    E = []
    for a in a_grid:
        d = fcc_shell_distances(a)
        assert len(d) == 42
        rho = O.eam(O.RHO_PD, d).sum()
        E.append(O.eam(O.F_PD, [rho])[0] + 0.5 * O.eam(O.PHI_PD, d).sum())
    E = np.array(E)
    amin = a_grid[np.argmin(E)]
    assert 3.86 <= amin <= 3.91
    assert -4.0 < E.min() < -3.9
