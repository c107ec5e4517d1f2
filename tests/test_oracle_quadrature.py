"""Pins of the 12-point phase-packet rule (quadrature.cpp:43-75, weight :39)."""
import itertools
import json
import os

import numpy as np

import oracle as O

GOLD = os.path.join(os.path.dirname(__file__), "golden")


def test_nodes_bit_exact_vs_reference_build():
    """The fixture holds what the reference's own gaussianQuadrature printed."""
    xi, _ = O.quadrature()
    ref = json.load(open(os.path.join(GOLD, "quadrature_ref.json")))
    want = np.array([[float.fromhex(t) for t in row] for row in ref["nodes_hex"]])
    assert want.shape == (12, 6)
    assert np.array_equal(xi, want)
    live = O.reference_quadrature()  # present when oracle/_ref was built here
    if live is not None:
        assert np.array_equal(xi, live)


def test_gaussian_moments_to_degree_3():
    """Against the weight pi^-3 exp(-|xi|^2) in R^6: E[1]=1, E[xi_a]=0,
    E[xi_a xi_b]=delta_ab/2, all third moments 0."""
    xi, w = O.quadrature()
    assert np.isclose(w.sum(), 1.0, rtol=0, atol=1e-15)
    assert np.allclose(w @ xi, 0, atol=1e-15)
    M2 = (xi * w[:, None]).T @ xi
    assert np.allclose(M2, 0.5 * np.eye(6), atol=1e-15)
    for a, b, c in itertools.product(range(6), repeat=3):
        assert abs(np.sum(w * xi[:, a] * xi[:, b] * xi[:, c])) < 1e-15


def test_swap_symmetry():
    """(xi1, xi2) -> (xi2, xi1) maps the rule onto itself: the reference's
    one-sided gradient (no scatter to j) is exact only because of this."""
    xi, _ = O.quadrature()
    sw = np.concatenate([xi[:, 3:], xi[:, :3]], axis=1)
    a = {tuple(r) for r in xi}
    b = {tuple(r) for r in sw}
    assert a == b
