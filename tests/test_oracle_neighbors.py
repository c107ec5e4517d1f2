"""Pins of the oracle neighbor sets (neighbors.cpp:74-96, minimizer.cpp:923-1005)."""
import numpy as np
import pytest

import oracle as O
from helpers import brute_neighbors, csr_rows, specimen


@pytest.mark.parametrize("n,periodic,disp", [(3, True, 0.02), (3, False, 0.05), (4, True, 0.15)])
def test_oracle_sets_equal_brute_force(n, periodic, disp):
    X, x, box, per, npd, nh = specimen(n, seed=7, disp=disp, periodic=periodic)
    orc = O.Oracle(npd, nh, box, per)
    orc.build_neighbors(X)
    want, num = brute_neighbors(X, npd, box, per)
    for k in want:
        rows = csr_rows(*orc.neighbors(k))
        assert len(rows) == len(want[k])
        for a, b in zip(rows, want[k]):
            assert np.array_equal(a, b), k
    assert np.array_equal(orc.pdpdnum(), num)


def test_perfect_fcc_shell_counts():
    """Appendix geometry: 42 Pd-Pd, 38 Pd-H, 38 H-Pd, 42 H-H, 12 HH1,
    PdPdnum (12, 6, 24) at a = 3.89, rc = 5.35 with periodic images."""
    X, x, box, per, npd, nh = specimen(4, disp=0.0)
    orc = O.Oracle(npd, nh, box, per)
    assert orc.build_neighbors(X) == 80
    counts = {k: np.unique(np.diff(orc.neighbors(k)[0])) for k in ["PdPd", "PdH", "HPd", "HH", "HH1"]}
    assert counts == {"PdPd": [42], "PdH": [38], "HPd": [38], "HH": [42], "HH1": [12]}
    assert (orc.pdpdnum() == [12, 6, 24]).all()


def test_lists_symmetric():
    X, x, box, per, npd, nh = specimen(3, seed=3, disp=0.08)
    orc = O.Oracle(npd, nh, box, per)
    orc.build_neighbors(X)
    pairs = set()
    for k, (oi, oj) in {"PdPd": (0, 0), "PdH": (0, npd), "HPd": (npd, 0), "HH": (npd, npd)}.items():
        for r, row in enumerate(csr_rows(*orc.neighbors(k))):
            for j in row:
                pairs.add((r + oi, int(j) + oj))
    assert all((j, i) in pairs for i, j in pairs)


def test_small_periodic_box_refused():
    X, x, box, per, npd, nh = specimen(2)  # 7.78 A < 2 rc
    orc = O.Oracle(npd, nh, box, per)
    with pytest.raises(ValueError):
        orc.build_neighbors(X)
