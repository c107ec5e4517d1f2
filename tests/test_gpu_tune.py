"""Launch-shape table (include/a2c_tune.h): the autotuner's output, and
results under tables other than the default -- parity with the oracle at the
same bars, and the same bits alone and in a batch under any one table."""
import numpy as np
import pytest

import a2c_amd as A
import oracle as O
from helpers import specimen

pytestmark = pytest.mark.gpu

LANES = (1, 2, 4, 8, 16)


@pytest.fixture
def restore_table():
    yield
    A.tune_set(None)


def test_autotune_returns_a_valid_table_and_leaves_the_default(restore_table):
    before = A.tune_get()
    table, rows = A.autotune(cells=[4, 6], target_sites=20000, reps=2)
    assert A.tune_get() == before  # apply=False
    assert len(table["max_sites"]) == 2 and table["max_sites"][-1] == 2**31 - 1
    assert 512 < table["max_sites"][0] < 1728  # geometric mean of 512 and 1,728 sites
    for k in table["lanes_mixed"] + table["lanes_fp64"]:
        assert k in LANES
    for r in rows:
        assert r["sites"] == 8 * r["cells"] ** 3 and r["copies"] >= 1
        for p in ("mixed", "fp64"):
            assert all(t > 0 for t in r["ns_per_site"][p]) and all(t > 0 for t in r["us_alone"][p])
    for i, r in enumerate(rows):
        # the fastest K alone among those within 5% of the best throughput,
        # or 25% for a lattice that alone under the best-throughput K gives
        # fewer waves than the device has SIMDs (1,024 on MI355X)
        for p in ("mixed", "fp64"):
            b, a = np.array(r["ns_per_site"][p]), np.array(r["us_alone"][p])
            k = LANES[int(np.argmin(b))]
            ok = np.flatnonzero(b <= (1.25 if r["sites"] * k < 64 * 1024 else 1.05) * b.min())
            k = LANES[int(ok[np.argmin(a[ok])])]
            assert table["lanes_" + p][i] == k
            assert r["latency"][p] == (k != LANES[int(np.argmin(b))])
    A.autotune(cells=[4], target_sites=1000, reps=1, apply=True)
    assert A.tune_get() != before or before["max_sites"] == [2**31 - 1]


@pytest.mark.parametrize("k", LANES)
def test_other_table_meets_the_bars_and_batches_bitwise(k, restore_table):
    A.tune_set({"max_sites": [2**31 - 1], "lanes_mixed": [k], "lanes_fp64": [k]})
    X, x, box, per, npd, nh = specimen(4, seed=2, disp=0.05, sigma=(0.08, 0.15), x=(0.0, 0.6))
    orc = O.Oracle(npd, nh, box, per, T=500.0, threads=8)
    orc.build_neighbors(X)
    fo, Go = orc.eval(X, x)
    for prec, frel, grel in (("fp64", 1e-10, 1e-8), ("mixed", 1e-7, 5e-5)):
        with A.FreeEntropy(npd, nh, box, per, T=500.0, precision=prec) as fe:
            fe.set_occupancy(x)
            fe.findNeighbors(X)
            f, G = fe.eval(X)
        assert abs(f - fo) <= frel * abs(fo)
        assert np.max(np.abs(G - Go)) <= grel * np.max(np.abs(Go))
        X2, x2, *_ = specimen(3, seed=5, disp=0.03, sigma=(0.08, 0.15), x=(0.1, 0.4))
        with A.Batch(precision=prec) as B:
            s1 = B.add(X, x, npd, nh, box, per, T=500.0)
            B.add(X2, x2, 4 * 27, 4 * 27, (3 * 3.89,) * 3, per, T=400.0)
            fb = B.eval([s1, 1], precision=prec)
            assert fb[0] == f and np.array_equal(B.get_state(s1)[1], G)


def test_table_change_reaches_an_existing_context(restore_table):
    X, x, box, per, npd, nh = specimen(4, seed=3, disp=0.03, sigma=(0.08, 0.15), x=(0.0, 0.6))
    with A.FreeEntropy(npd, nh, box, per, T=400.0, precision="mixed") as fe:
        fe.set_occupancy(x)
        fe.findNeighbors(X)
        fe.eval(X)
        A.tune_set({"max_sites": [2**31 - 1], "lanes_mixed": [2], "lanes_fp64": [2]})
        f1, G1 = fe.eval(X)  # same context, new lanes per site
    with A.FreeEntropy(npd, nh, box, per, T=400.0, precision="mixed") as fe2:
        fe2.set_occupancy(x)
        fe2.findNeighbors(X)
        f2, G2 = fe2.eval(X)
    assert f1 == f2 and np.array_equal(G1, G2)
