"""Neighbor rebuilds from the skin lists: from its second build on a context
keeps lists at rc + skin and, while no site has moved by skin/2, filters
them instead of walking the cells.  The filtered lists must be the sets,
rows, shell counts and HH1 of the oracle at the new positions, bit for bit,
and an evaluation on them must equal, bit for bit, the one of a fresh
context that walked the cells."""
import numpy as np
import pytest

import a2c_amd as A
import oracle as O
from helpers import specimen

pytestmark = pytest.mark.gpu

SKIN = 0.2  # the library's default (A2C_NBR_SKIN)


def moved(X, rng, step):
    Y = X.copy().reshape(-1, 4)
    Y[:, :3] += rng.normal(0.0, step, size=(Y.shape[0], 3))
    return Y.reshape(-1)


def same_lists(fe, orc):
    for w in ("PdPd", "PdH", "HPd", "HH", "HH1"):
        go, gi = fe.neighbors(w)
        oo, oi = orc.neighbors(w)
        assert np.array_equal(go, oo) and np.array_equal(gi, oi), w
    assert np.array_equal(fe.pdpdnum(), orc.pdpdnum())


def evaluate(fe, X):
    """(f, G), or the error code: a mixed evaluation with a pair closer than
    its tables reach (under 1 A) is non-finite, in either context alike."""
    try:
        return fe.eval(X)
    except A.A2CError as e:
        return e.code


def fresh_eval(npd, nh, box, per, x, X, precision):
    with A.FreeEntropy(npd, nh, box, per, T=500.0, precision=precision) as fe:
        fe.set_occupancy(x)
        st = fe.findNeighbors(X)
        assert not st["refreshed"]
        return evaluate(fe, X)


@pytest.mark.parametrize("periodic,disp", [(True, 0.03), (True, 0.15), (False, 0.05)])
def test_refreshed_lists_match_the_oracle_and_a_fresh_build(periodic, disp):
    n = 4
    X, x, box, per, npd, nh = specimen(n, seed=17, disp=disp, sigma=(0.08, 0.15), x=(0.05, 0.6), periodic=periodic)
    rng = np.random.default_rng(3)
    orc = O.Oracle(npd, nh, box, per, T=500.0, threads=8)
    print("skin lists:", "periodic" if periodic else "open", "displacement", disp, flush=True)
    with A.FreeEntropy(npd, nh, box, per, T=500.0, precision="mixed") as fe:
        fe.set_occupancy(x)
        st = fe.findNeighbors(X)
        assert not st["refreshed"]  # first build: the cell walk at rc
        X1 = moved(X, rng, 0.01)
        st = fe.findNeighbors(X1)  # second build: skin lists, then filtered
        assert st["refreshed"] and st["max_displacement"] == 0.0
        orc.build_neighbors(X1)
        same_lists(fe, orc)
        Xk = X1
        for k in range(3):  # small moves: filtered again
            Xk = moved(Xk, rng, 0.008)
            st = fe.findNeighbors(Xk)
            assert st["refreshed"] and 0.0 < st["max_displacement"] < SKIN / 2
            orc.build_neighbors(Xk)
            same_lists(fe, orc)
            for prec in ("mixed", "fp64"):
                fe.set_precision(prec)
                r, r0 = evaluate(fe, Xk), fresh_eval(npd, nh, box, per, x, Xk, prec)
                if isinstance(r0, int):
                    assert r == r0 and prec == "mixed", (k, prec, r, r0)
                    continue
                (f, G), (f0, G0) = r, r0
                assert f == f0 and np.array_equal(G, G0), (k, prec)
        Xb = moved(X1, rng, 0.12)  # beyond skin/2: new skin lists
        st = fe.findNeighbors(Xb)
        assert st["refreshed"] and st["max_displacement"] == 0.0
        orc.build_neighbors(Xb)
        same_lists(fe, orc)
        fe.set_precision("fp64")
        f, G = fe.eval(Xb)
        fo, Go = orc.eval(Xb, x)
        assert abs(f - fo) <= 1e-10 * abs(fo)


def test_smallest_periodic_box_that_holds_the_skin():
    """3 cells (11.67 A) per axis: 2 (rc + skin) = 11.1 A fits, so the third
    build at unchanged positions is filtered from the skin lists."""
    X, x, box, per, npd, nh = specimen(3, seed=5, disp=0.02)
    with A.FreeEntropy(npd, nh, box, per, T=500.0) as fe:
        fe.set_occupancy(x)
        for _ in range(3):
            st = fe.findNeighbors(X)
        assert st["refreshed"]
        orc = O.Oracle(npd, nh, box, per, T=500.0, threads=8)
        orc.build_neighbors(X)
        same_lists(fe, orc)
