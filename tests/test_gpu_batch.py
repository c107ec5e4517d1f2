"""Batched engine: many uneven lattices in one fused launch and a lockstep
LMVM.  Every slot's results must be bitwise the same alone and in the mix
(same kernels, per-slot reduction shapes), including a slot admitted while
others are mid-relaxation (continuous batching)."""
import numpy as np
import pytest

import a2c_amd as A
import oracle as O
from helpers import specimen

pytestmark = pytest.mark.gpu

LATS = [  # (cells, seed, T, x-range, periodic)
    (3, 1, 300.0, (0.10, 0.10), True),
    (4, 2, 500.0, (0.0, 0.6), True),
    ((4, 3, 3), 3, 400.0, (0.01, 0.6), True),
    (3, 4, 650.0, (0.2, 0.4), False),
]


def make(spec):
    n, seed, T, xr, per = spec
    X, x, box, per3, npd, nh = specimen(n, seed=seed, disp=0.03, sigma=(0.08, 0.15), x=xr, periodic=per)
    return dict(X=X, x=x, box=box, periodic=per3, n_pd=npd, n_h=nh, T=T)


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_fused_eval_matches_single_lattice_bitwise(precision):
    specs = [make(s) for s in LATS]
    with A.Batch(precision=precision, max_lattices=8) as B:
        slots = [B.add(**s) for s in specs]
        f = B.eval(slots, precision=precision)
        for sl, s, fb in zip(slots, specs, f):
            with A.FreeEntropy(s["n_pd"], s["n_h"], s["box"], s["periodic"], T=s["T"], precision=precision) as fe:
                fe.set_occupancy(s["x"])
                fe.findNeighbors(s["X"])
                f1, G1 = fe.eval(s["X"])
            Xb, Gb = B.get_state(sl)
            assert fb == f1 and np.array_equal(Gb, G1)
        # a subset in another order gives the same numbers
        f2 = B.eval(slots[::-2], precision=precision)
        assert np.array_equal(f2, f[::-2])


def test_fused_eval_matches_oracle():
    specs = [make(s) for s in LATS]
    with A.Batch(precision="fp64") as B:
        slots = [B.add(**s) for s in specs]
        f = B.eval(slots, precision="fp64")
    for s, fb in zip(specs, f):
        orc = O.Oracle(s["n_pd"], s["n_h"], s["box"], s["periodic"], T=s["T"], threads=8)
        orc.build_neighbors(s["X"])
        fo, _ = orc.eval(s["X"], s["x"])
        assert abs(fb - fo) <= 1e-10 * abs(fo)


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_lockstep_relaxation_same_alone_and_in_mix(precision):
    specs = [make(s) for s in LATS]
    opts = dict(gatol=1e-9, grtol=0.0, max_it=8000)  # the non-periodic lattice needs ~2000 iterations
    alone = []
    for s in specs:
        with A.FreeEntropy(s["n_pd"], s["n_h"], s["box"], s["periodic"], T=s["T"], precision=precision) as fe:
            fe.set_occupancy(s["x"])
            fe.findNeighbors(s["X"])
            alone.append(fe.minimizeFreeEntropy(s["X"], **opts))
    with A.Batch(precision=precision) as B:
        slots = [B.add(**s) for s in specs[:3]]
        for sl in slots:
            B.start(sl, **opts)
        B.step(3)  # the fourth lattice arrives while the others are running
        slots.append(B.add(**specs[3]))
        B.start(slots[3], **opts)
        B.run()
        for sl, (Xa, ra) in zip(slots, alone):
            rb = B.report(sl)
            Xb, _ = B.get_state(sl)
            assert rb == ra, (rb, ra)
            assert np.array_equal(Xb, Xa)
            assert rb["reason"] > 0


def test_admission_by_hbm_budget():
    s = make(LATS[1])
    need = A.Batch.footprint(s["n_pd"], s["n_h"])
    with A.Batch(precision="mixed", hbm_budget_bytes=int(need * 1.5)) as B:
        a = B.add(**s)
        assert B.mem_used() == need
        with pytest.raises(A.A2CError) as e:
            B.add(**s)
        assert e.value.code == A.A2C_ENOMEM
        B.remove(a)
        assert B.mem_used() == 0
        B.add(**s)


# ---------------------------------------------------------------- hipGraph replay
@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_graph_replay_matches_direct_issue_bitwise(precision):
    """Recurring launch sequences replayed from captured hipGraphs give the
    same relaxations, bit for bit, as direct issue of every launch."""
    specs = [make(s) for s in LATS[:3]]
    opts = dict(gatol=1e-9, grtol=0.0, max_it=3000)
    out = {}
    for graphs in (False, True):
        with A.Batch(precision=precision, graphs=graphs) as B:
            slots = [B.add(**s) for s in specs]
            for sl in slots:
                B.start(sl, **opts)
            B.run()
            out[graphs] = ([B.report(sl) for sl in slots], [B.get_state(sl)[0] for sl in slots], B.graph_stats())
    (r0, X0, st0), (r1, X1, st1) = out[False], out[True]
    assert st0["replays"] == 0 and st0["captures"] == 0
    # steady state runs from graphs: each captured sequence recurs several times
    assert st1["replays"] > 4 * st1["captures"] > 0, st1
    assert st1["replays"] > st1["direct"], st1
    for a, b, xa, xb in zip(r0, r1, X0, X1):
        assert a == b and a["reason"] > 0
        assert np.array_equal(xa, xb)


def test_graph_replays_of_one_evaluation_with_direct_work_between():
    """The same fused evaluation replayed several times, with other launches
    and synchronisations between the replays: every replay equals the
    directly issued result."""
    specs = [make(s) for s in LATS]
    with A.Batch(precision="mixed", graphs=False) as B0:
        s0 = [B0.add(**s) for s in specs]
        f_ref = B0.eval(s0)
        G_ref = [B0.get_state(sl)[1] for sl in s0]
    with A.Batch(precision="mixed", graphs=True) as B:
        slots = [B.add(**s) for s in specs]
        for rep in range(5):
            f = B.eval(slots)
            assert np.array_equal(f, f_ref), rep
            for sl, G in zip(slots, G_ref):
                assert np.array_equal(B.get_state(sl)[1], G), rep
            B.eval(slots[:2], precision="fp64")  # a different sequence between the replays
        st = B.graph_stats()
        assert st["captures"] >= 1 and st["replays"] >= 3, st


def test_single_lattice_minimizer_same_with_and_without_graphs(monkeypatch):
    s = make(LATS[1])
    res = []
    for env in ("1", "0"):
        monkeypatch.setenv("A2C_GRAPHS", env)
        with A.FreeEntropy(s["n_pd"], s["n_h"], s["box"], s["periodic"], T=s["T"], precision="mixed") as fe:
            fe.set_occupancy(s["x"])
            fe.findNeighbors(s["X"])
            res.append(fe.minimizeFreeEntropy(s["X"], gatol=1e-9, grtol=0.0))
    (Xa, ra), (Xb, rb) = res
    assert ra == rb and np.array_equal(Xa, Xb)


def test_stepping_subsets_gives_the_alone_trajectories():
    """Slots stepped in uneven subsets (as a size-tiered scheduler does) end
    with the same reports and states, bit for bit, as relaxed alone."""
    specs = [make(s) for s in LATS]
    opts = dict(gatol=1e-9, grtol=0.0, max_it=8000)
    alone = []
    for s in specs:
        with A.FreeEntropy(s["n_pd"], s["n_h"], s["box"], s["periodic"], T=s["T"], precision="mixed") as fe:
            fe.set_occupancy(s["x"])
            fe.findNeighbors(s["X"])
            alone.append(fe.minimizeFreeEntropy(s["X"], **opts))
    with A.Batch(precision="mixed") as B:
        slots = [B.add(**s) for s in specs]
        for sl in slots:
            B.start(sl, **opts)
        tick = 0
        while True:
            due = [sl for k, sl in enumerate(slots) if tick % (k + 1) == 0]  # slot k every (k+1)-th tick
            running = B.step_slots(due, 1 + tick % 3)
            tick += 1
            if running == 0:
                break
        for sl, (Xa, ra) in zip(slots, alone):
            assert B.report(sl) == ra
            assert np.array_equal(B.get_state(sl)[0], Xa)
    with A.Batch(precision="mixed") as B:
        with pytest.raises(A.A2CError):
            B.step_slots([5], 1)  # not a used slot
