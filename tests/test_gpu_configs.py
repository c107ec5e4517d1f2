"""Full-size parity of the HIP path on the four large workloads, against the
FP64 oracle (run on the host cores beside the GPU).

single 40^3 cells (512,000 sites): neighbor sets, PdPdnum and HH1 byte-
    identical to the oracle's cell-list build; the FP64 kernel within 1e-10
    of the oracle's objective (analytic derivatives) and 1e-8 of its largest
    gradient component; the mixed path within 1e-7 of the oracle's objective.
batched 64 lattices of 8^3..24^3 cells relaxed in lockstep: every slot
    converges; a sample of slots gives bitwise the same report and state
    when relaxed alone; their end states are stationary for the oracle.
serving requests from 6^3 to 48^3 cells (one of 48^3) served with continuous
    batching: each result is bitwise the one of the request relaxed alone,
    and a sample is stationary for the oracle.
dmd 64^3 cells (2,097,152 sites), two coupled steps: the chemical potential
    and the diffusion step match the oracle at each relaxed state, and the
    step's H balance is the clamp lift, with the net flux at round-off.
"""
import os

import numpy as np
import pytest

import a2c_amd as A
import oracle as O
from a2c_amd import serve, workloads

pytestmark = [pytest.mark.gpu, pytest.mark.timeout(600)]

THREADS = max(1, min(16, os.cpu_count() or 1))
LISTS = ["PdPd", "PdH", "HPd", "HH", "HH1"]


def say(*a):
    print(*a, flush=True)


def oracle_for(s, **kw):
    return O.Oracle(s["n_pd"], s["n_h"], s["box"], s["periodic"], T=s["T"], threads=THREADS, **kw)


def fe_for(s, precision):
    fe = A.FreeEntropy(s["n_pd"], s["n_h"], s["box"], s["periodic"], T=s["T"], precision=precision)
    fe.set_occupancy(s["x"])
    return fe


def test_single_40cubed_neighbors_and_evaluation():
    s = workloads.single(40)
    say("single: 40^3 cells, %d sites" % (s["n_pd"] + s["n_h"]))
    orc = oracle_for(s)
    mx = orc.build_neighbors(s["X"])
    with fe_for(s, "fp64") as fe:
        st = fe.findNeighbors(s["X"])
        assert st["max_neib"] == mx
        for k in LISTS:
            go, gi = fe.neighbors(k)
            oo, oi = orc.neighbors(k)
            assert np.array_equal(go, oo) and np.array_equal(gi, oi), k
        assert np.array_equal(fe.pdpdnum(), orc.pdpdnum())
        say("single: neighbor sets identical (%s entries)" % st["total"])
        f64, G64 = fe.eval(s["X"])
        fe.set_precision("mixed")
        f32, G32 = fe.eval(s["X"])
    fo, Go = orc.eval(s["X"], s["x"])
    say("single: f fp64 %.15g mixed %.15g oracle %.15g" % (f64, f32, fo))
    assert abs(f64 - fo) <= 1e-10 * abs(fo)
    assert np.max(np.abs(G64 - Go)) <= 1e-8 * np.max(np.abs(Go))
    assert abs(f32 - fo) <= 1e-7 * abs(fo)
    assert np.linalg.norm(G32 - Go) <= 1e-5 * np.linalg.norm(Go)


def stationary(s, Xr, rep):
    """The oracle's objective at the relaxed state equals the reported one, and
    its gradient norm is at most the device's final FP64 gradient norm plus
    1e-9 (the two FP64 evaluation orders differ by far less)."""
    orc = oracle_for(s)
    orc.build_neighbors(s["X"])
    fo, Go = orc.eval(Xr, s["x"])
    g = float(np.linalg.norm(Go))
    say("  oracle: f %.15g (device %.15g), |g| %.3e (device %.3e)" % (fo, rep["f"], g, rep["gnorm"]))
    assert abs(rep["f"] - fo) <= 1e-10 * abs(fo)
    assert g <= rep["gnorm"] + 1e-9


def test_batched_64_lattices_alone_equals_mix_and_stationary():
    specs = workloads.batched(64)
    opts = dict(gatol=1e-8, grtol=0.0)
    say("batched: 64 lattices, %d sites" % sum(s["n_pd"] + s["n_h"] for s in specs))
    with A.Batch(precision="mixed", max_lattices=64) as B:
        slots = [B.add(**s) for s in specs]
        for sl in slots:
            B.start(sl, **opts)
        B.run()
        reps = [B.report(sl) for sl in slots]
        states = [B.get_state(sl)[0] for sl in slots]
    assert all(r["reason"] > 0 for r in reps), [r["reason"] for r in reps]
    sizes = [s["n_pd"] + s["n_h"] for s in specs]
    sample = sorted({int(np.argmin(sizes)), int(np.argmax(sizes)), 17, 42})
    for i in sample:
        s = specs[i]
        say("batched: slot %d (%d sites) alone" % (i, sizes[i]))
        with fe_for(s, "mixed") as fe:
            fe.findNeighbors(s["X"])
            Xa, ra = fe.minimizeFreeEntropy(s["X"], **opts)
        assert ra == reps[i], (i, ra, reps[i])
        assert np.array_equal(Xa, states[i]), i
        stationary(s, states[i], reps[i])


def test_serving_mix_to_48cubed_equals_alone_and_stationary():
    reqs = serve.make_requests(8, rate=20.0)
    reqs[0]["n"], reqs[2]["n"] = 6, 48  # the extremes of the request sizes
    opts = {"gatol": 1e-8, "grtol": 0.0}
    say("serving: edges %s" % [r["n"] for r in reqs])
    done, wall = serve.serve([dict(r) for r in reqs], precision="mixed", opts=opts, keep_state=True)
    assert len(done) == len(reqs)
    overlap = sum(1 for d in done if any(o["admitted"] < d["admitted"] < o["finish"] for o in done if o is not d))
    assert overlap > 0
    for d in sorted(done, key=lambda r: r["id"]):
        assert d["report"]["reason"] > 0, d["report"]
        X, x, box, npd, nh = serve.materialize(d)
        s = dict(X=X, x=x, box=box, periodic=(1, 1, 1), n_pd=npd, n_h=nh, T=d["T"])
        say("serving: request %d (%d^3 cells) alone" % (d["id"], d["n"]))
        with fe_for(s, "mixed") as fe:
            fe.findNeighbors(X)
            Xa, ra = fe.minimizeFreeEntropy(X, **opts)
        assert ra == d["report"], (d["id"], ra, d["report"])
        assert np.array_equal(Xa, d["X"]), d["id"]
        if d["n"] in (6, 48):
            stationary(s, d["X"], d["report"])


def test_dmd_64cubed_two_steps_match_oracle():
    s = workloads.dmd(64)
    npd, nh = s["n_pd"], s["n_h"]
    X, x = s["X"].copy(), s["x"].copy()
    st = workloads.DMD_STEP
    prm = A.dmd_params(**st, lmvm=workloads.dmd_lmvm(npd + nh))
    g = np.zeros(nh)
    fh = np.zeros(nh, np.int32)
    orc = oracle_for(s)
    with fe_for(s, "mixed") as fe:
        fe.findNeighbors(X)
        for step in range(2):
            x_before, X_start = x.copy(), X.copy()
            say("dmd: step %d" % step)
            rep = A.dmd_step(fe, prm, X, x, g, fh)
            assert rep["lmvm_reason"] > 0, rep
            orc.build_neighbors(X_start)  # the step builds its lists at its start
            go, fo = orc.chem_potential(X, x_before, np.zeros(4), np.zeros(4), st["gamma_bd"], np.zeros(nh),
                                        np.zeros(nh, np.int32))
            err = np.max(np.abs(g - go) / np.maximum(np.abs(go), 1.0))
            say("dmd: step %d gamma max err %.3e, dH %.6e, lift %.6e, flux %.3e of %.3e" % (
                step, err, rep["total_h_after"] - rep["total_h_before"], rep["clamp_lift"], rep["flux_sum"],
                rep["flux_gross"]))
            assert err <= 1e-10
            assert np.array_equal(fh, fo)
            xo = orc.diffuse(st["dt"], st["nu"], st["Qm"], st["xLb"], st["xUb"], g, fh, x_before)
            assert np.max(np.abs(x - xo)) <= 1e-13
            assert abs(rep["flux_sum"]) <= 1e-12 * rep["flux_gross"]
            xu = orc.diffuse(st["dt"], st["nu"], st["Qm"], -1e300, 1e300, g, fh, x_before)
            assert rep["clamp_lift"] == pytest.approx((xo - xu).sum(), rel=1e-9, abs=1e-12)
            dH = rep["total_h_after"] - rep["total_h_before"]
            assert abs(dH - rep["clamp_lift"]) <= 1e-12 * rep["total_h_before"]
