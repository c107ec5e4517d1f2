"""Scheduling policies of the serving loop (a2c_amd.serve.due_set), on the
host: which running lattices a tick steps, and, in a simulation that charges
each tick the loop's own iteration-cost model, that no request starves."""
import pytest

from a2c_amd import serve as S


def req(i, n, arrival):
    return {"id": i, "n": n, "arrival": arrival}


def test_lockstep_steps_everything_and_unknown_policies_are_refused():
    running = {0: req(0, 6, 0.0), 3: req(1, 48, 0.1), 7: req(2, 20, 0.2)}
    assert sorted(S.due_set("lockstep", running, {})) == [0, 3, 7]
    with pytest.raises(ValueError):
        S.due_set("shortest", running, {})
    with pytest.raises(ValueError):
        S.serve([], policy="shortest")


def test_deadline_order_and_tick_budget():
    big = req(0, 48, 0.0)  # deadline 0 + 3 x 150 x 8.36 ms = 3.76 s
    small = [req(i, 6, 0.5 + 0.01 * i) for i in range(1, 6)]  # deadlines ~0.69 s
    running = {10: big, **{10 + r["id"]: r for r in small}}
    due = S.due_set("deadline", running, {})
    # the small requests come first; the big one does not fit behind them
    assert due == [11, 12, 13, 14, 15]
    # arrival order puts the big one first and lets small ones ride along
    due = S.due_set("fcfs", running, {})
    assert due[0] == 10 and sorted(due) == [10, 11, 12, 13, 14, 15]
    # once its deadline is near, later small arrivals queue behind the big one
    late = {10: big, **{20 + i: req(20 + i, 6, 3.7 + 0.001 * i) for i in range(3)}}
    assert S.due_set("deadline", late, {})[0] == 10


def test_tick_budget_charges_the_fixed_part_once():
    head = req(0, 40, 0.0)
    riders = {i: req(i, 6, 0.001 * i) for i in range(1, 400)}
    running = {0: head, **riders}
    due = S.due_set("fcfs", running, {})
    cost = lambda r: S.iter_cost(8 * r["n"] ** 3)
    cap = max((1.0 + S.RIDE_ALONG) * cost(head), S.TICK_MIN_S)
    spent = S.ITER_FIXED_S + sum(cost(running[s]) - S.ITER_FIXED_S for s in due)
    assert due[0] == 0 and spent <= cap
    nxt = max(due) + 1  # the first request left out would not have fitted
    assert nxt in running and spent + cost(running[nxt]) - S.ITER_FIXED_S > cap
    assert len(due) > 100  # small requests ride along by their per-site part


def test_fair_gives_equal_gpu_time():
    running = {0: req(0, 48, 0.0), 1: req(1, 6, 0.0), 2: req(2, 20, 0.0)}
    vt = {0: 0.0, 1: 0.0, 2: 0.0}
    for _ in range(5000):
        S.due_set("fair", running, vt)
    step = max(S.iter_cost(8 * r["n"] ** 3) for r in running.values())
    assert max(vt.values()) - min(vt.values()) <= step


def simulate(policy, reqs, iters):
    """Tick-by-tick run of the serving loop under the cost model: a tick
    lasts the fixed part plus the per-site parts of the lattices it steps.
    Returns each request's finish time."""
    t, pending, running, left, vt, finish = 0.0, list(reqs), {}, {}, {}, {}
    while pending or running:
        while pending and pending[0]["arrival"] <= t:
            r = pending.pop(0)
            running[r["id"]] = r
            left[r["id"]] = iters[r["id"]]
            vt[r["id"]] = min(vt.values()) if vt else 0.0
        if not running:
            t = pending[0]["arrival"]
            continue
        due = S.due_set(policy, running, vt)
        assert due and set(due) <= set(running)
        t += S.ITER_FIXED_S + sum(S.ITER_SITE_S * 8 * running[s]["n"] ** 3 for s in due)
        for s in due:
            left[s] -= 1
            if left[s] == 0:
                finish[s] = t
                running.pop(s)
                vt.pop(s)
    return finish


@pytest.mark.parametrize("policy", S.POLICIES)
def test_no_request_starves_under_a_stream_of_small_ones(policy):
    """A 48^3-cell request at t = 0 under 8 s of 6^3-cell arrivals, one per
    4 ms (under the deadline policy's tick budget the small ones alone load
    the server to about 60%)."""
    reqs = [req(0, 48, 0.0)] + [req(i, 6, 0.004 * i) for i in range(1, 2000)]
    iters = {0: 300, **{i: 150 for i in range(1, 2000)}}
    finish = simulate(policy, reqs, iters)
    assert len(finish) == len(reqs)
    big = S.iter_cost(8 * 48 ** 3)
    assert finish[0] >= 300 * big  # no faster than alone
    if policy == "deadline":
        # from its deadline on: small requests already past theirs finish first
        # (150 ticks of at most TICK_MIN_S), then it heads every tick, each
        # tick within its budget
        assert finish[0] <= S.deadline(reqs[0]) + 150 * S.TICK_MIN_S + 300 * (1.0 + S.RIDE_ALONG) * big
        assert finish[0] < simulate("fair", reqs, iters)[0]
    if policy == "fcfs":
        assert finish[0] <= 300 * (1.0 + S.RIDE_ALONG) * big
    if policy == "fair":
        small = sorted(finish[i] - reqs[i]["arrival"] for i in range(1, 2000))
        assert small[-1] < 0.5  # equal GPU time: no small request waits on the big one


@pytest.mark.parametrize("policy", S.POLICIES)
def test_incremental_scheduler_matches_due_set(policy):
    """The serving loop's Scheduler (costs and order kept from admission)
    picks what due_set picks from scratch, through admissions and
    retirements of requests of every size."""
    import numpy as np
    rng = np.random.default_rng(7)
    sched, running, vt, nxt, t = S.Scheduler(policy), {}, {}, 0, 0.0
    for tick in range(3000):
        if rng.random() < 0.3 or not running:  # an arrival
            t += rng.exponential(0.02)
            r = req(nxt, int(round(np.exp(rng.uniform(np.log(6), np.log(48))))), t)
            sched.add(nxt, r)
            running[nxt] = r
            vt[nxt] = min(vt.values()) if vt else 0.0
            nxt += 1
        due = sched.due()
        assert due == S.due_set(policy, running, vt), tick
        if policy == "fair":
            assert sched.vt == vt
        for sl in due:  # some of the stepped ones finish
            if rng.random() < 0.05:
                sched.remove(sl)
                running.pop(sl)
                vt.pop(sl)
