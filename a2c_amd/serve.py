"""Continuous-batching relaxation server on one GPU.

Requests (a lattice to relax) arrive as a Poisson stream.  The loop keeps
every admitted lattice resident in HBM in one a2c_batch and advances running
relaxations together, one LMVM iteration per tick (one fused launch per
evaluation round for all lattices stepped).  Between ticks the loop admits
the requests that have arrived, as long as their device footprint fits the
HBM budget, and retires the converged ones.  Latency = finish - arrival.

Which lattices a tick steps is the policy's choice (due_set):

* "deadline" (default): earliest deadline first, the deadline of a request
  being its arrival plus DEADLINE_SLACK times an estimate of its relaxation
  alone (EST_ITERS iterations at its iteration cost).  A tick steps the most
  urgent lattice and, behind it in deadline order, every lattice that fits a
  tick budget of (1 + RIDE_ALONG) times the head's iteration cost (at least
  TICK_MIN_S); the fixed part of an iteration (launches, host round trips)
  is paid once per tick, so a lattice that rides along costs only its
  per-site part.  Small requests finish early, and a large one is not
  overtaken by later arrivals once its deadline is near.
* "fcfs": the same with arrival order (deadline slack 0).
* "fair": stride scheduling: each lattice's virtual time advances by its
  iteration cost whenever it is stepped, and a tick steps every lattice
  within one smallest-iteration cost of the least virtual time (each request
  receives about equal GPU time).
* "lockstep": every running lattice on every tick.

A lattice's trajectory is the same whichever lattices share its ticks, so a
policy changes only when results arrive, not what they are.

Workload (defaults): edge n log-uniform in [6, 48] cells (n^3 cells, 8 sites
per cell), T ~ U(300, 700) K, mean H fraction ~ U(0, 0.6) with iid noise
+-0.05, displacements N(0, 0.02 A), sigma 0.10 A, periodic, seed 4 + request
id.  The comparator relaxes the same requests one at a time in FP64.

    python -m a2c_amd.serve --rates 5,10,20,40 --requests 200
"""
import argparse
import bisect
import json
import math
import sys
import time

import numpy as np

if __package__:
    from . import Batch, FreeEntropy, lattice
else:  # run as a script: python a2c_amd/serve.py
    import os
    import sys
    sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
    from a2c_amd import Batch, FreeEntropy, lattice


def make_requests(n, rate, seed=4, nmin=6, nmax=48):
    rng = np.random.default_rng(seed)
    t = np.cumsum(rng.exponential(1.0 / rate, size=n))
    reqs = []
    for i in range(n):
        r = np.random.default_rng(seed + 1000 + i)
        edge = int(round(math.exp(r.uniform(math.log(nmin), math.log(nmax)))))
        T = float(r.uniform(300.0, 700.0))
        xbar = float(r.uniform(0.0, 0.6))
        reqs.append({"id": i, "arrival": float(t[i]), "n": edge, "T": T, "xbar": xbar, "seed": seed + i})
    return reqs


def materialize(req):
    lo, hi = max(req["xbar"] - 0.05, 1e-3), min(req["xbar"] + 0.05, 1 - 1e-16)
    X, x, box, npd, nh = lattice(req["n"], seed=req["seed"], disp_std=0.02, sigma=(0.10, 0.10), x=(lo, hi),
                                 x_clip=(1e-3, 1 - 1e-16))
    return X, x, box, npd, nh


# Cost of one LMVM iteration of a lattice (mixed precision, in a batch): a
# fixed part (~50 launches and 3 host round trips) and a per-site part.
ITER_FIXED_S = 4e-4
ITER_SITE_S = 9e-9


def iter_cost(sites):
    return ITER_FIXED_S + ITER_SITE_S * sites


POLICIES = ("deadline", "fcfs", "fair", "lockstep")
DEADLINE_SLACK = 3.0  # deadline = arrival + slack x estimated relaxation alone
EST_ITERS = 150  # LMVM iterations of a typical request of this workload
RIDE_ALONG = 0.5  # tick budget: (1 + RIDE_ALONG) x the head's iteration cost ...
TICK_MIN_S = 5e-3  # ... and at least this


def deadline(req, slack=DEADLINE_SLACK):
    return req["arrival"] + slack * EST_ITERS * iter_cost(8 * req["n"] ** 3)


def due_set(policy, running, vt):
    """Slots one tick steps.  running: slot -> request (with "arrival" and
    "n"); vt: slot -> virtual time, advanced here for "fair"."""
    cost = {sl: iter_cost(8 * r["n"] ** 3) for sl, r in running.items()}
    if policy == "lockstep":
        return list(running)
    if policy == "fair":
        horizon = min(vt[sl] for sl in running) + min(cost.values())
        due = [sl for sl in running if vt[sl] <= horizon]
        for sl in due:
            vt[sl] += cost[sl]
        return due
    if policy not in ("deadline", "fcfs"):
        raise ValueError("unknown policy %r (one of %s)" % (policy, ", ".join(POLICIES)))
    slack = DEADLINE_SLACK if policy == "deadline" else 0.0
    order = sorted(running, key=lambda sl: (deadline(running[sl], slack), sl))
    cap = max((1.0 + RIDE_ALONG) * cost[order[0]], TICK_MIN_S)
    due, acc = [], ITER_FIXED_S
    for sl in order:
        c = cost[sl] - ITER_FIXED_S  # the fixed part is the tick's, paid once
        if not due or acc + c <= cap:
            due.append(sl)
            acc += c
    return due


class Scheduler:
    """due_set kept up to date incrementally, for the serving loop: a
    request's iteration cost and deadline are computed once at admission and
    the deadline (or arrival) order is a sorted list updated at admission and
    retirement, so a tick sorts nothing.  due() returns what due_set returns
    for the same running set (the walk stops once no running lattice can
    fit the budget any more)."""

    def __init__(self, policy):
        if policy not in POLICIES:
            raise ValueError("unknown policy %r (one of %s)" % (policy, ", ".join(POLICIES)))
        self.policy = policy
        self.slack = DEADLINE_SLACK if policy == "deadline" else 0.0
        self.cost = {}  # slot -> iteration cost, in admission order
        self.vt = {}  # slot -> virtual time ("fair")
        self.order = []  # sorted (deadline, slot) ("deadline", "fcfs")
        self.key = {}
        self.min_part = math.inf  # lower bound of every running lattice's per-site part

    def add(self, slot, req):
        c = iter_cost(8 * req["n"] ** 3)
        self.cost[slot] = c
        self.min_part = min(self.min_part, c - ITER_FIXED_S)
        if self.policy == "fair":
            self.vt[slot] = min(self.vt.values()) if self.vt else 0.0
        elif self.policy != "lockstep":
            k = (deadline(req, self.slack), slot)
            self.key[slot] = k
            bisect.insort(self.order, k)

    def remove(self, slot):
        del self.cost[slot]
        self.vt.pop(slot, None)
        k = self.key.pop(slot, None)
        if k is not None:
            del self.order[bisect.bisect_left(self.order, k)]

    def due(self):
        if self.policy == "lockstep":
            return list(self.cost)
        if self.policy == "fair":
            horizon = min(self.vt.values()) + min(self.cost.values())
            due = [sl for sl in self.cost if self.vt[sl] <= horizon]
            for sl in due:
                self.vt[sl] += self.cost[sl]
            return due
        head = self.order[0][1]
        cap = max((1.0 + RIDE_ALONG) * self.cost[head], TICK_MIN_S)
        due, acc = [], ITER_FIXED_S
        for _, sl in self.order:
            c = self.cost[sl] - ITER_FIXED_S
            if not due or acc + c <= cap:
                due.append(sl)
                acc += c
            if acc + self.min_part > cap:
                break
        return due


def serve(reqs, precision="mixed", budget=200 << 30, max_lattices=256, opts=None, log=None, policy="deadline",
          keep_state=False):
    """Serve reqs (sorted by arrival) under one of POLICIES.  keep_state:
    each finished request also carries its relaxed state ("X")."""
    if policy not in POLICIES:
        raise ValueError("unknown policy %r (one of %s)" % (policy, ", ".join(POLICIES)))
    opts = opts or {"gatol": 1e-8, "grtol": 0.0}
    sched = Scheduler(policy)
    pending = list(reqs)
    running = {}  # slot -> request
    done = []
    t0 = time.perf_counter()
    with Batch(precision=precision, max_lattices=max_lattices, hbm_budget_bytes=budget) as B:
        while pending or running:
            now = time.perf_counter() - t0
            # admit arrived requests while the footprint fits
            while pending and pending[0]["arrival"] <= now:
                req = pending[0]
                n_sites = 8 * req["n"] ** 3
                if B.mem_used() + Batch.footprint(n_sites // 2, n_sites // 2) > budget and running:
                    break
                X, x, box, npd, nh = req["state"] if "state" in req else materialize(req)
                slot = B.add(X, x, npd, nh, box, (1, 1, 1), T=req["T"])
                B.start(slot, **opts)
                sched.add(slot, req)
                req["admitted"] = time.perf_counter() - t0
                running[slot] = req
                pending.pop(0)
            if not running:
                time.sleep(max(0.0, pending[0]["arrival"] - (time.perf_counter() - t0)))
                continue
            due = sched.due()
            B.step_slots(due, 1)
            for slot in due:  # only a stepped relaxation can have finished
                rep = B.report(slot)
                if rep["reason"] != 0:
                    req = running.pop(slot)
                    req["finish"] = time.perf_counter() - t0
                    req["latency"] = req["finish"] - req["arrival"]
                    req["report"] = rep
                    if keep_state:
                        req["X"] = B.get_state(slot)[0]
                    done.append(req)
                    B.remove(slot)
                    sched.remove(slot)
                    if log:
                        log(req)
    wall = time.perf_counter() - t0
    return done, wall


def serve_one_at_a_time(reqs, precision="fp64", opts=None):
    """Comparator: requests in arrival order, one lattice at a time."""
    opts = opts or {"gatol": 1e-8, "grtol": 0.0}
    t0 = time.perf_counter()
    done = []
    for req in reqs:
        now = time.perf_counter() - t0
        if req["arrival"] > now:
            time.sleep(req["arrival"] - now)
        X, x, box, npd, nh = req["state"] if "state" in req else materialize(req)
        with FreeEntropy(npd, nh, box, (1, 1, 1), T=req["T"], precision=precision) as fe:
            fe.set_occupancy(x)
            fe.findNeighbors(X)
            _, rep = fe.minimizeFreeEntropy(X, **opts)
        r = dict(req)
        r["finish"] = time.perf_counter() - t0
        r["latency"] = r["finish"] - r["arrival"]
        r["report"] = rep
        done.append(r)
    return done, time.perf_counter() - t0


def service_times(reqs, precision="fp64", opts=None):
    """Relax the requests one at a time, back to back (no arrivals): the
    service time and report of each, in request order."""
    opts = opts or {"gatol": 1e-8, "grtol": 0.0}
    out = []
    for req in reqs:
        X, x, box, npd, nh = req["state"] if "state" in req else materialize(req)
        t0 = time.perf_counter()
        with FreeEntropy(npd, nh, box, (1, 1, 1), T=req["T"], precision=precision) as fe:
            fe.set_occupancy(x)
            fe.findNeighbors(X)
            _, rep = fe.minimizeFreeEntropy(X, **opts)
        out.append((time.perf_counter() - t0, rep))
    return out


def fifo_latencies(arrivals, service):
    """Latencies of a single first-come-first-served server with the given
    arrival times and service times (one at a time, no overlap)."""
    t, lat = 0.0, []
    for a, sv in zip(arrivals, service):
        t = max(t, a) + sv
        lat.append(t - a)
    return np.array(lat)


def sweep(rates, n_requests, fp64_sample=None, log=None, opts=None, policy="deadline"):
    """Serve the same n_requests (sizes, temperatures, occupancies) at each
    Poisson rate with continuous batching (mixed precision, FP64 endgame).
    The comparator relaxes the requests one at a time in FP64, back to back
    (all of them, or the first fp64_sample): the mean service time gives its
    capacity, and, when every request was relaxed, a first-come-first-served
    replay of each request's own measured service time on each rate's
    arrivals gives its p50/p99.  relaxed_f_rel_vs_fp64: the largest relative
    difference of the served objectives from the FP64 ones."""
    opts = opts or {"gatol": 1e-8, "grtol": 0.0}
    base = make_requests(n_requests, 1.0)
    states = [materialize(r) for r in base]  # the clients' payloads, built before the clock starts
    k = n_requests if fp64_sample is None else min(fp64_sample, n_requests)
    sv = service_times([dict(r, state=st) for r, st in zip(base[:k], states[:k])], opts=opts)
    f64 = [rep["f"] for _, rep in sv]
    svt = np.array([t for t, _ in sv])
    rows = []
    for rate in rates:
        reqs = make_requests(n_requests, rate)
        for r, st in zip(reqs, states):
            r["state"] = st
        done, wall = serve(reqs, opts=opts, log=log, policy=policy)
        row = {"rate": rate, "batched_mixed": summarize(done, wall)}
        byid = {d["id"]: d for d in done}
        row["relaxed_f_rel_vs_fp64"] = max(abs(byid[i]["report"]["f"] - f64[i]) / abs(f64[i]) for i in range(k))
        if k == n_requests:
            lat = fifo_latencies([r["arrival"] for r in reqs], svt)
            row["one_at_a_time_fp64"] = {"p50_s": float(np.percentile(lat, 50)), "p99_s": float(np.percentile(lat, 99))}
        rows.append(row)
        if log:
            log({"rate": rate, "summary": row["batched_mixed"]})
    sites = [8 * r["n"] ** 3 for r in base]
    return {
        "requests_per_rate": n_requests, "sites_mean": float(np.mean(sites)), "sites_max": int(max(sites)),
        "policy": policy, "rates": rows,
        "saturation_relaxations_per_s": max(r["batched_mixed"]["relaxations_per_s"] for r in rows),
        "one_at_a_time_fp64": {"relaxed": k, "service_s_mean": float(svt.mean()),
                               "relaxations_per_s": float(1.0 / svt.mean()),
                               "converged": sum(1 for _, rep in sv if rep["reason"] > 0)},
    }


def summarize(done, wall):
    lat = np.array([d["latency"] for d in done])
    conv = sum(1 for d in done if d["report"]["reason"] > 0)
    return {"requests": len(done), "converged": conv, "wall_s": wall, "relaxations_per_s": len(done) / wall,
            "p50_s": float(np.percentile(lat, 50)), "p99_s": float(np.percentile(lat, 99)),
            "sites": int(sum(8 * d["n"] ** 3 for d in done))}


def main():
    p = argparse.ArgumentParser()
    p.add_argument("--rates", default="5,10,20,40")
    p.add_argument("--requests", type=int, default=200)
    p.add_argument("--fp64-sample", type=int, default=None, help="requests relaxed in FP64 (default: all)")
    p.add_argument("--policy", default="deadline", choices=POLICIES)
    a = p.parse_args()
    note = lambda r: print(json.dumps(r) if "summary" in r else
                           "request %d (n=%d) done, latency %.3f s, %d iterations"
                           % (r["id"], r["n"], r["latency"], r["report"]["iterations"]), file=sys.stderr, flush=True)
    out = sweep([float(r) for r in a.rates.split(",")], a.requests, a.fp64_sample, log=note, policy=a.policy)
    print(json.dumps(out), flush=True)
    return out


if __name__ == "__main__":
    main()
