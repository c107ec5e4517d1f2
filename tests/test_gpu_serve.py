"""Continuous-batching server: requests arriving while others run are
admitted between iterations; every result equals the same request relaxed
alone (bitwise: per-slot kernels and reductions do not depend on the mix)."""
import numpy as np
import pytest

import a2c_amd as A
from a2c_amd import serve

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("policy", serve.POLICIES)
def test_served_results_equal_standalone(policy):
    reqs = serve.make_requests(6, rate=50.0, nmin=3, nmax=6)
    opts = {"gatol": 1e-8, "grtol": 0.0}
    print("serving 6 requests, policy", policy, flush=True)
    done, wall = serve.serve([dict(r) for r in reqs], precision="mixed", opts=opts, policy=policy)
    assert len(done) == len(reqs)
    overlap = sum(1 for d in done if any(o["admitted"] < d["admitted"] < o["finish"] for o in done if o is not d))
    assert overlap > 0  # some request was admitted while another was running
    for d in done:
        assert d["report"]["reason"] > 0
        X, x, box, npd, nh = serve.materialize(d)
        with A.FreeEntropy(npd, nh, box, (1, 1, 1), T=d["T"], precision="mixed") as fe:
            fe.set_occupancy(x)
            fe.findNeighbors(X)
            _, rep = fe.minimizeFreeEntropy(X, **opts)
        assert rep == d["report"]
    s = serve.summarize(done, wall)
    assert s["p99_s"] >= s["p50_s"] > 0
