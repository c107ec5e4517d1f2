"""a2c_fe_diffuse_device_noalloc, the diffusion step that allocates nothing
(its scratch is a library-static device array), against a2c_fe_diffuse on
the same inputs: the new occupancy, the balance and the context occupancy it
leaves behind are bitwise equal, on two calls in a row with the second on a
smaller lattice (the static scratch keeps the first call's contents); and
n_h beyond the static scratch is rejected before any device work."""
import ctypes as C

import numpy as np
import pytest

import a2c_amd as A
from helpers import specimen

pytestmark = pytest.mark.gpu

STEP = dict(dt=5e-13, nu=1e13, Qm=0.23, xLb=1e-3, xUb=1 - 1e-16)


def inputs(n, seed):
    X, x, box, per, npd, nh = specimen(n, seed=seed, disp=0.02, sigma=(0.08, 0.15), x=(1e-3, 0.6))
    x[::7] = 1e-3  # sites at the clamp: those that lose H are lifted back
    gamma = np.random.default_rng(seed).normal(0.0, 2.0, nh)
    full = np.zeros(nh, np.int32)
    full[::17] = 1  # reservoir sites keep their occupancy
    return X, x, box, per, npd, nh, gamma, full


def reference(X, x, box, per, npd, nh, gamma, full):
    """Two host-array steps: the first is the one compared, the second shows
    which occupancy the first left in the context."""
    with A.FreeEntropy(npd, nh, box, per, T=400.0) as fe:
        fe.set_occupancy(x)
        fe.findNeighbors(X)
        x1 = fe.integrateDiffusion(STEP["dt"], STEP["nu"], STEP["Qm"], STEP["xLb"], STEP["xUb"], gamma, full)
        b1 = fe.diffusion_balance()
        x2 = fe.integrateDiffusion(STEP["dt"], STEP["nu"], STEP["Qm"], STEP["xLb"], STEP["xUb"], gamma, full)
    return x1, b1, x2


def noalloc(X, x, box, per, npd, nh, gamma, full):
    with A.FreeEntropy(npd, nh, box, per, T=400.0) as fe:
        fe.set_occupancy(x)
        fe.findNeighbors(X)
        dg, dfh, dx, db = fe.dev_alloc(8 * nh), fe.dev_alloc(4 * nh), fe.dev_alloc(8 * nh), fe.dev_alloc(32)
        try:
            fe.upload(dg, gamma)
            fe.upload(dfh, full)
            print("noalloc call: n_h = %d" % nh, flush=True)
            rc = A.lib.a2c_fe_diffuse_device_noalloc(fe._ctx, STEP["dt"], STEP["nu"], STEP["Qm"], STEP["xLb"],
                                                     STEP["xUb"], dg, dfh, dx, db)
            fe._check(rc)
            fe.sync()
            x1 = fe.download(np.zeros(nh), dx)
            b = fe.download(np.zeros(4), db)
        finally:
            for p in (dg, dfh, dx, db):
                fe.dev_free(p)
        # the context now holds x1: a host-array step from it
        x2 = fe.integrateDiffusion(STEP["dt"], STEP["nu"], STEP["Qm"], STEP["xLb"], STEP["xUb"], gamma, full)
    return x1, dict(zip(("flux_sum", "flux_gross", "clamp_lift", "reservoir_supply"), b)), x2


def test_noalloc_diffusion_equals_host_entry_point_twice():
    for n, seed in ((5, 11), (3, 12)):  # the second call on a smaller lattice
        inp = inputs(n, seed)
        x1r, b1r, x2r = reference(*inp)
        x1, b1, x2 = noalloc(*inp)
        assert np.array_equal(x1, x1r), n
        assert b1 == b1r, (n, b1, b1r)
        assert np.array_equal(x2, x2r), n
        # the inputs exercise the clamp and the reservoirs; the balance closes
        assert b1r["clamp_lift"] > 0.0 and b1r["reservoir_supply"] != 0.0
        assert abs(b1r["flux_sum"] - b1r["reservoir_supply"]) <= 1e-12 * b1r["flux_gross"]


def test_noalloc_diffusion_rejects_sizes_beyond_its_scratch():
    nh = 1048576 + 1
    with A.FreeEntropy(0, nh) as fe:
        rc = A.lib.a2c_fe_diffuse_device_noalloc(fe._ctx, 1e-13, 1e13, 0.23, 1e-3, 1.0, None, None, None, None)
        assert rc == -1
        assert b"static scratch" in A.lib.a2c_fe_last_error(fe._ctx)
