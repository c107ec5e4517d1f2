"""CPU-side checks of the product library: it loads, exports every symbol the
public headers declare, validates descriptors before touching a GPU, and its
specimen generator is the documented one."""
import ctypes as C

import numpy as np
import pytest

import a2c_amd as A


def test_library_exports_every_declared_symbol():
    raw = C.CDLL(A.LIB_PATH)
    names = A.exported_symbols()
    assert len(names) >= 25
    missing = [n for n in names if not hasattr(raw, n)]
    assert not missing, missing


def test_philox_known_answers():
    """Random123 known-answer vectors for Philox4x32-10."""
    assert A.philox4x32_10([0, 0, 0, 0], [0, 0]) == [0x6627E8D5, 0xE169C58D, 0xBC57AC4C, 0x9B00DBD8]
    f = 0xFFFFFFFF
    assert A.philox4x32_10([f, f, f, f], [f, f]) == [0x408F276D, 0x41C83B0E, 0xA20BC7C6, 0x6D5451FD]
    assert A.philox4x32_10([0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344], [0xA4093822, 0x299F31D0]) == \
        [0xD16CFE09, 0x94FDCCEB, 0x5001E420, 0x24126EA1]


def test_specimen_layout():
    X, x, box, npd, nh = A.lattice((2, 3, 4), disp_std=0.0, sigma=(0.1, 0.1), x=(0.2, 0.2))
    assert npd == nh == 4 * 24 and X.size == 4 * (npd + nh) and x.size == nh
    assert np.allclose(box, [2 * 3.89, 3 * 3.89, 4 * 3.89])
    S = X.reshape(-1, 4)
    # site = 4*cell + basis, cell = (iz*ny + iy)*nx + ix; Pd block first
    assert np.allclose(S[1, :3], [0.5 * 3.89, 0.5 * 3.89, 0.0])
    assert np.allclose(S[4, :3], [3.89, 0, 0])           # cell 1 = ix 1
    assert np.allclose(S[8, :3], [0, 3.89, 0])           # cell 2 = iy 1
    assert np.allclose(S[npd + 0, :3], [0.5 * 3.89, 0, 0])  # first H: octahedral (1/2,0,0)
    assert np.allclose(S[:, 3], 0.1) and np.allclose(x, 0.2)


def test_specimen_statistics_and_determinism():
    X1, x1, *_ = A.lattice(8, seed=9, disp_std=0.03, sigma=(0.08, 0.15), x=(0.0, 0.6), x_clip=(0.0, 1.0))
    X2, x2, *_ = A.lattice(8, seed=9, disp_std=0.03, sigma=(0.08, 0.15), x=(0.0, 0.6), x_clip=(0.0, 1.0))
    assert np.array_equal(X1, X2) and np.array_equal(x1, x2)
    X0, *_ = A.lattice(8, seed=9, disp_std=0.0)
    disp = (X1 - X0).reshape(-1, 4)[:, :3].ravel()
    assert abs(disp.std() - 0.03) < 0.001 and abs(disp.mean()) < 0.001
    s = X1.reshape(-1, 4)[:, 3]
    assert s.min() >= 0.08 and s.max() < 0.15 and abs(s.mean() - 0.115) < 0.002
    assert x1.min() >= 0 and x1.max() < 0.6 and abs(x1.mean() - 0.3) < 0.01
    X3, *_ = A.lattice(8, seed=10, disp_std=0.03)
    assert not np.array_equal(X3, X1)


def test_slab_occupancy():
    X, x, box, npd, nh = A.lattice((16, 2, 2), x=(1e-3, 1e-3), slab_cells=8, slab_x=0.6)
    ix = (np.arange(nh) // 4) % 16
    assert np.all(x[ix < 8] == 0.6) and np.all(x[ix >= 8] == 1e-3)


def test_descriptor_validation_happens_before_device_use():
    d = A.FeDesc()
    A.lib.a2c_fe_desc_default(C.byref(d))
    assert d.rc == 5.35 and d.fd_eps == 1e-5 and d.max_neighbors == 256
    ctx = C.c_void_p()
    d.n_pd, d.n_h = 10, 10
    d.periodic[0], d.box[0] = 1, 5.0  # < 2 rc
    assert A.lib.a2c_fe_create(C.byref(d), C.byref(ctx)) == A.A2C_EINVAL
    d.periodic[0] = 0
    d.temperature = -1.0
    assert A.lib.a2c_fe_create(C.byref(d), C.byref(ctx)) == A.A2C_EINVAL
    d.temperature = 300.0
    d.max_neighbors = 100000
    assert A.lib.a2c_fe_create(C.byref(d), C.byref(ctx)) == A.A2C_EINVAL
    d.max_neighbors = 256
    d.quad_rule = 2  # only A2C_QUAD_R12 (0) and A2C_QUAD_GH3 (1) exist
    assert A.lib.a2c_fe_create(C.byref(d), C.byref(ctx)) == A.A2C_EINVAL


def test_mixed_tables_reproduce_the_closed_forms():
    """The FP32 cubic tables of the mixed path against the oracle's FP64
    closed forms and analytic derivatives, away from the C1 switch radii."""
    import oracle as O
    tab = A.eam_tables().astype(np.float64)
    h = 5.0 / tab.shape[1]
    fns = [O.RHO_PD, O.RHO_H, O.PHI_PD, O.PHI_PDH, O.PHI_H]
    r = np.linspace(1.6, 5.99, 3001)
    r = r[np.min(np.abs(r[:, None] - np.array([5.0, 5.05, 5.35])[None]), axis=1) > 0.02]
    k = np.floor((r - 1.0) / h).astype(int)
    u = (r - 1.0) / h - k
    for t, fn in enumerate(fns):
        c = tab[t, k]
        v = c[:, 0] + u * (c[:, 1] + u * (c[:, 2] + u * c[:, 3]))
        dv = (c[:, 1] + u * (2 * c[:, 2] + 3 * u * c[:, 3])) / h
        ev, ed = O.eam(fn, r), O.eam_d(fn, r)
        assert np.max(np.abs(v - ev)) <= 2e-7 * np.abs(ev).max(), t
        assert abs(np.mean(v - ev)) <= 1e-9 * np.abs(ev).max(), t
        assert np.max(np.abs(dv - ed)) <= 5e-6 * np.abs(ed).max(), t


def test_fp64_tables_reproduce_the_closed_forms():
    """The FP64 polynomial tables of the FP64 pair-owner passes against the
    oracle's FP64 closed forms and analytic derivatives over [1, 6) A, on
    both sides of the switch radii 5.0, 5.05 and 5.35 A (the radii themselves
    excluded: there the two sides take different branches).  Exponential
    functions (degree 8): values and slopes within 1e-13 of the function's
    largest magnitude.  Pd product polynomials (degree 11): within twice the
    FP64 rounding noise of the closed forms themselves, measured here as
    their change under a one-ulp change of r (~2e-11 of the largest
    magnitude near 5 A), plus 1e-13."""
    import oracle as O
    texp, tpd = A.eam_tables64()
    assert np.all(texp[:, :, 9] == 0.0)
    r = np.linspace(1.0, 6.0, 20001)[:-1]
    r = r[np.min(np.abs(r[:, None] - np.array([5.0, 5.05, 5.35])[None]), axis=1) > 1e-9]
    k = np.minimum(np.floor((r - 1.0) / 0.05).astype(int), 99)
    u = (r - 1.0) / 0.05 - k
    for tab, fns, deg in ((texp, [O.RHO_H, O.PHI_PDH, O.PHI_H], 8), (tpd, [O.RHO_PD, O.PHI_PD], 11)):
        for t, fn in enumerate(fns):
            c = tab[t, k]
            v = np.zeros_like(r)
            d = np.zeros_like(r)
            for j in range(deg, -1, -1):
                d = d * u + v
                v = v * u + c[:, j]
            d *= 20.0
            ev, ed = O.eam(fn, r), O.eam_d(fn, r)
            if deg == 8:
                vbar, dbar = 1e-13 * np.abs(ev).max(), 1e-13 * np.abs(ed).max()
            else:
                r1 = r * (1.0 + 2.0 ** -52)
                vbar = 2.0 * np.max(np.abs(O.eam(fn, r1) - ev)) + 1e-13 * np.abs(ev).max()
                dbar = 2.0 * np.max(np.abs(O.eam_d(fn, r1) - ed)) + 1e-13 * np.abs(ed).max()
                assert vbar <= 1e-10 * np.abs(ev).max() and dbar <= 1e-10 * np.abs(ed).max()
            assert np.max(np.abs(v - ev)) <= vbar, (deg, t)
            assert np.max(np.abs(d - ed)) <= dbar, (deg, t)


def test_serving_request_stream_is_seeded():
    from a2c_amd import serve
    a = serve.make_requests(20, rate=5.0)
    b = serve.make_requests(20, rate=5.0)
    assert a == b
    assert all(6 <= r["n"] <= 48 for r in a) and all(300 <= r["T"] <= 700 for r in a)
    t = [r["arrival"] for r in a]
    assert t == sorted(t) and 1.0 < 20 / t[-1] < 20.0


def test_batch_descriptor_graph_flag(monkeypatch):
    d = A.BatchDesc()
    A.lib.a2c_batch_desc_default(C.byref(d))
    assert d.graphs == 0
    monkeypatch.setenv("A2C_GRAPHS", "1")
    A.lib.a2c_batch_desc_default(C.byref(d))
    assert d.graphs == 1
    d.graphs = 2  # only 0 and 1; rejected before any device use
    b = C.c_void_p()
    assert A.lib.a2c_batch_create(C.byref(d), C.byref(b)) == A.A2C_EINVAL


def test_tune_table_get_set_and_validation():
    """The lanes-per-site table is host state: no GPU needed to read, set or
    reject it."""
    t0 = A.tune_get()
    assert t0["max_sites"][-1] == 2**31 - 1 and all(k in (1, 2, 4, 8, 16) for k in t0["lanes_mixed"] + t0["lanes_fp64"])
    try:
        A.tune_set({"max_sites": [1000, 2**31 - 1], "lanes_mixed": [4, 2], "lanes_fp64": [8, 1]})
        assert A.tune_get() == {"max_sites": [1000, 2**31 - 1], "lanes_mixed": [4, 2], "lanes_fp64": [8, 1]}
        for bad in ({"max_sites": [10], "lanes_mixed": [3], "lanes_fp64": [1]},             # K not a power of 2
                    {"max_sites": [10], "lanes_mixed": [32], "lanes_fp64": [1]},            # K above 16
                    {"max_sites": [100, 50], "lanes_mixed": [1, 1], "lanes_fp64": [1, 1]},  # not increasing
                    {"max_sites": [], "lanes_mixed": [], "lanes_fp64": []}):
            with pytest.raises(A.A2CError):
                A.tune_set(bad)
        assert A.tune_get()["lanes_mixed"] == [4, 2]  # a rejected table changes nothing
    finally:
        A.tune_set(None)
    assert A.tune_get() == t0
