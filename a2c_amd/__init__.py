"""a2c_amd: MI355X free-entropy engine for Gaussian phase packets (Pd-H EAM).

Python host mirror over the C-ABI of liba2c_fe.so (include/a2c_fe.h), loaded
with ctypes.  The method names follow the reference's C++ interface:

    FreeEntropy.formFunctionAndGradient  Minimizer::formFunctionAndGradient
                                         (minimizer.cpp:83-369)
    FreeEntropy.findNeighbors            findNeighbors (neighbors.cpp:8-113) +
                                         Minimizer::findLocalNeighbors
                                         (minimizer.cpp:833-1144)
    FreeEntropy.calculateChemicalPotential  minimizer.cpp:588-802
    FreeEntropy.integrateDiffusion       integrate.cpp:9-87
    FreeEntropy.minimizeFreeEntropy      Minimizer::minimizeFreeEntropy
                                         (minimizer.cpp:548-586, TAOBLMVM)

Errors are raised as A2CError carrying the library status code and message
(the reference exits the process instead).  There is no CPU fallback: if the
HIP library is missing, importing this package raises.
"""
import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "liba2c_fe.so")

A2C_OK = 0
A2C_EINVAL = -1
A2C_ENOMEM = -2
A2C_EHIP = -3
A2C_ENBR_OVERFLOW = -4
A2C_ENONFINITE = -5
A2C_ESTATE = -6
A2C_ENOTCONV = -7

DERIV = {"analytic": 0, "fd": 1}
PREC = {"fp64": 0, "mixed": 1}
QUAD = {"r12": 0, "gh3": 1}  # 12-point rule of quadrature.cpp:43-75; 27-node Gauss-Hermite
LIST_IDS = {"PdPd": 0, "PdH": 1, "HPd": 2, "HH": 3, "HH1": 4}


class A2CError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__("a2c error %d: %s" % (code, msg))
        self.code = code


class FeDesc(C.Structure):
    _fields_ = [
        ("n_pd", C.c_int32), ("n_h", C.c_int32),
        ("box", C.c_double * 3), ("periodic", C.c_int32 * 3),
        ("rc", C.c_double), ("a_pd", C.c_double), ("temperature", C.c_double),
        ("quad_rule", C.c_int32), ("deriv_mode", C.c_int32), ("fd_eps", C.c_double),
        ("precision", C.c_int32), ("device", C.c_int32), ("max_neighbors", C.c_int32),
    ]


class EamTabulated(C.Structure):
    _fields_ = [
        ("n_rho", C.c_int32), ("d_rho", C.c_double), ("n_r", C.c_int32), ("d_r", C.c_double),
        ("F_pd", C.POINTER(C.c_double)), ("F_h", C.POINTER(C.c_double)),
        ("f_pd", C.POINTER(C.c_double)), ("f_h", C.POINTER(C.c_double)),
        ("rphi_pdpd", C.POINTER(C.c_double)), ("rphi_pdh", C.POINTER(C.c_double)),
        ("rphi_hh", C.POINTER(C.c_double)),
    ]


TABLE_KEYS = ("F_pd", "F_h", "f_pd", "f_h", "rphi_pdpd", "rphi_pdh", "rphi_hh")


class NbrStats(C.Structure):
    _fields_ = [("max_neib", C.c_int32), ("width", C.c_int32 * 5), ("total", C.c_int64 * 5),
                ("refreshed", C.c_int32), ("max_displacement", C.c_double)]


class LmvmOpts(C.Structure):
    _fields_ = [
        ("m", C.c_int32), ("max_it", C.c_int32), ("max_funcs", C.c_int32),
        ("gatol", C.c_double), ("grtol", C.c_double), ("gttol", C.c_double),
        ("sigma_lb", C.c_double), ("mixed_switch", C.c_int32), ("mixed_gtol", C.c_double),
        ("sigma_ub", C.c_double), ("precond", C.c_int32),
    ]


class LmvmReport(C.Structure):
    _fields_ = [
        ("reason", C.c_int32), ("iterations", C.c_int32), ("evals", C.c_int32),
        ("evals_fp64", C.c_int32), ("f", C.c_double), ("gnorm", C.c_double),
    ]


class Timing(C.Structure):
    _fields_ = [
        ("n_kernels", C.c_int32), ("names", (C.c_char * 32) * 8),
        ("avg_ms", C.c_double * 8), ("total_ms", C.c_double),
    ]


class BatchDesc(C.Structure):
    _fields_ = [
        ("device", C.c_int32), ("precision", C.c_int32), ("deriv_mode", C.c_int32), ("fd_eps", C.c_double),
        ("max_lattices", C.c_int32), ("hbm_budget_bytes", C.c_size_t), ("max_neighbors", C.c_int32),
        ("graphs", C.c_int32),
    ]


TUNE_MAX = 8


class TuneTable(C.Structure):
    _fields_ = [("n_classes", C.c_int32), ("max_sites", C.c_int32 * TUNE_MAX),
                ("lanes_mixed", C.c_int32 * TUNE_MAX), ("lanes_fp64", C.c_int32 * TUNE_MAX)]


class AutotuneOpts(C.Structure):
    _fields_ = [("device", C.c_int32), ("n_sizes", C.c_int32), ("cells", C.c_int32 * TUNE_MAX),
                ("target_sites", C.c_int64), ("reps", C.c_int32), ("apply", C.c_int32)]


class AutotuneReport(C.Structure):
    _fields_ = [("sites", C.c_int32 * TUNE_MAX), ("copies", C.c_int32 * TUNE_MAX),
                ("ns_per_site", C.c_double * 5 * 2 * TUNE_MAX), ("us_alone", C.c_double * 5 * 2 * TUNE_MAX),
                ("latency", C.c_int32 * 2 * TUNE_MAX)]


class DmdParams(C.Structure):
    _fields_ = [
        ("dt", C.c_double), ("nu", C.c_double), ("Qm", C.c_double), ("xLb", C.c_double), ("xUb", C.c_double),
        ("gamma_bd", C.c_double), ("coe_pd", C.c_double * 4), ("coe_h", C.c_double * 4),
        ("rebuild_neighbors", C.c_int32), ("lmvm", LmvmOpts),
    ]


class DmdReport(C.Structure):
    _fields_ = [
        ("f", C.c_double), ("lmvm_reason", C.c_int32), ("lmvm_iterations", C.c_int32), ("lmvm_evals", C.c_int32),
        ("total_h_before", C.c_double), ("total_h_after", C.c_double), ("xH", C.c_double),
        ("xH_interior", C.c_double), ("flux_sum", C.c_double), ("flux_gross", C.c_double),
        ("clamp_lift", C.c_double), ("reservoir_supply", C.c_double),
    ]


class DiffusionBalance(C.Structure):
    _fields_ = [("flux_sum", C.c_double), ("flux_gross", C.c_double), ("clamp_lift", C.c_double),
                ("reservoir_supply", C.c_double)]


class LatticeDesc(C.Structure):
    _fields_ = [
        ("nx", C.c_int32), ("ny", C.c_int32), ("nz", C.c_int32), ("a", C.c_double),
        ("seed", C.c_uint64), ("lattice_id", C.c_uint32), ("disp_std", C.c_double),
        ("sigma_lo", C.c_double), ("sigma_hi", C.c_double),
        ("x_lo", C.c_double), ("x_hi", C.c_double),
        ("x_clip_lo", C.c_double), ("x_clip_hi", C.c_double),
        ("slab_cells", C.c_int32), ("slab_x", C.c_double),
    ]


def _load():
    if not os.path.exists(LIB_PATH):
        raise ImportError("liba2c_fe.so not built (run a2c_amd/build.py): %s" % LIB_PATH)
    lib = C.CDLL(LIB_PATH)
    P, I, D, VP = C.POINTER, C.c_int, C.c_double, C.c_void_p
    dp = P(C.c_double)
    ip = P(C.c_int32)
    sig = {
        "a2c_fe_desc_default": (None, [P(FeDesc)]),
        "a2c_fe_create": (I, [P(FeDesc), P(VP)]),
        "a2c_fe_destroy": (None, [VP]),
        "a2c_fe_last_error": (C.c_char_p, [VP]),
        "a2c_fe_set_precision": (I, [VP, C.c_int32]),
        "a2c_fe_set_occupancy": (I, [VP, dp]),
        "a2c_fe_set_potential_tables": (I, [VP, P(EamTabulated)]),
        "a2c_fe_set_potential_file": (I, [VP, C.c_char_p, C.c_char_p, C.c_char_p]),
        "a2c_fe_clear_potential": (I, [VP]),
        "a2c_fe_set_quadrature": (I, [VP, C.c_int32, dp, dp]),
        "a2c_eam_setfl_read": (I, [C.c_char_p, C.c_char_p, C.c_char_p, ip, dp, ip, dp, dp] + [dp] * 7),
        "a2c_fe_build_neighbors": (I, [VP, dp, P(NbrStats)]),
        "a2c_fe_build_neighbors_device": (I, [VP, VP, P(NbrStats)]),
        "a2c_fe_get_neighbors": (I, [VP, C.c_int32, ip, ip]),
        "a2c_fe_get_pdpdnum": (I, [VP, ip]),
        "a2c_fe_density": (I, [VP, dp, dp]),
        "a2c_fe_eval": (I, [VP, dp, dp, dp]),
        "a2c_fe_eval_device": (I, [VP, VP, VP, VP]),
        "a2c_tao_objgrad": (I, [VP, dp, dp, dp, VP]),
        "a2c_fe_chem_potential": (I, [VP, dp, dp, dp, D, dp, ip]),
        "a2c_fe_diffuse": (I, [VP, D, D, D, D, D, dp, ip, dp]),
        "a2c_fe_diffusion_balance": (I, [VP, P(DiffusionBalance)]),
        "a2c_fe_diffuse_device_noalloc": (I, [VP, D, D, D, D, D, VP, VP, VP, VP]),
        "a2c_lmvm_opts_default": (None, [P(LmvmOpts)]),
        "a2c_fe_minimize": (I, [VP, dp, P(LmvmOpts), dp, P(LmvmReport)]),
        "a2c_dev_alloc": (I, [VP, C.c_size_t, P(VP)]),
        "a2c_dev_free": (I, [VP, VP]),
        "a2c_dev_upload": (I, [VP, VP, VP, C.c_size_t]),
        "a2c_dev_download": (I, [VP, VP, VP, C.c_size_t]),
        "a2c_fe_sync": (I, [VP]),
        "a2c_dev_copy": (I, [VP, VP, VP, C.c_size_t]),
        "a2c_fe_time_eval": (I, [VP, VP, VP, VP, C.c_int32, P(Timing)]),
        "a2c_fe_time_diffuse": (I, [VP, D, D, D, D, D, VP, VP, VP, C.c_int32, P(Timing)]),
        "a2c_lattice_desc_default": (None, [P(LatticeDesc)]),
        "a2c_lattice_fcc_pdh": (I, [P(LatticeDesc), dp, dp, dp]),
        "a2c_philox4x32_10": (None, [P(C.c_uint32), P(C.c_uint32), P(C.c_uint32)]),
        "a2c_eam_tables": (I, [P(C.c_float), C.c_int32]),
        "a2c_eam_tables64": (I, [dp, C.c_int32]),
        "a2c_enforce_boundary_conditions": (I, [C.c_int32, ip, D, D, dp, dp, ip]),
        "a2c_macro_results": (I, [C.c_int32, dp, C.c_int32, ip, dp, dp]),
        "a2c_read_results": (I, [C.c_char_p, C.c_int32, C.c_int32, D, D, dp, dp, dp, ip]),
        "a2c_write_results": (I, [C.c_char_p, C.c_int32, C.c_int32, dp, dp, dp, ip, D, C.c_int32]),
        "a2c_write_xyz": (I, [C.c_char_p, C.c_int32, C.c_int32, dp, dp, dp, dp, dp, ip, D]),
        "a2c_dmd_params_default": (None, [P(DmdParams)]),
        "a2c_dmd_step": (I, [VP, P(DmdParams), C.c_int32, ip, C.c_int32, ip, dp, dp, dp, ip, P(DmdReport)]),
        "a2c_batch_desc_default": (None, [P(BatchDesc)]),
        "a2c_batch_create": (I, [P(BatchDesc), P(VP)]),
        "a2c_batch_destroy": (None, [VP]),
        "a2c_batch_last_error": (C.c_char_p, [VP]),
        "a2c_batch_footprint": (C.c_size_t, [C.c_int32, C.c_int32, C.c_int32]),
        "a2c_batch_mem_used": (C.c_size_t, [VP]),
        "a2c_batch_graph_stats": (I, [VP, C.POINTER(C.c_int64), C.POINTER(C.c_int64), C.POINTER(C.c_int64)]),
        "a2c_batch_add": (I, [VP, P(FeDesc), dp, dp, P(C.c_int32)]),
        "a2c_batch_remove": (I, [VP, C.c_int32]),
        "a2c_batch_set_state": (I, [VP, C.c_int32, dp]),
        "a2c_batch_get_state": (I, [VP, C.c_int32, dp, dp]),
        "a2c_batch_eval": (I, [VP, C.c_int32, ip, C.c_int32, dp]),
        "a2c_batch_lmvm_start": (I, [VP, C.c_int32, P(LmvmOpts)]),
        "a2c_batch_lmvm_step": (I, [VP, C.c_int32, P(C.c_int32)]),
        "a2c_batch_lmvm_step_slots": (I, [VP, C.c_int32, ip, C.c_int32, P(C.c_int32)]),
        "a2c_batch_lmvm_report": (I, [VP, C.c_int32, P(LmvmReport)]),
        "a2c_tune_get": (None, [P(TuneTable)]),
        "a2c_tune_set": (I, [P(TuneTable)]),
        "a2c_autotune_opts_default": (None, [P(AutotuneOpts)]),
        "a2c_autotune": (I, [P(AutotuneOpts), P(TuneTable), P(AutotuneReport)]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    return lib


lib = _load()


def _dptr(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def _iptr(a):
    return a.ctypes.data_as(C.POINTER(C.c_int32))


def _f64(a, n=None):
    a = np.ascontiguousarray(a, dtype=np.float64)
    if n is not None and a.size != n:
        raise ValueError("expected %d doubles, got %d" % (n, a.size))
    return a


# --------------------------------------------------------------------------
# specimens
# --------------------------------------------------------------------------
def eam_setfl_read(path, host="Pd", solute="H"):
    """The tables a DYNAMO setfl file gives the engine (a2c_eam_setfl_read):
    dict with n_rho, d_rho, n_r, d_r, cutoff and the seven arrays."""
    n_rho, n_r = C.c_int32(), C.c_int32()
    d_rho, d_r, cut = C.c_double(), C.c_double(), C.c_double()
    args = [path.encode(), host.encode(), solute.encode(), C.byref(n_rho), C.byref(d_rho), C.byref(n_r),
            C.byref(d_r), C.byref(cut)]
    rc = lib.a2c_eam_setfl_read(*args, *([None] * 7))
    if rc != 0:
        raise A2CError(rc, "cannot read setfl file %s" % path)
    out = {k: np.zeros(n_rho.value if k.startswith("F_") else n_r.value) for k in TABLE_KEYS}
    rc = lib.a2c_eam_setfl_read(*args, *[_dptr(out[k]) for k in TABLE_KEYS])
    if rc != 0:
        raise A2CError(rc, "cannot read setfl file %s" % path)
    out.update(n_rho=n_rho.value, d_rho=d_rho.value, n_r=n_r.value, d_r=d_r.value, cutoff=cut.value)
    return out


def lattice(n, a=3.89, seed=1, lattice_id=0, disp_std=0.02, sigma=(0.10, 0.10),
            x=(0.10, 0.10), x_clip=(1e-3, 1.0 - 1e-16), slab_cells=0, slab_x=0.0):
    """Seeded FCC Pd + octahedral-H specimen (include/a2c_lattice.h).

    n: cells per axis (int or 3-tuple).  Returns (X[4N], x[nH], box[3], n_pd, n_h).
    """
    nx, ny, nz = (n, n, n) if np.isscalar(n) else n
    d = LatticeDesc()
    lib.a2c_lattice_desc_default(C.byref(d))
    d.nx, d.ny, d.nz, d.a = nx, ny, nz, a
    d.seed, d.lattice_id, d.disp_std = seed, lattice_id, disp_std
    d.sigma_lo, d.sigma_hi = sigma
    d.x_lo, d.x_hi = x
    d.x_clip_lo, d.x_clip_hi = x_clip
    d.slab_cells, d.slab_x = slab_cells, slab_x
    nper = 4 * nx * ny * nz
    X = np.empty(8 * nper, dtype=np.float64)
    xo = np.empty(nper, dtype=np.float64)
    box = np.empty(3, dtype=np.float64)
    rc = lib.a2c_lattice_fcc_pdh(C.byref(d), _dptr(X), _dptr(xo), _dptr(box))
    if rc != 0:
        raise A2CError(rc, "bad lattice descriptor")
    return X, xo, box, nper, nper


def eam_tables64():
    """FP64 polynomial tables of the FP64 pair-owner passes: (exp, pd) with
    exp an array (3, 100, 10) for f_H, phi_PdH, phi_H (degree 8, one pad) and
    pd an array (2, 100, 12) for f_Pd, phi_Pd (degree 11); interval k is
    [1 + 0.05 k, 1 + 0.05 (k+1)), value sum_j c_j u^j, u = (r - 1)/0.05 - k."""
    n = lib.a2c_eam_tables64(None, 0)
    out = np.zeros(n)
    lib.a2c_eam_tables64(_dptr(out), n)
    return out[:3000].reshape(3, 100, 10), out[3000:].reshape(2, 100, 12)


def eam_tables():
    """FP32 cubic tables of the mixed path: array (5, NT, 4); table order
    f_Pd, f_H, phi_Pd, phi_PdH, phi_H; NT intervals of h = 5/NT A cover
    [1, 6) A, interval k being [1 + h k, 1 + h (k+1))."""
    n = lib.a2c_eam_tables(None, 0)
    out = np.zeros(n, dtype=np.float32)
    lib.a2c_eam_tables(out.ctypes.data_as(C.POINTER(C.c_float)), n)
    return out.reshape(5, n // 20, 4)


def philox4x32_10(ctr, key):
    c = (C.c_uint32 * 4)(*ctr)
    k = (C.c_uint32 * 2)(*key)
    o = (C.c_uint32 * 4)()
    lib.a2c_philox4x32_10(c, k, o)
    return list(o)


# --------------------------------------------------------------------------
# engine
# --------------------------------------------------------------------------
def _nbr_stats(st):
    return {"max_neib": st.max_neib, "width": list(st.width), "total": list(st.total),
            "refreshed": bool(st.refreshed), "max_displacement": st.max_displacement}


class FreeEntropy:
    """One lattice on one GPU (a2c_fe_ctx)."""

    def __init__(self, n_pd, n_h, box=(0.0, 0.0, 0.0), periodic=(0, 0, 0), rc=5.35, a_pd=3.89,
                 T=300.0, deriv="analytic", eps=1e-5, precision="fp64", device=0, max_neighbors=0, quad="r12"):
        d = FeDesc()
        lib.a2c_fe_desc_default(C.byref(d))
        d.n_pd, d.n_h = n_pd, n_h
        d.quad_rule = QUAD[quad]
        for c in range(3):
            d.box[c] = float(box[c])
            d.periodic[c] = int(bool(periodic[c]))
        d.rc, d.a_pd, d.temperature = rc, a_pd, T
        d.deriv_mode, d.fd_eps = DERIV[deriv], eps
        d.precision, d.device, d.max_neighbors = PREC[precision], device, max_neighbors
        self.desc = d
        self.n_pd, self.n_h, self.N = n_pd, n_h, n_pd + n_h
        self._ctx = C.c_void_p()
        rc_ = lib.a2c_fe_create(C.byref(d), C.byref(self._ctx))
        if rc_ != 0:
            raise A2CError(rc_, "a2c_fe_create failed")

    # -- plumbing --
    def close(self):
        if self._ctx:
            lib.a2c_fe_destroy(self._ctx)
            self._ctx = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def _check(self, rc):
        if rc != 0:
            msg = lib.a2c_fe_last_error(self._ctx)
            raise A2CError(rc, msg.decode() if msg else "")

    def set_precision(self, precision):
        self._check(lib.a2c_fe_set_precision(self._ctx, PREC[precision]))

    # -- reference-named API --
    def set_potential_tables(self, tables):
        """Tabulated potential (include/a2c_fe.h a2c_eam_tabulated): dict with
        n_rho, d_rho, n_r, d_r and the arrays F_pd, F_h, f_pd, f_h, rphi_pdpd,
        rphi_pdh, rphi_hh (pair functions as r*phi)."""
        arrs = {k: np.ascontiguousarray(tables[k], dtype=np.float64) for k in TABLE_KEYS}
        t = EamTabulated()
        t.n_rho, t.d_rho = int(tables["n_rho"]), float(tables["d_rho"])
        t.n_r, t.d_r = int(tables["n_r"]), float(tables["d_r"])
        for k in TABLE_KEYS:
            n = t.n_rho if k.startswith("F_") else t.n_r
            if arrs[k].size != n:
                raise ValueError("%s: %d values, the grid has %d" % (k, arrs[k].size, n))
            setattr(t, k, _dptr(arrs[k]))
        self._check(lib.a2c_fe_set_potential_tables(self._ctx, C.byref(t)))

    def set_potential_file(self, path, host="Pd", solute="H"):
        """Tabulated potential from a DYNAMO setfl file."""
        self._check(lib.a2c_fe_set_potential_file(self._ctx, path.encode(), host.encode(), solute.encode()))

    def clear_potential(self):
        self._check(lib.a2c_fe_clear_potential(self._ctx))

    def set_quadrature(self, xi=None, w=None):
        """Node table (K x 6 xi, K weights; swap-symmetric); None restores the
        descriptor's rule."""
        if xi is None:
            self._check(lib.a2c_fe_set_quadrature(self._ctx, 0, None, None))
            return
        xi = np.ascontiguousarray(xi, dtype=np.float64).reshape(-1, 6)
        w = np.ascontiguousarray(w, dtype=np.float64)
        if len(w) != len(xi):
            raise ValueError("one weight per node")
        self._check(lib.a2c_fe_set_quadrature(self._ctx, len(w), _dptr(xi), _dptr(w)))

    def set_occupancy(self, x):
        x = _f64(x, self.n_h)
        self._x = x
        self._check(lib.a2c_fe_set_occupancy(self._ctx, _dptr(x)))

    def findNeighbors(self, X):
        X = _f64(X, 4 * self.N)
        st = NbrStats()
        self._check(lib.a2c_fe_build_neighbors(self._ctx, _dptr(X), C.byref(st)))
        return _nbr_stats(st)

    build_neighbors = findNeighbors

    def findNeighbors_device(self, dX):
        """Neighbor build from device-resident positions (a device pointer)."""
        st = NbrStats()
        self._check(lib.a2c_fe_build_neighbors_device(self._ctx, C.c_void_p(dX), C.byref(st)))
        return _nbr_stats(st)

    def neighbors(self, which):
        """CSR (offsets, indices) of one list: 'PdPd', 'PdH', 'HPd', 'HH' or 'HH1'."""
        w = LIST_IDS[which]
        rows = self.n_pd if w < 2 else self.n_h
        off = np.zeros(rows + 1, dtype=np.int32)
        self._check(lib.a2c_fe_get_neighbors(self._ctx, w, _iptr(off), None))
        idx = np.zeros(max(int(off[-1]), 1), dtype=np.int32)
        self._check(lib.a2c_fe_get_neighbors(self._ctx, w, _iptr(off), _iptr(idx)))
        return off, idx[: off[-1]]

    def pdpdnum(self):
        out = np.zeros((max(self.n_pd, 1), 3), dtype=np.int32)
        self._check(lib.a2c_fe_get_pdpdnum(self._ctx, _iptr(out)))
        return out[: self.n_pd]

    def density(self, X):
        X = _f64(X, 4 * self.N)
        rho = np.zeros(max(self.N, 1))
        self._check(lib.a2c_fe_density(self._ctx, _dptr(X), _dptr(rho)))
        return rho[: self.N]

    def formFunctionAndGradient(self, X):
        X = _f64(X, 4 * self.N)
        G = np.zeros(max(4 * self.N, 1))
        f = C.c_double()
        self._check(lib.a2c_fe_eval(self._ctx, _dptr(X), C.byref(f), _dptr(G)))
        return f.value, G[: 4 * self.N]

    eval = formFunctionAndGradient

    def calculateChemicalPotential(self, X, coe_pd, coe_h, gamma_bd, gamma, full_h):
        """Updates gamma (float64[n_h]) and full_h (int32[n_h]) in place."""
        X = _f64(X, 4 * self.N)
        cp, ch = _f64(coe_pd, 4), _f64(coe_h, 4)
        assert gamma.dtype == np.float64 and full_h.dtype == np.int32
        self._check(lib.a2c_fe_chem_potential(self._ctx, _dptr(X), _dptr(cp), _dptr(ch), gamma_bd,
                                              _dptr(gamma), _iptr(full_h)))
        return gamma, full_h

    def integrateDiffusion(self, dt, nu, Qm, xLb, xUb, gamma, full_h):
        """One diffusion step from the context occupancy; returns the new x."""
        g = _f64(gamma, self.n_h)
        fh = np.ascontiguousarray(full_h, dtype=np.int32)
        x = np.array(self._x, dtype=np.float64)
        self._check(lib.a2c_fe_diffuse(self._ctx, dt, nu, Qm, xLb, xUb, _dptr(g), _iptr(fh), _dptr(x)))
        self._x = x
        return x

    def diffusion_balance(self):
        """Balance of the last integrateDiffusion: applied increments (equal
        to the reservoirs' supply up to rounding), their scale (sum of |edge
        terms|), the clamp lift and the reservoirs' supply."""
        b = DiffusionBalance()
        self._check(lib.a2c_fe_diffusion_balance(self._ctx, C.byref(b)))
        return {k: getattr(b, k) for k, _ in DiffusionBalance._fields_}

    def minimizeFreeEntropy(self, X0, **opts):
        """Bounded LMVM relaxation; returns (X, report dict)."""
        o = LmvmOpts()
        lib.a2c_lmvm_opts_default(C.byref(o))
        for k, v in opts.items():
            setattr(o, k, v)
        X0 = _f64(X0, 4 * self.N)
        Xo = np.zeros_like(X0)
        rep = LmvmReport()
        self._check(lib.a2c_fe_minimize(self._ctx, _dptr(X0), C.byref(o), _dptr(Xo), C.byref(rep)))
        return Xo, {k: getattr(rep, k) for k, _ in LmvmReport._fields_}

    # -- device-resident helpers --
    def dev_alloc(self, nbytes):
        p = C.c_void_p()
        self._check(lib.a2c_dev_alloc(self._ctx, nbytes, C.byref(p)))
        return p.value

    def dev_free(self, p):
        self._check(lib.a2c_dev_free(self._ctx, C.c_void_p(p)))

    def upload(self, dptr, arr):
        arr = np.ascontiguousarray(arr)
        self._check(lib.a2c_dev_upload(self._ctx, C.c_void_p(dptr), arr.ctypes.data_as(C.c_void_p), arr.nbytes))

    def download(self, arr, dptr):
        self._check(lib.a2c_dev_download(self._ctx, arr.ctypes.data_as(C.c_void_p), C.c_void_p(dptr), arr.nbytes))
        return arr

    def dev_copy(self, dst, src, nbytes):
        self._check(lib.a2c_dev_copy(self._ctx, C.c_void_p(dst), C.c_void_p(src), nbytes))

    def eval_device(self, dX, dG, df):
        self._check(lib.a2c_fe_eval_device(self._ctx, C.c_void_p(dX), C.c_void_p(dG), C.c_void_p(df)))

    def sync(self):
        self._check(lib.a2c_fe_sync(self._ctx))

    def time_eval(self, dX, dG, df, iters):
        t = Timing()
        self._check(lib.a2c_fe_time_eval(self._ctx, C.c_void_p(dX), C.c_void_p(dG), C.c_void_p(df), iters,
                                         C.byref(t)))
        kern = {t.names[k].value.decode(): t.avg_ms[k] for k in range(t.n_kernels)}
        return {"kernels_ms": kern, "total_ms": t.total_ms}

    def time_diffuse(self, dt, nu, Qm, xLb, xUb, d_gamma, d_full_h, d_x, iters):
        """iters diffusion steps on device arrays, each from the context
        occupancy (unchanged), timed per kernel with HIP events."""
        t = Timing()
        self._check(lib.a2c_fe_time_diffuse(self._ctx, dt, nu, Qm, xLb, xUb, C.c_void_p(d_gamma),
                                            C.c_void_p(d_full_h), C.c_void_p(d_x), iters, C.byref(t)))
        kern = {t.names[k].value.decode(): t.avg_ms[k] for k in range(t.n_kernels)}
        return {"kernels_ms": kern, "total_ms": t.total_ms}


# --------------------------------------------------------------------------
# DMD helpers (include/a2c_dmd.h)
# --------------------------------------------------------------------------
def _i32(a):
    return np.ascontiguousarray(a, dtype=np.int32)


def enforceBoundaryConditions(h_subsurf, gamma_bd, xUb, x0, gamma, full_h):
    """boundaryconditions.cpp:30-55; updates x0, gamma, full_h in place."""
    sub = _i32(h_subsurf)
    assert x0.dtype == np.float64 and gamma.dtype == np.float64 and full_h.dtype == np.int32
    rc = lib.a2c_enforce_boundary_conditions(len(sub), _iptr(sub), gamma_bd, xUb, _dptr(x0), _dptr(gamma),
                                             _iptr(full_h))
    if rc:
        raise A2CError(rc, "enforceBoundaryConditions")


def calculateMacroResults(x, h_surf=()):
    """calculateresults.cpp:8-78 -> (mean H fraction, interior mean)."""
    x = _f64(x)
    surf = _i32(h_surf)
    a, b = C.c_double(), C.c_double()
    rc = lib.a2c_macro_results(len(x), _dptr(x), len(surf), _iptr(surf), C.byref(a), C.byref(b))
    if rc:
        raise A2CError(rc, "calculateMacroResults")
    return a.value, b.value


def readResults(path, n_pd, n_h, xLb=1e-3, xUb=1 - 1e-16):
    """readresults.cpp:10-55 -> (X, x, gamma, full_h)."""
    X, x, g = np.zeros(4 * (n_pd + n_h)), np.zeros(max(n_h, 1)), np.zeros(max(n_h, 1))
    fh = np.zeros(max(n_h, 1), dtype=np.int32)
    rc = lib.a2c_read_results(path.encode(), n_pd, n_h, xLb, xUb, _dptr(X), _dptr(x), _dptr(g), _iptr(fh))
    if rc:
        raise A2CError(rc, "cannot read restart file %s" % path)
    return X, x[:n_h], g[:n_h], fh[:n_h]


def writeResults(path, n_pd, n_h, X, x, gamma, full_h, time=0.0, step=0):
    X, x, g = _f64(X), _f64(x), _f64(gamma)
    fh = _i32(full_h)
    rc = lib.a2c_write_results(path.encode(), n_pd, n_h, _dptr(X), _dptr(x), _dptr(g), _iptr(fh), time, step)
    if rc:
        raise A2CError(rc, "cannot write restart file %s" % path)


def writeXYZ(path, n_pd, n_h, X, X0, x, gamma, box, periodic=(1, 1, 1), time=0.0):
    """Extended-XYZ snapshot (Output.cpp:332-494 layout)."""
    X, X0, x, g, b = _f64(X), _f64(X0), _f64(x), _f64(gamma), _f64(box, 3)
    per = _i32(periodic)
    rc = lib.a2c_write_xyz(path.encode(), n_pd, n_h, _dptr(X), _dptr(X0), _dptr(x), _dptr(g), _dptr(b), _iptr(per),
                           time)
    if rc:
        raise A2CError(rc, "cannot write %s" % path)


def dmd_params(**kw):
    p = DmdParams()
    lib.a2c_dmd_params_default(C.byref(p))
    for k, v in kw.items():
        if k == "lmvm":
            for kk, vv in v.items():
                setattr(p.lmvm, kk, vv)
        elif k in ("coe_pd", "coe_h"):
            for i in range(4):
                getattr(p, k)[i] = v[i]
        else:
            setattr(p, k, v)
    return p


def dmd_step(fe, params, X, x, gamma, full_h, h_subsurf=(), h_surf=()):
    """One DMD step (include/a2c_dmd.h); X, x, gamma, full_h updated in place."""
    sub, surf = _i32(h_subsurf), _i32(h_surf)
    rep = DmdReport()
    rc = lib.a2c_dmd_step(fe._ctx, C.byref(params), len(sub), _iptr(sub), len(surf), _iptr(surf), _dptr(X),
                          _dptr(x), _dptr(gamma), _iptr(full_h), C.byref(rep))
    fe._check(rc)
    fe._x = x.copy()
    return {k: getattr(rep, k) for k, _ in DmdReport._fields_}


def lmvm_opts(**kw):
    o = LmvmOpts()
    lib.a2c_lmvm_opts_default(C.byref(o))
    for k, v in kw.items():
        setattr(o, k, v)
    return o


def fe_desc(n_pd, n_h, box=(0.0, 0.0, 0.0), periodic=(0, 0, 0), rc=5.35, a_pd=3.89, T=300.0, max_neighbors=0,
            quad="r12"):
    d = FeDesc()
    lib.a2c_fe_desc_default(C.byref(d))
    d.n_pd, d.n_h = n_pd, n_h
    d.quad_rule = QUAD[quad]
    for c in range(3):
        d.box[c] = float(box[c])
        d.periodic[c] = int(bool(periodic[c]))
    d.rc, d.a_pd, d.temperature, d.max_neighbors = rc, a_pd, T, max_neighbors
    return d


class Batch:
    """Many lattices on one GPU (include/a2c_batch.h): fused evaluation and a
    lockstep LMVM with admission/retirement between iterations."""

    def __init__(self, precision="mixed", max_lattices=64, hbm_budget_bytes=200 << 30, device=0,
                 deriv="analytic", eps=1e-5, graphs=None):
        d = BatchDesc()
        lib.a2c_batch_desc_default(C.byref(d))
        d.device, d.precision, d.deriv_mode, d.fd_eps = device, PREC[precision], DERIV[deriv], eps
        if graphs is not None:
            d.graphs = int(bool(graphs))
        d.max_lattices, d.hbm_budget_bytes = max_lattices, hbm_budget_bytes
        self._b = C.c_void_p()
        rc = lib.a2c_batch_create(C.byref(d), C.byref(self._b))
        if rc != 0:
            raise A2CError(rc, "a2c_batch_create failed")
        self.sizes = {}

    def _check(self, rc):
        if rc != 0:
            msg = lib.a2c_batch_last_error(self._b)
            raise A2CError(rc, msg.decode() if msg else "")

    def close(self):
        if self._b:
            lib.a2c_batch_destroy(self._b)
            self._b = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    @staticmethod
    def footprint(n_pd, n_h, m=5):
        return lib.a2c_batch_footprint(n_pd, n_h, m)

    def mem_used(self):
        return lib.a2c_batch_mem_used(self._b)

    def graph_stats(self):
        """Launch sequences so far: {"replays", "captures", "direct"}."""
        v = [C.c_int64() for _ in range(3)]
        self._check(lib.a2c_batch_graph_stats(self._b, *[C.byref(x) for x in v]))
        return dict(zip(("replays", "captures", "direct"), (x.value for x in v)))

    def add(self, X, x, n_pd, n_h, box=(0, 0, 0), periodic=(0, 0, 0), T=300.0, rc=5.35, a_pd=3.89, quad="r12"):
        d = fe_desc(n_pd, n_h, box, periodic, rc, a_pd, T, quad=quad)
        X = _f64(X, 4 * (n_pd + n_h))
        x = _f64(x, n_h)
        slot = C.c_int32()
        self._check(lib.a2c_batch_add(self._b, C.byref(d), _dptr(X), _dptr(x), C.byref(slot)))
        self.sizes[slot.value] = n_pd + n_h
        return slot.value

    def remove(self, slot):
        self._check(lib.a2c_batch_remove(self._b, slot))
        self.sizes.pop(slot, None)

    def set_state(self, slot, X):
        X = _f64(X, 4 * self.sizes[slot])
        self._check(lib.a2c_batch_set_state(self._b, slot, _dptr(X)))

    def get_state(self, slot):
        n = 4 * self.sizes[slot]
        X, G = np.zeros(n), np.zeros(n)
        self._check(lib.a2c_batch_get_state(self._b, slot, _dptr(X), _dptr(G)))
        return X, G

    def eval(self, slots, precision="mixed"):
        ids = np.ascontiguousarray(slots, dtype=np.int32)
        f = np.zeros(max(len(ids), 1))
        self._check(lib.a2c_batch_eval(self._b, len(ids), _iptr(ids), PREC[precision], _dptr(f)))
        return f[: len(ids)]

    def start(self, slot, **opts):
        o = lmvm_opts(**opts)
        self._check(lib.a2c_batch_lmvm_start(self._b, slot, C.byref(o)))

    def step(self, max_iters=1):
        n = C.c_int32()
        self._check(lib.a2c_batch_lmvm_step(self._b, max_iters, C.byref(n)))
        return n.value

    def step_slots(self, slots, max_iters=1):
        """Advance only the listed slots; returns the running count of the batch."""
        ids = np.ascontiguousarray(slots, dtype=np.int32)
        n = C.c_int32()
        self._check(lib.a2c_batch_lmvm_step_slots(self._b, len(ids), _iptr(ids), max_iters, C.byref(n)))
        return n.value

    def report(self, slot):
        r = LmvmReport()
        self._check(lib.a2c_batch_lmvm_report(self._b, slot, C.byref(r)))
        return {k: getattr(r, k) for k, _ in LmvmReport._fields_}

    def run(self, max_iters=1 << 20, progress=None, every=50):
        """Step every running slot to completion; progress(iterations, running)
        is called every `every` iterations when given."""
        if progress is None:
            while self.step(max_iters) > 0:
                pass
            return
        done = 0
        while True:
            running = self.step(every)
            done += every
            progress(done, running)
            if running == 0:
                return


# --------------------------------------------------------------------------
# launch-shape table (include/a2c_tune.h)
# --------------------------------------------------------------------------
def _table_dict(t):
    n = t.n_classes
    return {"max_sites": list(t.max_sites[:n]), "lanes_mixed": list(t.lanes_mixed[:n]),
            "lanes_fp64": list(t.lanes_fp64[:n])}


def tune_get():
    """The lanes-per-site table in effect: {"max_sites", "lanes_mixed", "lanes_fp64"}."""
    t = TuneTable()
    lib.a2c_tune_get(C.byref(t))
    return _table_dict(t)


def tune_set(table=None):
    """Install a table (dict as returned by tune_get), or the default when None."""
    if table is None:
        rc = lib.a2c_tune_set(None)
    else:
        t = TuneTable()
        n = len(table["max_sites"])
        if not 1 <= n <= TUNE_MAX or len(table["lanes_mixed"]) != n or len(table["lanes_fp64"]) != n:
            raise A2CError(A2C_EINVAL, "malformed tune table")
        t.n_classes = n
        for c in range(n):
            t.max_sites[c] = min(int(table["max_sites"][c]), 2**31 - 1)
            t.lanes_mixed[c] = int(table["lanes_mixed"][c])
            t.lanes_fp64[c] = int(table["lanes_fp64"][c])
        rc = lib.a2c_tune_set(C.byref(t))
    if rc != 0:
        raise A2CError(rc, "a2c_tune_set rejected the table")


def autotune(cells=None, target_sites=None, reps=None, apply=False, device=0):
    """Measure lanes per site for each lattice size (include/a2c_tune.h).
    Returns (table, report) where report lists, per size, sites, copies, ns
    per site of a fused launch of the copies and microseconds per evaluation
    of one lattice alone, for K = 1, 2, 4, 8, 16 in mixed and FP64 precision, and
    whether the lone-lattice time chose K ("latency")."""
    o = AutotuneOpts()
    lib.a2c_autotune_opts_default(C.byref(o))
    o.device, o.apply = device, int(bool(apply))
    if cells is not None:
        o.n_sizes = len(cells)
        for i, c in enumerate(cells):
            o.cells[i] = int(c)
    if target_sites is not None:
        o.target_sites = int(target_sites)
    if reps is not None:
        o.reps = int(reps)
    t, r = TuneTable(), AutotuneReport()
    rc = lib.a2c_autotune(C.byref(o), C.byref(t), C.byref(r))
    if rc != 0:
        raise A2CError(rc, "a2c_autotune failed")
    rows = []
    for i in range(o.n_sizes):
        rows.append({"cells": o.cells[i], "sites": r.sites[i], "copies": r.copies[i],
                     "ns_per_site": {"mixed": list(r.ns_per_site[i][0]), "fp64": list(r.ns_per_site[i][1])},
                     "us_alone": {"mixed": list(r.us_alone[i][0]), "fp64": list(r.us_alone[i][1])},
                     "latency": {"mixed": bool(r.latency[i][0]), "fp64": bool(r.latency[i][1])}})
    return _table_dict(t), rows


def exported_symbols():
    """Symbols include/*.h declares (checked by the CPU test suite)."""
    import re
    root = os.path.dirname(_HERE)
    names = []
    for h in ("a2c_fe.h", "a2c_lattice.h", "a2c_batch.h", "a2c_dmd.h", "a2c_tune.h"):
        txt = open(os.path.join(root, "include", h)).read()
        names += re.findall(r"\b(a2c_\w+)\s*\(", txt)
    return sorted(set(names))
