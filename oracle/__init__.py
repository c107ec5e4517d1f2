"""FP64 CPU oracle of the a2c free-entropy hot path (TEST INFRASTRUCTURE).

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may use
this package, and only as the checker.  It wraps oracle/liba2c_oracle.so, a
C++17 restatement of the reference's algorithms (file:line citations in
a2c_oracle.cpp), built with `make -C oracle`.

Pinning status (what checks each function besides its own formula) is listed
in a2c_oracle.h; nothing here is "parity unpinned" except the optimizer,
which has no counterpart in the reference tree (PETSc TAO is not vendored).
"""
import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liba2c_oracle.so")
REF_QUAD = os.path.join(HERE, "_ref", "quadrature_ref")

F_PD, F_H, RHO_PD, RHO_H, PHI_PD, PHI_PDH, PHI_H = range(7)
FUNC_NAMES = ["F_Pd", "F_H", "f_Pd", "f_H", "phi_Pd", "phi_PdH", "phi_H"]
DERIV = {"analytic": 0, "fd": 1}
POT = {"eam": 0, "test": 1, "test4": 2, "tab": 3}
QUAD = {"r12": 0, "gh3": 1, "table": 2}


def build(reference=None):
    """make -C oracle (and oracle/_ref when the reference tree is present)."""
    subprocess.run(["make", "-s", "-C", HERE], check=True)
    ref = reference or "/root/reference"
    if os.path.isdir(ref):
        subprocess.run(["make", "-s", "-C", HERE, "ref", "REF=" + ref], check=True)


class Params(C.Structure):
    _fields_ = [
        ("n_pd", C.c_int32), ("n_h", C.c_int32), ("box", C.c_double * 3), ("periodic", C.c_int32 * 3),
        ("rc", C.c_double), ("a_pd", C.c_double), ("T", C.c_double), ("deriv_mode", C.c_int32),
        ("eps", C.c_double), ("potential", C.c_int32), ("threads", C.c_int32), ("quad_rule", C.c_int32),
    ]


def _load():
    if not os.path.exists(LIB_PATH):
        build()
    lib = C.CDLL(LIB_PATH)
    P, I, D, VP = C.POINTER, C.c_int, C.c_double, C.c_void_p
    dp, ip = P(C.c_double), P(C.c_int32)
    sig = {
        "or_eam": (D, [I, D]),
        "or_eam_d": (D, [I, D, I, D]),
        "or_eam_constants": (I, [I, dp, I]),
        "or_quadrature": (I, [dp, dp]),
        "or_nbr_build": (VP, [P(Params), dp]),
        "or_nbr_get": (I, [VP, I, P(ip), P(ip)]),
        "or_nbr_pdpdnum": (ip, [VP]),
        "or_nbr_max": (I, [VP]),
        "or_nbr_free": (None, [VP]),
        "or_density": (I, [P(Params), VP, dp, dp, dp]),
        "or_eval": (I, [P(Params), VP, dp, dp, dp, dp]),
        "or_chem_potential": (I, [P(Params), VP, dp, dp, dp, dp, D, dp, ip]),
        "or_diffuse": (I, [P(Params), VP, D, D, D, D, D, dp, ip, dp, dp]),
        "or_set_tabulated": (I, [I, D, I, D, P(dp)]),
        "or_set_quad_table": (I, [I, dp, dp]),
        "or_tab_eval": (I, [I, D, dp, I, D, dp, dp]),
    }
    for k, (r, a) in sig.items():
        f = getattr(lib, k)
        f.restype, f.argtypes = r, a
    return lib


lib = _load()


def _d(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def _i(a):
    return a.ctypes.data_as(C.POINTER(C.c_int32))


def eam(fn, x):
    return np.array([lib.or_eam(fn, float(v)) for v in np.atleast_1d(x)])


def eam_d(fn, x, deriv="analytic", eps=1e-5):
    return np.array([lib.or_eam_d(fn, float(v), DERIV[deriv], eps) for v in np.atleast_1d(x)])


def eam_constants(fn):
    out = np.zeros(64)
    n = lib.or_eam_constants(fn, _d(out), 64)
    return out[:n]


def quadrature():
    xi = np.zeros(72)
    w = np.zeros(12)
    lib.or_quadrature(_d(xi), _d(w))
    return xi.reshape(12, 6), w


_keep = []  # the arrays behind the process-wide tables stay alive


def set_tabulated(tables):
    """Tables of the "tab" potential (process-wide): dict with n_rho, d_rho,
    n_r, d_r and F_pd, F_h, f_pd, f_h, rphi_pdpd, rphi_pdh, rphi_hh."""
    arrs = [np.ascontiguousarray(tables[k], dtype=np.float64)
            for k in ("F_pd", "F_h", "f_pd", "f_h", "rphi_pdpd", "rphi_pdh", "rphi_hh")]
    ptrs = (C.POINTER(C.c_double) * 7)(*[_d(a) for a in arrs])
    assert lib.or_set_tabulated(int(tables["n_rho"]), float(tables["d_rho"]), int(tables["n_r"]),
                                float(tables["d_r"]), ptrs) == 0


def set_quad_table(xi, w):
    """Nodes of the "table" rule (process-wide): xi K x 6, w K."""
    xi = np.ascontiguousarray(xi, dtype=np.float64).reshape(-1, 6)
    w = np.ascontiguousarray(w, dtype=np.float64)
    assert lib.or_set_quad_table(len(w), _d(xi), _d(w)) == 0


def tab_eval(y, h, x, emb=False):
    """One table on its own: (value, slope) at x."""
    y = np.ascontiguousarray(y, dtype=np.float64)
    v, d = C.c_double(), C.c_double()
    assert lib.or_tab_eval(len(y), h, _d(y), int(emb), float(x), C.byref(v), C.byref(d)) == 0
    return v.value, d.value


class Oracle:
    """One lattice: neighbor sets, densities, objective/gradient, chemical
    potential and diffusion, in FP64 on the CPU."""

    def __init__(self, n_pd, n_h, box=(0, 0, 0), periodic=(0, 0, 0), rc=5.35, a_pd=3.89, T=300.0,
                 deriv="analytic", eps=1e-5, potential="eam", threads=1, quad="r12"):
        p = Params()
        p.n_pd, p.n_h = n_pd, n_h
        for c in range(3):
            p.box[c] = float(box[c])
            p.periodic[c] = int(bool(periodic[c]))
        p.rc, p.a_pd, p.T = rc, a_pd, T
        p.deriv_mode, p.eps, p.potential, p.threads = DERIV[deriv], eps, POT[potential], threads
        p.quad_rule = QUAD[quad]
        self.p = p
        self.n_pd, self.n_h, self.N = n_pd, n_h, n_pd + n_h
        self._nb = None

    def __del__(self):
        if getattr(self, "_nb", None):
            lib.or_nbr_free(self._nb)

    def build_neighbors(self, X):
        X = np.ascontiguousarray(X, dtype=np.float64)
        if self._nb:
            lib.or_nbr_free(self._nb)
        self._nb = lib.or_nbr_build(C.byref(self.p), _d(X))
        if not self._nb:
            raise ValueError("oracle neighbor build refused the box (periodic box < 2 rc?)")
        return lib.or_nbr_max(self._nb)

    def neighbors(self, which):
        w = {"PdPd": 0, "PdH": 1, "HPd": 2, "HH": 3, "HH1": 4}[which]
        off = C.POINTER(C.c_int32)()
        idx = C.POINTER(C.c_int32)()
        rows = lib.or_nbr_get(self._nb, w, C.byref(off), C.byref(idx))
        o = np.ctypeslib.as_array(off, shape=(rows + 1,)).copy()
        i = np.ctypeslib.as_array(idx, shape=(max(int(o[-1]), 1),)).copy()[: o[-1]] if o[-1] else np.zeros(0, np.int32)
        return o, i

    def pdpdnum(self):
        p = lib.or_nbr_pdpdnum(self._nb)
        return np.ctypeslib.as_array(p, shape=(self.n_pd * 3,)).copy().reshape(self.n_pd, 3) if self.n_pd else np.zeros((0, 3), np.int32)

    def density(self, X, x):
        X = np.ascontiguousarray(X, dtype=np.float64)
        x = np.ascontiguousarray(x, dtype=np.float64)
        rho = np.zeros(self.N)
        assert lib.or_density(C.byref(self.p), self._nb, _d(X), _d(x), _d(rho)) == 0
        return rho

    def eval(self, X, x):
        X = np.ascontiguousarray(X, dtype=np.float64)
        x = np.ascontiguousarray(x, dtype=np.float64)
        G = np.zeros(4 * self.N)
        f = C.c_double()
        assert lib.or_eval(C.byref(self.p), self._nb, _d(X), _d(x), C.byref(f), _d(G)) == 0
        return f.value, G

    def chem_potential(self, X, x, coe_pd, coe_h, gamma_bd, gamma, full_h):
        X = np.ascontiguousarray(X, dtype=np.float64)
        x = np.ascontiguousarray(x, dtype=np.float64)
        cp = np.ascontiguousarray(coe_pd, dtype=np.float64)
        ch = np.ascontiguousarray(coe_h, dtype=np.float64)
        g = np.array(gamma, dtype=np.float64)
        fh = np.array(full_h, dtype=np.int32)
        assert lib.or_chem_potential(C.byref(self.p), self._nb, _d(X), _d(x), _d(cp), _d(ch), gamma_bd, _d(g), _i(fh)) == 0
        return g, fh

    def diffuse(self, dt, nu, Qm, xLb, xUb, gamma, full_h, x0):
        g = np.ascontiguousarray(gamma, dtype=np.float64)
        fh = np.ascontiguousarray(full_h, dtype=np.int32)
        x0 = np.ascontiguousarray(x0, dtype=np.float64)
        x = x0.copy()
        assert lib.or_diffuse(C.byref(self.p), self._nb, dt, nu, Qm, xLb, xUb, _d(g), _i(fh), _d(x0), _d(x)) == 0
        return x


def reference_quadrature():
    """Nodes printed by the reference's own gaussianQuadrature (oracle/_ref),
    or None when that binary was not built."""
    if not os.path.exists(REF_QUAD):
        return None
    out = subprocess.run([REF_QUAD], capture_output=True, text=True, check=True).stdout.split("\n")
    k = int(out[0])
    rows = [[float.fromhex(t) for t in line.split()] for line in out[1 : 1 + k]]
    return np.array(rows)
