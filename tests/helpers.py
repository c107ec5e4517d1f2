"""Shared test helpers: specimens and independent numpy restatements used as
cross-checks (brute-force neighbor sets, closed-form test-potential energy)."""
import numpy as np

import a2c_amd as A

KB = 8.6173324e-5


def specimen(n, seed=1, disp=0.02, sigma=(0.10, 0.10), x=(0.10, 0.10), periodic=True, **kw):
    X, xo, box, npd, nh = A.lattice(n, seed=seed, disp_std=disp, sigma=sigma, x=x, **kw)
    per = (1, 1, 1) if periodic else (0, 0, 0)
    return X, xo, box, per, npd, nh


def brute_neighbors(X, npd, box, periodic, rc=5.35, a_pd=3.89):
    """All-pairs FP64 sets with the reference's expression
    ((dx*dx + dy*dy) + dz*dz) <= rc*rc, d = q_j - q_i (numpy does not contract)."""
    q = X.reshape(-1, 4)[:, :3]
    N = q.shape[0]
    d = q[None, :, :] - q[:, None, :]
    for c in range(3):
        if periodic[c]:
            d[:, :, c] = d[:, :, c] - box[c] * np.rint(d[:, :, c] / box[c])
    d2 = d[:, :, 0] * d[:, :, 0] + d[:, :, 1] * d[:, :, 1] + d[:, :, 2] * d[:, :, 2]
    mask = d2 <= rc * rc
    np.fill_diagonal(mask, False)
    sh1, sh2 = a_pd * a_pd * 0.75, a_pd * a_pd * 1.25
    out = {k: [] for k in ["PdPd", "PdH", "HPd", "HH", "HH1"]}
    num = np.zeros((npd, 3), dtype=np.int32)
    for s in range(N):
        js = np.nonzero(mask[s])[0]
        pdn = js[js < npd]
        hn = js[js >= npd] - npd
        if s < npd:
            out["PdPd"].append(pdn)
            out["PdH"].append(hn)
            dd = d2[s, pdn]
            num[s] = [(dd < sh1).sum(), ((dd > sh1) & (dd < sh2)).sum(), 0]
            num[s, 2] = len(pdn) - num[s, 0] - num[s, 1]
        else:
            out["HPd"].append(pdn)
            out["HH"].append(hn)
            out["HH1"].append(hn[d2[s, hn + npd] < sh1])
    return out, num


def csr_rows(off, idx):
    return [idx[off[r]:off[r + 1]] for r in range(len(off) - 1)]


def sampled_tables(n_rho=10001, d_rho=0.005, n_r=8001, d_r=1e-3):
    """The compiled-in Pd-H functions sampled on uniform grids, in the layout
    of a tabulated potential (pair functions as r*phi), through the oracle's
    closed forms."""
    import oracle as O
    rho = np.arange(n_rho) * d_rho
    r = np.arange(n_r) * d_r
    return {
        "n_rho": n_rho, "d_rho": d_rho, "n_r": n_r, "d_r": d_r,
        "F_pd": O.eam(O.F_PD, rho), "F_h": O.eam(O.F_H, rho),
        "f_pd": O.eam(O.RHO_PD, r), "f_h": O.eam(O.RHO_H, r),
        "rphi_pdpd": r * O.eam(O.PHI_PD, r), "rphi_pdh": r * O.eam(O.PHI_PDH, r), "rphi_hh": r * O.eam(O.PHI_H, r),
    }


def write_setfl(path, elements, F, f, rphi, n_rho, d_rho, n_r, d_r, cutoff):
    """A DYNAMO setfl file: F[e], f[e] per element, rphi[(i, j)] for i >= j
    (element indices), five values a line, every digit of each double."""
    def block(v):
        v = np.asarray(v, dtype=np.float64)
        return "".join(" ".join("%.17g" % x for x in v[k:k + 5]) + "\n" for k in range(0, len(v), 5))
    out = ["written by the tests\n", "tabulated Pd-H\n", "third comment line\n",
           "%d %s\n" % (len(elements), " ".join(elements)),
           "%d %.17g %d %.17g %.17g\n" % (n_rho, d_rho, n_r, d_r, cutoff)]
    for e in elements:
        out.append("1 1.0 3.89 fcc\n")
        out.append(block(F[e]))
        out.append(block(f[e]))
    for i in range(len(elements)):
        for j in range(i + 1):
            out.append(block(rphi[(elements[i], elements[j])]))
    with open(path, "w") as fh:
        fh.write("".join(out))


def product_rule_2pt():
    """6-D two-point Gauss-Hermite product rule for the weight exp(-|xi|^2):
    nodes +-1/sqrt(2) per coordinate, 64 nodes of weight 1/64 (exact to
    degree 3 in each coordinate; swap-symmetric)."""
    v = 1.0 / np.sqrt(2.0)
    xi = np.array([[v if (k >> c) & 1 else -v for c in range(6)] for k in range(64)])
    return xi, np.full(64, 1.0 / 64)
