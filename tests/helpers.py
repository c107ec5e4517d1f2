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
