"""Writes the golden fixtures under tests/golden/ from SciPy, an implementation
independent of this project and of the reference:

* expm_cases.npz   upper-Hessenberg H (k = 1..60, ||H||_1 from 1e-3 to 1e3)
                   and scipy.linalg.expm(H)
* expmv_cases.npz  Gershgorin-negative sparse A (CSR), v, t and
                   scipy.sparse.linalg.expm_multiply (Al-Mohy & Higham 2011)

Run from the repository root:  python tests/golden/make_golden.py
"""
import os

import numpy as np
import scipy.linalg
import scipy.sparse as sp
import scipy.sparse.linalg as spla

HERE = os.path.dirname(os.path.abspath(__file__))


def hessenberg_cases():
    rng = np.random.default_rng(20240601)
    mats, exps, ks = [], [], []
    for k in [1, 2, 3, 8, 20, 40, 60]:
        for target in [1e-3, 1e-2, 0.2, 0.9, 2.0, 5.0, 40.0, 1e3]:
            H = np.triu(rng.standard_normal((k, k)), -1)
            H *= target / np.abs(H).sum(0).max()
            if target > 10:  # shift left so exp(H) stays representable
                H -= np.eye(k) * (target / 2)
            E = scipy.linalg.expm(H)
            if not np.all(np.isfinite(E)):
                continue
            pad = np.zeros((60, 60))
            pad[:k, :k] = H
            mats.append(pad)
            pad = np.zeros((60, 60))
            pad[:k, :k] = E
            exps.append(pad)
            ks.append(k)
    return np.array(ks), np.array(mats), np.array(exps)


def random_negative(n, density, rng):
    cnt = int(density * n * n)
    r = rng.integers(0, n, cnt)
    c = rng.integers(0, n, cnt)
    keep = r != c
    A = sp.coo_matrix((rng.uniform(-1, 1, keep.sum()), (r[keep], c[keep])), shape=(n, n)).tocsr()
    A.sum_duplicates()
    A = (A + sp.diags(-1.0 - np.abs(A).sum(1).A1)).tocsr()
    A.sort_indices()
    return A


def expmv_cases():
    rng = np.random.default_rng(77)
    out = {}
    for i, (n, dens, t) in enumerate([(300, 2e-2, 0.5), (800, 1e-2, 1.0), (1500, 5e-3, 2.0)]):
        A = random_negative(n, dens, rng)
        v = rng.standard_normal(n)
        y = spla.expm_multiply(t * A, v)
        out[f"c{i}_indptr"] = A.indptr.astype(np.int32)
        out[f"c{i}_indices"] = A.indices.astype(np.int32)
        out[f"c{i}_data"] = A.data
        out[f"c{i}_v"] = v
        out[f"c{i}_t"] = np.array(t)
        out[f"c{i}_y"] = y
    out["count"] = np.array(3)
    return out


if __name__ == "__main__":
    ks, H, E = hessenberg_cases()
    np.savez_compressed(os.path.join(HERE, "expm_cases.npz"), k=ks, H=H, E=E)
    np.savez_compressed(os.path.join(HERE, "expmv_cases.npz"), **expmv_cases())
    print("wrote", len(ks), "expm cases and 3 expmv cases")
