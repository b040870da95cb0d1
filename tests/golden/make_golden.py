"""Write the small golden fixtures in tests/golden/golden_small.npz.

Each case is a small random square CSR matrix (rows may be unsorted, may
carry duplicate columns, explicit zeros and empty rows) together with the
products A*A and A*A^T computed by dense brute force in numpy:

  structure: pattern(C) = boolean product of the stored patterns (structural
             semantics: an entry that cancels to 0.0 stays stored);
  values:    dense float64 matmul (a different summation order from the CPU
             Gustavson loop, so checks use |c - c_ref| <= 1e-12 * sum|a*b|).

Run from the repo root: python tests/golden/make_golden.py
"""
from pathlib import Path

import numpy as np

OUT = Path(__file__).resolve().parent / "golden_small.npz"


def random_case(rng, n, density, unsorted=False, dups=False, zeros=False, empty_rows=0):
    rows, cols, vals = [], [], []
    for i in range(n):
        k = rng.binomial(n, density)
        cs = rng.choice(n, size=k, replace=False)
        if dups and k > 0 and rng.random() < 0.5:
            cs = np.concatenate([cs, cs[:1]])
        v = rng.uniform(-1, 1, size=len(cs))
        if zeros and len(v):
            v[rng.random(len(v)) < 0.2] = 0.0
        if not unsorted:
            o = np.argsort(cs, kind="stable")
            cs, v = cs[o], v[o]
        rows.append(np.full(len(cs), i)); cols.append(cs); vals.append(v)
    if empty_rows:
        drop = set(rng.choice(n, size=empty_rows, replace=False).tolist())
        keep = [i not in drop for i in range(n)]
        rows = [r if k else r[:0] for r, k in zip(rows, keep)]
        cols = [c if k else c[:0] for c, k in zip(cols, keep)]
        vals = [v if k else v[:0] for v, k in zip(vals, keep)]
    rowptr = np.zeros(n + 1, np.int64)
    rowptr[1:] = np.cumsum([len(c) for c in cols])
    col = np.concatenate(cols).astype(np.int32) if n else np.zeros(0, np.int32)
    val = np.concatenate(vals).astype(np.float64) if n else np.zeros(0)
    return rowptr.astype(np.int32), col, val


def dense(n, rowptr, col, val):
    D = np.zeros((n, n)); P = np.zeros((n, n), dtype=bool)
    for i in range(n):
        for p in range(rowptr[i], rowptr[i + 1]):
            D[i, col[p]] += val[p]; P[i, col[p]] = True
    return D, P


def to_csr(Cd, Cp):
    n = Cd.shape[0]
    rp = np.zeros(n + 1, np.int64); cols, vals = [], []
    for i in range(n):
        js = np.nonzero(Cp[i])[0]
        cols.append(js); vals.append(Cd[i, js]); rp[i + 1] = rp[i] + len(js)
    return rp, np.concatenate(cols).astype(np.int32), np.concatenate(vals)


def main():
    rng = np.random.default_rng(20240601)
    specs = [
        dict(n=1, density=1.0), dict(n=7, density=0.3), dict(n=16, density=0.2, unsorted=True),
        dict(n=24, density=0.15, dups=True), dict(n=31, density=0.1, zeros=True, empty_rows=5),
        dict(n=40, density=0.3, unsorted=True, dups=True, zeros=True, empty_rows=3),
        dict(n=33, density=0.02),
    ]
    blob = {"ncases": np.int64(len(specs))}
    for t, sp in enumerate(specs):
        n = sp.pop("n"); dens = sp.pop("density")
        rp, col, val = random_case(rng, n, dens, **sp)
        D, P = dense(n, rp, col, val)
        Pi = P.astype(np.int64)
        blob[f"c{t}_n"] = np.int64(n)
        blob[f"c{t}_rowptr"], blob[f"c{t}_col"], blob[f"c{t}_val"] = rp, col, val
        for op, (Bd, Bp) in {"AA": (D, Pi), "AAt": (D.T, Pi.T)}.items():
            Cd = D @ Bd
            Cp = (Pi @ Bp) > 0
            crp, ccol, cval = to_csr(Cd, Cp)
            blob[f"c{t}_{op}_rowptr"], blob[f"c{t}_{op}_col"], blob[f"c{t}_{op}_val"] = crp, ccol, cval
    np.savez_compressed(OUT, **blob)
    print("wrote", OUT)


if __name__ == "__main__":
    main()
