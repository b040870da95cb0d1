"""Synthetic workloads of the batched and served configurations.

Matrices come from the library's seeded C++ generators (banded,
Erdos-Renyi, stencils, R-MAT); this module only draws their parameters.
"""
from __future__ import annotations

import numpy as np

from .engine import HostMatrix, banded, erdos_renyi, laplace2d, rmat, stencil27


def batch_specs(count: int = 4096, seed: int = 1, nmin: int = 64, nmax: int = 16384):
    """The uneven batch as generator calls: n log-uniform in [nmin, nmax];
    even index random-banded (half-bandwidth U{1..32}, 70 % band fill), odd
    index Erdos-Renyi (mean row nnz U(2, 20)); each matrix A*A or A*A^T with
    probability 1/2.  FP64 values U(-1,1).  Returns (callables that build the
    matrices, ops), so callers can generate them in parallel."""
    rng = np.random.default_rng(seed)
    specs = []
    for k in range(count):
        n = int(np.floor(np.exp(rng.uniform(np.log(nmin), np.log(nmax)))))
        s = seed * 1_000_003 + k
        if k % 2 == 0:
            bw = int(rng.integers(1, 33))
            specs.append(lambda n=n, bw=bw, s=s: banded(n, bw, 0.7, s))
        else:
            mu = float(rng.uniform(2.0, 20.0))
            specs.append(lambda n=n, mu=mu, s=s: erdos_renyi(n, mu, s))
    ops = ["AA" if rng.random() < 0.5 else "AAt" for _ in range(count)]
    return specs, ops


def generate(specs, workers: int = 16) -> list:
    """Run generator callables on a thread pool (the C++ generators release
    the interpreter lock), keeping their order."""
    from concurrent.futures import ThreadPoolExecutor
    with ThreadPoolExecutor(max(1, workers)) as ex:
        return list(ex.map(lambda f: f(), specs))


def serve_job(rng: np.random.Generator, nmin: int = 1000, nmax: int = 2_000_000, seed: int = 0):
    """One job of the served stream: n log-uniform in [nmin, nmax] drawn
    from the Laplacian / 27-point stencil / R-MAT families (grid sides from
    n^(1/2), n^(1/3); R-MAT scale ceil(log2 n)), op A*A or A*A^T.
    Returns (matrix, op, family)."""
    n = float(np.exp(rng.uniform(np.log(nmin), np.log(nmax))))
    fam = int(rng.integers(0, 3))
    op = "AA" if rng.random() < 0.5 else "AAt"
    if fam == 0:
        g = max(2, int(round(np.sqrt(n))))
        return laplace2d(g), op, "laplace2d"
    if fam == 1:
        g = max(2, int(round(n ** (1.0 / 3.0))))
        return stencil27(g, 0.1 if rng.random() < 0.5 else 0.0, seed), op, "stencil27"
    scale = max(6, int(np.ceil(np.log2(n))))
    return truncate(rmat(scale, 16, seed=seed), int(n)), op, "rmat"


def truncate(m: HostMatrix, n: int) -> HostMatrix:
    """Leading n x n block of m (rows and columns below n)."""
    if n >= m.n:
        return m
    rp = np.asarray(m.rowptr, np.int64)
    keep_rows = rp[n]
    col = np.asarray(m.col[:keep_rows])
    val = np.asarray(m.val[:keep_rows])
    rows = np.repeat(np.arange(n), np.diff(rp[:n + 1]))
    sel = col < n
    r, c, v = rows[sel], col[sel], val[sel]
    out_rp = np.zeros(n + 1, np.int64)
    np.add.at(out_rp, r + 1, 1)
    return HostMatrix(n, np.cumsum(out_rp).astype(np.int32), c.astype(np.int32), v.astype(np.float64), dict(m.meta))
