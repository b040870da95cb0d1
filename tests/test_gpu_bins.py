"""Every row bin of both passes receives rows under test, and every such run
matches the oracle.

Matrices are built so that each kind of output row lands where it is meant
to: a query row has nA entries pointing at distinct source rows of length L
whose columns lie in a window of `spread` columns, so its product count is
nA * L, its mean B-row length is L (lanes per row) and its output span is the
spread (window or list when narrow, word hash or register sort when
scattered, workgroup or heavy bins when long).  Each matrix runs under the
default tuning and with the register-sort bins switched off; the bin report
of every run is summed and printed as a coverage table."""
import sys

import numpy as np
import pytest

import appt_spgemm_amd as ap
from tests._util import assert_product, ref

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def mix(n, specs, rows_per_spec, seed):
    """Square matrix of order n from specs (nA, L, spread): for each spec,
    `rows_per_spec` query rows with nA entries on distinct source rows of L
    entries each, the source rows' columns uniform in one window of `spread`
    columns (the whole range when spread >= n).  Other rows are empty."""
    rng = np.random.default_rng(seed)
    rows = {}
    nxt = 0
    for nA, L, spread in specs:
        m = max(nA, 64)
        src = np.arange(nxt, nxt + m)
        nxt += m
        lo = 0 if spread >= n else int(rng.integers(0, n - spread))
        w = min(spread, n)
        for k in src:
            rows[int(k)] = np.sort(rng.choice(w, size=L, replace=False)) + lo
        for _ in range(rows_per_spec):
            rows[nxt] = np.sort(rng.choice(src, size=nA, replace=False))
            nxt += 1
    assert nxt <= n
    lens = np.zeros(n, np.int64)
    for r, c in rows.items():
        lens[r] = len(c)
    rp = np.zeros(n + 1, np.int64)
    rp[1:] = np.cumsum(lens)
    col = np.concatenate([rows[r] for r in sorted(rows)]).astype(np.int32)
    return ap.HostMatrix(n, rp.astype(np.int32), col, rng.uniform(-1, 1, len(col)))


# (nA, L): lanes per row G = 8, 16, 32, 64 for products classes <= 128, <= 512, <= 2048
CLASS_G = {0: [(8, 8), (8, 12), (4, 24), (2, 48)],
           1: [(40, 8), (24, 14), (12, 28), (6, 60)],
           2: [(150, 8), (80, 15), (40, 30), (20, 60)]}
HASH_G = {0: [(8, 8), (8, 12), (4, 24), (2, 48)],
          1: [(40, 8), (24, 14), (12, 28), (6, 60)],
          2: [(80, 8), (40, 16), (20, 32), (10, 64)]}


def clustered():
    specs = []
    for c, spread in ((0, 512), (1, 1024), (2, 1800)):      # few words: word lists
        specs += [(nA, L, spread) for nA, L in CLASS_G[c]]
    specs += [(nA, L, 8000) for c in (0, 1) for nA, L in CLASS_G[c]]  # > 64 words: no list
    specs += [(150, 8, 8000), (80, 15, 8000), (40, 30, 8000), (20, 60, 8000)]
    specs += [(150, 8, 1000)]                               # class 2 in <= 32 words: 8-lane list
    specs += [(60, 60, 6000)]                               # past the classes, > 64 words
    specs += [(40, 40, 96)]                                 # <= 8 span words
    return "clustered", mix(300_000, specs, 60, 1)


def scattered():
    specs = [(nA, L, 10 ** 9) for c in (0, 1, 2) for nA, L in HASH_G[c]]
    specs += [(k, 12, 10 ** 9) for k in (1, 2, 4, 8, 16, 32, 64, 128)]  # register-sort sizes
    return "scattered", mix(400_000, specs, 60, 2)


def wide():
    # outputs over 2M columns: workgroup word hash (2400 products) and heavy rows
    return "wide", mix(2_000_000, [(40, 60, 10 ** 9), (500, 60, 10 ** 9)], 8, 3)


# row templates off: these matrices test the other bins (their patterns do
# not repeat, so templates would not be taken anyway)
TUNINGS = {"default": {"row_templates": -1}, "no-sort": {"sort_max_products": -1, "row_templates": -1}}


def test_every_bin_receives_rows():
    sym_tot, num_tot = {}, {}
    mats = [clustered(), scattered(), wide()]
    for tname, tuning in TUNINGS.items():
        eng = ap.Engine(0, tuning=tuning)
        try:
            for mname, m in mats:
                for op in ("AA",) if mname != "clustered" else ("AA", "AAtS"):
                    say("case bins", tname, mname, op, "n", m.n, "nnz", m.nnz)
                    got = eng.multiply(m, op)
                    assert_product(got, ref(m, op), exact=False)
                    br = eng.bin_report()
                    for k, v in br["symbolic"].items():
                        sym_tot[k] = sym_tot.get(k, 0) + v[0]
                    for k, v in br["numeric"].items():
                        num_tot[k] = num_tot.get(k, 0) + v[0]
        finally:
            eng.close()
    names = [ap.bin_name(b) for b in range(1, ap.NUM_BINS)]
    say("bin coverage (rows summed over all runs):")
    say(f"{'bin':16s} {'symbolic':>10s} {'numeric':>10s}")
    for nm in names:
        say(f"{nm:16s} {sym_tot.get(nm, 0):10d} {num_tot.get(nm, 0):10d}")
    # list bins are numeric only; template rows are covered by test_gpu_templates
    sym_bins = [nm for nm in names if not nm.startswith("list/") and nm != "template"]
    miss_sym = [nm for nm in sym_bins if sym_tot.get(nm, 0) == 0]
    miss_num = [nm for nm in names if num_tot.get(nm, 0) == 0 and nm != "template"]
    assert not miss_sym and not miss_num, f"unreached: symbolic {miss_sym}, numeric {miss_num}"


@pytest.mark.parametrize("op", ["AA", "AAt", "AAtS"])
def test_int64_row_pointers(op):
    """rowptr64_min_nnz forces 64-bit row pointers on small results (the
    layout every result of 2^31 entries or more takes)."""
    eng = ap.Engine(0, tuning={"rowptr64_min_nnz": 1})
    try:
        cases = [("stencil", ap.stencil27(12, 0.1, 3)), ("rmat", ap.rmat(11, 16, seed=4)),
                 ("er", ap.erdos_renyi(5000, 6.0, 2)),
                 ("empty_rows", ap.HostMatrix(6, np.array([0, 2, 2, 3, 3, 5, 5], np.int32),
                                              np.array([1, 4, 0, 2, 5], np.int32),
                                              np.array([1.0, 2.0, 3.0, 4.0, 5.0])))]
        for name, m in cases:
            say("case int64 rowptr", name, op)
            A = eng.upload(m)
            try:
                R = eng.compute(A, op)
                try:
                    assert R.res.rowptr_is_64 == 1
                    got = R.download()
                    assert_product(got, ref(m, op), exact=False)
                    sub = R.download_rows(1, m.n - 1)
                    rp = got.rowptr
                    np.testing.assert_array_equal(sub.rowptr, rp[1:m.n] - rp[1])
                    np.testing.assert_array_equal(sub.col, got.col[rp[1]:rp[m.n - 1]])
                    R.eliminate_zeros()
                    assert R.res.rowptr_is_64 == 1
                    z = R.download()
                    keep = got.val != 0.0
                    np.testing.assert_array_equal(z.col, got.col[keep])
                    np.testing.assert_array_equal(z.val, got.val[keep])
                finally:
                    R.release()
                if op != "AAtS":
                    R = eng.compute_rows(A, op, m.n // 3, m.n)
                    try:
                        assert R.res.rowptr_is_64 == 1
                        h = R.download()
                        full = ref(m, op)
                        lo, hi = full[0][m.n // 3], full[0][m.n]
                        np.testing.assert_array_equal(h.rowptr, full[0][m.n // 3:] - lo)
                        np.testing.assert_array_equal(h.col, full[1][lo:hi])
                    finally:
                        R.release()
            finally:
                A.free()
    finally:
        eng.close()
