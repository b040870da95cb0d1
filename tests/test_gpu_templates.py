"""Row templates: rows whose relative pattern, and whose B rows' patterns,
repeat take the template bin (no symbolic pass).  Their results must equal
the oracle bit for bit and equal an engine with templates switched off."""
import sys

import numpy as np
import pytest

import appt_spgemm_amd as ap
from tests._util import assert_product, ref

pytestmark = pytest.mark.gpu
OPS = ["AA", "AAt", "AAtS"]


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def engs():
    # templates looked for at any size (the default skips operands under 2^18 entries)
    on, off = ap.Engine(0, tuning={"row_templates": 1}), ap.Engine(0, tuning={"row_templates": -1})
    yield on, off
    on.close()
    off.close()


def cases():
    yield "laplace2d 64", ap.laplace2d(64), True
    yield "stencil27 g=14", ap.stencil27(14, 0.0, 5), True
    # wide enough (56-column lines) that template rows are walked in tiles of
    # fewer lines than a plane
    yield "stencil27 g=56", ap.stencil27(56, 0.0, 6), True
    yield "stencil27 g=14 drop", ap.stencil27(14, 0.1, 5), False
    yield "banded full 3000 bw 8", ap.banded(3000, 8, 1.0, 1), True
    yield "banded 0.7", ap.banded(3000, 8, 0.7, 2), False
    yield "erdos-renyi", ap.erdos_renyi(4000, 6.0, 3), False
    # hub rows of more than 64 entries: their B-row metadata comes from the
    # pass the class pass launches for long rows (values accumulate in
    # workgroup bins: within tolerance, not bit for bit)
    yield "rmat 12 long rows", ap.rmat(12, 16, seed=3), False
    # a stencil with one row's columns shuffled (same set, other storage order):
    # it and the rows that reference it leave the template
    m = ap.stencil27(10, 0.0, 7)
    col = m.col.copy(); val = m.val.copy()
    s, e = m.rowptr[555], m.rowptr[556]
    col[s:e] = col[s:e][::-1]; val[s:e] = val[s:e][::-1]
    yield "stencil27 one reversed row", ap.HostMatrix(m.n, m.rowptr, col, val), True


@pytest.mark.parametrize("op", OPS)
def test_templates_match_oracle(engs, op):
    on, off = engs
    for name, m, expect in cases():
        say("case templates", name, op)
        got = on.multiply(m, op)
        rows = on.bin_report()["numeric"].get("template", (0, 0))[0]
        say("   template rows", rows, "of", m.n)
        exact = (name != "stencil27 one reversed row" or op == "AA") and not name.startswith("rmat")
        assert_product(got, ref(m, op), exact=exact)
        if expect:
            assert rows > m.n // 4
        plain = off.multiply(m, op)
        assert "template" not in off.bin_report()["numeric"]
        np.testing.assert_array_equal(got.rowptr, plain.rowptr)
        np.testing.assert_array_equal(got.col, plain.col)
        if op != "AAtS" and exact:
            np.testing.assert_array_equal(got.val, plain.val)


@pytest.mark.parametrize("op", ["AA", "AAt"])
def test_templates_in_row_panels(engs, op):
    on, _ = engs
    m = ap.stencil27(16, 0.0, 3)
    want = ref(m, op)
    A = on.upload(m)
    try:
        rp = on.symbolic_rowptr(A, op)
        np.testing.assert_array_equal(rp, want[0])
        for r0, r1 in ((0, 1000), (1000, 2777), (2777, m.n)):
            R = on.compute_rows(A, op, r0, r1)
            try:
                assert on.bin_report()["numeric"].get("template", (0, 0))[0] > 0
                h = R.download()
            finally:
                R.release()
            lo, hi = want[0][r0], want[0][r1]
            np.testing.assert_array_equal(h.rowptr, want[0][r0:r1 + 1] - lo)
            np.testing.assert_array_equal(h.col, want[1][lo:hi])
            np.testing.assert_array_equal(h.val, want[2][lo:hi])
    finally:
        A.free()
