"""Sweep harness host logic: the stdout protocol parsers, the checksum match
rule, the products bound and the repetition policy (bench/run_compare.py:23-43,69;
bench/run_dispatch.py:19-20), plus one small GPU sweep through the tools."""
import csv

import numpy as np
import pytest

import appt_spgemm_amd as ap
import oracle
from appt_spgemm_amd import sweep

OUT = """  __chk s=1.000000e+00 as=2.000000e+00 sq=3.000000e+00
RESULT,t,AA,10,20,30,0.5000
  __chk s=-4.000000e+00 as=5.000000e+00 sq=6.000000e+00
RESULT,t,AAt,10,20,31,0.2500
"""


def test_parse_results_pairs_checksums_in_order():
    r = sweep.parse_results(OUT)
    assert r["AA"] == (30, 0.5, (1.0, 2.0, 3.0))
    assert r["AAt"] == (31, 0.25, (-4.0, 5.0, 6.0))
    assert sweep.parse_results("nothing") == {}


def test_parse_dispatch_line():
    d = sweep.parse_dispatch("DISPATCH,t,AA,ours,dcv=0.123,fill=4.5,C_nnz=77,desc_ms=0.0100,run_ms=1.5000,"
                             "total_ms=1.5100,rule=cusparse,bins=3\n")
    assert d["rule"] == "cusparse" and d["C_nnz"] == 77 and d["desc_ms"] == 0.01 and d["total_ms"] == 1.51
    assert sweep.parse_dispatch("RESULT,x") is None


def test_match_rule():
    a = (10, 1.0, (1.0, 2.0, 3.0))
    assert sweep.match(a, (10, 9.0, (1.0 + 5e-7, 2.0, 3.0)))
    assert not sweep.match(a, (10, 1.0, (1.0 + 2e-6, 2.0, 3.0)))
    assert not sweep.match(a, (11, 1.0, (1.0, 2.0, 3.0)))
    # absolute near zero: max(1, |b|) in the denominator
    assert sweep.rel_ok(1e-7, 0.0) and not sweep.rel_ok(2e-6, 0.0)


def test_products_bound_matches_oracle():
    for m in (ap.laplace2d(20), ap.erdos_renyi(500, 5.0, 1), ap.rmat(9, 8, seed=2)):
        aa, aat = sweep.products(m)
        assert aa == int(np.sum(oracle.row_products(m.n, m.rowptr, m.col, m.rowptr)))
        trp = oracle.transpose(m.n, m.rowptr, m.col, m.val)[0]
        assert aat == int(np.sum(oracle.row_products(m.n, m.rowptr, m.col, trp)))


def test_reps_policy():
    assert [sweep.reps_for(x) for x in (0, 199_999, 200_000, 1_999_999, 2_000_000)] == [20, 20, 8, 8, 3]


def test_dir_suite_reads_csr1_and_npz(tmp_path):
    import scipy.sparse as sp
    m = ap.laplace2d(8)
    ap.csr1_write(str(tmp_path / "a.csr"), m)
    sp.save_npz(tmp_path / "b.npz", m.to_scipy())
    s = dict(sweep.dir_suite(tmp_path))
    for k in ("a", "b"):
        x = s[k]()
        np.testing.assert_array_equal(x.rowptr, m.rowptr)
        np.testing.assert_array_equal(x.col, m.col)
        np.testing.assert_array_equal(x.val, m.val)


@pytest.mark.gpu
def test_small_sweep_matches_vendor(tmp_path):
    suite = [("laplace2d_64", lambda: ap.laplace2d(64)), ("stencil27_12", lambda: ap.stencil27(12, 0.1, 1)),
             ("rmat_10", lambda: ap.rmat(10, 8, seed=1)), ("er_3000", lambda: ap.erdos_renyi(3000, 6.0, 2))]
    print("case sweep compare/symabl/dispatch", flush=True)
    rows = list(csv.DictReader(open(sweep.compare(suite, tmp_path))))
    assert len(rows) == 8 and all(r["status"] == "ok" and r["chk_match"] == "1" for r in rows)
    rows = list(csv.DictReader(open(sweep.symabl(suite, tmp_path))))
    assert len(rows) == 4 and all(r["match"] == "1" for r in rows)
    rows = list(csv.DictReader(open(sweep.dispatch(suite, tmp_path))))
    assert len(rows) == 8 and all(r["rule"] in ("ours", "cusparse") for r in rows)
