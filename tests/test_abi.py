"""C-ABI library: loads, exports every declared symbol, host-side helpers
(generators, CSR1 files, task parsing) work without a GPU."""
import ctypes as C
import re
import struct
from pathlib import Path

import numpy as np
import pytest

import appt_spgemm_amd as ap
from appt_spgemm_amd._abi import ApptError, OPS, lib
from tests._util import rows_sorted

HEADER = Path(__file__).resolve().parent.parent / "include" / "appt_spgemm.h"


def declared_functions():
    text = HEADER.read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(appt_[a-z0-9_]+)\s*\(", text)))


def test_every_declared_symbol_is_exported():
    L = lib()
    names = declared_functions()
    assert len(names) >= 20
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing


def test_task_strings():
    out = C.c_int()
    for t, v in OPS.items():
        assert lib().appt_op_from_task(t.encode(), C.byref(out)) == 0 and out.value == v
    assert lib().appt_op_from_task(b"AtA", C.byref(out)) == 1


def test_laplace_generator():
    m = ap.laplace2d(64)
    assert (m.n, m.nnz) == (4096, 20224)
    assert rows_sorted(m.rowptr, m.col)
    assert set(np.unique(m.val).tolist()) == {-1.0, 4.0}


@pytest.mark.parametrize("g", [4, 9])
def test_stencil_generator(g):
    m = ap.stencil27(g, 0.0, 1)
    assert m.nnz == (3 * g - 2) ** 3 and rows_sorted(m.rowptr, m.col)
    assert np.all((m.val >= -1) & (m.val < 1))
    u = ap.stencil27(g, 0.1, 1)
    rows = np.repeat(np.arange(u.n), np.diff(u.rowptr))
    assert np.all(np.isin(np.arange(u.n), rows[rows == u.col]))  # diagonal kept
    offd = m.nnz - m.n
    kept = u.nnz - u.n
    assert abs(kept / offd - 0.9) < 0.05
    A = u.to_scipy()
    assert (A != A.T).nnz > 0  # unsymmetric


def test_generators_deterministic_and_canonical():
    for f in (lambda s: ap.rmat(10, 16, seed=s), lambda s: ap.erdos_renyi(1000, 8.0, s),
              lambda s: ap.banded(1000, 12, 0.7, s), lambda s: ap.stencil27(6, 0.1, s)):
        a, b, c = f(3), f(3), f(4)
        np.testing.assert_array_equal(a.col, b.col)
        np.testing.assert_array_equal(a.val, b.val)
        assert not (a.nnz == c.nnz and np.array_equal(a.col, c.col) and np.array_equal(a.val, c.val))
        assert rows_sorted(a.rowptr, a.col)  # strictly ascending: duplicates merged


def test_rmat_shape():
    m = ap.rmat(12, 16, seed=1)
    assert m.n == 4096 and 0.5 * 16 * 4096 < m.nnz <= 16 * 4096
    deg = np.diff(m.rowptr)
    assert deg.max() > 20 * deg.mean()  # power law
    assert np.all((m.val >= 0) & (m.val < 1))


def test_banded_shape():
    m = ap.banded(500, 7, 0.7, 9)
    rows = np.repeat(np.arange(m.n), np.diff(m.rowptr))
    assert np.abs(rows - m.col).max() <= 7
    assert np.all(np.isin(np.arange(m.n), rows[rows == m.col]))


def test_csr1_roundtrip_and_layout(tmp_path):
    m = ap.stencil27(5, 0.1, 2)
    p = tmp_path / "a.csr"
    ap.csr1_write(str(p), m)
    raw = p.read_bytes()
    # "CSR1", int32 n, int32 pad, int64 nnz, indptr, indices, data (src/export_bin.py:6-8)
    assert raw[:4] == b"CSR1"
    n, pad, nnz = struct.unpack("<iiq", raw[4:20])
    assert (n, pad, nnz) == (m.n, 0, m.nnz)
    assert len(raw) == 20 + 4 * (n + 1) + 12 * nnz
    np.testing.assert_array_equal(np.frombuffer(raw[20:20 + 4 * (n + 1)], "<i4"), m.rowptr)
    r = ap.csr1_read(str(p))
    np.testing.assert_array_equal(r.rowptr, m.rowptr)
    np.testing.assert_array_equal(r.col, m.col)
    np.testing.assert_array_equal(r.val, m.val)


def test_csr1_rejects_bad_files(tmp_path):
    m = ap.laplace2d(4)
    good = tmp_path / "g.csr"
    ap.csr1_write(str(good), m)
    raw = good.read_bytes()
    (tmp_path / "magic.csr").write_bytes(b"CSR2" + raw[4:])
    (tmp_path / "trunc.csr").write_bytes(raw[:-3])
    bad = bytearray(raw)
    bad[20 + 4 * (m.n + 1):20 + 4 * (m.n + 1) + 4] = struct.pack("<i", 99)
    (tmp_path / "range.csr").write_bytes(bytes(bad))
    for name, status in (("magic.csr", "APPT_ERR_IO"), ("trunc.csr", "APPT_ERR_IO"),
                         ("range.csr", "APPT_ERR_INVALID"), ("missing.csr", "APPT_ERR_IO")):
        with pytest.raises(ApptError) as e:
            ap.csr1_read(str(tmp_path / name))
        assert e.value.status_name == status


def test_generator_argument_errors():
    with pytest.raises(ApptError) as e:
        ap.stencil27(10, 1.5, 0)
    assert e.value.status_name == "APPT_ERR_INVALID"
    with pytest.raises(ApptError):
        ap.rmat(12, 16, a=0.9, b=0.2, c=0.2)


def test_panel_split():
    # panels never exceed the byte bound unless a single row does
    rp = np.concatenate([[0], np.cumsum(np.array([5, 0, 300, 7, 7, 7, 1, 0, 0, 40], np.int64))])
    parts = ap.panels(rp, max_bytes=12 * 20 + 8 * 4)
    assert parts[0][0] == 0 and parts[-1][1] == len(rp) - 1
    assert all(a < b for a, b in parts) and all(parts[k][1] == parts[k + 1][0] for k in range(len(parts) - 1))
    for a, b in parts:
        size = 12 * (rp[b] - rp[a]) + 8 * (b - a + 1)
        assert size <= 12 * 20 + 8 * 4 or b == a + 1
