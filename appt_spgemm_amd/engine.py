"""Python mirror of the reference's engine interface over the C ABI.

Reference interface (Wendy-Hamlet/appt_spgemm):
  run_ours(task, dOff, dCol, dVal, n, nnz, reps) -> EngResult
      src/spgemm_engine.cuh:132-133, EngResult :129, tasks "AA" | "AAt" | "AAtS" :134
  load_bin / write_bin (CSR1)            src/spgemm_engine.cuh:22-32, src/export_bin.py:20-29
  dispatcher descriptors                  bench/dispatch.cu:26-38

All compute runs in ``libappt_spgemm.so`` on the GPU; this module only moves
host arrays and handles.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, field

import numpy as np

from ._abi import (JOB_STATES, OPS, STATUS_NAMES, BatchBuffers, BatchInfo, BinReport, CsrResult, CsrView, Descriptors,
                   EngResult, HostCsr, JobInfo, Options, PhaseTimes, ServerOptions, ServerStats, Tuning, check,
                   lib)


# ---------------------------------------------------------------------------
# host matrices
# ---------------------------------------------------------------------------
@dataclass
class HostMatrix:
    """Square CSR matrix in host memory (int32 indices, fp64 values)."""
    n: int
    rowptr: np.ndarray
    col: np.ndarray
    val: np.ndarray
    meta: dict = field(default_factory=dict)

    @property
    def nnz(self) -> int:
        return int(self.col.shape[0])

    def _as_c(self) -> HostCsr:
        self.rowptr = np.ascontiguousarray(self.rowptr, dtype=np.int32)
        self.col = np.ascontiguousarray(self.col, dtype=np.int32)
        self.val = np.ascontiguousarray(self.val, dtype=np.float64)
        return HostCsr(self.n, self.nnz, self.rowptr.ctypes.data_as(C.POINTER(C.c_int32)),
                       self.col.ctypes.data_as(C.POINTER(C.c_int32)),
                       self.val.ctypes.data_as(C.POINTER(C.c_double)))

    def to_scipy(self):
        import scipy.sparse as sp
        return sp.csr_matrix((self.val, self.col, self.rowptr), shape=(self.n, self.n))


def _take(h: HostCsr, meta: dict) -> HostMatrix:
    """Copy a library-owned host CSR into numpy and free it."""
    n, nnz = int(h.n), int(h.nnz)
    rp = np.ctypeslib.as_array(h.rowptr, shape=(n + 1,)).copy()
    col = np.ctypeslib.as_array(h.col, shape=(max(nnz, 1),))[:nnz].copy()
    val = np.ctypeslib.as_array(h.val, shape=(max(nnz, 1),))[:nnz].copy()
    lib().appt_host_csr_free(C.byref(h))
    return HostMatrix(n, rp, col, val, meta)


def laplace2d(g: int) -> HostMatrix:
    h = HostCsr()
    check(lib().appt_gen_laplace2d(g, C.byref(h)), "appt_gen_laplace2d")
    return _take(h, {"gen": "laplace2d", "g": g})


def stencil27(g: int, drop_p: float = 0.0, seed: int = 0) -> HostMatrix:
    h = HostCsr()
    check(lib().appt_gen_stencil27(g, drop_p, seed, C.byref(h)), "appt_gen_stencil27")
    return _take(h, {"gen": "stencil27", "g": g, "drop_p": drop_p, "seed": seed})


def rmat(scale: int, edge_factor: int = 16, a: float = 0.57, b: float = 0.19, c: float = 0.19,
         seed: int = 1) -> HostMatrix:
    h = HostCsr()
    check(lib().appt_gen_rmat(scale, edge_factor, a, b, c, seed, C.byref(h)), "appt_gen_rmat")
    return _take(h, {"gen": "rmat", "scale": scale, "edge_factor": edge_factor, "abc": (a, b, c),
                     "seed": seed})


def banded(n: int, bw: int, fill: float = 0.7, seed: int = 0) -> HostMatrix:
    h = HostCsr()
    check(lib().appt_gen_banded(n, bw, fill, seed, C.byref(h)), "appt_gen_banded")
    return _take(h, {"gen": "banded", "n": n, "bw": bw, "fill": fill, "seed": seed})


def erdos_renyi(n: int, mean_deg: float, seed: int = 0) -> HostMatrix:
    h = HostCsr()
    check(lib().appt_gen_erdos_renyi(n, mean_deg, seed, C.byref(h)), "appt_gen_erdos_renyi")
    return _take(h, {"gen": "erdos_renyi", "n": n, "mean_deg": mean_deg, "seed": seed})


def csr1_read(path: str) -> HostMatrix:
    """load_bin (src/spgemm_engine.cuh:22-32) with structural validation."""
    h = HostCsr()
    check(lib().appt_csr1_read(str(path).encode(), C.byref(h)), "appt_csr1_read")
    return _take(h, {"path": str(path)})


def csr1_write(path: str, m: HostMatrix) -> None:
    """write_bin (src/export_bin.py:20-29).  Writes rows as given."""
    h = m._as_c()
    check(lib().appt_csr1_write(str(path).encode(), C.byref(h)), "appt_csr1_write")


def bin_name(b: int) -> str | None:
    """Name of row bin ``b`` (``appt_bin_name``), None past the last bin."""
    p = lib().appt_bin_name(b)
    return p.decode() if p else None


def _num_bins() -> int:
    b = 0
    while lib().appt_bin_name(b):
        b += 1
    return b


# ---------------------------------------------------------------------------
# device engine
# ---------------------------------------------------------------------------
class DeviceMatrix:
    """Device CSR uploaded through the engine (owned; freed by ``free``)."""

    def __init__(self, engine: "Engine", view: CsrView):
        self._engine = engine
        self.view = view

    @property
    def n(self) -> int:
        return int(self.view.n)

    @property
    def nnz(self) -> int:
        return int(self.view.nnz)

    def free(self) -> None:
        if self.view.rowptr:
            check(lib().appt_device_free_view(self._engine.handle, C.byref(self.view)), "free_view")


class DeviceResult:
    """Device CSR result; buffers belong to the engine until ``release``."""

    def __init__(self, engine: "Engine", res: CsrResult):
        self._engine = engine
        self.res = res

    @property
    def n(self) -> int:
        return int(self.res.n)

    @property
    def nnz(self) -> int:
        return int(self.res.nnz)

    def download(self) -> HostMatrix:
        n, nnz = self.n, self.nnz
        rp = np.empty(n + 1, dtype=np.int64)
        col = np.empty(max(nnz, 1), dtype=np.int32)
        val = np.empty(max(nnz, 1), dtype=np.float64)
        check(lib().appt_csr_download(self._engine.handle if self._engine else None, C.byref(self.res),
                                      rp.ctypes.data_as(C.POINTER(C.c_int64)),
                                      col.ctypes.data_as(C.POINTER(C.c_int32)),
                                      val.ctypes.data_as(C.POINTER(C.c_double))), "csr_download")
        return HostMatrix(n, rp, col[:nnz], val[:nnz])

    def download_rows(self, r0: int, r1: int) -> HostMatrix:
        """Rows [r0, r1) as a host matrix (r1 - r0 rows, global columns)."""
        L = lib()
        h = self._engine.handle if self._engine else None
        rp = np.empty(r1 - r0 + 1, dtype=np.int64)
        check(L.appt_csr_download_rows(h, C.byref(self.res), r0, r1, rp.ctypes.data_as(C.POINTER(C.c_int64)),
                                       None, None), "csr_download_rows")
        cnt = int(rp[-1])
        col = np.empty(max(cnt, 1), dtype=np.int32)
        val = np.empty(max(cnt, 1), dtype=np.float64)
        check(L.appt_csr_download_rows(h, C.byref(self.res), r0, r1, rp.ctypes.data_as(C.POINTER(C.c_int64)),
                                       col.ctypes.data_as(C.POINTER(C.c_int32)),
                                       val.ctypes.data_as(C.POINTER(C.c_double))), "csr_download_rows")
        return HostMatrix(r1 - r0, rp, col[:cnt], val[:cnt])

    def digest(self, engine: "Engine | None" = None) -> tuple:
        """(structure, values) 64-bit digests computed on the device."""
        eng = engine or self._engine
        out = (C.c_uint64 * 2)()
        check(lib().appt_csr_digest(eng.handle, C.byref(self.res), out), "csr_digest")
        return int(out[0]), int(out[1])

    def eliminate_zeros(self) -> "DeviceResult":
        """Numeric semantics: drop entries whose value is exactly zero (in place)."""
        check(lib().appt_csr_eliminate_zeros(self._engine.handle, C.byref(self.res)), "csr_eliminate_zeros")
        return self

    def release(self) -> None:
        if self.res._priv:
            check(lib().appt_csr_release(self._engine.handle, C.byref(self.res)), "csr_release")


class Engine:
    """One engine per host thread; owns a HIP stream and a device memory pool."""

    def __init__(self, device: int = 0, hbm_budget_bytes: int = 0, tuning: dict | None = None,
                 stream: int | None = None, timing: bool = False):
        opt = Options()
        opt.device = device
        opt.stream = stream
        opt.hbm_budget_bytes = hbm_budget_bytes
        if tuning:
            t = Tuning()
            for k, v in tuning.items():
                if k in ("f_class_max", "group_for_len", "class_lds_bytes"):
                    arr = getattr(t, k)
                    for i, x in enumerate(v):
                        arr[i] = int(x)
                else:
                    setattr(t, k, int(v))
            opt.tuning = t
        h = C.c_void_p()
        check(lib().appt_engine_create(C.byref(opt), C.byref(h)), "appt_engine_create")
        self.handle = h
        if timing:
            self.set_timing(True)

    def close(self) -> None:
        if self.handle:
            check(lib().appt_engine_destroy(self.handle), "appt_engine_destroy")
            self.handle = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def set_timing(self, on: bool) -> None:
        check(lib().appt_engine_set_timing(self.handle, 1 if on else 0), "set_timing")

    def phase_times(self) -> dict:
        t = PhaseTimes()
        check(lib().appt_engine_phase_times(self.handle, C.byref(t)), "phase_times")
        return {f: getattr(t, f) for f, _ in PhaseTimes._fields_}

    def bin_report(self) -> dict:
        """Rows per symbolic / numeric bin of the last compute call, by bin name
        (bins that received no rows are left out)."""
        r = BinReport()
        check(lib().appt_engine_bin_report(self.handle, C.byref(r)), "bin_report")
        out = {"symbolic": {}, "numeric": {}}
        for b in range(1, r.n_bins):
            name = lib().appt_bin_name(b).decode()
            if r.sym_rows[b]:
                out["symbolic"][name] = (int(r.sym_rows[b]), int(r.sym_metric_max[b]))
            if r.num_rows[b]:
                out["numeric"][name] = (int(r.num_rows[b]), int(r.num_metric_max[b]))
        return out

    def device_bytes(self) -> int:
        return int(lib().appt_engine_device_bytes(self.handle))

    def trim(self) -> None:
        check(lib().appt_engine_trim(self.handle), "trim")

    def upload(self, m: HostMatrix) -> DeviceMatrix:
        v = CsrView()
        h = m._as_c()
        check(lib().appt_device_upload(self.handle, C.byref(h), C.byref(v)), "device_upload")
        return DeviceMatrix(self, v)

    def check(self, A: DeviceMatrix) -> None:
        """Structure check of a device CSR on the GPU; raises on invalid input."""
        check(lib().appt_csr_check(self.handle, C.byref(A.view)), "csr_check")

    def compute(self, A: DeviceMatrix, op: str = "AA") -> DeviceResult:
        r = CsrResult()
        check(lib().appt_spgemm_compute(self.handle, C.byref(A.view), OPS[op], C.byref(r)),
              f"appt_spgemm_compute({op})")
        return DeviceResult(self, r)

    def compute_rows(self, A: DeviceMatrix, op: str, r0: int, r1: int) -> DeviceResult:
        """Row panel [r0, r1) of op(A) (global column indices).  After
        symbolic_rowptr(A, op) on this engine, heavy rows take their counts
        from that plan: A must not change in between, and a matrix uploaded
        into the freed buffers of a planned one must be planned again."""
        r = CsrResult()
        check(lib().appt_spgemm_compute_rows(self.handle, C.byref(A.view), OPS[op], r0, r1, C.byref(r)),
              f"appt_spgemm_compute_rows({op}, {r0}, {r1})")
        return DeviceResult(self, r)

    def symbolic_rowptr(self, A: DeviceMatrix, op: str = "AA") -> np.ndarray:
        """Exact int64 row pointers of op(A) from the symbolic phase; the
        engine keeps the rows' symbolic results for compute_rows on A."""
        rp = np.empty(A.n + 1, dtype=np.int64)
        check(lib().appt_spgemm_symbolic_rowptr(self.handle, C.byref(A.view), OPS[op],
                                                rp.ctypes.data_as(C.POINTER(C.c_int64))), "symbolic_rowptr")
        return rp

    def multiply(self, m: HostMatrix, op: str = "AA", eliminate_zeros: bool = False) -> HostMatrix:
        """Upload, compute, download, release: C = op(m) as a host matrix
        (structural semantics; numeric semantics with eliminate_zeros)."""
        A = self.upload(m)
        try:
            R = self.compute(A, op)
            try:
                if eliminate_zeros:
                    R.eliminate_zeros()
                return R.download()
            finally:
                R.release()
        finally:
            A.free()

    def run_timed(self, task: str, A: DeviceMatrix, reps: int) -> EngResult:
        out = EngResult()
        check(lib().appt_engine_run_timed(self.handle, task.encode(), C.byref(A.view), reps,
                                          C.byref(out)), "run_timed")
        return out

    def descriptors(self, A: DeviceMatrix, op: str = "AA") -> dict:
        d = Descriptors()
        check(lib().appt_dispatch_descriptors(self.handle, C.byref(A.view), OPS[op], C.byref(d)),
              "dispatch_descriptors")
        return {f: getattr(d, f) for f, _ in Descriptors._fields_}


def _tuning_dict(t: Tuning) -> dict:
    return {"f_class_max": [int(x) for x in t.f_class_max if x > 0], "wg_lds_bytes": int(t.wg_lds_bytes),
            "lds_budget_bytes": int(t.lds_budget_bytes), "bitmap_words_factor": int(t.bitmap_words_factor),
            "group_for_len": [int(x) for x in t.group_for_len[:3]], "heavy_window_cols": int(t.heavy_window_cols),
            "sort_span_factor": int(t.sort_span_factor),
            "sort_span_factor_symbolic": int(t.sort_span_factor_symbolic),
            "transpose_radix_min_nnz": int(t.transpose_radix_min_nnz),
            "mirror_radix_min_nnz": int(t.mirror_radix_min_nnz),
            "class_lds_bytes": [int(x) for x in t.class_lds_bytes if x > 0],
            "sort_max_products": int(t.sort_max_products), "rowptr64_min_nnz": int(t.rowptr64_min_nnz),
            "row_templates": int(t.row_templates), "gram": int(t.gram)}


def panels(rowptr: np.ndarray, max_bytes: int) -> list:
    """Split rows into contiguous panels whose C (4 or 8 B per row pointer,
    12 B per entry) stays within max_bytes; a single row larger than that is
    a panel of its own."""
    rowptr = np.asarray(rowptr, np.int64)
    n = len(rowptr) - 1
    out, r0 = [], 0
    while r0 < n:
        # largest r1 with 12 * (rowptr[r1] - rowptr[r0]) + 8 * (r1 - r0 + 1) <= max_bytes
        lo, hi = r0 + 1, n
        while lo < hi:
            mid = (lo + hi + 1) // 2
            if 12 * (rowptr[mid] - rowptr[r0]) + 8 * (mid - r0 + 1) <= max_bytes:
                lo = mid
            else:
                hi = mid - 1
        out.append((r0, lo))
        r0 = lo
    return out


def pack(mats: list) -> HostMatrix:
    """Packed batch layout: global row pointers, column indices local to each matrix."""
    nnz = np.array([m.nnz for m in mats], dtype=np.int64)
    base = np.concatenate([[0], np.cumsum(nnz)])
    parts = [np.asarray(m.rowptr[:-1], np.int64) + base[k] for k, m in enumerate(mats)]
    rp = np.concatenate(parts + [np.array([base[-1]], np.int64)]) if mats else np.zeros(1, np.int64)
    col = np.concatenate([np.asarray(m.col, np.int32) for m in mats]) if mats else np.zeros(0, np.int32)
    val = np.concatenate([np.asarray(m.val, np.float64) for m in mats]) if mats else np.zeros(0)
    return HostMatrix(int(sum(m.n for m in mats)), rp.astype(np.int32), col, val)


class Batch:
    """Many matrices (A_k*A_k or A_k*A_k^T each) as one block-diagonal product:
    planned once, then replayed from a captured HIP graph (``appt_batch_*``)."""

    def __init__(self, engine: Engine, ns, nnzs, ops):
        self._engine = engine
        self.count = len(ns)
        n_arr = (C.c_int32 * max(self.count, 1))(*[int(x) for x in ns])
        z_arr = (C.c_int64 * max(self.count, 1))(*[int(x) for x in nnzs])
        o_arr = (C.c_int * max(self.count, 1))(*[OPS[o] for o in ops])
        h = C.c_void_p()
        check(lib().appt_batch_create(engine.handle, self.count, n_arr, z_arr, o_arr, C.byref(h)),
              "appt_batch_create")
        self.handle = h

    @classmethod
    def from_host(cls, engine: Engine, mats: list, ops: list) -> "Batch":
        b = cls(engine, [m.n for m in mats], [m.nnz for m in mats], ops)
        b.upload(pack(mats))
        return b

    def upload(self, packed: HostMatrix) -> None:
        h = packed._as_c()
        check(lib().appt_batch_upload(self.handle, C.byref(h)), "appt_batch_upload")

    def gather(self, mats: list) -> None:
        arr = (CsrView * max(len(mats), 1))(*[m.view for m in mats])
        check(lib().appt_batch_gather(self.handle, arr), "appt_batch_gather")

    def plan(self, graph: bool = True) -> None:
        check(lib().appt_batch_plan(self.handle, 1 if graph else 0), "appt_batch_plan")

    def autotune(self, reps: int = 2) -> dict:
        t = Tuning()
        check(lib().appt_batch_autotune(self.handle, reps, C.byref(t)), "appt_batch_autotune")
        return _tuning_dict(t)

    def execute(self) -> None:
        check(lib().appt_batch_execute(self.handle), "appt_batch_execute")

    def result(self, k: int) -> DeviceResult:
        r = CsrResult()
        check(lib().appt_batch_result(self.handle, k, C.byref(r)), "appt_batch_result")
        return DeviceResult(self._engine, r)

    def download(self, k: int) -> HostMatrix:
        return self.result(k).download()

    def info(self) -> dict:
        i = BatchInfo()
        check(lib().appt_batch_info_get(self.handle, C.byref(i)), "appt_batch_info_get")
        d = {f: getattr(i, f) for f, _ in BatchInfo._fields_ if f != "tuning"}
        d["tuning"] = _tuning_dict(i.tuning)
        return d

    def close(self) -> None:
        if self.handle:
            check(lib().appt_batch_destroy(self.handle), "appt_batch_destroy")
            self.handle = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()


def _tuning_struct(tuning: dict | None) -> Tuning:
    t = Tuning()
    for k, v in (tuning or {}).items():
        if k in ("f_class_max", "group_for_len", "class_lds_bytes"):
            arr = getattr(t, k)
            for i, x in enumerate(v):
                arr[i] = int(x)
        else:
            setattr(t, k, int(v))
    return t


class Server:
    """Request loop over device-resident jobs (``appt_server_*``): every byte
    the slots' pools hold is charged against the HBM budget; a job is admitted
    for a bound of its working set before it stages its operands and for its
    exact C once the symbolic phase has fixed nnz(C); several jobs in flight
    on their own streams, small waiting jobs run together as one batch."""

    def __init__(self, device: int = 0, slots: int = 4, hbm_budget_bytes: int = 200 * 10 ** 9,
                 batch_max_jobs: int = 64, batch_row_limit: int = 65536, admission_wait_ms: int = 10000,
                 tuning: dict | None = None):
        o = ServerOptions()
        o.device = device
        o.slots = slots
        o.hbm_budget_bytes = hbm_budget_bytes
        o.batch_max_jobs = batch_max_jobs
        o.batch_row_limit = batch_row_limit
        o.admission_wait_ms = admission_wait_ms
        o.tuning = _tuning_struct(tuning)
        h = C.c_void_p()
        check(lib().appt_server_create(C.byref(o), C.byref(h)), "appt_server_create")
        self.handle = h

    def submit(self, A: DeviceMatrix, op: str = "AA") -> int:
        j = C.c_uint64()
        check(lib().appt_server_submit(self.handle, C.byref(A.view), OPS[op], C.byref(j)), "appt_server_submit")
        return int(j.value)

    @staticmethod
    def _info(i: JobInfo) -> dict:
        d = {f: getattr(i, f) for f, _ in JobInfo._fields_ if f != "C"}
        d["state"] = JOB_STATES.get(i.state, str(i.state))
        d["status_name"] = STATUS_NAMES.get(i.status, str(i.status))
        d["nnz"] = int(i.C.nnz)
        return d

    def poll(self, job: int) -> dict:
        i = JobInfo()
        check(lib().appt_server_poll(self.handle, job, C.byref(i)), "appt_server_poll")
        return self._info(i)

    def wait(self, job: int) -> dict:
        i = JobInfo()
        check(lib().appt_server_wait(self.handle, job, C.byref(i)), "appt_server_wait")
        return self._info(i)

    def result(self, job: int) -> DeviceResult:
        """Device view of a finished job's result (valid until ``release``)."""
        i = JobInfo()
        check(lib().appt_server_wait(self.handle, job, C.byref(i)), "appt_server_wait")
        check(i.status, f"job {job}")
        return DeviceResult(None, i.C)

    def download(self, job: int) -> HostMatrix:
        i = JobInfo()
        check(lib().appt_server_wait(self.handle, job, C.byref(i)), "appt_server_wait")
        check(i.status, f"job {job}")
        return DeviceResult(None, i.C).download()

    def release(self, job: int) -> None:
        check(lib().appt_server_release(self.handle, job), "appt_server_release")

    def stats(self) -> dict:
        s = ServerStats()
        check(lib().appt_server_stats_get(self.handle, C.byref(s)), "appt_server_stats_get")
        return {f: getattr(s, f) for f, _ in ServerStats._fields_}

    def close(self) -> None:
        if self.handle:
            check(lib().appt_server_destroy(self.handle), "appt_server_destroy")
            self.handle = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()


def run_ours(task: str, dOff: int, dCol: int, dVal: int, n: int, nnz: int, reps: int) -> EngResult:
    """run_ours(task, dOff, dCol, dVal, n, nnz, reps) of src/spgemm_engine.cuh:132-133:
    device pointers in, EngResult {ms, nnz, s, as, sq} out (mean ms over reps)."""
    out = EngResult()
    check(lib().appt_spgemm_run_timed(task.encode(), dOff, dCol, dVal, n, nnz, reps, C.byref(out)),
          f"run_ours({task})")
    return out
