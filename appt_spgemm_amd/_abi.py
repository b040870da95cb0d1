"""ctypes declarations of ``include/appt_spgemm.h`` and the library loader.

The loader fails loudly: there is no fallback path.  If ``libappt_spgemm.so``
is missing, build it with ``python -m appt_spgemm_amd.build`` (or
``__graft_entry__.build()``).
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path

PKG = Path(__file__).resolve().parent
LIB_PATH = PKG / "libappt_spgemm.so"

APPT_OK = 0
STATUS_NAMES = {
    0: "APPT_OK", 1: "APPT_ERR_INVALID", 2: "APPT_ERR_HIP", 3: "APPT_ERR_OOM",
    4: "APPT_ERR_ADMISSION", 5: "APPT_ERR_IO", 6: "APPT_ERR_UNSUPPORTED",
    7: "APPT_ERR_OVERFLOW", 8: "APPT_ERR_NO_DEVICE",
}
OPS = {"AA": 0, "AAt": 1, "AAtS": 2}
MAX_CLASSES = 8


class CsrView(C.Structure):
    _fields_ = [("n", C.c_int32), ("nnz", C.c_int64), ("rowptr", C.c_void_p),
                ("col", C.c_void_p), ("val", C.c_void_p)]


class CsrResult(C.Structure):
    _fields_ = [("n", C.c_int32), ("nnz", C.c_int64), ("rowptr_is_64", C.c_int32),
                ("rowptr", C.c_void_p), ("col", C.c_void_p), ("val", C.c_void_p),
                ("_priv", C.c_void_p)]


class EngResult(C.Structure):
    """struct EngResult { float ms; int64_t nnz; double s, as, sq; }
    (src/spgemm_engine.cuh:129)."""
    _fields_ = [("ms", C.c_float), ("nnz", C.c_int64), ("s", C.c_double),
                ("as_", C.c_double), ("sq", C.c_double)]


class Tuning(C.Structure):
    _fields_ = [("f_class_max", C.c_int32 * MAX_CLASSES), ("wg_lds_bytes", C.c_int32),
                ("lds_budget_bytes", C.c_int32), ("bitmap_words_factor", C.c_int32),
                ("group_for_len", C.c_int32 * 4), ("heavy_window_cols", C.c_int32),
                ("sort_span_factor", C.c_int32), ("sort_span_factor_symbolic", C.c_int32),
                ("transpose_radix_min_nnz", C.c_int32), ("mirror_radix_min_nnz", C.c_int32),
                ("class_lds_bytes", C.c_int32 * MAX_CLASSES), ("sort_max_products", C.c_int32),
                ("rowptr64_min_nnz", C.c_int32), ("row_templates", C.c_int32), ("gram", C.c_int32)]


class Options(C.Structure):
    _fields_ = [("device", C.c_int32), ("stream", C.c_void_p),
                ("hbm_budget_bytes", C.c_int64), ("tuning", Tuning)]


class PhaseTimes(C.Structure):
    _fields_ = [("transpose_ms", C.c_float), ("prepass_ms", C.c_float),
                ("symbolic_ms", C.c_float), ("scan_ms", C.c_float),
                ("numeric_ms", C.c_float), ("mirror_ms", C.c_float),
                ("total_ms", C.c_float), ("products", C.c_int64), ("nnz_c", C.c_int64),
                ("n_bins_used", C.c_int32), ("sync_points", C.c_int32)]


MAX_BINS = 64


class BinReport(C.Structure):
    _fields_ = [("n_bins", C.c_int32), ("sym_rows", C.c_int32 * MAX_BINS), ("num_rows", C.c_int32 * MAX_BINS),
                ("sym_metric_max", C.c_int32 * MAX_BINS), ("num_metric_max", C.c_int32 * MAX_BINS)]


class Descriptors(C.Structure):
    _fields_ = [("n", C.c_int32), ("nnz", C.c_int64), ("deg_mean", C.c_double),
                ("deg_cv", C.c_double), ("deg_max", C.c_int32), ("fill_aa", C.c_double),
                ("fill_aat", C.c_double), ("psym", C.c_double), ("products_aa", C.c_int64),
                ("products_aat", C.c_int64), ("max_row_products_aa", C.c_int32),
                ("bandwidth", C.c_int32), ("pick_ours", C.c_int32)]


class HostCsr(C.Structure):
    _fields_ = [("n", C.c_int32), ("nnz", C.c_int64), ("rowptr", C.POINTER(C.c_int32)),
                ("col", C.POINTER(C.c_int32)), ("val", C.POINTER(C.c_double))]


class BatchBuffers(C.Structure):
    _fields_ = [("n_total", C.c_int32), ("nnz", C.c_int64), ("rowptr", C.c_void_p),
                ("col", C.c_void_p), ("val", C.c_void_p)]


class BatchInfo(C.Structure):
    _fields_ = [("count", C.c_int32), ("n_total", C.c_int32), ("nnz_a", C.c_int64),
                ("nnz_c", C.c_int64), ("products", C.c_int64), ("graph", C.c_int32),
                ("bins_used", C.c_int32), ("tuning", Tuning)]


class ServerOptions(C.Structure):
    _fields_ = [("device", C.c_int32), ("slots", C.c_int32), ("hbm_budget_bytes", C.c_int64),
                ("batch_max_jobs", C.c_int32), ("batch_row_limit", C.c_int32),
                ("admission_wait_ms", C.c_int32), ("tuning", Tuning)]


class JobInfo(C.Structure):
    _fields_ = [("state", C.c_int32), ("status", C.c_int32), ("group", C.c_int32),
                ("queue_ms", C.c_double), ("run_ms", C.c_double), ("latency_ms", C.c_double),
                ("products", C.c_int64), ("reserved_bytes", C.c_int64), ("C", CsrResult)]


class ServerStats(C.Structure):
    _fields_ = [("submitted", C.c_int64), ("completed", C.c_int64), ("refused", C.c_int64),
                ("failed", C.c_int64), ("p50_ms", C.c_double), ("p99_ms", C.c_double),
                ("mean_ms", C.c_double), ("gflops", C.c_double), ("jobs_per_s", C.c_double),
                ("bytes_reserved", C.c_int64), ("peak_bytes_reserved", C.c_int64),
                ("groups_run", C.c_int64), ("device_bytes", C.c_int64), ("peak_device_bytes", C.c_int64),
                ("overruns", C.c_int64)]


JOB_STATES = {0: "queued", 1: "running", 2: "done", 3: "released"}

P = C.POINTER
_SIGNATURES = {
    "appt_last_error": (C.c_char_p, []),
    "appt_op_from_task": (C.c_int, [C.c_char_p, P(C.c_int)]),
    "appt_engine_create": (C.c_int, [P(Options), P(C.c_void_p)]),
    "appt_engine_destroy": (C.c_int, [C.c_void_p]),
    "appt_engine_set_timing": (C.c_int, [C.c_void_p, C.c_int]),
    "appt_engine_phase_times": (C.c_int, [C.c_void_p, P(PhaseTimes)]),
    "appt_engine_bin_report": (C.c_int, [C.c_void_p, P(BinReport)]),
    "appt_bin_name": (C.c_char_p, [C.c_int32]),
    "appt_engine_trim": (C.c_int, [C.c_void_p]),
    "appt_engine_device_bytes": (C.c_int64, [C.c_void_p]),
    "appt_spgemm_compute": (C.c_int, [C.c_void_p, P(CsrView), C.c_int, P(CsrResult)]),
    "appt_spgemm_compute_noalloc": (C.c_int, [C.c_void_p, P(CsrView), C.c_int, C.c_void_p, C.c_void_p,
                                              C.c_void_p, C.c_int64, P(C.c_int64)]),
    "appt_csr_release": (C.c_int, [C.c_void_p, P(CsrResult)]),
    "appt_csr_eliminate_zeros": (C.c_int, [C.c_void_p, P(CsrResult)]),
    "appt_spgemm_compute_rows": (C.c_int, [C.c_void_p, P(CsrView), C.c_int, C.c_int32, C.c_int32,
                                           P(CsrResult)]),
    "appt_spgemm_symbolic_rowptr": (C.c_int, [C.c_void_p, P(CsrView), C.c_int, P(C.c_int64)]),
    "appt_spgemm_run_timed": (C.c_int, [C.c_char_p, C.c_void_p, C.c_void_p, C.c_void_p,
                                        C.c_int32, C.c_int64, C.c_int32, P(EngResult)]),
    "appt_engine_run_timed": (C.c_int, [C.c_void_p, C.c_char_p, P(CsrView), C.c_int32, P(EngResult)]),
    "appt_dispatch_descriptors": (C.c_int, [C.c_void_p, P(CsrView), C.c_int, P(Descriptors)]),
    "appt_host_csr_free": (None, [P(HostCsr)]),
    "appt_csr1_read": (C.c_int, [C.c_char_p, P(HostCsr)]),
    "appt_csr1_write": (C.c_int, [C.c_char_p, P(HostCsr)]),
    "appt_gen_laplace2d": (C.c_int, [C.c_int32, P(HostCsr)]),
    "appt_gen_stencil27": (C.c_int, [C.c_int32, C.c_double, C.c_uint64, P(HostCsr)]),
    "appt_gen_rmat": (C.c_int, [C.c_int32, C.c_int32, C.c_double, C.c_double, C.c_double,
                                C.c_uint64, P(HostCsr)]),
    "appt_gen_banded": (C.c_int, [C.c_int32, C.c_int32, C.c_double, C.c_uint64, P(HostCsr)]),
    "appt_gen_erdos_renyi": (C.c_int, [C.c_int32, C.c_double, C.c_uint64, P(HostCsr)]),
    "appt_batch_create": (C.c_int, [C.c_void_p, C.c_int32, P(C.c_int32), P(C.c_int64), P(C.c_int),
                                    P(C.c_void_p)]),
    "appt_batch_destroy": (C.c_int, [C.c_void_p]),
    "appt_batch_inputs": (C.c_int, [C.c_void_p, P(BatchBuffers)]),
    "appt_batch_upload": (C.c_int, [C.c_void_p, P(HostCsr)]),
    "appt_batch_gather": (C.c_int, [C.c_void_p, P(CsrView)]),
    "appt_batch_plan": (C.c_int, [C.c_void_p, C.c_int32]),
    "appt_batch_autotune": (C.c_int, [C.c_void_p, C.c_int32, P(Tuning)]),
    "appt_batch_execute": (C.c_int, [C.c_void_p]),
    "appt_batch_result": (C.c_int, [C.c_void_p, C.c_int32, P(CsrResult)]),
    "appt_batch_info_get": (C.c_int, [C.c_void_p, P(BatchInfo)]),
    "appt_server_create": (C.c_int, [P(ServerOptions), P(C.c_void_p)]),
    "appt_server_destroy": (C.c_int, [C.c_void_p]),
    "appt_server_submit": (C.c_int, [C.c_void_p, P(CsrView), C.c_int, P(C.c_uint64)]),
    "appt_server_poll": (C.c_int, [C.c_void_p, C.c_uint64, P(JobInfo)]),
    "appt_server_wait": (C.c_int, [C.c_void_p, C.c_uint64, P(JobInfo)]),
    "appt_server_release": (C.c_int, [C.c_void_p, C.c_uint64]),
    "appt_server_stats_get": (C.c_int, [C.c_void_p, P(ServerStats)]),
    "appt_device_upload": (C.c_int, [C.c_void_p, P(HostCsr), P(CsrView)]),
    "appt_csr_check": (C.c_int, [C.c_void_p, P(CsrView)]),
    "appt_device_free_view": (C.c_int, [C.c_void_p, P(CsrView)]),
    "appt_csr_download_rows": (C.c_int, [C.c_void_p, P(CsrResult), C.c_int32, C.c_int32, P(C.c_int64),
                                          P(C.c_int32), P(C.c_double)]),
    "appt_csr_digest": (C.c_int, [C.c_void_p, P(CsrResult), P(C.c_uint64)]),
    "appt_csr_download": (C.c_int, [C.c_void_p, P(CsrResult), P(C.c_int64), P(C.c_int32),
                                     P(C.c_double)]),
}

_lib = None


class ApptError(RuntimeError):
    def __init__(self, status: int, where: str, detail: str = ""):
        self.status = status
        self.status_name = STATUS_NAMES.get(status, str(status))
        super().__init__(f"{where}: {self.status_name}" + (f" ({detail})" if detail else ""))


def lib() -> C.CDLL:
    """Load ``libappt_spgemm.so`` (once) and declare every exported signature."""
    global _lib
    if _lib is not None:
        return _lib
    if not LIB_PATH.exists():
        raise ImportError(f"{LIB_PATH} is not built; run `python -m appt_spgemm_amd.build` "
                          "(the HIP library is the only implementation: there is no fallback)")
    L = C.CDLL(str(LIB_PATH))
    for name, (res, args) in _SIGNATURES.items():
        fn = getattr(L, name)
        fn.restype = res
        fn.argtypes = args
    _lib = L
    return L


def check(status: int, where: str) -> None:
    if status != APPT_OK:
        detail = lib().appt_last_error()
        raise ApptError(status, where, detail.decode() if detail else "")
