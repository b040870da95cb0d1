"""MI355X (gfx950) SpGEMM engine: C = A*A, A*A^T (plain and symmetric-Gram)
for int32/fp64 CSR, behind the C ABI in include/appt_spgemm.h."""
from ._abi import ApptError, lib  # noqa: F401
from .engine import (Batch, DeviceMatrix, DeviceResult, Engine, HostMatrix, Server, banded, bin_name,  # noqa: F401
                     csr1_read, csr1_write, erdos_renyi, laplace2d, pack, panels, rmat, run_ours, stencil27)


def __getattr__(name):
    # NUM_BINS comes from the library, loaded on first use (never at import,
    # so the build can import the package before the library exists)
    if name == "NUM_BINS":
        from .engine import _num_bins
        return _num_bins()
    raise AttributeError(name)

__all__ = ["ApptError", "NUM_BINS", "Batch", "DeviceMatrix", "DeviceResult", "Engine", "HostMatrix", "Server", "banded", "bin_name",
           "csr1_read", "csr1_write", "erdos_renyi", "laplace2d", "lib", "pack", "panels", "rmat", "run_ours",
           "stencil27"]
