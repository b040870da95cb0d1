"""Build the gfx950 shared library ``libappt_spgemm.so`` in-tree.

HIP sources are compiled with hipcc for gfx950 only; host-only C++ sources
with g++ against the HIP runtime headers; everything is linked with hipcc.
The object files go to ``appt_spgemm_amd/_build/``.
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import shutil
import subprocess
import sys
from pathlib import Path

PKG = Path(__file__).resolve().parent
CSRC = PKG / "csrc"
OUT = PKG / "_build"
LIB = PKG / "libappt_spgemm.so"
BIN = PKG / "bin"
CLI_TOOLS = ["appt_spgemm", "appt_dispatch", "appt_rocsparse"]
CLI_LIBS = {"appt_rocsparse": ["-lrocsparse"]}  # vendor-library context timing only
ROCM = Path(os.environ.get("ROCM_PATH", "/opt/rocm"))
ARCH = "gfx950"

HIP_FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC", "-Wall",
    "-Wno-unused-function", "-Wno-unused-variable",
    # never let the compiler form packed fp32 ops (v_pk_*_f32)
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    # fp64 atomics to coarse-grained device memory map to hardware atomics
    "-munsafe-fp-atomics",
    # value parity with the CPU Gustavson order: no contraction of a*b + c
    "-ffp-contract=off",
]
CXX_FLAGS = ["-O2", "-std=c++17", "-fPIC", "-Wall", "-D__HIP_PLATFORM_AMD__",
             f"-I{ROCM / 'include'}"]

HIP_SOURCES = ["kernels.hip", "templates.hip", "descriptors.hip", "batch.hip", "noalloc.hip", "gram.hip"]
CXX_SOURCES = ["engine.cpp", "batch.cpp", "server.cpp", "capi.cpp", "gen.cpp"]


def _hipcc() -> str:
    p = ROCM / "bin" / "hipcc"
    return str(p) if p.exists() else (shutil.which("hipcc") or "hipcc")


def _run(cmd: list[str]) -> None:
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"build step failed ({r.returncode}): {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    # hipcc's host pass warns about the packed-fp32 feature name; that is expected
    err = "\n".join(l for l in r.stderr.splitlines() if "packed-fp32-ops" not in l and l.strip())
    if err and os.environ.get("APPT_BUILD_VERBOSE"):
        print(err, file=sys.stderr)


def build(verbose: bool = False) -> Path:
    """Compile every source and link ``libappt_spgemm.so``; returns its path."""
    OUT.mkdir(exist_ok=True)
    hipcc = _hipcc()
    jobs = []
    for src in HIP_SOURCES:
        obj = OUT / (Path(src).stem + ".hip.o")
        jobs.append([hipcc, *HIP_FLAGS, "-c", str(CSRC / src), "-o", str(obj)])
    for src in CXX_SOURCES:
        obj = OUT / (Path(src).stem + ".o")
        jobs.append(["g++", *CXX_FLAGS, "-c", str(CSRC / src), "-o", str(obj)])
    workers = max(1, min(len(jobs), os.cpu_count() or 1, 8))
    with cf.ThreadPoolExecutor(workers) as ex:
        for f in [ex.submit(_run, j) for j in jobs]:
            f.result()
    objs = [str(OUT / (Path(s).stem + ".hip.o")) for s in HIP_SOURCES] + \
           [str(OUT / (Path(s).stem + ".o")) for s in CXX_SOURCES]
    tmp = LIB.with_suffix(".so.tmp")
    _run([hipcc, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", str(tmp), *objs,
          f"-L{ROCM / 'lib'}", "-lamdhip64", f"-Wl,-rpath,{ROCM / 'lib'}"])
    os.replace(tmp, LIB)
    BIN.mkdir(exist_ok=True)
    cli_jobs = []
    for tool in CLI_TOOLS:
        exe = BIN / tool
        cli_jobs.append(["g++", *CXX_FLAGS, str(CSRC / "cli" / f"{tool}.cpp"), "-o", str(exe) + ".tmp",
                         f"-L{PKG}", "-lappt_spgemm", f"-Wl,-rpath,$ORIGIN/..", f"-L{ROCM / 'lib'}",
                         *CLI_LIBS.get(tool, []), "-lamdhip64", f"-Wl,-rpath,{ROCM / 'lib'}"])
    with cf.ThreadPoolExecutor(len(cli_jobs)) as ex:
        for f in [ex.submit(_run, j) for j in cli_jobs]:
            f.result()
    for tool in CLI_TOOLS:
        os.replace(str(BIN / tool) + ".tmp", BIN / tool)
    if verbose:
        print(f"built {LIB}")
    return LIB


if __name__ == "__main__":
    build(verbose=True)
