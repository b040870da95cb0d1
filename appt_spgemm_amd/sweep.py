"""Suite sweeps over the command-line tools, in the reference's methodology.

  compare   engine vs the vendor library (rocSPARSE) on A*A and A*A^T: C's nnz
            must be equal and the checksums (sum v, sum |v|, sum v^2) agree
            within 1e-6 relative; speedup = vendor ms / engine ms
            (bench/run_compare.py:45-91)
  symabl    A*A^T vs the symmetric upper-triangle form AAtS: same nnz and
            checksums, speedup of AAtS over AAt (bench/run_symabl.py:23-59)
  dispatch  the dispatcher tool over the compare rows: realised time of the
            reference's rule (descriptor overhead + the chosen path) against
            engine-always, vendor-always and the per-task best
            (bench/run_dispatch.py:9-34)

Matrices: a synthetic suite (the seeded generators; the reference's
SuiteSparse groups are not shipped), or every ``*.csr`` (CSR1) / ``*.npz``
(scipy CSR) file of a directory.  Each matrix is written once as CSR1 into
the work directory, the tools run as subprocesses with a 600 s limit, and the
rows go to CSV files there.

  python -m appt_spgemm_amd.sweep compare|symabl|dispatch|all [--suite synthetic|DIR]
         [--out DIR] [--limit N] [--cap PRODUCTS]
"""
from __future__ import annotations

import argparse
import csv
import math
import re
import subprocess
import sys
from pathlib import Path

import numpy as np

from . import engine as E

BIN = Path(__file__).resolve().parent / "bin"
TOOL = {"ours": BIN / "appt_spgemm", "vendor": BIN / "appt_rocsparse", "dispatch": BIN / "appt_dispatch"}
CAP = 300_000_000          # products bound above which a matrix is skipped (bench/run_compare.py:17)
TIMEOUT = 600              # seconds per tool run (bench/run_compare.py:71-75)
RTOL = 1e-6                # checksum agreement (bench/run_compare.py:42-43)


def synthetic_suite() -> list:
    """(tag, generator) pairs over the generator families at sizes from the
    sub-millisecond regime to ~10^8 products."""
    s = []
    for g in (64, 256, 1024):
        s.append((f"laplace2d_{g}", lambda g=g: E.laplace2d(g)))
    for g, d in ((32, 0.0), (48, 0.1), (64, 0.0), (96, 0.1)):
        s.append((f"stencil27_{g}_{d}", lambda g=g, d=d: E.stencil27(g, d, 1)))
    for sc in (12, 14, 16):
        s.append((f"rmat_{sc}", lambda sc=sc: E.rmat(sc, 16, seed=1)))
    for n, bw in ((100_000, 4), (100_000, 32), (1_000_000, 8)):
        s.append((f"banded_{n}_{bw}", lambda n=n, bw=bw: E.banded(n, bw, 0.7, 2)))
    for n, deg in ((100_000, 4.0), (200_000, 16.0), (1_000_000, 8.0)):
        s.append((f"er_{n}_{int(deg)}", lambda n=n, deg=deg: E.erdos_renyi(n, deg, 3)))
    return s


def dir_suite(d: Path) -> list:
    out = []
    for p in sorted(d.iterdir()):
        if p.suffix == ".csr":
            out.append((p.stem, lambda p=p: E.csr1_read(str(p))))
        elif p.suffix == ".npz":
            out.append((p.stem, lambda p=p: _from_scipy(p)))
    return out


def _from_scipy(p: Path) -> E.HostMatrix:
    import scipy.sparse as sp
    A = sp.load_npz(p).astype(np.float64).tocsr()
    if A.shape[0] != A.shape[1]:
        raise ValueError(f"{p}: not square")
    A.sort_indices()
    return E.HostMatrix(A.shape[0], A.indptr.astype(np.int32), A.indices.astype(np.int32), A.data)


def products(m: E.HostMatrix) -> tuple[int, int]:
    """Products of A*A and of A*A^T: sum over entries a_ik of nnz(row k), and
    sum over columns of count^2 (bench/run_compare.py:23-27)."""
    rp = np.asarray(m.rowptr, np.int64)
    col = np.asarray(m.col, np.int64)
    deg = np.diff(rp)
    cc = np.bincount(col, minlength=m.n).astype(np.int64)
    return int(deg[col].sum()), int((cc * cc).sum())


def reps_for(nnz: int) -> int:
    """20 / 8 / 3 timed repetitions by nnz(A) (bench/run_compare.py:69)."""
    return 20 if nnz < 200_000 else (8 if nnz < 2_000_000 else 3)


def parse_results(out: str) -> dict:
    """{task: (C_nnz, ms, (s, as, sq))}, pairing the i-th checksum line with the
    i-th RESULT line (bench/run_compare.py:30-40)."""
    chks = re.findall(r"__chk s=(\S+) as=(\S+) sq=(\S+)", out)
    res, ci = {}, 0
    for ln in out.splitlines():
        if ln.startswith("RESULT,"):
            f = ln.split(",")
            chk = tuple(float(x) for x in chks[ci]) if ci < len(chks) else (0.0, 0.0, 0.0)
            res[f[2]] = (int(f[5]), float(f[6]), chk)
            ci += 1
    return res


_DISPATCH = re.compile(r"DISPATCH,[^,]+,[^,]+,(\w+),dcv=([\d.]+),fill=([\d.]+),C_nnz=(\d+),desc_ms=([\d.]+),"
                       r"run_ms=([\d.]+),total_ms=([\d.]+),rule=(\w+)")


def parse_dispatch(out: str) -> dict | None:
    m = _DISPATCH.search(out)
    if not m:
        return None
    return {"choice": m.group(1), "dcv": float(m.group(2)), "fill": float(m.group(3)), "C_nnz": int(m.group(4)),
            "desc_ms": float(m.group(5)), "run_ms": float(m.group(6)), "total_ms": float(m.group(7)),
            "rule": m.group(8)}


def rel_ok(a: float, b: float, rt: float = RTOL) -> bool:
    return abs(a - b) <= rt * max(1.0, abs(b))


def match(a: tuple, b: tuple) -> bool:
    """Same nnz(C) and all three checksums within RTOL."""
    return a[0] == b[0] and all(rel_ok(x, y) for x, y in zip(a[2], b[2]))


def _run(tool: str, tag: str, path: Path, task: str, reps: int) -> str:
    r = subprocess.run([str(TOOL[tool]), tag, str(path), task, str(reps)], capture_output=True, text=True,
                       timeout=TIMEOUT)
    if r.returncode != 0:
        print(f"  {tool} {tag} {task} exited {r.returncode}: {r.stderr.strip()[:200]}", flush=True)
    return r.stdout


def _prepare(suite: list, work: Path, limit: int | None):
    work.mkdir(parents=True, exist_ok=True)
    for tag, make in suite[:limit]:
        path = work / f"{tag}.csr"
        m = make() if not path.exists() else E.csr1_read(str(path))
        if not path.exists():
            E.csr1_write(str(path), m)
        yield tag, m, path


def compare(suite: list, work: Path, limit: int | None = None, cap: int = CAP) -> Path:
    work.mkdir(parents=True, exist_ok=True)
    out = work / "compare.csv"
    n_ok = n_bad = n_skip = 0
    with open(out, "w", newline="") as fh:
        w = csv.writer(fh)
        w.writerow(["tag", "n", "nnz", "task", "C_nnz_ours", "C_nnz_vendor", "ms_ours", "ms_vendor", "speedup",
                    "chk_match", "status"])
        for tag, m, path in _prepare(suite, work, limit):
            nnz = int(m.rowptr[-1])
            if max(products(m)) > cap:
                for t in ("AA", "AAt"):
                    w.writerow([tag, m.n, nnz, t, "", "", "", "", "", "", "skip_oom"])
                n_skip += 1
                print(f"SKIP {tag}", flush=True)
                continue
            reps = reps_for(nnz)
            for t in ("AA", "AAt"):
                try:
                    ours = parse_results(_run("ours", tag, path, t, reps)).get(t)
                    vend = parse_results(_run("vendor", tag, path, t, reps)).get(t)
                except subprocess.TimeoutExpired:
                    w.writerow([tag, m.n, nnz, t, "", "", "", "", "", "", "timeout"])
                    print(f"TIMEOUT {tag} {t}", flush=True)
                    continue
                if not ours or not vend:
                    w.writerow([tag, m.n, nnz, t, "", "", "", "", "", "", "fail"])
                    print(f"  FAIL {tag} {t}", flush=True)
                    continue
                ok = match(ours, vend)
                spd = vend[1] / ours[1] if ours[1] > 0 else 0.0
                w.writerow([tag, m.n, nnz, t, ours[0], vend[0], f"{ours[1]:.4f}", f"{vend[1]:.4f}", f"{spd:.2f}",
                            int(ok), "ok"])
                n_ok += 1
                if not ok:
                    n_bad += 1
                    print(f"  MISMATCH {tag} {t}: ours {ours} vs vendor {vend}", flush=True)
                print(f"OK {tag:24s} {t:4s} ours {ours[1]:9.4f} ms  vendor {vend[1]:9.4f} ms  {spd:6.2f}x",
                      flush=True)
    print(f"== compare: ok_tasks={n_ok} mismatches={n_bad} skipped={n_skip} -> {out}", flush=True)
    return out


def symabl(suite: list, work: Path, limit: int | None = None, cap: int = CAP) -> Path:
    work.mkdir(parents=True, exist_ok=True)
    out = work / "symabl.csv"
    spds, nbad = [], 0
    with open(out, "w", newline="") as fh:
        w = csv.writer(fh)
        w.writerow(["tag", "n", "nnz", "C_nnz", "ms_aat", "ms_aats", "speedup", "match"])
        for tag, m, path in _prepare(suite, work, limit):
            nnz = int(m.rowptr[-1])
            if products(m)[1] > cap:
                continue
            reps = reps_for(nnz)
            try:
                a = parse_results(_run("ours", tag, path, "AAt", reps)).get("AAt")
                b = parse_results(_run("ours", tag, path, "AAtS", reps)).get("AAtS")
            except subprocess.TimeoutExpired:
                continue
            if not a or not b:
                continue
            ok = match(b, a)
            spd = a[1] / b[1] if b[1] > 0 else 0.0
            w.writerow([tag, m.n, nnz, a[0], f"{a[1]:.4f}", f"{b[1]:.4f}", f"{spd:.3f}", int(ok)])
            if ok:
                spds.append((nnz, spd))
            else:
                nbad += 1
                print(f"MISMATCH {tag}: AAt {a} vs AAtS {b}", flush=True)
            print(f"{tag:24s} AAt {a[1]:9.4f} ms  AAtS {b[1]:9.4f} ms  {spd:6.3f}x", flush=True)
    if spds:
        s = sorted(x for _, x in spds)
        gm = math.exp(sum(math.log(x) for x in s) / len(s))
        print(f"== AAtS ablation: {len(s)} matrices, mismatches={nbad}; AAtS/AAt geomean {gm:.3f}x "
              f"median {s[len(s) // 2]:.3f}x wins {sum(x >= 1.0 for x in s)}/{len(s)} "
              f"range [{s[0]:.2f}, {s[-1]:.2f}] -> {out}", flush=True)
    return out


def dispatch(suite: list, work: Path, limit: int | None = None, cap: int = CAP) -> Path:
    """Realised time of the reference's dispatcher rule: its verdict picks the
    engine or the vendor path, charged its descriptor time plus the chosen
    path's time from the compare rows."""
    cmp_csv = work / "compare.csv"
    if not cmp_csv.exists():
        compare(suite, work, limit, cap)
    rows = [r for r in csv.DictReader(open(cmp_csv)) if r["status"] == "ok"]
    out = work / "dispatch.csv"
    tot = desc_tot = t_vend = t_ours = t_best = 0.0
    n_ours = 0
    with open(out, "w", newline="") as fh:
        w = csv.writer(fh)
        w.writerow(["tag", "task", "rule", "desc_ms", "run_ms", "total_ms", "ms_ours", "ms_vendor"])
        for r in rows:
            tag, task = r["tag"], r["task"]
            d = parse_dispatch(_run("dispatch", tag, work / f"{tag}.csr", task, reps_for(int(r["nnz"]))))
            if d is None:
                print(f"FAIL {tag} {task}", flush=True)
                continue
            mo, mv = float(r["ms_ours"]), float(r["ms_vendor"])
            run = mo if d["rule"] == "ours" else mv
            total = d["desc_ms"] + run
            n_ours += d["rule"] == "ours"
            tot += total
            desc_tot += d["desc_ms"]
            t_vend += mv
            t_ours += mo
            t_best += min(mo, mv)
            w.writerow([tag, task, d["rule"], f"{d['desc_ms']:.4f}", f"{run:.4f}", f"{total:.4f}", f"{mo:.4f}",
                        f"{mv:.4f}"])
    if rows and tot > 0:
        print(f"tasks={len(rows)}  rule picked engine {n_ours}, vendor {len(rows) - n_ours}")
        print(f"vendor-always : {t_vend:9.2f} ms (1.00x)")
        print(f"engine-always : {t_ours:9.2f} ms ({t_vend / t_ours:.2f}x)")
        print(f"rule dispatch : {tot:9.2f} ms ({t_vend / tot:.2f}x), descriptors {desc_tot:.2f} ms "
              f"= {100 * desc_tot / tot:.1f}%")
        print(f"best-of-both  : {t_best:9.2f} ms ({t_vend / t_best:.2f}x) -> {out}", flush=True)
    return out


def main(argv=None) -> int:
    p = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    p.add_argument("mode", choices=["compare", "symabl", "dispatch", "all"])
    p.add_argument("--suite", default="synthetic", help="'synthetic' or a directory of .csr/.npz files")
    p.add_argument("--out", default="sweep_out", help="work directory (CSR1 files and CSVs)")
    p.add_argument("--limit", type=int, default=None)
    p.add_argument("--cap", type=float, default=CAP, help="skip matrices with more products than this")
    a = p.parse_args(argv)
    for t in TOOL.values():
        if not t.exists():
            print(f"missing tool {t}: build the package first (python -m appt_spgemm_amd.build)", file=sys.stderr)
            return 1
    suite = synthetic_suite() if a.suite == "synthetic" else dir_suite(Path(a.suite))
    work = Path(a.out)
    cap = int(a.cap)
    work.mkdir(parents=True, exist_ok=True)
    if a.mode in ("compare", "all"):
        compare(suite, work, a.limit, cap)
    if a.mode in ("symabl", "all"):
        symabl(suite, work, a.limit, cap)
    if a.mode in ("dispatch", "all"):
        dispatch(suite, work, a.limit, cap)
    return 0


if __name__ == "__main__":
    sys.exit(main())
