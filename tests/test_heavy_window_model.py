"""CPU model of the heavy-row window sequence: every window of the numeric
heavy kernel advances and the windows emit each output once, for every
heavy_window_cols tuning (the shape rules are the product's own header)."""
import shutil
import subprocess
import tempfile
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent
SRC = ROOT / "tests" / "models" / "heavy_window_model.cpp"
INC = ROOT / "appt_spgemm_amd" / "csrc"


@pytest.mark.skipif(shutil.which("g++") is None, reason="needs g++")
def test_heavy_window_sequence_terminates():
    with tempfile.TemporaryDirectory() as d:
        exe = Path(d) / "model"
        subprocess.run(["g++", "-O2", "-std=c++17", f"-I{INC}", str(SRC), "-o", str(exe)], check=True)
        r = subprocess.run([str(exe)], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    assert r.stdout.strip().splitlines()[-1].startswith("OK")
