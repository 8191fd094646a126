"""The mean removal without allocation (cfa_pressure_remove_mean_noalloc, a fixed device array
instead of a workspace) through the C ABI against cfa_pressure_remove_mean: two calls in a row on
one GPU, the second on a smaller mesh, bit for bit, and a mesh beyond the array refused with
CFA_E_SIZE (tests/native/parity.cpp case noalloc)."""
import subprocess
from pathlib import Path

import pytest

BIN = Path(__file__).resolve().parent / "native" / "build" / "parity"

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("level", [1, 3])
def test_no_alloc_mean_removal(level):
    print(f"[case] parity noalloc level {level}", flush=True)
    r = subprocess.run([str(BIN), "noalloc", str(level)], capture_output=True, text=True, timeout=300)
    print(r.stdout)
    assert r.returncode == 0 and "ALL PASSED" in r.stdout, r.stdout[-3000:] + r.stderr[-2000:]
