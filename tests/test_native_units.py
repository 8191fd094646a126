"""The C++ unit tests (cylinder-fem-amr_amd/csrc/app/test_cfa.cpp, the checks of the reference's
tests/test_cfa.cu) run as one program on the GPU: the device NVB forest's adaptation checks
(conformity, level cap, angle bound, exact transfer of linear fields, refine/coarsen round trip,
minimum level, retransfer) and the solver checks. Its host part runs without a device: the
sliced-ELL layout of the ILU(0) pattern's split rows."""
import subprocess

import pytest

import cfa


def run_units(part):
    r = subprocess.run([str(cfa.TEST_BIN), part], capture_output=True, text=True, timeout=600)
    print(r.stdout)
    print(r.stderr)
    assert r.returncode == 0 and "ALL PASSED" in r.stdout, r.stderr[-2000:]


@pytest.mark.gpu
def test_units_device():
    print("[case] C++ unit tests", flush=True)
    run_units("gpu")


def test_units_host():
    print("[case] C++ unit tests, host part", flush=True)
    run_units("host")
