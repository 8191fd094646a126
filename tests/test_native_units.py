"""The C++ unit tests (cylinder-fem-amr_amd/csrc/app/test_cfa.cpp, the checks of the reference's
tests/test_cfa.cu) run as one program: host-side mesh adaptation checks here, device checks on
the GPU."""
import subprocess

import pytest

import cfa


def run_units(part):
    r = subprocess.run([str(cfa.TEST_BIN), part], capture_output=True, text=True, timeout=600)
    print(r.stdout)
    print(r.stderr)
    assert r.returncode == 0 and "ALL PASSED" in r.stdout, r.stderr[-2000:]


def test_units_host():
    run_units("cpu")


@pytest.mark.gpu
def test_units_device():
    run_units("gpu")
