"""The C-ABI libraries load without a GPU and export every function their headers declare."""
import ctypes

import pytest

import cfa


@pytest.mark.parametrize("header,lib", [("cfa_hip.h", cfa.HIP_LIB), ("cfa_host.h", cfa.HOST_LIB)])
def test_header_symbols_exported(header, lib):
    cfa.load()
    names = cfa.header_functions(cfa.INCLUDE_DIR / header)
    assert len(names) >= (40 if header == "cfa_hip.h" else 10)
    so = ctypes.CDLL(str(lib), mode=ctypes.RTLD_GLOBAL)
    missing = [n for n in names if not hasattr(so, n)]
    assert not missing, f"{lib.name} lacks {missing}"


def test_error_strings():
    hip, _ = cfa.load()
    assert b"argument" in hip.cfa_error_string(-1).lower()
    assert hip.cfa_error_string(0)


def test_cli_rejects_bad_options():
    """Validation failures exit with EXIT_FAILURE before any device call
    (src/cfa/options.cuh:190-289)."""
    import subprocess
    r = subprocess.run([str(cfa.CLI), "--dt", "-1"], capture_output=True, text=True, timeout=60)
    assert r.returncode == 1
    r = subprocess.run([str(cfa.CLI), "--bogus"], capture_output=True, text=True, timeout=60)
    assert r.returncode == 1 and "Unknown option" in r.stderr
    r = subprocess.run([str(cfa.CLI), "--help"], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0 and "--amr-interval" in r.stdout
