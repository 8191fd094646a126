"""ORACLE loader — test infrastructure only (tests/, __graft_entry__.smoke(), bench.py's
cpu_baseline leg). ctypes binding of oracle/build/liboracle.so, the serial FP64 CPU restatement
of the reference's per-step projection solve (see cfa_oracle.hpp for what each part follows).
"""
from __future__ import annotations

import ctypes
from pathlib import Path

import numpy as np

ORACLE_DIR = Path(__file__).resolve().parent
LIB = ORACLE_DIR / "build" / "liboracle.so"
RUNNER = ORACLE_DIR / "build" / "oracle_run"
ROW_COLUMNS = ("step", "time", "C_D", "C_L", "delta_p", "nodes", "elements", "p_front", "p_back",
               "C_D_surface", "C_L_surface", "it_uv", "it_p")

_lib = None


def load():
    global _lib
    if _lib is None:
        if not LIB.exists():
            raise RuntimeError(f"{LIB} is missing: make -C {ORACLE_DIR}")
        _lib = ctypes.CDLL(str(LIB))
        d = ctypes.POINTER(ctypes.c_double)
        _lib.oracle_run_fixed.argtypes = [ctypes.c_int, ctypes.c_double, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                          ctypes.c_double, ctypes.c_double, ctypes.c_char_p, ctypes.c_int, d, d, d]
        _lib.oracle_run_fixed.restype = ctypes.c_int
        _lib.oracle_run_params.argtypes = [ctypes.c_int, ctypes.c_double, ctypes.c_int, ctypes.c_double,
                                           ctypes.c_double, ctypes.c_double, d]
        _lib.oracle_run_params.restype = ctypes.c_int
        _lib.oracle_run_outlet_dirichlet.argtypes = [ctypes.c_int, ctypes.c_double, ctypes.c_int, ctypes.c_int,
                                                     ctypes.c_int, ctypes.c_double, d]
        _lib.oracle_run_outlet_dirichlet.restype = ctypes.c_int
        _lib.oracle_run_state.argtypes = [ctypes.c_char_p, ctypes.c_int, ctypes.c_double, ctypes.c_double,
                                          ctypes.c_int, d, d]
        _lib.oracle_run_state.restype = ctypes.c_int
    return _lib


def run_fixed(steps: int, dt: float, uniform_level: int, min_level: int = 2, mms: bool = False,
              mom_tol: float = 1e-7, ppe_tol: float = 1e-7, prefix: str | None = None, dump_systems: int = 0):
    """Fixed-mesh run of the restated time loop. Returns (rows (steps x 13), mms[9], wall seconds);
    it_uv = it_u + 1000 it_v. With a prefix the CSV files are written, and with dump_systems = N
    the momentum and pressure systems of step N (the reference's --dump-systems layout)."""
    lib = load()
    rows = np.zeros((steps, len(ROW_COLUMNS)))
    mms_out = np.zeros(9)
    wall = np.zeros(1)
    d = lambda a: a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))  # noqa: E731
    k = lib.oracle_run_fixed(steps, dt, uniform_level, min_level, 1 if mms else 0, mom_tol, ppe_tol,
                             prefix.encode() if prefix else None, dump_systems, d(rows), d(mms_out), d(wall))
    return rows[:k], mms_out, float(wall[0])


def run_params(steps: int, dt: float, uniform_level: int, nu: float, u_max: float, tol: float = 1e-12):
    """DFG flow on a fixed uniform mesh from the initial condition with viscosity nu and inflow
    peak u_max, solvers at relative tolerance tol. Returns rows (steps x 13)."""
    lib = load()
    rows = np.zeros((steps, len(ROW_COLUMNS)))
    k = lib.oracle_run_params(steps, dt, uniform_level, nu, u_max, tol,
                              rows.ctypes.data_as(ctypes.POINTER(ctypes.c_double)))
    return rows[:k]


def run_outlet_dirichlet(steps: int, dt: float, uniform_level: int, min_level: int = 2, case_2d3: bool = False,
                         tol: float = 1e-7):
    """Fixed-mesh DFG run with the inflow parabola on the outlet and the pressure increment pinned
    at one outlet node (--outlet dirichlet), solvers at relative tolerance tol. Returns rows."""
    lib = load()
    rows = np.zeros((steps, len(ROW_COLUMNS)))
    k = lib.oracle_run_outlet_dirichlet(steps, dt, uniform_level, min_level, 1 if case_2d3 else 0, tol,
                                        rows.ctypes.data_as(ctypes.POINTER(ctypes.c_double)))
    return rows[:k]


def run_state(path: str, steps: int, dt: float, tol: float = 1e-8, threads: int = 0):
    """Fixed-mesh continuation of a --dump-state file for `steps` steps on `threads` OpenMP
    threads (0: the runtime default). Returns (rows (steps x 13), wall seconds of the loop)."""
    lib = load()
    rows = np.zeros((steps, len(ROW_COLUMNS)))
    wall = np.zeros(1)
    d = lambda a: a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))  # noqa: E731
    k = lib.oracle_run_state(str(path).encode(), steps, dt, tol, threads, d(rows), d(wall))
    return rows[:k], float(wall[0])
