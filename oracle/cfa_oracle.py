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
        ip = ctypes.POINTER(ctypes.c_int)
        for name in ("oracle_pcg_amg", "oracle_pcg_jacobi", "oracle_bicgstab_jacobi"):
            fn = getattr(_lib, name)
            fn.argtypes = [ctypes.c_int, ip, ip, d, d, d, ctypes.c_double, ctypes.c_int, d]
            fn.restype = ctypes.c_int
        _lib.oracle_bicgstab2_ilu.argtypes = [ctypes.c_int, ip, ip, d, d, d, d, d, ctypes.c_double, ctypes.c_int, d]
        _lib.oracle_bicgstab2_ilu.restype = ctypes.c_int
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


def _d(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def _i(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_int))


def _csr(rp, ci, val):
    return (np.ascontiguousarray(rp, np.int32), np.ascontiguousarray(ci, np.int32),
            np.ascontiguousarray(val, np.float64))


def _single_solve(name: str, rp, ci, val, b, x0, tol: float, max_iter: int):
    rp, ci, val = _csr(rp, ci, val)
    b = np.ascontiguousarray(b, np.float64)
    x = np.array(x0, np.float64, copy=True)
    st = np.zeros(4)
    getattr(load(), name)(len(b), _i(rp), _i(ci), _d(val), _d(b), _d(x), tol, max_iter, _d(st))
    return x, {"iterations": int(st[0]), "converged": bool(st[1]), "rel_residual": float(st[2]),
               "amg_levels": int(st[3])}


def pcg_amg(rp, ci, val, b, x0, tol: float = 1e-8, max_iter: int = 4000):
    """SA-AMG set-up on the CSR matrix and PCG from x0 (src/cfa/linalg.cuh:172-256, 864-1432).
    Returns (x, stats)."""
    return _single_solve("oracle_pcg_amg", rp, ci, val, b, x0, tol, max_iter)


def pcg_jacobi(rp, ci, val, b, x0, tol: float = 1e-8, max_iter: int = 4000):
    """Jacobi-preconditioned CG from x0 (--pressure-precond jacobi). Returns (x, stats)."""
    return _single_solve("oracle_pcg_jacobi", rp, ci, val, b, x0, tol, max_iter)


def bicgstab_jacobi(rp, ci, val, b, x0, tol: float = 1e-7, max_iter: int = 200):
    """Jacobi-preconditioned BiCGSTAB from x0 (src/cfa/linalg.cuh:261-347). Returns (x, stats)."""
    x, st = _single_solve("oracle_bicgstab_jacobi", rp, ci, val, b, x0, tol, max_iter)
    del st["amg_levels"]
    return x, st


def bicgstab2_ilu(rp, ci, val, bu, bv, u0, v0, tol: float = 1e-8, max_iter: int = 150):
    """Greedy multicolour ILU(0) and the two-component BiCGSTAB2 from (u0, v0)
    (src/cfa/linalg.cuh:361-779). Returns (u, v, stats)."""
    rp, ci, val = _csr(rp, ci, val)
    bu = np.ascontiguousarray(bu, np.float64)
    bv = np.ascontiguousarray(bv, np.float64)
    u = np.array(u0, np.float64, copy=True)
    v = np.array(v0, np.float64, copy=True)
    st = np.zeros(7)
    load().oracle_bicgstab2_ilu(len(bu), _i(rp), _i(ci), _d(val), _d(bu), _d(bv), _d(u), _d(v), tol, max_iter,
                                _d(st))
    return u, v, {"it_u": int(st[0]), "it_v": int(st[1]), "converged_u": bool(st[2]), "converged_v": bool(st[3]),
                  "rel_u": float(st[4]), "rel_v": float(st[5]), "colours": int(st[6])}


def uniform_mesh(levels: int) -> dict:
    """The reference's base mesh refined `levels` times by the oracle's NVB restatement (Morton
    order): the keys of the product's cfa.uniform_mesh (without the probes)."""
    lib = load()
    lib.oracle_mesh_uniform.argtypes = [ctypes.c_int]
    lib.oracle_mesh_uniform.restype = ctypes.c_void_p
    lib.oracle_mesh_free.argtypes = [ctypes.c_void_p]
    lib.oracle_mesh_sizes.argtypes = [ctypes.c_void_p, ctypes.POINTER(ctypes.c_int)]
    lib.oracle_mesh_arrays.argtypes = [ctypes.c_void_p] + [ctypes.POINTER(ctypes.c_double)] * 2 + \
        [ctypes.POINTER(ctypes.c_int)] * 5
    h = lib.oracle_mesh_uniform(levels)
    try:
        sizes = np.zeros(6, np.int32)
        lib.oracle_mesh_sizes(h, _i(sizes))
        n, ne, nnz, ncyl = (int(v) for v in sizes[:4])
        m = {"n": n, "ne": ne, "nnz": nnz, "n_cyl": ncyl, "max_level": int(sizes[4]), "min_angle_deg": sizes[5] / 1.0e6,
             "x": np.zeros(n), "y": np.zeros(n), "bc": np.zeros(n, np.int32), "ev": np.zeros(3 * ne, np.int32),
             "rp": np.zeros(n + 1, np.int32), "ci": np.zeros(nnz, np.int32), "cyl": np.zeros(3 * ncyl, np.int32)}
        lib.oracle_mesh_arrays(h, _d(m["x"]), _d(m["y"]), _i(m["bc"]), _i(m["ev"]), _i(m["rp"]), _i(m["ci"]),
                               _i(m["cyl"]))
        return m
    finally:
        lib.oracle_mesh_free(h)
