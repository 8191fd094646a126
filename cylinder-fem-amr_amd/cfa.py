"""ctypes binding of the gfx950 library (include/cfa_hip.h, libcfa_hip.so) and of its C++ host
layer (include/cfa_host.h, libcfa_host.so).

The host layer keeps the reference's program and solver interfaces (src/cylinder_fem_amr.cu
main, src/cfa/*.cuh classes); this module exposes the in-process entry points a Python caller
needs: run the program with the reference's flags, read the per-step history at full precision
and the MMS figures, and build the reference's meshes. The libraries are built in-tree by
`make -C cylinder-fem-amr_amd` (or __graft_entry__.build()); loading fails loudly when they are
missing, there is no fallback.
"""
from __future__ import annotations

import ctypes
import os
import re
from pathlib import Path

import numpy as np

PKG_DIR = Path(__file__).resolve().parent
BUILD_DIR = PKG_DIR / "build"
INCLUDE_DIR = PKG_DIR.parent / "include"
HIP_LIB = BUILD_DIR / "libcfa_hip.so"
HOST_LIB = BUILD_DIR / "libcfa_host.so"
CLI = BUILD_DIR / "cylinder_fem_amr"
TEST_BIN = BUILD_DIR / "test_cfa"

# columns of cfa_last_run_history (the forces CSV columns of src/cylinder_fem_amr.cu:236, then
# iteration counts and the elapsed time at the end of the step)
HISTORY_COLUMNS = ("step", "time", "C_D", "C_L", "delta_p", "nodes", "elements", "p_front", "p_back",
                   "C_D_surface", "C_L_surface", "it_u", "it_v", "it_p", "elapsed_s")
STATS_FIELDS = ("wall_s", "momentum_s", "pressure_s", "amr_s", "amg_setup_s", "steps", "nodes", "elements",
                "nonconverged", "it_u", "it_v", "it_p", "C_D", "C_L", "delta_p")
MMS_FIELDS = ("err_u_L2", "err_u_H1", "err_p_L2", "norm_u_L2", "norm_u_H1", "norm_p_L2", "err_p_cyl_max",
              "err_p_probe_max", "h_max")

_hip = None
_host = None


class LibraryMissing(RuntimeError):
    pass


def load():
    """Loads libcfa_hip.so and libcfa_host.so (RTLD_GLOBAL, so the host layer resolves the
    kernel library) and declares the signatures used here."""
    global _hip, _host
    if _host is not None:
        return _hip, _host
    for lib in (HIP_LIB, HOST_LIB):
        if not lib.exists():
            raise LibraryMissing(f"{lib} is missing: build the HIP library first "
                                 f"(make -C {PKG_DIR}, or __graft_entry__.build())")
    _hip = ctypes.CDLL(str(HIP_LIB), mode=ctypes.RTLD_GLOBAL)
    _host = ctypes.CDLL(str(HOST_LIB), mode=ctypes.RTLD_GLOBAL)
    c_int, c_double, c_void_p = ctypes.c_int, ctypes.c_double, ctypes.c_void_p
    dptr = ctypes.POINTER(c_double)
    iptr = ctypes.POINTER(c_int)
    _host.cfa_main.argtypes = [c_int, ctypes.POINTER(ctypes.c_char_p)]
    _host.cfa_main.restype = c_int
    _host.cfa_last_run_stats.argtypes = [dptr]
    _host.cfa_last_run_stats.restype = None
    _host.cfa_last_run_history.argtypes = [dptr, c_int]
    _host.cfa_last_run_history.restype = c_int
    _host.cfa_last_run_kernels.argtypes = [ctypes.c_char_p, c_int, dptr, c_int]
    _host.cfa_last_run_kernels.restype = c_int
    _host.cfa_last_run_mms.argtypes = [dptr]
    _host.cfa_last_run_mms.restype = None
    _host.cfa_mesh_uniform.argtypes = [c_int]
    _host.cfa_mesh_uniform.restype = c_void_p
    _host.cfa_mesh_free.argtypes = [c_void_p]
    _host.cfa_mesh_free.restype = None
    _host.cfa_mesh_sizes.argtypes = [c_void_p, iptr]
    _host.cfa_mesh_sizes.restype = None
    _host.cfa_mesh_arrays.argtypes = [c_void_p, dptr, dptr, iptr, iptr, iptr, iptr, iptr]
    _host.cfa_mesh_arrays.restype = None
    _host.cfa_mesh_incidence.argtypes = [c_void_p, iptr, iptr, iptr]
    _host.cfa_mesh_incidence.restype = None
    _host.cfa_mesh_probes.argtypes = [c_void_p, iptr, dptr]
    _host.cfa_mesh_probes.restype = None
    _host.cfa_workload_pressure.argtypes = [c_int, c_int, c_double, c_int, dptr]
    _host.cfa_workload_pressure.restype = c_int
    _host.cfa_workload_momentum.argtypes = [c_int, c_double, c_int, dptr]
    _host.cfa_workload_momentum.restype = c_int
    _hip.cfa_error_string.argtypes = [c_int]
    _hip.cfa_error_string.restype = ctypes.c_char_p
    return _hip, _host


def header_functions(header: str | os.PathLike) -> list[str]:
    """Names of the functions a header declares (prototypes at the start of a line)."""
    text = Path(header).read_text()
    text = re.sub(r"/\*.*?\*/", " ", text, flags=re.S)
    names = re.findall(r"^\s*(?:const\s+)?[A-Za-z_][A-Za-z0-9_]*\s*\**\s+\**\s*(cfa_[A-Za-z0-9_]+)\s*\(", text,
                       flags=re.M)
    return sorted(set(names))


def _dptr(a: np.ndarray):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def _iptr(a: np.ndarray):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_int))


def run(args: list[str]) -> int:
    """Runs the program in-process with the reference's command-line flags (argv[0] added)."""
    _, host = load()
    argv = [b"cylinder_fem_amr"] + [str(a).encode() for a in args]
    arr = (ctypes.c_char_p * len(argv))(*argv)
    return host.cfa_main(len(argv), arr)


def last_stats() -> dict:
    _, host = load()
    out = np.zeros(16)
    host.cfa_last_run_stats(_dptr(out))
    return dict(zip(STATS_FIELDS, out[:15].tolist()))


def last_history() -> np.ndarray:
    """Per-step history of the last run, shape (steps, 15), columns HISTORY_COLUMNS."""
    _, host = load()
    rows = host.cfa_last_run_history(None, 0)
    out = np.zeros((rows, len(HISTORY_COLUMNS)))
    if rows:
        host.cfa_last_run_history(_dptr(out), rows)
    return out


def last_kernels() -> dict:
    """--kernel-timing results of the last run: name -> {name, bytes, us, launches}."""
    _, host = load()
    count = host.cfa_last_run_kernels(None, 0, None, 0)
    buf = ctypes.create_string_buffer(65536)
    out = np.zeros((max(count, 1), 3))
    host.cfa_last_run_kernels(buf, len(buf), _dptr(out), count)
    names = buf.value.decode().split("\n") if count else []
    return {nm: {"name": nm, "bytes": float(out[k, 0]), "us": float(out[k, 1]), "launches": int(out[k, 2])}
            for k, nm in enumerate(names)}


PRESSURE_FIELDS = ("n", "nnz", "ne", "mesh_s", "amg_setup_ms", "amg_levels", "solves", "solve_ms", "iterations",
                   "ms_per_solve", "us_per_iteration", "max_rel_residual", "amg_setup_cold_ms",
                   "fp32_ms_per_solve", "fp32_iterations", "fp32_max_rel_residual")
MOMENTUM_FIELDS = ("n", "nnz", "ne", "mesh_s", "colouring_ms", "colours", "fp64_factor_ms", "fp64_solve_ms",
                   "fp64_it_u", "fp64_it_v", "fp32_factor_ms", "fp32_solve_ms", "fp32_it_u", "fp32_it_v",
                   "rel_diff_fp32_fp64", "us_per_fp64_iteration")


def workload_pressure(level: int, nrhs: int = 32, tol: float = 1e-8, kernel_reps: int = 0) -> dict:
    _, host = load()
    out = np.zeros(len(PRESSURE_FIELDS))
    rc = host.cfa_workload_pressure(level, nrhs, tol, kernel_reps, _dptr(out))
    if rc != 0:
        raise RuntimeError(f"cfa_workload_pressure failed: {rc}")
    return dict(zip(PRESSURE_FIELDS, out.tolist()))


AMG_PHASES = ("download", "dinv_rho", "aggregation", "p_smoothing", "transpose", "galerkin", "upload", "coarse_inverse",
              "remap", "device_levels", "front_rounds", "sorted_products")


def workload_amg_phases() -> dict:
    """AMG set-up phase times (ms) of the last pressure workload and the levels built on the device."""
    _, host = load()
    host.cfa_workload_amg_phases.argtypes = [ctypes.POINTER(ctypes.c_double)]
    host.cfa_workload_amg_phases.restype = ctypes.c_int
    out = np.zeros(len(AMG_PHASES))
    host.cfa_workload_amg_phases(_dptr(out))
    return dict(zip(AMG_PHASES, out.tolist()))


def workload_momentum(level: int, tol: float = 1e-8, kernel_reps: int = 0) -> dict:
    _, host = load()
    out = np.zeros(16)
    rc = host.cfa_workload_momentum(level, tol, kernel_reps, _dptr(out))
    if rc != 0:
        raise RuntimeError(f"cfa_workload_momentum failed: {rc}")
    return dict(zip(MOMENTUM_FIELDS, out.tolist()))


def workload_keep_systems(on: bool = True) -> None:
    """Keep host copies of the following workload systems (cfa_workload_keep_systems)."""
    _, host = load()
    host.cfa_workload_keep_systems.argtypes = [ctypes.c_int]
    host.cfa_workload_keep_systems.restype = ctypes.c_int
    host.cfa_workload_keep_systems(1 if on else 0)


def workload_system(which: str) -> dict:
    """The kept system of the last pressure or momentum workload: rp, ci, tag, val, b, x0 (one row
    per right-hand side) and x (pressure: one solution per right-hand side; momentum: u, v with FP64
    factors, then u, v with FP32 factors)."""
    _, host = load()
    k = {"pressure": 0, "momentum": 1}[which]
    ip = ctypes.POINTER(ctypes.c_int)
    host.cfa_workload_system_sizes.argtypes = [ctypes.c_int, ip, ip, ip]
    host.cfa_workload_system_sizes.restype = ctypes.c_int
    host.cfa_workload_system.argtypes = [ctypes.c_int, ip, ip, ip] + [ctypes.POINTER(ctypes.c_double)] * 4 + [ip]
    host.cfa_workload_system.restype = ctypes.c_int
    n, nnz, cols = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
    _check(host.cfa_workload_system_sizes(k, ctypes.byref(n), ctypes.byref(nnz), ctypes.byref(cols)),
           "cfa_workload_system_sizes")
    n, nnz, cols = n.value, nnz.value, cols.value
    rp = np.zeros(n + 1, np.int32)
    ci = np.zeros(nnz, np.int32)
    tag = np.zeros(n, np.int32)
    val = np.zeros(nnz)
    b = np.zeros((cols, n))
    x0 = np.zeros((cols, n))
    x = np.zeros((cols if k == 0 else 4, n))
    iters = np.zeros(x.shape[0], np.int32)
    _check(host.cfa_workload_system(k, _iptr(rp), _iptr(ci), _iptr(tag), _dptr(val), _dptr(b), _dptr(x0), _dptr(x),
                                    _iptr(iters)), "cfa_workload_system")
    return {"n": n, "nnz": nnz, "rp": rp, "ci": ci, "tag": tag, "val": val, "b": b, "x0": x0, "x": x,
            "iters": iters}


DUMP_MOMENTUM_FIELDS = ("n", "nnz", "colours", "colouring_ms", "factor_ms", "solve_ms", "it_u", "it_v",
                        "converged_u", "converged_v", "rel_dev_from_dumped")
DUMP_PRESSURE_FIELDS = ("n", "nnz", "amg_levels", "amg_setup_ms", "solve_ms", "iterations", "converged",
                        "rel_residual")


def read_system(path: str | os.PathLike) -> dict:
    """Parses a --dump-systems file on the host (the layout of src/cylinder_fem_amr.cu:331-352,
    375-378 for <prefix>_momentum.bin and :426-438 for <prefix>_pressure.bin; the reference reads
    it the same way in scripts/momentum_study/study.py load()). Keys: n, nnz, rp, ci, val, then
    for momentum b_u, b_v, u0, v0, x, y, bc_tag and, when present, u, v; for pressure b, x0."""
    raw = Path(path).read_bytes()
    n, nnz = (int(k) for k in np.frombuffer(raw, np.int32, 2, 0))
    base = 8 + 4 * (n + 1) + 12 * nnz
    momentum = len(raw) in (base + 52 * n, base + 68 * n)
    if n <= 0 or nnz <= 0 or not (momentum or len(raw) == base + 16 * n):
        raise ValueError(f"{path}: size {len(raw)} fits neither dump layout for n={n}, nnz={nnz}")
    off = 8
    out = {"n": n, "nnz": nnz}

    def take(key, dtype, count):
        nonlocal off
        a = np.frombuffer(raw, dtype, count, off).copy()
        off += a.nbytes
        out[key] = a

    take("rp", np.int32, n + 1)
    take("ci", np.int32, nnz)
    take("val", np.float64, nnz)
    if momentum:
        for key in ("b_u", "b_v", "u0", "v0", "x", "y"):
            take(key, np.float64, n)
        take("bc_tag", np.int32, n)
        if off < len(raw):
            take("u", np.float64, n)
            take("v", np.float64, n)
    else:
        take("b", np.float64, n)
        take("x0", np.float64, n)
    return out


def _declare_dump_functions(host):
    c_int, c_double, c_char_p = ctypes.c_int, ctypes.c_double, ctypes.c_char_p
    dptr = ctypes.POINTER(c_double)
    host.cfa_dumped_momentum_solve.argtypes = [c_char_p, c_double, c_int, c_int, c_int, dptr, dptr, dptr]
    host.cfa_dumped_momentum_solve.restype = c_int
    host.cfa_dumped_pressure_solve.argtypes = [c_char_p, c_double, c_int, c_int, c_int, c_int, dptr, dptr]
    host.cfa_dumped_pressure_solve.restype = c_int
    host.cfa_dump_header.argtypes = [c_char_p, ctypes.POINTER(c_int), ctypes.POINTER(c_int)]
    host.cfa_dump_header.restype = c_int


def _check(rc: int, what: str):
    if rc != 0:
        hip, _ = load()
        raise RuntimeError(f"{what} failed: {hip.cfa_error_string(rc).decode()} ({rc})")


def solve_dumped_momentum(path: str | os.PathLike, tol: float = 1e-7, max_iter: int = 150, precision: str = "fp64",
                          kernel_reps: int = 0):
    """BiCGSTAB2 with multicolour ILU(0) on a dumped momentum system, on the GPU (the library's own
    solver classes). Returns (stats dict, u, v)."""
    _, host = load()
    _declare_dump_functions(host)
    n, nnz = ctypes.c_int(0), ctypes.c_int(0)
    _check(host.cfa_dump_header(str(path).encode(), ctypes.byref(n), ctypes.byref(nnz)), "cfa_dump_header")
    u, v, out = np.zeros(n.value), np.zeros(n.value), np.zeros(11)
    _check(host.cfa_dumped_momentum_solve(str(path).encode(), tol, max_iter, {"fp64": 0, "mixed": 1}[precision],
                                          kernel_reps, _dptr(u), _dptr(v), _dptr(out)), "cfa_dumped_momentum_solve")
    return dict(zip(DUMP_MOMENTUM_FIELDS, out.tolist())), u, v


def solve_dumped_pressure(path: str | os.PathLike, tol: float = 1e-7, max_iter: int = 4000, precision: str = "fp64",
                          from_zero: bool = False, kernel_reps: int = 0):
    """SA-AMG PCG on K_p of a dumped pressure system, on the GPU. Returns (stats dict, x)."""
    _, host = load()
    _declare_dump_functions(host)
    n, nnz = ctypes.c_int(0), ctypes.c_int(0)
    _check(host.cfa_dump_header(str(path).encode(), ctypes.byref(n), ctypes.byref(nnz)), "cfa_dump_header")
    x, out = np.zeros(n.value), np.zeros(8)
    _check(host.cfa_dumped_pressure_solve(str(path).encode(), tol, max_iter, {"fp64": 0, "mixed": 1}[precision],
                                          1 if from_zero else 0, kernel_reps, _dptr(x), _dptr(out)),
           "cfa_dumped_pressure_solve")
    return dict(zip(DUMP_PRESSURE_FIELDS, out.tolist())), x


ADJOINT_STATS = ("forward_ms", "reverse_ms", "recomputed_steps", "adjoint_pcg_iterations",
                 "adjoint_bicgstab2_iterations")


class Adjoint:
    """Frozen uniform-mesh window of the program's time step with its discrete adjoint (C ABI
    cfa_adjoint_*, cfa_step_fwd, cfa_window_forward, cfa_window_adjoint). theta = (nu, U_max);
    a state is 6 n doubles [u | v | u_old | v_old | p | dp]; the window functional is the mean of
    weights[0] C_D + weights[1] C_L over the window's steps."""

    def __init__(self, level: int, dt: float = 1e-3, tol: float = 1e-12):
        _, host = load()
        c_int, c_double, vp = ctypes.c_int, ctypes.c_double, ctypes.c_void_p
        dptr, iptr = ctypes.POINTER(c_double), ctypes.POINTER(c_int)
        host.cfa_adjoint_create.argtypes = [c_int, c_double, c_double]
        host.cfa_adjoint_create.restype = vp
        host.cfa_adjoint_destroy.argtypes = [vp]
        host.cfa_adjoint_destroy.restype = None
        host.cfa_adjoint_sizes.argtypes = [vp, iptr, iptr, iptr]
        host.cfa_adjoint_sizes.restype = c_int
        host.cfa_adjoint_initial_state.argtypes = [vp, dptr]
        host.cfa_adjoint_initial_state.restype = c_int
        host.cfa_step_fwd.argtypes = [vp, dptr, c_int, dptr, dptr, dptr]
        host.cfa_step_fwd.restype = c_int
        host.cfa_window_forward.argtypes = [vp, dptr, c_int, c_int, dptr, dptr, dptr, dptr, dptr]
        host.cfa_window_forward.restype = c_int
        host.cfa_window_adjoint.argtypes = [vp, dptr, c_int, c_int, c_int, dptr, dptr, dptr, dptr, dptr, dptr, dptr]
        host.cfa_window_adjoint.restype = c_int
        host.cfa_adjoint_set_shape.argtypes = [vp, dptr, c_int]
        host.cfa_adjoint_set_shape.restype = c_int
        self._host = host
        self._ctx = host.cfa_adjoint_create(level, dt, tol)
        if not self._ctx:
            raise RuntimeError(f"cfa_adjoint_create(level={level}) failed")
        sizes = np.zeros(3, dtype=np.int32)
        host.cfa_adjoint_sizes(self._ctx, _iptr(sizes[0:1]), _iptr(sizes[1:2]), _iptr(sizes[2:3]))
        self.n, self.ne, self.nnz = (int(k) for k in sizes)

    def close(self):
        if self._ctx:
            self._host.cfa_adjoint_destroy(self._ctx)
            self._ctx = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @staticmethod
    def _vec(a, size=None):
        a = np.ascontiguousarray(a, dtype=np.float64)
        if size is not None and a.size != size:
            raise ValueError(f"expected {size} values, got {a.size}")
        return a

    def initial_state(self) -> np.ndarray:
        out = np.zeros(6 * self.n)
        _check(self._host.cfa_adjoint_initial_state(self._ctx, _dptr(out)), "cfa_adjoint_initial_state")
        return out

    def step(self, theta, step: int, state):
        th = self._vec(theta, 2)
        s_in = self._vec(state, 6 * self.n)
        s_out, forces = np.zeros(6 * self.n), np.zeros(2)
        _check(self._host.cfa_step_fwd(self._ctx, _dptr(th), step, _dptr(s_in), _dptr(s_out), _dptr(forces)),
               "cfa_step_fwd")
        return s_out, forces

    def forward(self, theta, first_step: int, steps: int, state0, weights=(1.0, 0.0)):
        """Returns (J, final state, history (steps x 2) of C_D, C_L)."""
        th, w = self._vec(theta, 2), self._vec(weights, 2)
        s0 = self._vec(state0, 6 * self.n)
        s_out, J, hist = np.zeros(6 * self.n), np.zeros(1), np.zeros((steps, 2))
        _check(self._host.cfa_window_forward(self._ctx, _dptr(th), first_step, steps, _dptr(w), _dptr(s0),
                                             _dptr(s_out), _dptr(J), _dptr(hist)), "cfa_window_forward")
        return float(J[0]), s_out, hist

    N_MODES = 64

    def set_shape(self, eps):
        """Radial modes of the hole boundary (64 values; zeros = the reference geometry)."""
        e = self._vec(eps, self.N_MODES)
        _check(self._host.cfa_adjoint_set_shape(self._ctx, _dptr(e), self.N_MODES), "cfa_adjoint_set_shape")

    def gradient(self, theta, first_step: int, steps: int, state0, checkpoint_every: int = 20,
                 weights=(1.0, 0.0), state_gradient: bool = False, shape_gradient: bool = False):
        """Returns (J, dJ/dtheta [2], dJ/dstate0 [6 n] or None, stats dict); with shape_gradient the
        dict also holds dJ_dshape [64]."""
        th, w = self._vec(theta, 2), self._vec(weights, 2)
        s0 = self._vec(state0, 6 * self.n)
        J, g, st = np.zeros(1), np.zeros(2), np.zeros(5)
        gs = np.zeros(6 * self.n) if state_gradient else None
        gsh = np.zeros(self.N_MODES) if shape_gradient else None
        _check(self._host.cfa_window_adjoint(self._ctx, _dptr(th), first_step, steps, checkpoint_every, _dptr(w),
                                             _dptr(s0), _dptr(J), _dptr(g),
                                             _dptr(gsh) if gsh is not None else None,
                                             _dptr(gs) if gs is not None else None, _dptr(st)),
               "cfa_window_adjoint")
        stats = dict(zip(ADJOINT_STATS, st.tolist()))
        if gsh is not None:
            stats["dJ_dshape"] = gsh
        return float(J[0]), g, gs, stats


def last_mms() -> dict:
    _, host = load()
    out = np.zeros(9)
    host.cfa_last_run_mms(_dptr(out))
    return dict(zip(MMS_FIELDS, out.tolist()))


def uniform_mesh(levels: int) -> dict:
    """The reference's base mesh refined `levels` times (NVB, Morton order): arrays and sizes."""
    _, host = load()
    h = host.cfa_mesh_uniform(levels)
    try:
        sizes = np.zeros(6, dtype=np.int32)
        host.cfa_mesh_sizes(h, _iptr(sizes))
        n, ne, nnz, ncyl = (int(v) for v in sizes[:4])
        m = {"n": n, "ne": ne, "nnz": nnz, "n_cyl": ncyl, "max_level": int(sizes[4]),
             "min_angle_deg": sizes[5] / 1.0e6,
             "x": np.zeros(n), "y": np.zeros(n), "bc": np.zeros(n, np.int32), "ev": np.zeros(3 * ne, np.int32),
             "rp": np.zeros(n + 1, np.int32), "ci": np.zeros(nnz, np.int32), "cyl": np.zeros(3 * ncyl, np.int32)}
        host.cfa_mesh_arrays(h, _dptr(m["x"]), _dptr(m["y"]), _iptr(m["bc"]), _iptr(m["ev"]), _iptr(m["rp"]),
                             _iptr(m["ci"]), _iptr(m["cyl"]))
        ab = np.zeros(4, np.int32)
        w = np.zeros(2)
        host.cfa_mesh_probes(h, _iptr(ab), _dptr(w))
        m["probes"] = (ab, w)
        return m
    finally:
        host.cfa_mesh_free(h)
