#!/usr/bin/env python3
"""Benchmark of the per-step projection solve on one MI355X.

Workload (default): the adaptive obstacle flow with NVB refinement capped at level 6 (levels 2-6,
element budget 600k, remesh every 20 steps overlapped by 10, look-ahead, coarsening, dt 5e-4,
parabolic inflow of mean 1, Re 100), the momentum BiCGSTAB and pressure PCG run to a relative
residual of 1e-8. A step is one time step: momentum predictor, pressure increment, correction,
forces, and the remesh work that falls on it (the launch of a remesh, and at its join the upload of
the new mesh, the solution transfer, the scheme-consistent re-projection and the AMG set-up).

The run goes through an untimed spin-up (--spinup, 1000 steps by default, so the window sees a
developed wake and its refined mesh), then --warmup W untimed steps, then exactly --steps K timed
steps, then an untimed tail of overlap + 1 steps (a remesh launched in the window is only allowed
when its join falls before the last step). Any 20 consecutive steps hold one remesh launch and one
join. The program synchronises the device at the end of every step (the forces are read back), and
the timed region runs from the end of step spin-up + W to the end of step spin-up + W + K on the
host's steady clock. The whole run is done twice: with the FP32 ILU(0) factors and FP32 SA-AMG
cycle inside the FP64 Krylov solvers (`value`), and in FP64 (`fp64`); the force coefficients of
the two runs are compared step by step while their meshes are identical.

After the FP64 run the main kernels are launched alone on the final mesh and state
(--kernel-timing) and timed with HIP events on the solver stream ("kernels"). The roofline object
reports the dominant kernel family's level-0 member, the damped-Jacobi sweep of the SA-AMG
V-cycle, measured the same way on a fixed shape so that the hardware-counter passes can profile
exactly that launch (--roofline-only). The CPU baseline is the oracle (the FP64 CPU restatement of
the step, OpenMP) continuing the FP64 run from the state at the start of the timed window, on all
cores this process may use and with the reference's thread rule min(8, nproc / 2).

Prints one JSON line.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
from pathlib import Path

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT / "cylinder-fem-amr_amd"))

HBM_PEAK_GBS = 8000.0
DOMINANT = "level-0 Jacobi sweep ("   # the FP64 sweep's timing, whichever layout the level carries
ROOFLINE_LEVEL = 7
TRAFFIC_FILE = ROOT / "profiles" / "jacobi_level0_uniform_L7_counters.txt"

WORKLOADS = {
    # runs/paper/d2_A2_dt05/cmd.txt (--steps set by the benchmark)
    "adaptive": ["--dt", "0.0005", "--amr-interval", "20", "--amr-overlap", "10", "--amr-tol-vort", "0.05",
                 "--amr-tol-interp", "0.005", "--max-level", "5", "--max-elems", "400000", "--min-level", "2"],
    # runs/paper/mms_L3/cmd.txt
    "mms_u3": ["--mms", "--uniform-level", "3", "--dt", "0.00025"],
    # the adaptive flow above with refinement capped at level 6 (300k-500k triangles)
    "adaptive_l6": ["--dt", "0.0005", "--amr-interval", "20", "--amr-overlap", "10", "--amr-tol-vort", "0.05",
                    "--amr-tol-interp", "0.005", "--max-level", "6", "--max-elems", "600000", "--min-level", "2"],
}
DESCRIPTION = {
    "adaptive": "adaptive obstacle flow, NVB levels 2-5, remesh every 20 steps, dt 5e-4, Re 100",
    "mms_u3": "manufactured solution on uniform level 3 (29152 nodes), dt 2.5e-4",
    "adaptive_l6": "adaptive obstacle flow, NVB levels 2-6 (element budget 600k), remesh every 20 steps, dt 5e-4, "
                   "Re 100",
}


def host_cores() -> dict:
    """Threads this process may use, the machine's physical cores (lscpu) and its CPU model."""
    import subprocess
    avail = len(os.sched_getaffinity(0))
    phys, model = None, ""
    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=30).stdout
        kv = {k.strip(): v.strip() for k, v in (ln.split(":", 1) for ln in out.splitlines() if ":" in ln)}
        phys = int(kv.get("Core(s) per socket", "0")) * int(kv.get("Socket(s)", "0")) or None
        model = kv.get("Model name", "")
    except (OSError, ValueError, subprocess.SubprocessError):
        pass
    omp = os.environ.get("OMP_NUM_THREADS")
    share = int(omp) if omp and omp.isdigit() and int(omp) > 0 else avail
    return {"affinity": avail, "physical": phys, "omp_num_threads": omp, "model": model,
            "threads": max(1, min(share, avail, phys or avail))}


def cpu_baseline(state_file: str, dt: float, tol: float, seconds_target: float = 12.0) -> dict:
    """The oracle (the FP64 CPU restatement of the step, OpenMP row loops) continuing the benchmark
    run from the state at the start of the timed window (mesh and fields written by --dump-state),
    on a fixed mesh for a bounded number of steps: once on every core this process may use and once
    with the reference's host thread rule min(8, nproc / 2) (src/cylinder_fem_amr.cu:138)."""
    sys.path.insert(0, str(ROOT / "oracle"))
    import cfa_oracle
    cores = host_cores()
    runs = {}
    for label, threads in (("all_cores", cores["threads"]), ("reference_rule", max(1, min(8, cores["affinity"] // 2)))):
        rows, wall = cfa_oracle.run_state(state_file, 2, dt, tol, threads)
        per_step = wall / max(len(rows), 1)
        steps = max(2, min(200, int(seconds_target / max(per_step, 1e-6))))
        rows, wall = cfa_oracle.run_state(state_file, steps, dt, tol, threads)
        runs[label] = {"value": round(len(rows) / wall, 4), "threads": threads, "steps": len(rows),
                       "seconds": round(wall, 2),
                       "nodes": int(rows[0, cfa_oracle.ROW_COLUMNS.index("nodes")]) if len(rows) else 0}
    main = runs["all_cores"]
    return {"value": main["value"], "unit": "steps/s", "cores": main["threads"], "kind": "port",
            "sample": f"oracle (FP64 CPU restatement, OpenMP) continuing the FP64 run from the state at the start "
                      f"of the timed window, fixed mesh of {main['nodes']} nodes, dt {dt:g}, tol {tol:g}: "
                      f"{main['steps']} steps in {main['seconds']} s on {main['threads']} threads (the window's "
                      f"remesh work is not in the sample)",
            "reference_thread_rule": runs["reference_rule"], "host": cores}


def trace_summary(path: str, reps: int, top: int = 25) -> int:
    """Per-kernel dispatch statistics of a rocprofv3 kernel trace (CSV), and the mean duration of
    the last `reps` dispatches of each kernel (the --kernel-timing pass, which launches each timed
    kernel alone `reps` times after the run) for comparison with the HIP-event figures."""
    import csv
    from collections import defaultdict
    durs = defaultdict(list)
    with open(path, newline="") as f:
        for row in csv.DictReader(f):
            durs[row["Kernel_Name"]].append(int(row["End_Timestamp"]) - int(row["Start_Timestamp"]))
    total = sum(sum(v) for v in durs.values())
    print(f"kernel dispatches {sum(len(v) for v in durs.values())}, busy {total / 1e6:.3f} ms")
    print(f"{'total ms':>10} {'share':>6} {'calls':>8} {'mean us':>9} {'last-' + str(reps) + ' mean us':>16}  kernel")
    for name, v in sorted(durs.items(), key=lambda kv: -sum(kv[1]))[:top]:
        tail = v[-reps:] if reps > 0 else v
        print(f"{sum(v) / 1e6:10.3f} {100.0 * sum(v) / total:5.1f}% {len(v):8d} {sum(v) / len(v) / 1e3:9.2f} "
              f"{sum(tail) / len(tail) / 1e3:16.2f}  {name[:120]}")
    return 0


def counter_summary(path: str, reps: int, top: int = 25) -> int:
    """Mean value per dispatch of each counter in a rocprofv3 --pmc CSV, over all dispatches of a
    kernel and over its last `reps` dispatches (the --kernel-timing pass)."""
    import csv
    from collections import defaultdict
    vals = defaultdict(lambda: defaultdict(list))
    with open(path, newline="") as f:
        for row in csv.DictReader(f):
            vals[row["Counter_Name"]][row["Kernel_Name"]].append(float(row["Counter_Value"]))
    for counter, per_kernel in vals.items():
        print(f"counter {counter}")
        print(f"{'calls':>8} {'mean':>14} {'last-' + str(reps) + ' mean':>16}  kernel")
        for name, v in sorted(per_kernel.items(), key=lambda kv: -sum(kv[1]))[:top]:
            tail = v[-reps:] if reps > 0 else v
            print(f"{len(v):8d} {sum(v) / len(v):14.1f} {sum(tail) / len(tail):16.1f}  {name[:120]}")
    return 0


def solver_workloads(level: int, nrhs: int, tol: float, reps: int) -> int:
    """The isolated solver workloads on the uniform level-`level` mesh: SA-AMG PCG on the pressure
    Laplacian (nrhs seeded right-hand sides) and ILU-BiCGSTAB2 on the momentum system with FP64 and
    FP32 factors; one JSON line with times per solve and per iteration and the kernel timings."""
    import cfa
    p = cfa.workload_pressure(level, nrhs, tol, reps)
    p["amg_phases_ms"] = cfa.workload_amg_phases()
    pk = cfa.last_kernels()
    m = cfa.workload_momentum(level, tol, reps)
    mk = cfa.last_kernels()
    ks = {}
    for k, v in list(pk.items()) + list(mk.items()):
        gbs = v["bytes"] / (v["us"] * 1e3)
        ks[k] = {"us": round(v["us"], 3), "bytes": v["bytes"], "GB/s": round(gbs, 1), "frac": round(gbs / HBM_PEAK_GBS, 4)}
    print(json.dumps({"workload": f"frozen uniform level-{level} mesh, tol {tol:g}", "pressure": p, "momentum": m,
                      "kernels": ks}), flush=True)
    return 0


def read_traffic() -> dict | None:
    """HBM bytes per launch of the roofline kernel from the committed counter summary (two
    rocprofv3 --pmc passes of `bench.py --roofline-only`: FETCH_SIZE and WRITE_SIZE in KiB per
    dispatch, each scaled by the same counter over a 1 GiB device copy)."""
    if not TRAFFIC_FILE.exists():
        return None
    vals = {}
    for line in TRAFFIC_FILE.read_text().splitlines():
        if ":" in line and not line.startswith("#"):
            k, v = line.split(":", 1)
            try:
                vals[k.strip()] = float(v.split()[0])
            except (ValueError, IndexError):
                pass
    need = ("kernel_fetch_kib", "kernel_write_kib", "copy_fetch_kib", "copy_write_kib", "copy_bytes")
    if not all(k in vals for k in need):
        return None
    half = vals["copy_bytes"] / 2.0   # the copy reads and writes half of its bytes each
    f_scale = half / (1024.0 * vals["copy_fetch_kib"])
    w_scale = half / (1024.0 * vals["copy_write_kib"])
    return {"bytes": 1024.0 * (vals["kernel_fetch_kib"] * f_scale + vals["kernel_write_kib"] * w_scale),
            "fetch_scale": f_scale, "write_scale": w_scale}


def roofline_measurement(reps: int, tol: float) -> tuple[dict, dict]:
    """The level-0 smoother of the SA-AMG V-cycle on the uniform level-7 pressure matrix (7,347,712 rows,
    51.4 M entries: 881 MB per sweep, above the 256 MB Infinity Cache, so the sweep streams HBM), launched
    `reps` times between HIP events on the solver stream (cfa_workload_pressure timings)."""
    import cfa
    cfa.workload_pressure(ROOFLINE_LEVEL, 1, tol, max(reps, 1))
    k = cfa.last_kernels()
    dom = [v for name, v in k.items() if name.startswith(DOMINANT)]
    if len(dom) != 1:
        raise RuntimeError(f"expected one timing named {DOMINANT}..., got {sorted(k)}")
    return dom[0], k


def roofline_object(dom: dict) -> dict:
    achieved = dom["bytes"] / (dom["us"] * 1e3)
    tr = read_traffic()
    return {
        "kernel": f"{dom['name']} (uniform level-{ROOFLINE_LEVEL} pressure hierarchy)",
        "bound": "hbm",
        "achieved": round(achieved, 1),
        "peak": HBM_PEAK_GBS,
        "unit": "GB/s",
        "frac": round(achieved / HBM_PEAK_GBS, 4),
        "traffic": round(tr["bytes"]) if tr else None,
        "bytes_per_launch": dom["bytes"],
        "us_per_launch": round(dom["us"], 3),
        "measured_on": "HIP events on the solver stream, the kernel launched alone",
    }


def remesh_events(flags: list[str], total: int, first: int, last: int) -> dict:
    """Remesh launches and joins in steps first..last of a run of `total` steps (the program's rule:
    a remesh is launched at the end of every interval-th step unless its join, `overlap` steps
    later, would fall at or after the last step; the join is the end of step launch + overlap)."""
    f = {flags[i]: flags[i + 1] for i in range(len(flags) - 1) if flags[i].startswith("--")}
    interval, overlap = int(f.get("--amr-interval", 0)), int(f.get("--amr-overlap", 0))
    launches = [s for s in range(interval, total, interval) if not (overlap and s + overlap >= total)] if interval else []
    joins = [s + overlap for s in launches] if overlap else list(launches)
    return {"launch_steps": [s for s in launches if first <= s <= last],
            "join_steps": [s for s in joins if first <= s <= last]}


def window_summary(hist, first: int, last: int) -> dict:
    """Steps/s over the window of steps first..last (1-based, from the end of step first - 1 to the
    end of step last on the host's steady clock) and what the window ran on."""
    import numpy as np
    col = {c: i for i, c in enumerate(__import__("cfa").HISTORY_COLUMNS)}
    t0 = hist[first - 2, col["elapsed_s"]] if first >= 2 else 0.0
    timed = hist[last - 1, col["elapsed_s"]] - t0
    w = hist[first - 1:last]
    k = last - first + 1
    return {"value": round(k / timed, 3), "ms_per_step": round(1000.0 * timed / k, 4), "seconds": timed,
            "nodes_range": [int(w[:, col["nodes"]].min()), int(w[:, col["nodes"]].max())],
            "elements_range": [int(w[:, col["elements"]].min()), int(w[:, col["elements"]].max())],
            "mean_iters_u_v_p": [round(float(w[:, col[c]].mean()), 2) for c in ("it_u", "it_v", "it_p")],
            "mesh_changes": int(np.count_nonzero(np.diff(hist[first - 2:last, col["elements"]]))) if first >= 2 else 0}


def force_deviation(h64, hmx, first: int, last: int) -> dict:
    """Per-step |dC_D|, |dC_L| between the FP64 and the mixed-precision runs: over the window, over
    the whole run up to the first step whose mesh differs, and from step 100 on (after the impulsive
    start, C_D near 276 at step 1)."""
    import numpy as np
    col = {c: i for i, c in enumerate(__import__("cfa").HISTORY_COLUMNS)}
    same = (h64[:, col["nodes"]] == hmx[:, col["nodes"]]) & (h64[:, col["elements"]] == hmx[:, col["elements"]])
    upto = int(np.argmin(same)) if not same.all() else len(same)
    out = {"steps_on_identical_meshes": upto, "window_on_identical_meshes": bool(upto >= last)}
    for c in ("C_D", "C_L"):
        d = np.abs(h64[:upto, col[c]] - hmx[:upto, col[c]])
        out[f"max_abs_d{c}"] = float(d.max()) if upto else None
        out[f"max_abs_d{c}_from_step_100"] = float(d[99:].max()) if upto > 99 else None
        out[f"max_abs_d{c}_window"] = float(d[first - 1:last].max()) if upto >= last else None
    out["target"] = "|dC_D|, |dC_L| <= 1e-8"
    return out


def adjoint_cost(level: int, spinup: int, window: int, every: int, tol: float) -> dict:
    """Discrete adjoint of the window functional J = mean C_D over `window` steps on the frozen
    uniform level-`level` mesh after a `spinup`-step FP64 run from the DFG initial condition, with
    respect to all 66 parameters (nu, U_max and the 64 radial modes of the hole boundary): time of
    the forward window, time of the gradient (forward pass with checkpoints every `every` steps,
    segment recomputation and reverse sweep), and their ratio."""
    import time
    import cfa
    theta = [1.0e-3, 1.5]
    a = cfa.Adjoint(level, dt=1e-3, tol=tol)
    t = time.perf_counter()
    _, state, _ = a.forward(theta, 1, spinup, a.initial_state())
    spin_s = time.perf_counter() - t
    t = time.perf_counter()
    J, _, _ = a.forward(theta, spinup + 1, window, state)
    fwd_s = time.perf_counter() - t
    t = time.perf_counter()
    J2, g, _, st = a.gradient(theta, spinup + 1, window, state, every, shape_gradient=True)
    grad_s = time.perf_counter() - t
    a.close()
    return {
        "workload": f"uniform level-{level} mesh ({a.n} nodes, {a.ne} elements), {spinup}-step spin-up, "
                    f"{window}-step window, checkpoints every {every}, tol {tol:g}",
        "parameters": 2 + len(st["dJ_dshape"]),
        "J": J, "dJ_dnu": float(g[0]), "dJ_dUmax": float(g[1]),
        "dJ_dshape_norm": float(sum(v * v for v in st["dJ_dshape"]) ** 0.5),
        "forward_s": round(fwd_s, 4), "gradient_s": round(grad_s, 4),
        "adjoint_over_forward": round(grad_s / fwd_s, 3),
        "reverse_over_forward": round(st["reverse_ms"] / 1000.0 / fwd_s, 3),
        "spinup_s": round(spin_s, 3),
        "adjoint_iterations_pcg_bicgstab2": [int(st["adjoint_pcg_iterations"]), int(st["adjoint_bicgstab2_iterations"])],
        "J_check": abs(J2 - J) <= 1e-12 * abs(J),
    }


def main() -> int:
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--spinup", type=int, default=None,
                    help="untimed steps before the warm-up (default: 1000 for the adaptive workloads, 0 otherwise)")
    ap.add_argument("--workload", choices=sorted(WORKLOADS), default="adaptive_l6")
    ap.add_argument("--tol", type=float, default=1e-8)
    ap.add_argument("--kernel-reps", type=int, default=200)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-graphs", action="store_true", help="launch Krylov iterations kernel by kernel")
    ap.add_argument("--trace-summary", metavar="CSV", help="summarise a rocprofv3 kernel trace and exit")
    ap.add_argument("--counter-summary", metavar="CSV", help="summarise a rocprofv3 --pmc CSV and exit")
    ap.add_argument("--solver-workloads", type=int, metavar="LEVEL",
                    help="run the isolated solver workloads on a uniform mesh of this level and exit")
    ap.add_argument("--nrhs", type=int, default=32)
    ap.add_argument("--no-adjoint", action="store_true", help="skip the adjoint/forward timing")
    ap.add_argument("--adjoint-level", type=int, default=4)
    ap.add_argument("--adjoint-spinup", type=int, default=1000)
    ap.add_argument("--adjoint-window", type=int, default=200)
    ap.add_argument("--adjoint-every", type=int, default=20)
    ap.add_argument("--roofline-only", action="store_true",
                    help="time the roofline kernel alone (the launch the counter passes profile) and exit")
    args = ap.parse_args()
    if args.trace_summary:
        return trace_summary(args.trace_summary, args.kernel_reps)
    if args.counter_summary:
        return counter_summary(args.counter_summary, args.kernel_reps)
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU (the path's sharding is a design note only); "
              f"--gpus {args.gpus} is not supported", file=sys.stderr)
        return 2
    if args.solver_workloads is not None:
        return solver_workloads(args.solver_workloads, args.nrhs, args.tol, args.kernel_reps)
    if args.roofline_only:
        dom, ks = roofline_measurement(args.kernel_reps, args.tol)
        out = {"roofline": roofline_object(dom),
               "kernels": {k: {"us": round(v["us"], 3), "bytes": v["bytes"]} for k, v in ks.items()}}
        print(json.dumps(out), flush=True)
        return 0
    if args.steps < 1 or args.warmup < 0:
        print("bench.py: --steps must be >= 1 and --warmup >= 0", file=sys.stderr)
        return 2
    if args.no_graphs:
        os.environ["CFA_KRYLOV_GRAPHS"] = "0"
    import tempfile
    import cfa

    base = WORKLOADS[args.workload]
    adaptive = "--amr-interval" in base
    spinup = args.spinup if args.spinup is not None else (1000 if adaptive else 0)
    first = spinup + args.warmup + 1                 # first timed step
    last = spinup + args.warmup + args.steps         # last timed step
    # untimed tail: a remesh launched inside the window is only allowed when its join falls
    # before the last step of the run
    tail = int(base[base.index("--amr-overlap") + 1]) + 1 if adaptive else 0
    total = last + tail
    dt = float(base[base.index("--dt") + 1])
    tmp = tempfile.mkdtemp(prefix="cfa_bench_")
    common = base + ["--steps", str(total), "--mom-tol", str(args.tol), "--ppe-tol", str(args.tol), "--quiet",
                     "--no-csv", "--prefix", os.path.join(tmp, "run")]
    hist, summary = {}, {}
    kernels = {}
    for prec in ("fp64", "mixed"):
        flags = common + ["--precision", prec]
        if prec == "fp64":
            flags += ["--dump-state", str(first - 1), "--kernel-timing", str(args.kernel_reps)]
        rc = cfa.run(flags)
        if rc != 0:
            print(f"bench.py: {prec} run failed with status {rc}", file=sys.stderr)
            return 1
        hist[prec] = cfa.last_history()
        summary[prec] = window_summary(hist[prec], first, last)
        if prec == "fp64":
            kernels = cfa.last_kernels()
    mx, f64 = summary["mixed"], summary["fp64"]
    events = remesh_events(common, total, first, last)
    line = {
        "metric": "steps/s at 1e-8 rel. residual, FP32-precond vs FP64 GPU; adjoint/forward time",
        "value": mx["value"],
        "unit": "steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": mx["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64 Krylov, residuals, coarse solve and forces; f32 ILU(0) factors and SA-AMG cycle",
        "data": "synthetic (deterministic: the reference's inflow set-up, no seeds)",
        "config": {
            "workload": f"{args.workload}: {DESCRIPTION[args.workload]}, tol {args.tol:g}; untimed spin-up of "
                        f"{spinup} steps and warm-up of {args.warmup}, timed steps {first}-{last}, untimed tail "
                        f"of {tail}; value = FP32-preconditioned run, fp64 = the same window in FP64",
            "precision": "mixed",
            "spinup_steps": spinup,
            "timed_steps": [first, last],
            "remesh_in_window": events,
            "nodes_range": mx["nodes_range"],
            "elements_range": mx["elements_range"],
            "mean_iters_u_v_p": mx["mean_iters_u_v_p"],
            "parallelism": "1 GPU",
        },
        "fp64": {k: f64[k] for k in ("value", "ms_per_step", "nodes_range", "elements_range", "mean_iters_u_v_p")},
        "mixed_precision": dict(force_deviation(hist["fp64"], hist["mixed"], first, last),
                                speedup_vs_fp64=round(mx["value"] / f64["value"], 3)),
        "roofline": None,
        "kernels": {k: {"us": round(v["us"], 3), "GB/s": round(v["bytes"] / (v["us"] * 1e3), 1)}
                    for k, v in kernels.items()},
        "cpu_baseline": None,
    }
    if not args.no_cpu_baseline:
        line["cpu_baseline"] = cpu_baseline(os.path.join(tmp, "run_state.bin"), dt, args.tol)
    dom, _ = roofline_measurement(args.kernel_reps, args.tol)
    line["roofline"] = roofline_object(dom)
    if not args.no_adjoint:
        line["adjoint"] = adjoint_cost(args.adjoint_level, args.adjoint_spinup, args.adjoint_window,
                                       args.adjoint_every, args.tol)
    print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
