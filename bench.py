#!/usr/bin/env python3
"""Benchmark of the per-step projection solve on one MI355X.

Workload (default): the adaptive obstacle flow of the reference's runs/paper/d2_A2_dt05 —
parabolic inflow (mean 1, Re 100), NVB levels 2-5 with coarsening and look-ahead, remesh every
20 steps (overlapped by 10), dt 5e-4, FP64 throughout — with the momentum BiCGSTAB and pressure
PCG run to a relative residual of 1e-8. A step is one time step (momentum predictor, pressure
increment, correction, forces, and the remesh work that falls on it). `--warmup W` steps run
first, untimed; then exactly `--steps K` steps are timed. The program synchronises the device
at the end of every step (the forces are read back), and the timed region runs from the end of
step W to the end of step W + K on the host's steady clock.

After the timed run the main kernels are launched alone on the final mesh and state
(--kernel-timing) and timed with HIP events on the solver stream ("kernels"). The roofline object
reports the dominant kernel family's level-0 member, the damped-Jacobi sweep of the SA-AMG
V-cycle, measured the same way on a fixed shape (the pressure matrix of the uniform level-4
mesh, 115,648 rows, the size of the adaptive run's late meshes), so that the hardware-counter
passes can profile exactly that launch (--roofline-only). The CPU baseline is the oracle (the
serial FP64 CPU restatement of the same step) on a bounded sample, run on this host.

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
DOMINANT = "k_csr_op<jacobi> level 0"
ROOFLINE_LEVEL = 4
TRAFFIC_FILE = ROOT / "profiles" / "jacobi_level0_uniform_L4_counters.txt"

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


def cpu_baseline(seconds_target: float = 15.0) -> dict:
    """The oracle on a bounded sample: fixed uniform level-3 mesh (29,152 nodes, the size of the
    adaptive run's early meshes), DFG inflow, dt 5e-4, tolerances 1e-8, serial FP64."""
    sys.path.insert(0, str(ROOT / "oracle"))
    import cfa_oracle
    steps = 5
    rows, _, wall = cfa_oracle.run_fixed(steps, 5e-4, 3, 2, mms=False, mom_tol=1e-8, ppe_tol=1e-8)
    per_step = wall / max(len(rows), 1)
    steps = max(5, min(400, int(seconds_target / max(per_step, 1e-6))))
    rows, _, wall = cfa_oracle.run_fixed(steps, 5e-4, 3, 2, mms=False, mom_tol=1e-8, ppe_tol=1e-8)
    return {"value": round(len(rows) / wall, 3), "unit": "steps/s", "cores": 1, "kind": "port",
            "sample": f"oracle (serial FP64 CPU restatement) on a fixed uniform level-3 mesh (29152 nodes), "
                      f"DFG inflow, dt 5e-4, tol 1e-8, {len(rows)} steps in {wall:.1f} s"}


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
    dispatch, each scaled by the same counter over a 64 MiB device copy)."""
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
    """The level-0 smoother of the SA-AMG V-cycle on the uniform level-4 pressure matrix, launched
    `reps` times between HIP events on the solver stream (cfa_workload_pressure timings)."""
    import cfa
    cfa.workload_pressure(ROOFLINE_LEVEL, 1, tol, max(reps, 1))
    k = cfa.last_kernels()
    return k[DOMINANT], k


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


def mixed_comparison(flags: list[str], hist64, warmup: int, total: int) -> dict:
    """The same workload with FP32 ILU factors and FP32 SA-AMG cycle inside the FP64 Krylov solvers:
    steps/s over the same timed steps and the force deviation from the FP64 run, step by step
    while the two runs still have the same mesh (remesh decisions are made from the solution, so
    the adaptive trajectories may part)."""
    import cfa
    import numpy as np
    mixed_flags = [f if f != "fp64" else "mixed" for f in flags]
    if cfa.run(mixed_flags) != 0:
        return {"error": "mixed-precision run failed"}
    h = cfa.last_history()
    col = {c: i for i, c in enumerate(cfa.HISTORY_COLUMNS)}
    t0 = h[warmup - 1, col["elapsed_s"]] if warmup > 0 else 0.0
    timed = h[total - 1, col["elapsed_s"]] - t0
    same = (h[:, col["nodes"]] == hist64[:, col["nodes"]]) & (h[:, col["elements"]] == hist64[:, col["elements"]])
    upto = int(np.argmin(same)) if not same.all() else len(same)
    d_cd = np.abs(h[:upto, col["C_D"]] - hist64[:upto, col["C_D"]])
    d_cl = np.abs(h[:upto, col["C_L"]] - hist64[:upto, col["C_L"]])
    return {
        "value": round((total - warmup) / timed, 3),
        "unit": "steps/s",
        "ms_per_step": round(1000.0 * timed / (total - warmup), 4),
        "steps_on_identical_meshes": upto,
        "max_abs_dC_D": float(d_cd.max()) if upto else None,
        "max_abs_dC_L": float(d_cl.max()) if upto else None,
        # the same after the impulsive start (C_D near 276 at step 1, where a 1e-10 relative
        # difference is already 3e-8 absolute)
        "max_abs_dC_D_from_step_100": float(d_cd[99:].max()) if upto > 99 else None,
        "max_abs_dC_L_from_step_100": float(d_cl[99:].max()) if upto > 99 else None,
        "mean_iters_u_v_p": [round(float(h[warmup:total, col[c]].mean()), 2) for c in ("it_u", "it_v", "it_p")],
    }


def adjoint_cost(level: int, spinup: int, window: int, every: int, tol: float) -> dict:
    """Discrete adjoint of the window functional J = mean C_D over `window` steps on the frozen
    uniform level-`level` mesh after a `spinup`-step FP64 run from the DFG initial condition,
    theta = (nu, U_max): time of the forward window, time of the gradient (forward pass with
    checkpoints every `every` steps, segment recomputation and reverse sweep), and their ratio."""
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
    J2, g, _, st = a.gradient(theta, spinup + 1, window, state, every)
    grad_s = time.perf_counter() - t
    a.close()
    return {
        "workload": f"uniform level-{level} mesh ({a.n} nodes, {a.ne} elements), {spinup}-step spin-up, "
                    f"{window}-step window, checkpoints every {every}, tol {tol:g}",
        "J": J, "dJ_dnu": float(g[0]), "dJ_dUmax": float(g[1]),
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
    ap.add_argument("--steps", type=int, default=2000)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--workload", choices=sorted(WORKLOADS), default="adaptive")
    ap.add_argument("--precision", choices=["fp64", "mixed"], default="fp64")
    ap.add_argument("--tol", type=float, default=1e-8)
    ap.add_argument("--kernel-reps", type=int, default=200)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-graphs", action="store_true", help="launch Krylov iterations kernel by kernel")
    ap.add_argument("--trace-summary", metavar="CSV", help="summarise a rocprofv3 kernel trace and exit")
    ap.add_argument("--counter-summary", metavar="CSV", help="summarise a rocprofv3 --pmc CSV and exit")
    ap.add_argument("--solver-workloads", type=int, metavar="LEVEL",
                    help="run the isolated solver workloads on a uniform mesh of this level and exit")
    ap.add_argument("--nrhs", type=int, default=32)
    ap.add_argument("--no-mixed", action="store_true", help="skip the mixed-precision run of the same workload")
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
    if args.solver_workloads is not None:
        return solver_workloads(args.solver_workloads, args.nrhs, args.tol, args.kernel_reps)
    if args.roofline_only:
        dom, ks = roofline_measurement(args.kernel_reps, args.tol)
        out = {"roofline": roofline_object(dom),
               "kernels": {k: {"us": round(v["us"], 3), "bytes": v["bytes"]} for k, v in ks.items()}}
        print(json.dumps(out), flush=True)
        return 0
    if args.gpus != 1:
        print(f"bench.py: the solver runs on one GPU; --gpus {args.gpus} is not supported", file=sys.stderr)
        return 2
    if args.steps < 1 or args.warmup < 0:
        print("bench.py: --steps must be >= 1 and --warmup >= 0", file=sys.stderr)
        return 2
    if args.no_graphs:
        os.environ["CFA_KRYLOV_GRAPHS"] = "0"
    import cfa

    total = args.warmup + args.steps
    flags = WORKLOADS[args.workload] + [
        "--steps", str(total), "--mom-tol", str(args.tol), "--ppe-tol", str(args.tol), "--precision", args.precision,
        "--quiet", "--no-csv", "--kernel-timing", str(args.kernel_reps)]
    rc = cfa.run(flags)
    if rc != 0:
        print(f"bench.py: run failed with status {rc}", file=sys.stderr)
        return 1
    hist = cfa.last_history()
    col = {c: i for i, c in enumerate(cfa.HISTORY_COLUMNS)}
    t_end = hist[total - 1, col["elapsed_s"]]
    t_start = hist[args.warmup - 1, col["elapsed_s"]] if args.warmup > 0 else 0.0
    timed = t_end - t_start
    window = hist[args.warmup:total]
    kernels = cfa.last_kernels()
    line = {
        "metric": "steps/s at 1e-8 rel. residual, FP32-precond vs FP64 GPU; adjoint/forward time",
        "value": round(args.steps / timed, 3),
        "unit": "steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(1000.0 * timed / args.steps, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if args.precision == "fp64" else "f64 (FP32 preconditioner factors)",
        "data": "synthetic (deterministic: the reference's inflow / manufactured-solution set-up, no seeds)",
        "config": {
            "workload": f"{args.workload}: {DESCRIPTION[args.workload]}, tol {args.tol:g}, steps "
                        f"{args.warmup + 1}-{total}",
            "precision": args.precision,
            "nodes_range": [int(window[:, col["nodes"]].min()), int(window[:, col["nodes"]].max())],
            "elements_range": [int(window[:, col["elements"]].min()), int(window[:, col["elements"]].max())],
            "mean_iters_u_v_p": [round(float(window[:, col[c]].mean()), 2) for c in ("it_u", "it_v", "it_p")],
            "parallelism": "1 GPU",
        },
        "roofline": None,
        "kernels": {k: {"us": round(v["us"], 3), "GB/s": round(v["bytes"] / (v["us"] * 1e3), 1)}
                    for k, v in kernels.items()},
        "cpu_baseline": None if args.no_cpu_baseline else cpu_baseline(),
    }
    dom, _ = roofline_measurement(args.kernel_reps, args.tol)
    line["roofline"] = roofline_object(dom)
    if not args.no_mixed and args.precision == "fp64":
        line["mixed_precision"] = mixed_comparison(flags, hist, args.warmup, total)
        if "value" in line["mixed_precision"]:
            line["mixed_precision"]["speedup_vs_fp64"] = round(line["mixed_precision"]["value"] / line["value"], 3)
    if not args.no_adjoint:
        line["adjoint"] = adjoint_cost(args.adjoint_level, args.adjoint_spinup, args.adjoint_window,
                                       args.adjoint_every, args.tol)
    print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
