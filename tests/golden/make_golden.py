"""Copies the reference's stored run outputs used as golden data into tests/golden/.

The files are data written by the reference's own runs (forces histories, MMS error tables);
they are read here from a checkout of the reference and committed as fixtures, so the tests never
need the reference itself. Usage: python tests/golden/make_golden.py /path/to/reference
"""
import os
import shutil
import sys

RUNS = {
    # MMS convergence runs (src/cylinder_fem_amr.cu --mms), runs/paper/mms_L*/cmd.txt
    "mms_L1": ["cylinder_fem_amr_forces.csv", "cylinder_fem_amr_mms.csv", "cmd.txt"],
    "mms_L2": ["cylinder_fem_amr_forces.csv", "cylinder_fem_amr_mms.csv", "cmd.txt"],
    "mms_L3": ["cylinder_fem_amr_forces.csv", "cylinder_fem_amr_mms.csv", "cmd.txt"],
    "mms_L4": ["cylinder_fem_amr_mms.csv", "cmd.txt"],
    "mms_L5": ["cylinder_fem_amr_mms.csv", "cmd.txt"],
}
# adaptive DFG 2D-2 run (levels 2-5, dt 5e-4): only its first 2000 steps are kept
ADAPTIVE = ("d2_A2_dt05", 2000)
# uniform level-3 DFG 2D-2 run to t = 8 s (runs/v14/timing/U3_r1): the force history from
# t >= 5.9 s (its statistics window was the last 1.5 s) and the statistics the run wrote
STATS_RUN = (os.path.join("v14", "timing", "U3_r1"), 5.9)
# --outlet dirichlet runs (runs/review, source = the reference's src/): the first 200 steps of a
# uniform level-4 DFG 2D-2 and 2D-3 run, and of the adaptive A1 2D-2 run (levels 2-4)
OUTLET_RUNS = (("d2_U4_outdir", 200), ("d3_U4_outdir", 200), ("d2_A1_outdir", 200))


def main(ref):
    here = os.path.dirname(os.path.abspath(__file__))
    for run, files in RUNS.items():
        for f in files:
            src = os.path.join(ref, "runs", "paper", run, f)
            dst = os.path.join(here, f"{run}_{f.replace('cylinder_fem_amr_', '')}")
            shutil.copyfile(src, dst)
    run, rows = ADAPTIVE
    src = os.path.join(ref, "runs", "paper", run, "cylinder_fem_amr_forces.csv")
    with open(src) as fi, open(os.path.join(here, f"{run}_forces_first{rows}.csv"), "w") as fo:
        for k, line in enumerate(fi):
            if k > rows:
                break
            fo.write(line)
    shutil.copyfile(os.path.join(ref, "runs", "paper", run, "cmd.txt"), os.path.join(here, f"{run}_cmd.txt"))
    # the same MMS U3 case run by a second binary of the reference (its run-to-run noise)
    shutil.copyfile(os.path.join(ref, "runs", "review", "mms_L3_dt2", "cylinder_fem_amr_mms.csv"),
                    os.path.join(here, "mms_L3_dt2_mms.csv"))
    run, t_from = STATS_RUN
    src = os.path.join(ref, "runs", run, "cylinder_fem_amr_forces.csv")
    with open(src) as fi, open(os.path.join(here, "U3_r1_forces_tail.csv"), "w") as fo:
        header = fi.readline().strip().split(",")
        cols = [header.index(c) for c in ("time", "C_D", "C_L", "delta_p")]
        fo.write("time,C_D,C_L,delta_p\n")
        for line in fi:
            v = line.strip().split(",")
            if float(v[1]) >= t_from:
                fo.write(",".join(v[c] for c in cols) + "\n")
    shutil.copyfile(os.path.join(ref, "runs", run, "cylinder_fem_amr_stats.csv"), os.path.join(here, "U3_r1_stats.csv"))
    for run, rows in OUTLET_RUNS:
        src = os.path.join(ref, "runs", "review", run, "cylinder_fem_amr_forces.csv")
        with open(src) as fi, open(os.path.join(here, f"{run}_forces_first{rows}.csv"), "w") as fo:
            for k, line in enumerate(fi):
                if k > rows:
                    break
                fo.write(line)
        shutil.copyfile(os.path.join(ref, "runs", "review", run, "cmd.txt"), os.path.join(here, f"{run}_cmd.txt"))


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
