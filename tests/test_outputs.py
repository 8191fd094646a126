"""End-of-run DFG 2D-2 statistics against the reference's stored run: the force history of
runs/v14/timing/U3_r1 (uniform level 3, dt 1e-3, to t = 8 s, --stats-window 1.5 as set by
scripts/timing.sh:30; the rows from t = 5.9 s cover the window) fed to the program's statistics,
compared with the _stats.csv that run wrote. The history is stored with 10 significant digits, so the values agree to ~1e-9."""
import ctypes

import numpy as np

import cfa
from conftest import read_csv


def test_dfg_statistics_match_reference(golden):
    _, host = cfa.load()
    h = read_csv(golden / "U3_r1_forces_tail.csv")
    ref = {row[0]: (float(row[1]), float(row[2]))
           for row in (line.strip().split(",") for line in open(golden / "U3_r1_stats.csv").readlines()[1:])}
    d = ctypes.POINTER(ctypes.c_double)
    host.cfa_force_statistics.argtypes = [ctypes.c_int, d, d, d, d, ctypes.c_double, d]
    host.cfa_force_statistics.restype = ctypes.c_int
    cols = [np.ascontiguousarray(h[c], dtype=float) for c in ("time", "C_D", "C_L", "delta_p")]
    out = np.zeros(13)
    assert host.cfa_force_statistics(len(cols[0]), *(c.ctypes.data_as(d) for c in cols), 1.5,
                                     out.ctypes.data_as(d)) == 0
    assert int(out[0]) == int(ref["periods"][0]) == 2
    for k, key in enumerate(("cd_max", "cd_min", "cl_max", "cl_min", "St", "dp_half")):
        value, spread = ref[key]
        assert abs(out[1 + k] - value) <= 1e-8 * abs(value), key
        assert abs(out[7 + k] - spread) <= 1e-8, key
