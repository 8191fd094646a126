// ORACLE (test infrastructure only): C entry points for the Python tests (ctypes). Arrays are
// caller-owned host buffers; each function restates one reference routine (see the C++ side).
#include <cstring>

#ifdef _OPENMP
#include <omp.h>
#endif

#include "cfa_oracle.hpp"

using namespace oracle;

extern "C" {

// Fixed-mesh run. rows: steps x 13 doubles (step, t, C_D, C_L, dP, nodes, elements, p_front,
// p_back, C_D_surface, C_L_surface, it_u + 1000 it_v, it_p); mms: 9 doubles (6 norms, cyl, probe,
// h_max). With a prefix the forces/MMS CSV files are written, and with dump_systems > 0 the
// systems of that step (<prefix>_momentum.bin, <prefix>_pressure.bin). Returns the number of rows.
int oracle_run_fixed(int steps, double dt, int uniform_level, int min_level, int mms, double mom_tol,
                     double ppe_tol, const char* prefix, int dump_systems, double* rows, double* mms_out,
                     double* wall_s) {
    RunConfig c;
    c.steps = steps;
    c.dt = dt;
    c.uniform_level = uniform_level;
    c.min_level = min_level;
    c.mms = mms != 0;
    c.mom_tol = mom_tol;
    c.ppe_tol = ppe_tol;
    c.quiet = true;
    if (prefix) c.prefix = prefix;
    c.dump_systems = dump_systems;
    RunResult r = run_fixed(c);
    for (size_t k = 0; k < r.rows.size(); ++k) {
        const StepRow& s = r.rows[k];
        double* o = rows + 13 * k;
        o[0] = s.step;  o[1] = s.t;  o[2] = s.cd;  o[3] = s.cl;  o[4] = s.dp;  o[5] = s.nodes;
        o[6] = s.elems;  o[7] = s.p_front;  o[8] = s.p_back;  o[9] = s.cd_s;  o[10] = s.cl_s;
        o[11] = s.it_u + 1000.0 * s.it_v;  o[12] = s.it_p;
    }
    if (mms_out) {
        for (int k = 0; k < 6; ++k) mms_out[k] = r.mms[k];
        mms_out[6] = r.cyl_max;  mms_out[7] = r.probe_max;  mms_out[8] = r.h_max;
    }
    if (wall_s) *wall_s = r.wall_s;
    return (int)r.rows.size();
}

// Fixed-mesh DFG flow (no MMS) from the initial condition with the step's viscosity and inflow
// peak as parameters, both solvers at relative tolerance tol: rows as oracle_run_fixed.
int oracle_run_params(int steps, double dt, int uniform_level, double nu, double u_max, double tol, double* rows) {
    RunConfig c;
    c.steps = steps;
    c.dt = dt;
    c.uniform_level = uniform_level;
    c.min_level = 0;
    c.nu = nu;
    c.u_max = u_max;
    c.mom_tol = tol;
    c.ppe_tol = tol;
    c.quiet = true;
    RunResult r = run_fixed(c);
    for (size_t k = 0; k < r.rows.size(); ++k) {
        const StepRow& s = r.rows[k];
        double* o = rows + 13 * k;
        o[0] = s.step;  o[1] = s.t;  o[2] = s.cd;  o[3] = s.cl;  o[4] = s.dp;  o[5] = s.nodes;
        o[6] = s.elems;  o[7] = s.p_front;  o[8] = s.p_back;  o[9] = s.cd_s;  o[10] = s.cl_s;
        o[11] = s.it_u + 1000.0 * s.it_v;  o[12] = s.it_p;
    }
    return (int)r.rows.size();
}

// Fixed-mesh DFG run with the outlet parabola (--outlet dirichlet), both solvers at relative
// tolerance tol, case 2D-2 or (case_2d3) 2D-3: rows as oracle_run_fixed.
int oracle_run_outlet_dirichlet(int steps, double dt, int uniform_level, int min_level, int case_2d3, double tol,
                                double* rows) {
    RunConfig c;
    c.steps = steps;
    c.dt = dt;
    c.uniform_level = uniform_level;
    c.min_level = min_level;
    c.case_2d3 = case_2d3 != 0;
    c.outlet_dirichlet = true;
    c.mom_tol = tol;
    c.ppe_tol = tol;
    c.quiet = true;
    RunResult r = run_fixed(c);
    for (size_t k = 0; k < r.rows.size(); ++k) {
        const StepRow& s = r.rows[k];
        double* o = rows + 13 * k;
        o[0] = s.step;  o[1] = s.t;  o[2] = s.cd;  o[3] = s.cl;  o[4] = s.dp;  o[5] = s.nodes;
        o[6] = s.elems;  o[7] = s.p_front;  o[8] = s.p_back;  o[9] = s.cd_s;  o[10] = s.cl_s;
        o[11] = s.it_u + 1000.0 * s.it_v;  o[12] = s.it_p;
    }
    return (int)r.rows.size();
}

// Continues a run on a fixed mesh from a --dump-state file for `steps` steps (DFG inflow, 2D-2,
// default flags), both solvers at relative tolerance tol, on `threads` OpenMP threads (<= 0: the
// runtime's default): rows as oracle_run_fixed, wall seconds of the time loop in *wall_s.
int oracle_run_state(const char* path, int steps, double dt, double tol, int threads, double* rows, double* wall_s) {
#ifdef _OPENMP
    const int prev = omp_get_max_threads();
    if (threads > 0) omp_set_num_threads(threads);
#endif
    RunConfig c;
    c.steps = steps;
    c.dt = dt;
    c.mom_tol = tol;
    c.ppe_tol = tol;
    c.quiet = true;
    RunResult r = run_state_file(c, path);
#ifdef _OPENMP
    omp_set_num_threads(prev);
#endif
    for (size_t k = 0; k < r.rows.size(); ++k) {
        const StepRow& s = r.rows[k];
        double* o = rows + 13 * k;
        o[0] = s.step;  o[1] = s.t;  o[2] = s.cd;  o[3] = s.cl;  o[4] = s.dp;  o[5] = s.nodes;
        o[6] = s.elems;  o[7] = s.p_front;  o[8] = s.p_back;  o[9] = s.cd_s;  o[10] = s.cl_s;
        o[11] = s.it_u + 1000.0 * s.it_v;  o[12] = s.it_p;
    }
    if (wall_s) *wall_s = r.wall_s;
    return (int)r.rows.size();
}

}  // extern "C"
