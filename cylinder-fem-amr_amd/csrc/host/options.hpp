// Command-line options of the driver: every flag and default of src/cfa/options.cuh:8-289, plus
// solver tolerances and the preconditioner precision, whose defaults reproduce the reference's
// hard-coded values (momentum 1e-7, pressure 1e-7, re-projection 1e-8; FP64).
#pragma once

#include <string>

#include "amr.hpp"

namespace cfa {

struct RunOptions {
    int num_steps = 100;
    double dt = Sim::DT_DEFAULT;
    int amr_interval = 10;
    int init_amr = 2;
    int print_interval = 10;
    int vtk_interval = 0;
    double stats_window = 2.0;
    bool rotational = true;
    double lookahead = 1.0;
    bool bdf2 = true;
    double supg_tau_time = 1.0e-3;
    bool amg = true;
    bool momentum_ilu = true;
    bool skew_convection = true;
    int amg_rebuild = 50;
    int amr_overlap = 5;
    int amg_remap = 0;
    int uniform_level = 0;
    bool mms = false;
    bool case_2d3 = false;
    bool relative_tol = true;
    bool rotational_from_dp = true;
    int remesh_projection = 4;
    bool outlet_dirichlet = false;
    bool reproject_shift = false;
    int reproject_test = 0;
    double pressure_stab = 0.3;
    int dump_systems = 0;
    std::string prefix = "cylinder_fem_amr";
    AmrParams amr;
    // options beyond the reference's flags (defaults = its hard-coded behaviour)
    double mom_tol = 1.0e-7;    // momentum BiCGSTAB (src/cylinder_fem_amr.cu:358)
    double ppe_tol = 1.0e-7;    // pressure PCG (src/cylinder_fem_amr.cu:457)
    double proj_tol = 1.0e-8;   // re-projection PCG (src/cfa/device.cuh:388,449)
    bool mixed = false;         // FP32 ILU factors and AMG cycle inside the FP64 Krylov solvers
    bool quiet = false;         // no per-step console lines
    bool write_csv = true;
    int kernel_timing = 0;      // > 0: time the main kernels this many launches each on the final state
    int dump_state = 0;         // > 0: write the mesh and solution state at the end of this step
};

void print_usage(const char* prog);
// false: exit successfully (--help); invalid options abort with EXIT_FAILURE
bool parse_options(int argc, char** argv, RunOptions& opt);

}  // namespace cfa
