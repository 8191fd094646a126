// Option parsing with the reference's flags, value checks and messages (src/cfa/options.cuh:54-289).
#include "options.hpp"

#include <cstdio>
#include <cstdlib>
#include <iostream>

namespace cfa {

void print_usage(const char* prog) {
    RunOptions d;
    std::cout
        << "Usage: " << prog << " [options]\n"
        << "       " << prog << " [num_steps] [amr_interval] [max_amr_level] [amr_tol_vort]   (legacy)\n\n"
        << "  --steps N           number of time steps                          (" << d.num_steps << ")\n"
        << "  --dt S              time step [s]                                 (" << d.dt << ")\n"
        << "  --amr-interval N    adapt every N steps, 0 disables AMR           (" << d.amr_interval << ")\n"
        << "  --init-amr N        adaptation cycles on the initial condition    (" << d.init_amr << ")\n"
        << "  --max-level L       max AMR levels above the base mesh            (" << d.amr.max_level << ")\n"
        << "  --amr-tol-vort U    disturbance-vorticity criterion tolerance [m/s], 0 = off   (" << d.amr.tol_vort << ")\n"
        << "  --amr-tol-interp U  disturbance interpolation-error tolerance [m/s], 0 = off (" << d.amr.tol_interp << ")\n"
        << "  --coarsen-safety F  coarsen where merged score < F (refine at 1)  (" << d.amr.coarsen_safety << ")\n"
        << "  --lookahead F       advective look-ahead, in AMR intervals, 0=off (" << d.lookahead << ")\n"
        << "  --buffer N          refinement buffer layers                      (" << d.amr.buffer_layers << ")\n"
        << "  --max-elems N       element budget                                (" << d.amr.max_elems << ")\n"
        << "  --rotational-div increment|velocity     div u* in the rotational term from the pressure\n"
        << "                      increment (increment) or from u* directly (velocity)\n"
        << "  --case 2d2|2d3      DFG benchmark: periodic Re = 100 (2d2), or inflow 1.5 sin(pi t/8)\n"
        << "                      from rest over 0 <= t <= 8 s (2d3; use --steps 8000)\n"
        << "  --amr-tol-scale inflow|fixed            2d3: tolerances follow the inflow magnitude (inflow)\n"
        << "  --mms               manufactured-solution verification on the fixed (uniform) mesh;\n"
        << "                      prints the L2/H1 errors at the final time (<prefix>_mms.csv)\n"
        << "  --uniform-level L   uniformly refine the base mesh L levels and disable AMR (0)\n"
        << "  --min-level L       AMR never goes below level L (start from a uniform level-L mesh) (" << d.amr.min_level << ")\n"
        << "  --keep-radius R     never coarsen within R [m] of the cylinder centre, 0 = off (" << d.amr.keep_radius << ")\n"
        << "  --amr-overlap K     build each new mesh on the host while the GPU takes K more steps on\n"
        << "                      the old one, then transfer the state of that step; K < interval, 0 = off (" << d.amr_overlap << ")\n"
        << "  --amr-tol-radius R  tolerances grow as (d/R)^P beyond distance d > R [m] from the\n"
        << "                      cylinder centre, 0 = uniform tolerances (" << d.amr.tol_radius << ")\n"
        << "  --amr-tol-power P   exponent P of that growth (" << d.amr.tol_power << ")\n"
        << "  --print-interval N  console diagnostics every N steps             (" << d.print_interval << ")\n"
        << "  --vtk-interval N    write a VTK snapshot every N steps, 0 = final (" << d.vtk_interval << ")\n"
        << "  --pressure-update rotational|standard   incremental projection variant (rotational)\n"
        << "  --time-scheme bdf2|euler                time discretization (bdf2)\n"
        << "  --pressure-stab F   orthogonal-subscale pressure stabilization factor, 0 = off (0.3)\n"
        << "  --supg-tau-time T   fixed time scale in the SUPG tau [s], 0 = use dt (0.001)\n"
        << "  --pressure-precond amg|jacobi           PPE preconditioner (amg)\n"
        << "  --momentum-precond ilu|jacobi           momentum preconditioner: multicolour ILU(0) (ilu)\n"
        << "  --convection-form skew|convective       Galerkin convection operator (skew)\n"
        << "  --amg-rebuild N     rebuild the AMG hierarchy at least every N steps (50)\n"
        << "  --amg-remap N       after a remesh, carry the AMG hierarchy over to the new mesh for up\n"
        << "                      to N remeshes in a row before a full rebuild; 0 = always rebuild (0)\n"
        << "  --remesh-projection scheme|scheme-current|both|current|none\n"
        << "                      velocity re-projection after a remesh (scheme)\n"
        << "  --reproject-shift on|off                after a scheme re-projection, the next pressure step\n"
        << "                      uses S(p^n + phi) with phi the potential of the u^n re-projection (off)\n"
        << "  --reproject-test N  fixed mesh: re-project every N steps without remeshing (0 = off)\n"
        << "  --outlet pressure|dirichlet             outlet: p = 0 with nu du/dn = 0 (pressure), or the\n"
        << "                      inflow parabola, pressure pinned at one node (pressure)\n"
        << "  --stats-window S    time window for force statistics [s]          (" << d.stats_window << ")\n"
        << "  --dump-systems N    write the momentum and pressure systems of step N to\n"
        << "                      <prefix>_momentum.bin and <prefix>_pressure.bin (solver studies)\n"
        << "  --prefix NAME       output file prefix                            (" << d.prefix << ")\n"
        << "  --mom-tol T         momentum BiCGSTAB relative tolerance          (" << d.mom_tol << ")\n"
        << "  --ppe-tol T         pressure PCG relative tolerance               (" << d.ppe_tol << ")\n"
        << "  --proj-tol T        re-projection PCG relative tolerance          (" << d.proj_tol << ")\n"
        << "  --precision fp64|mixed                  ILU factors and AMG cycle in FP64 or FP32 (fp64)\n"
        << "  --quiet             no per-step console lines\n"
        << "  --no-csv            write no CSV files (the history stays available in-process)\n"
        << "  --kernel-timing R   after the run, time the main kernels R launches each (HIP events)\n"
        << "  --dump-state N      write the mesh and the solution state at the end of step N to\n"
        << "                      <prefix>_state.bin (restarts and CPU comparisons)\n";
}

static double parse_double(const char* s, const char* name) {
    char* end = nullptr;
    const double val = std::strtod(s, &end);
    if (end == s || *end != '\0') {
        std::fprintf(stderr, "Invalid value '%s' for %s\n", s, name);
        std::exit(EXIT_FAILURE);
    }
    return val;
}

static int parse_int(const char* s, const char* name) {
    char* end = nullptr;
    const long val = std::strtol(s, &end, 10);
    if (end == s || *end != '\0') {
        std::fprintf(stderr, "Invalid value '%s' for %s\n", s, name);
        std::exit(EXIT_FAILURE);
    }
    return (int)val;
}

bool parse_options(int argc, char** argv, RunOptions& opt) {
    int positional = 0;
    for (int i = 1; i < argc; ++i) {
        const std::string a = argv[i];
        auto value = [&]() -> const char* {
            if (i + 1 >= argc) {
                std::fprintf(stderr, "Missing value for %s\n", a.c_str());
                std::exit(EXIT_FAILURE);
            }
            return argv[++i];
        };
        // two-way choice: val must be one of on / off names
        auto choice = [&](const char* flag, const char* yes, const char* no, bool& target) {
            const std::string v = value();
            if (v == yes) target = true;
            else if (v == no) target = false;
            else {
                std::fprintf(stderr, "Invalid value '%s' for %s\n", v.c_str(), flag);
                std::exit(EXIT_FAILURE);
            }
        };
        if (a == "-h" || a == "--help") { print_usage(argv[0]); return false; }
        else if (a == "--steps") opt.num_steps = parse_int(value(), "--steps");
        else if (a == "--dt") opt.dt = parse_double(value(), "--dt");
        else if (a == "--amr-interval") opt.amr_interval = parse_int(value(), "--amr-interval");
        else if (a == "--init-amr") opt.init_amr = parse_int(value(), "--init-amr");
        else if (a == "--max-level") opt.amr.max_level = parse_int(value(), "--max-level");
        else if (a == "--amr-tol-vort") opt.amr.tol_vort = parse_double(value(), "--amr-tol-vort");
        else if (a == "--amr-tol-interp") opt.amr.tol_interp = parse_double(value(), "--amr-tol-interp");
        else if (a == "--lookahead") opt.lookahead = parse_double(value(), "--lookahead");
        else if (a == "--coarsen-safety") opt.amr.coarsen_safety = parse_double(value(), "--coarsen-safety");
        else if (a == "--buffer") opt.amr.buffer_layers = parse_int(value(), "--buffer");
        else if (a == "--max-elems") opt.amr.max_elems = parse_int(value(), "--max-elems");
        else if (a == "--print-interval") opt.print_interval = parse_int(value(), "--print-interval");
        else if (a == "--vtk-interval") opt.vtk_interval = parse_int(value(), "--vtk-interval");
        else if (a == "--stats-window") opt.stats_window = parse_double(value(), "--stats-window");
        else if (a == "--pressure-update") choice("--pressure-update", "rotational", "standard", opt.rotational);
        else if (a == "--time-scheme") choice("--time-scheme", "bdf2", "euler", opt.bdf2);
        else if (a == "--pressure-stab") opt.pressure_stab = parse_double(value(), "--pressure-stab");
        else if (a == "--supg-tau-time") opt.supg_tau_time = parse_double(value(), "--supg-tau-time");
        else if (a == "--amg-rebuild") opt.amg_rebuild = parse_int(value(), "--amg-rebuild");
        else if (a == "--amr-overlap") opt.amr_overlap = parse_int(value(), "--amr-overlap");
        else if (a == "--amg-remap") opt.amg_remap = parse_int(value(), "--amg-remap");
        else if (a == "--convection-form") choice("--convection-form", "skew", "convective", opt.skew_convection);
        else if (a == "--pressure-precond") choice("--pressure-precond", "amg", "jacobi", opt.amg);
        else if (a == "--momentum-precond") choice("--momentum-precond", "ilu", "jacobi", opt.momentum_ilu);
        else if (a == "--remesh-projection") {
            const std::string v = value();
            if (v == "none") opt.remesh_projection = 0;
            else if (v == "current") opt.remesh_projection = 1;
            else if (v == "both") opt.remesh_projection = 2;
            else if (v == "scheme-current") opt.remesh_projection = 3;
            else if (v == "scheme") opt.remesh_projection = 4;
            else {
                std::fprintf(stderr, "Invalid value '%s' for --remesh-projection\n", v.c_str());
                std::exit(EXIT_FAILURE);
            }
        }
        else if (a == "--reproject-shift") choice("--reproject-shift", "on", "off", opt.reproject_shift);
        else if (a == "--reproject-test") opt.reproject_test = parse_int(value(), "--reproject-test");
        else if (a == "--outlet") choice("--outlet", "dirichlet", "pressure", opt.outlet_dirichlet);
        else if (a == "--mms") opt.mms = true;
        else if (a == "--amr-tol-scale") choice("--amr-tol-scale", "inflow", "fixed", opt.relative_tol);
        else if (a == "--case") choice("--case", "2d3", "2d2", opt.case_2d3);
        else if (a == "--rotational-div") choice("--rotational-div", "increment", "velocity", opt.rotational_from_dp);
        else if (a == "--uniform-level") opt.uniform_level = parse_int(value(), "--uniform-level");
        else if (a == "--min-level") opt.amr.min_level = parse_int(value(), "--min-level");
        else if (a == "--keep-radius") opt.amr.keep_radius = parse_double(value(), "--keep-radius");
        else if (a == "--amr-tol-radius") opt.amr.tol_radius = parse_double(value(), "--amr-tol-radius");
        else if (a == "--amr-tol-power") opt.amr.tol_power = parse_double(value(), "--amr-tol-power");
        else if (a == "--dump-systems") opt.dump_systems = parse_int(value(), "--dump-systems");
        else if (a == "--prefix") opt.prefix = value();
        else if (a == "--mom-tol") opt.mom_tol = parse_double(value(), "--mom-tol");
        else if (a == "--ppe-tol") opt.ppe_tol = parse_double(value(), "--ppe-tol");
        else if (a == "--proj-tol") opt.proj_tol = parse_double(value(), "--proj-tol");
        else if (a == "--precision") choice("--precision", "mixed", "fp64", opt.mixed);
        else if (a == "--quiet") opt.quiet = true;
        else if (a == "--no-csv") opt.write_csv = false;
        else if (a == "--kernel-timing") opt.kernel_timing = parse_int(value(), "--kernel-timing");
        else if (a == "--dump-state") opt.dump_state = parse_int(value(), "--dump-state");
        else if (a.rfind("--", 0) == 0) {
            std::fprintf(stderr, "Unknown option %s (see --help)\n", a.c_str());
            std::exit(EXIT_FAILURE);
        } else {
            switch (positional++) {
                case 0: opt.num_steps = parse_int(a.c_str(), "num_steps"); break;
                case 1: opt.amr_interval = parse_int(a.c_str(), "amr_interval"); break;
                case 2: opt.amr.max_level = parse_int(a.c_str(), "max_amr_level"); break;
                case 3: opt.amr.tol_vort = parse_double(a.c_str(), "amr_tol_vort"); break;
                default:
                    std::fprintf(stderr, "Too many positional arguments (see --help)\n");
                    std::exit(EXIT_FAILURE);
            }
        }
    }
    auto require = [](bool cond, const char* msg) {
        if (!cond) {
            std::fprintf(stderr, "Invalid options: %s\n", msg);
            std::exit(EXIT_FAILURE);
        }
    };
    require(opt.num_steps > 0, "--steps must be > 0");
    require(opt.dt > 0.0, "--dt must be > 0");
    require(opt.kernel_timing >= 0, "--kernel-timing must be >= 0");
    require(opt.dump_state >= 0, "--dump-state must be >= 0");
    require(opt.amr_interval >= 0 && opt.init_amr >= 0, "AMR intervals must be >= 0");
    require(opt.amr.max_level >= 0 && opt.amr.max_level <= 12, "--max-level must be in [0, 12]");
    require(opt.amr.tol_vort > 0.0 || opt.amr.tol_interp > 0.0,
            "at least one of --amr-tol-vort / --amr-tol-interp must be > 0");
    require(opt.lookahead >= 0.0, "--lookahead must be >= 0");
    require(opt.pressure_stab >= 0.0, "--pressure-stab must be >= 0");
    require(opt.supg_tau_time >= 0.0, "--supg-tau-time must be >= 0");
    require(opt.amg_rebuild > 0, "--amg-rebuild must be > 0");
    require(opt.amg_remap >= 0, "--amg-remap must be >= 0");
    require(opt.amr_overlap >= 0, "--amr-overlap must be >= 0");
    require(opt.amr_interval == 0 || opt.amr_overlap < opt.amr_interval,
            "--amr-overlap must be smaller than --amr-interval (set --amr-overlap 0 for a synchronous remesh)");
    require(opt.amr.tol_radius >= 0.0 && opt.amr.tol_power >= 0.0, "--amr-tol-radius and --amr-tol-power must be >= 0");
    require(opt.uniform_level >= 0, "--uniform-level must be >= 0");
    require(opt.reproject_test >= 0, "--reproject-test must be >= 0");
    require(opt.reproject_test == 0 || opt.uniform_level > 0 || opt.mms,
            "--reproject-test needs a fixed mesh (--uniform-level or --mms)");
    require(!opt.reproject_shift || opt.remesh_projection >= 3, "--reproject-shift needs a scheme re-projection");
    require(!(opt.mms && opt.outlet_dirichlet), "--mms uses the default outlet (its exact pressure vanishes there)");
    require(opt.amr.min_level >= 0 && opt.amr.min_level <= opt.amr.max_level, "--min-level must be in [0, max level]");
    if (opt.uniform_level > 0 || opt.mms) { opt.amr_interval = 0; opt.init_amr = 0; }
    require(opt.amr.coarsen_safety > 0.0 && opt.amr.coarsen_safety < 1.0, "--coarsen-safety must be in (0, 1)");
    require(opt.amr.buffer_layers >= 0, "--buffer must be >= 0");
    require(opt.amr.max_elems > 0, "--max-elems must be > 0");
    require(opt.print_interval > 0 && opt.vtk_interval >= 0, "output intervals must be valid");
    require(opt.mom_tol > 0.0 && opt.ppe_tol > 0.0 && opt.proj_tol > 0.0, "solver tolerances must be > 0");
    return true;
}

}  // namespace cfa
