// ORACLE (test infrastructure only): command-line runner for fixed-mesh cases, with the
// reference's flag names for the options it supports (src/cfa/options.cuh:134-289).
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>

#include "cfa_oracle.hpp"

int main(int argc, char** argv) {
    oracle::RunConfig c;
    for (int i = 1; i < argc; ++i) {
        std::string a = argv[i];
        auto val = [&]() -> const char* {
            if (i + 1 >= argc) { std::fprintf(stderr, "missing value for %s\n", a.c_str()); std::exit(1); }
            return argv[++i];
        };
        if (a == "--steps") c.steps = std::atoi(val());
        else if (a == "--dt") c.dt = std::atof(val());
        else if (a == "--uniform-level") c.uniform_level = std::atoi(val());
        else if (a == "--min-level") c.min_level = std::atoi(val());
        else if (a == "--mms") c.mms = true;
        else if (a == "--outlet") c.outlet_dirichlet = std::string(val()) == "dirichlet";
        else if (a == "--case") c.case_2d3 = std::string(val()) == "2d3";
        else if (a == "--time-scheme") c.bdf2 = std::string(val()) == "bdf2";
        else if (a == "--pressure-update") c.rotational = std::string(val()) == "rotational";
        else if (a == "--pressure-stab") c.pressure_stab = std::atof(val());
        else if (a == "--supg-tau-time") c.supg_tau_time = std::atof(val());
        else if (a == "--amg-rebuild") c.amg_rebuild = std::atoi(val());
        else if (a == "--mom-tol") c.mom_tol = std::atof(val());
        else if (a == "--ppe-tol") c.ppe_tol = std::atof(val());
        else if (a == "--print-interval") c.print_interval = std::atoi(val());
        else if (a == "--prefix") c.prefix = val();
        else if (a == "--quiet") c.quiet = true;
        else { std::fprintf(stderr, "unknown option %s\n", a.c_str()); return 1; }
    }
    if (c.uniform_level > 0 || c.mms) { /* fixed mesh: AMR off, as the reference forces */ }
    oracle::RunResult r = oracle::run_fixed(c);
    std::printf("[oracle] %zu steps in %.2f s (%.2f steps/s), non-converged solves %d\n", r.rows.size(), r.wall_s,
                r.rows.size() / r.wall_s, r.nonconverged);
    if (c.mms)
        std::printf("[oracle MMS] ||u-u_h||_L2 = %.6e | |u-u_h|_H1 = %.6e | ||p-p_h||_L2 = %.6e | cyl %.6e | probes %.6e\n",
                    r.mms[0], r.mms[1], r.mms[2], r.cyl_max, r.probe_max);
    return r.nonconverged == 0 ? 0 : 2;
}
