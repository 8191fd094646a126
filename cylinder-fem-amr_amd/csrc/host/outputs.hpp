// Run outputs of the program besides the forces CSV: the DFG benchmark statistics of the force
// history (src/cfa/diagnostics.cuh:47-210 ForceHistory, printed and written to
// <prefix>_stats.csv) and legacy-VTK snapshots of the solution (diagnostics.cuh:215-291).
#pragma once

#include <string>
#include <vector>

#include "amr.hpp"

namespace cfa {

class ForceHistory {
public:
    void add(double t, double cd, double cl, double dp) {
        t_.push_back(t);
        cd_.push_back(cd);
        cl_.push_back(cl);
        dp_.push_back(dp);
    }
    // DFG 2D-2 statistics over the last `window` seconds after the 2 s start-up (Schaefer & Turek
    // 1996): per lift period (from one lift maximum to the next) C_D max/min, C_L max/min,
    // Strouhal number and delta_p half a period after the lift maximum; mean and spread over the
    // complete periods. Printed, and written to csv_path.
    void report(double window, const std::string& csv_path) const;
    // DFG 2D-3 statistics: C_D and C_L maxima with their times, delta_p at t = 8 s.
    void report_2d3(const std::string& csv_path) const;

    struct Stats {
        int periods = 0;
        double value[6] = {0, 0, 0, 0, 0, 0};
        double spread[6] = {0, 0, 0, 0, 0, 0};
    };
    Stats statistics(double window, size_t& first) const;

private:
    std::vector<double> t_, cd_, cl_, dp_;
};

// ASCII legacy VTK unstructured grid: velocity, pressure, vorticity at the nodes; level,
// generation and the AMR indicators per element.
void write_vtk(const std::string& filename, const Mesh2D& mesh, const std::vector<double>& u,
               const std::vector<double>& v, const std::vector<double>& p, const std::vector<double>& vort,
               const std::vector<double>& elem_err, const AmrParams& amr);

}  // namespace cfa
