// DFG statistics of the force history and VTK snapshots (src/cfa/diagnostics.cuh:47-291).
#include "outputs.hpp"

#include <algorithm>
#include <cmath>
#include <fstream>
#include <iomanip>
#include <iostream>

namespace cfa {
namespace {

// Extremum of the parabola through samples i-1, i, i+1: (time, value), the offset clamped to
// half a sample on either side.
struct Peak { double t, v; };
Peak parabola_peak(const std::vector<double>& t, const std::vector<double>& y, size_t i) {
    const double a = y[i - 1], b = y[i], c = y[i + 1];
    const double curv = a - 2.0 * b + c;
    double off = curv != 0.0 ? 0.5 * (a - c) / curv : 0.0;
    off = std::min(0.5, std::max(-0.5, off));
    return {t[i] + off * (t[i + 1] - t[i]), b - 0.25 * (a - c) * off};
}

// piecewise-linear value at tq, held constant outside the samples
double value_at(const std::vector<double>& t, const std::vector<double>& y, double tq) {
    const auto it = std::lower_bound(t.begin(), t.end(), tq);
    if (it == t.begin()) return y.front();
    if (it == t.end()) return y.back();
    const size_t k = (size_t)(it - t.begin());
    const double w = (tq - t[k - 1]) / (t[k] - t[k - 1]);
    return (1.0 - w) * y[k - 1] + w * y[k];
}

const char* const kRule = "------------------------------------------------------------------------\n";

}  // namespace

ForceHistory::Stats ForceHistory::statistics(double window, size_t& first) const {
    Stats S;
    const double startup = 2.0;
    first = 0;
    const size_t n = t_.size();
    if (n == 0 || t_.back() < startup + 0.5) return S;
    const double from = std::max(t_.back() - window, startup);
    while (first < n && t_[first] < from) ++first;
    if (n - first < 5) return S;
    double mean = 0.0;
    for (size_t i = first; i < n; ++i) mean += cl_[i];
    mean /= (double)(n - first);
    // periods are delimited by the upward crossings of the mean lift; the lift maximum of each
    // stretch between two crossings starts a period
    std::vector<size_t> rise;
    for (size_t i = first + 1; i < n; ++i)
        if (cl_[i - 1] < mean && cl_[i] >= mean) rise.push_back(i);
    std::vector<Peak> lift_max;
    for (size_t k = 0; k + 1 < rise.size(); ++k) {
        size_t top = rise[k];
        for (size_t i = rise[k]; i < rise[k + 1]; ++i)
            if (cl_[i] > cl_[top]) top = i;
        if (top > 0 && top + 1 < n) lift_max.push_back(parabola_peak(t_, cl_, top));
    }
    std::vector<double> q[6];
    for (size_t k = 0; k + 1 < lift_max.size(); ++k) {
        const double t0 = lift_max[k].t, t1 = lift_max[k + 1].t, f = 1.0 / (t1 - t0);
        double cd_hi = -1e300, cd_lo = 1e300, cl_lo = 1e300;
        for (size_t i = 1; i + 1 < n; ++i) {
            if (t_[i] < t0 || t_[i] >= t1) continue;
            if (cd_[i] >= cd_[i - 1] && cd_[i] > cd_[i + 1]) cd_hi = std::max(cd_hi, parabola_peak(t_, cd_, i).v);
            if (cd_[i] <= cd_[i - 1] && cd_[i] < cd_[i + 1]) cd_lo = std::min(cd_lo, parabola_peak(t_, cd_, i).v);
            if (cl_[i] <= cl_[i - 1] && cl_[i] < cl_[i + 1]) cl_lo = std::min(cl_lo, parabola_peak(t_, cl_, i).v);
        }
        if (cd_hi < -1e299 || cd_lo > 1e299 || cl_lo > 1e299) continue;
        q[0].push_back(cd_hi);
        q[1].push_back(cd_lo);
        q[2].push_back(lift_max[k].v);
        q[3].push_back(cl_lo);
        q[4].push_back(2.0 * Sim::R_CYL * f / Sim::U_MEAN);
        q[5].push_back(value_at(t_, dp_, t0 + 0.5 / f));
    }
    S.periods = (int)q[0].size();
    for (int k = 0; k < 6 && S.periods > 0; ++k) {
        double sum = 0.0, lo = 1e300, hi = -1e300;
        for (double x : q[k]) {
            sum += x;
            lo = std::min(lo, x);
            hi = std::max(hi, x);
        }
        S.value[k] = sum / S.periods;
        S.spread[k] = hi - lo;
    }
    return S;
}

void ForceHistory::report(double window, const std::string& csv_path) const {
    std::cout << kRule;
    size_t first = 0;
    const Stats S = statistics(window, first);
    if (S.periods == 0) {
        std::cout << "  Force statistics: n/a (statistics start after the 2 s start-up transient and"
                     " need two complete lift periods; run to t >= 6 s for DFG 2D-2 values)\n"
                  << kRule;
        return;
    }
    // reference values: FEATFLOW Q2/P1disc level 6, dt 1/1600, same definitions; DFG 1996 ranges
    static const char* const names[6] = {"C_D,max", "C_D,min", "C_L,max", "C_L,min", "St", "dp(t0+1/2f)"};
    static const double ref[6] = {3.22739, 3.16426, 0.98658, -1.02129, 0.30184, 2.48520};
    static const char* const range[6] = {"[3.22, 3.24]", "", "[0.99, 1.01]", "", "[0.295, 0.305]", "[2.46, 2.50]"};
    std::cout << std::fixed << std::setprecision(5) << "  Force statistics over t in [" << t_[first] << ", "
              << t_.back() << "] s, " << S.periods << " lift period(s)\n"
              << "                 value     spread     reference   rel. error   DFG 1996\n";
    for (int k = 0; k < 6; ++k)
        std::cout << "    " << std::left << std::setw(12) << names[k] << std::right << std::setw(9) << S.value[k]
                  << "  " << std::scientific << std::setprecision(1) << std::setw(8) << S.spread[k] << "  "
                  << std::fixed << std::setprecision(5) << std::setw(10) << ref[k] << "  " << std::showpos
                  << std::setprecision(3) << std::setw(8) << 100.0 * (S.value[k] - ref[k]) / std::fabs(ref[k])
                  << " %" << std::noshowpos << std::setprecision(5) << "   " << range[k] << "\n";
    std::cout << kRule;
    std::ofstream f(csv_path);
    if (!f) return;
    static const char* const keys[6] = {"cd_max", "cd_min", "cl_max", "cl_min", "St", "dp_half"};
    f << "quantity,value,spread,reference\n" << std::setprecision(10);
    for (int k = 0; k < 6; ++k) f << keys[k] << "," << S.value[k] << "," << S.spread[k] << "," << ref[k] << "\n";
    f << "periods," << S.periods << ",0,0\n";
}

void ForceHistory::report_2d3(const std::string& csv_path) const {
    std::cout << kRule;
    const size_t n = t_.size();
    if (n < 3 || t_.back() < 8.0 - 1.0e-9) {
        std::cout << "  DFG 2D-3 statistics need t_end = 8 s\n" << kRule;
        return;
    }
    size_t first = 1;   // past the impulsive start of the pressure
    while (first < n && t_[first] < 1.0) ++first;
    size_t id = first, il = first;
    for (size_t i = first; i + 1 < n; ++i) {
        if (cd_[i] > cd_[id]) id = i;
        if (cl_[i] > cl_[il]) il = i;
    }
    const Peak pd = parabola_peak(t_, cd_, id), pl = parabola_peak(t_, cl_, il);
    // reference values: John 2004 (delta_p = p(a1) - p(a2) as in the DFG definition)
    static const double ref[5] = {3.93625, 2.950921575, 5.693125, 0.47795, -0.1116};
    static const char* const names[5] = {"t(C_D,max)", "C_D,max", "t(C_L,max)", "C_L,max", "dp(8 s)"};
    const double val[5] = {pd.t, pd.v, pl.t, pl.v, value_at(t_, dp_, 8.0)};
    std::cout << std::fixed << std::setprecision(6) << "  DFG 2D-3 (reference: John 2004)\n";
    for (int k = 0; k < 5; ++k)
        std::cout << "    " << std::left << std::setw(11) << names[k] << std::right << std::setw(12) << val[k]
                  << "   ref " << std::setw(12) << ref[k] << "   rel. error " << std::showpos << std::setprecision(3)
                  << 100.0 * (val[k] - ref[k]) / std::fabs(ref[k]) << " %" << std::noshowpos << std::setprecision(6)
                  << "\n";
    std::cout << kRule;
    std::ofstream f(csv_path);
    if (!f) return;
    static const char* const keys[5] = {"t_cd_max", "cd_max", "t_cl_max", "cl_max", "dp_8"};
    f << "quantity,value,reference\n" << std::setprecision(10);
    for (int k = 0; k < 5; ++k) f << keys[k] << "," << val[k] << "," << ref[k] << "\n";
}

void write_vtk(const std::string& filename, const Mesh2D& mesh, const std::vector<double>& u,
               const std::vector<double>& v, const std::vector<double>& p, const std::vector<double>& vort,
               const std::vector<double>& elem_err, const AmrParams& amr) {
    std::ofstream o(filename);
    if (!o) {
        std::fprintf(stderr, "[Output] cannot open %s for writing\n", filename.c_str());
        return;
    }
    const int n = mesh.num_nodes, ne = mesh.num_elems;
    o << std::setprecision(10) << "# vtk DataFile Version 3.0\n2D FEM-AMR Cylinder Flow Solution (HIP)\nASCII\n"
      << "DATASET UNSTRUCTURED_GRID\nPOINTS " << n << " double\n";
    for (int i = 0; i < n; ++i) o << mesh.x[i] << " " << mesh.y[i] << " 0.0\n";
    o << "\nCELLS " << ne << " " << 4 * ne << "\n";
    for (const Triangle& t : mesh.elems) o << "3 " << t.v[0] << " " << t.v[1] << " " << t.v[2] << "\n";
    o << "\nCELL_TYPES " << ne << "\n";
    for (int e = 0; e < ne; ++e) o << "5\n";   // triangle
    o << "\nPOINT_DATA " << n << "\nVECTORS velocity double\n";
    for (int i = 0; i < n; ++i) o << u[i] << " " << v[i] << " 0.0\n";
    auto node_scalars = [&](const char* name, const std::vector<double>& f) {
        o << "\nSCALARS " << name << " double 1\nLOOKUP_TABLE default\n";
        for (int i = 0; i < n; ++i) o << f[i] << "\n";
    };
    node_scalars("pressure", p);
    node_scalars("vorticity", vort);
    o << "\nCELL_DATA " << ne << "\nSCALARS amr_level int 1\nLOOKUP_TABLE default\n";
    for (const Triangle& t : mesh.elems) o << t.level << "\n";
    o << "\nSCALARS nvb_generation int 1\nLOOKUP_TABLE default\n";
    for (const Triangle& t : mesh.elems) o << t.gen << "\n";
    const bool have = (int)elem_err.size() >= 2 * ne;
    o << "\nSCALARS indicator_vorticity double 1\nLOOKUP_TABLE default\n";
    for (int e = 0; e < ne; ++e) o << (have ? elem_err[e] : 0.0) << "\n";
    o << "\nSCALARS indicator_interp double 1\nLOOKUP_TABLE default\n";
    for (int e = 0; e < ne; ++e) o << (have ? elem_err[ne + e] : 0.0) << "\n";
    // refinement score max(eta_v / tol_v, eta_i / tol_i): > 1 refines, < coarsen_safety coarsens
    o << "\nSCALARS error_indicator double 1\nLOOKUP_TABLE default\n";
    for (int e = 0; e < ne; ++e) {
        const double sv = have && amr.tol_vort > 0.0 ? elem_err[e] / amr.tol_vort : 0.0;
        const double si = have && amr.tol_interp > 0.0 ? elem_err[ne + e] / amr.tol_interp : 0.0;
        o << std::max(sv, si) << "\n";
    }
}

}  // namespace cfa
