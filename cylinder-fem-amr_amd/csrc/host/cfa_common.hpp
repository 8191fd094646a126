// Shared host-side definitions: channel geometry, physical constants, boundary tags, timers.
//
// Follows src/cfa/common.cuh:48-74 (SimConfig, inflow_umax, BoundaryTag). The numbers are the
// DFG 2D-2/2D-3 benchmark data and must be identical, so they are spelled out again here.
#pragma once

#include <chrono>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <string>
#include <utility>
#include <vector>

namespace cfa {

// Geometry and fluid data (src/cfa/common.cuh:48-60).
struct Sim {
    static constexpr double L_CHANNEL = 2.20;
    static constexpr double H_CHANNEL = 0.41;
    static constexpr double X_CYL = 0.20;
    static constexpr double Y_CYL = 0.20;
    static constexpr double R_CYL = 0.05;
    static constexpr double U_MAX = 1.50;
    static constexpr double U_MEAN = 1.00;
    static constexpr double NU = 1.0e-3;
    static constexpr double RHO = 1.00;
    static constexpr double DT_DEFAULT = 1.0e-3;
    static constexpr double PI = 3.14159265358979323846;
};

// Peak inflow at time t: U_MAX (2D-2) or U_MAX sin(pi t / 8) (2D-3); src/cfa/common.cuh:63-65.
inline double inflow_peak(double t, bool ramp_2d3) {
    return ramp_2d3 ? Sim::U_MAX * std::sin(Sim::PI * t / 8.0) : Sim::U_MAX;
}

// Node classes (src/cfa/common.cuh:68-74). The integer values travel to the device.
enum BcTag : int {
    BC_INTERIOR = 0,
    BC_INLET = 1,
    BC_OUTLET = 2,
    BC_WALL = 3,
    BC_CYLINDER = 4
};

using Clock = std::chrono::steady_clock;
inline double ms_between(Clock::time_point a, Clock::time_point b) {
    return std::chrono::duration<double, std::milli>(b - a).count();
}
inline double ms_since(Clock::time_point t0) { return ms_between(t0, Clock::now()); }

// Undirected edge key: smaller vertex id in the high word (src/cfa/mesh.cuh:19-22).
inline uint64_t edge_key(int a, int b) {
    if (a > b) std::swap(a, b);
    return (static_cast<uint64_t>(static_cast<uint32_t>(a)) << 32) | static_cast<uint32_t>(b);
}

[[noreturn]] inline void fatal(const char* msg) {
    std::fprintf(stderr, "%s\n", msg);
    std::exit(EXIT_FAILURE);
}

}  // namespace cfa
