// common.hpp -- helpers shared by the drop-in programs.
#pragma once
#include <algorithm>
#include <array>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <string>
#include <vector>

#include "../../include/ccamd.h"

namespace ccprog {

// Exit with the library's message when a call fails (the reference programs have no error paths
// for the step itself; a failed device call ends the run with status 1).
inline void check(cc_status st, const char* what) {
    if (st != CC_OK) {
        std::fprintf(stderr, "%s failed: %s\n", what, cc_last_error());
        std::exit(1);
    }
}

// Cell centres x0 + (i + 0.5) * dx, dx = (x1 - x0) / n (M2:935-937, M3:1140-1143, EH:650-653, NS:101-109).
inline std::vector<double> centres(double x0, double x1, int n) {
    const double d = (x1 - x0) / n;
    std::vector<double> c((size_t)n);
    for (int i = 0; i < n; i++) c[(size_t)i] = x0 + (i + 0.5) * d;
    return c;
}

// Time-step collapse guard (no reference counterpart).  The reference's HLLD divides its star states by
// d = rho (S - u)(S - S_M) - Bn^2 with only |d| > 1e-14 as a guard (M2:424-442, M3:466-484).  Near the
// Alfven degeneracy a face can have |d| small but above the guard (1e-9 is seen in the 2D default run
// after a one-ulp change of the initial state): its star states reach O(100), a cell's density goes
// negative and is floored to 1e-10, and dt drops by five orders of magnitude, after which the loop
// would crawl for hours.  trip() is true when dt falls below `ratio` times the largest dt of the last
// 64 steps (the last step, clamped to t_end, excepted); ratio <= 0 disables the guard.
struct DtGuard {
    double ratio = 1e-3;
    std::array<double, 64> ring{};
    int n = 0;
    bool trip(double dt, double t, double t_end) {
        double ref = 0.0;
        for (int k = 0; k < std::min(n, 64); ++k) ref = std::max(ref, ring[(size_t)k]);
        ring[(size_t)(n++ % 64)] = dt;
        return ratio > 0.0 && n > 8 && t < t_end && dt < ratio * ref;
    }
    double scale() const {
        double ref = 0.0;
        for (int k = 0; k < std::min(n, 64); ++k) ref = std::max(ref, ring[(size_t)k]);
        return ref;
    }
};

// The cell of a frame with the largest signal-speed bound |v| + sqrt(gamma P / rho + B^2 / rho), from
// the frame's rho, P, vabs and B planes (ncell cells each); returns its index.
inline size_t fastest_cell(const double* rho, const double* P, const double* vabs, const double* Bx, const double* By,
                           const double* Bz, size_t ncell, double gamma) {
    size_t best = 0;
    double smax = -1.0;
    for (size_t q = 0; q < ncell; ++q) {
        const double r = std::max(rho[q], 1e-300);
        const double s = vabs[q] + std::sqrt(std::max(0.0, gamma * P[q] / r + (Bx[q] * Bx[q] + By[q] * By[q] + Bz[q] * Bz[q]) / r));
        if (s > smax || std::isnan(s)) { smax = s; best = q; if (std::isnan(s)) break; }
    }
    return best;
}

}  // namespace ccprog
