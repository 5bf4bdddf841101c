// common.hpp -- helpers shared by the drop-in programs.
#pragma once
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

}  // namespace ccprog
