// sorsched.cpp -- the tile schedules of the LDS-tiled SOR (host only): the tile levels of the rectangular
// and sheared wavefronts and the dependencies of their dataflow launch (see poisson2d.hpp).
#include <algorithm>
#include <vector>

#include "poisson2d.hpp"

namespace cc {

// Dependencies of the dataflow schedule: for each tile of `flat` (level order), the nearest non-empty
// tiles below it along TS, TJ and TI, reached through empty tile positions (CSR).
void sor_flow_deps(const std::vector<int>& flat, std::vector<int>& off, std::vector<int>& idx) {
    const int n = (int)flat.size() / 3;
    int mTS = 0, mTJ = 0, mTI = 0;
    for (int t = 0; t < n; ++t) {
        mTS = std::max(mTS, flat[3 * t] + 1); mTJ = std::max(mTJ, flat[3 * t + 1] + 1); mTI = std::max(mTI, flat[3 * t + 2] + 1);
    }
    auto key = [&](int TS, int TJ, int TI) { return ((size_t)TS * mTJ + TJ) * mTI + TI; };
    std::vector<int> at((size_t)mTS * mTJ * mTI, -1);  // tile index at a position, -1 if empty
    for (int t = 0; t < n; ++t) at[key(flat[3 * t], flat[3 * t + 1], flat[3 * t + 2])] = t;
    // deps of a position, memoised; positions are visited in increasing (TS, TJ, TI) so the lower ones are done
    std::vector<std::vector<int>> memo((size_t)mTS * mTJ * mTI);
    for (int TS = 0; TS < mTS; ++TS)
        for (int TJ = 0; TJ < mTJ; ++TJ)
            for (int TI = 0; TI < mTI; ++TI) {
                std::vector<int> d;
                const int nb[3][3] = {{TS - 1, TJ, TI}, {TS, TJ - 1, TI}, {TS, TJ, TI - 1}};
                for (const auto& q : nb) {
                    if (q[0] < 0 || q[1] < 0 || q[2] < 0) continue;
                    const size_t kq = key(q[0], q[1], q[2]);
                    if (at[kq] >= 0) d.push_back(at[kq]);
                    else d.insert(d.end(), memo[kq].begin(), memo[kq].end());
                }
                std::sort(d.begin(), d.end());
                d.erase(std::unique(d.begin(), d.end()), d.end());
                // keep the maximal ones: a tile below another dependency (coordinate-wise) is waited for
                // through it (every tile waits, transitively, for all non-empty tiles below it)
                std::vector<int> m;
                for (int x : d) {
                    bool dominated = false;
                    for (int y : d)
                        if (y != x && flat[3 * x] <= flat[3 * y] && flat[3 * x + 1] <= flat[3 * y + 1] &&
                            flat[3 * x + 2] <= flat[3 * y + 2]) { dominated = true; break; }
                    if (!dominated) m.push_back(x);
                }
                memo[key(TS, TJ, TI)] = std::move(m);
            }
    off.assign(1, 0);
    idx.clear();
    for (int t = 0; t < n; ++t) {
        const auto& d = memo[key(flat[3 * t], flat[3 * t + 1], flat[3 * t + 2])];
        idx.insert(idx.end(), d.begin(), d.end());
        off.push_back((int)idx.size());
    }
}

// The tile levels of the tiled schedule: (TS, TJ, TI) of every tile that updates at least one cell,
// grouped by level TS + TJ + TI (flat, 3 ints per tile), and the level offsets (in tiles).
// W is the tile width in cells (kTileW in the kernels).
void sor_levels_w(int W, int WJ, int S, int Nx, int Ny, int iters, std::vector<int>& flat, std::vector<int>& level_off) {
    const int nTS = (iters + S - 1) / S;
    const int nTI = (Nx + iters - 1 + W - 1) / W;
    const int nTJ = (Ny + iters - 1 + WJ - 1) / WJ;
    std::vector<std::vector<int>> lv(nTS + nTI + nTJ);
    for (int TS = 0; TS < nTS; ++TS) {
        const int s0 = S * TS, s1 = std::min(s0 + S, iters);
        for (int TJ = 0; TJ < nTJ; ++TJ)
            for (int TI = 0; TI < nTI; ++TI) {
                bool nonempty = false;
                for (int s = s0; s < s1 && !nonempty; ++s) {
                    const int ia = W * TI - s, ib = ia + W - 1, ja = WJ * TJ - s, jb = ja + WJ - 1;
                    nonempty = ib >= 0 && ia <= Nx - 1 && jb >= 0 && ja <= Ny - 1;
                }
                if (nonempty) {
                    auto& v = lv[TS + TJ + TI];
                    v.push_back(TS); v.push_back(TJ); v.push_back(TI);
                }
            }
    }
    flat.clear();
    level_off.assign(1, 0);
    for (auto& v : lv) {
        if (v.empty()) continue;
        flat.insert(flat.end(), v.begin(), v.end());
        level_off.push_back((int)flat.size() / 3);
    }
}

// The sheared schedule: (TS, TK, TI) of every sheared tile (S sweeps x W values of i + s x K values
// of i + j + 2s) that updates at least one cell, grouped by level TS + TK + TI.
void sor_levels_sheared(int W, int K, int S, int Nx, int Ny, int iters, std::vector<int>& flat,
                        std::vector<int>& level_off) {
    const int nTS = (iters + S - 1) / S;
    const int nTI = (Nx + iters - 1 + W - 1) / W;
    const int nTK = (Nx + Ny + 2 * iters - 3 + K - 1) / K;
    std::vector<std::vector<int>> lv(nTS + nTK + nTI);
    for (int TS = 0; TS < nTS; ++TS) {
        const int s0 = S * TS, s1 = std::min(s0 + S, iters);
        for (int TK = 0; TK < nTK; ++TK)
            for (int TI = 0; TI < nTI; ++TI) {
                bool nonempty = false;
                for (int s = s0; s < s1 && !nonempty; ++s) {
                    const int imin = std::max(0, W * TI - s), imax = std::min(Nx - 1, W * TI + W - 1 - s);
                    if (imin > imax) continue;
                    const int jmin = std::max(0, K * TK - imax - 2 * s), jmax = std::min(Ny - 1, K * TK + K - 1 - imin - 2 * s);
                    nonempty = jmin <= jmax;
                }
                if (nonempty) {
                    auto& v = lv[TS + TK + TI];
                    v.push_back(TS); v.push_back(TK); v.push_back(TI);
                }
            }
    }
    flat.clear();
    level_off.assign(1, 0);
    for (auto& v : lv) {
        if (v.empty()) continue;
        flat.insert(flat.end(), v.begin(), v.end());
        level_off.push_back((int)flat.size() / 3);
    }
}

void sor_levels(const SorTile& t, int Nx, int Ny, int iters, std::vector<int>& flat, std::vector<int>& level_off) {
    if (t.sheared) sor_levels_sheared(kSorTileW, t.WJ, t.S, Nx, Ny, iters, flat, level_off);
    else sor_levels_w(kSorTileW, t.WJ, t.S, Nx, Ny, iters, flat, level_off);
}

}  // namespace cc

extern "C" {

cc_status cc_sor_flow_schedule(int Nx, int Ny, int iters, int W, int WJ, int S, int sheared, int* ntiles, int* ndeps,
                               int* tiles, int* dep_off, int* dep_idx) {
    if (Nx < 1 || Ny < 1 || iters < 1 || W < 1 || WJ < 1 || S < 1 || !ntiles || !ndeps ||
        (double)(Nx + iters) * (Ny + iters) * iters / ((double)W * WJ * S) > 1.0e7) {
        cc::set_error(CC_ERR_INVALID_ARG, "cc_sor_flow_schedule: bad grid, tile or output");
        return CC_ERR_INVALID_ARG;
    }
    std::vector<int> flat, off, doff, didx;
    if (sheared) cc::sor_levels_sheared(W, WJ, S, Nx, Ny, iters, flat, off);
    else cc::sor_levels_w(W, WJ, S, Nx, Ny, iters, flat, off);
    cc::sor_flow_deps(flat, doff, didx);
    const int n = (int)flat.size() / 3, nd = (int)didx.size();
    if (tiles && dep_off && dep_idx) {
        if (*ntiles < n || *ndeps < nd) {
            cc::set_error(CC_ERR_INVALID_ARG, "cc_sor_flow_schedule: output arrays too small");
            return CC_ERR_INVALID_ARG;
        }
        std::copy(flat.begin(), flat.end(), tiles);
        std::copy(doff.begin(), doff.end(), dep_off);
        std::copy(didx.begin(), didx.end(), dep_idx);
    }
    *ntiles = n;
    *ndeps = nd;
    return CC_OK;
}

}  // extern "C"
