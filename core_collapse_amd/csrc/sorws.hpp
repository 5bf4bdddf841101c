// sorws.hpp -- lexicographic SOR on clamped grids with wave-streamed tiles; see sorws.hip.
#pragma once
#include <vector>

#include "common.hpp"
#include "profiler.hpp"

namespace cc {

constexpr int kWsPD = 6;    // steps the sweep-0 loads are issued ahead

struct WsArgs {
    int Nx, Ny, nsw, NB, pitch, ntiles;
    double kx, ky, om;
    void* D;           // iterate (FP64 potential), diagonal layout D[k * pitch + r] = cell (k - r, r)
    const void* C;     // source term (FP64), diagonal layout
    void* rec;         // per tile and local step: lane 63's values of all SW sweeps (iterate, then source term)
    size_t rec_stride; // elements per tile
    unsigned rec_bytes; // bytes per tile
    int* state;        // [0] tile counter, [1] error flag, [2 + tile] progress (global time of its next step)
    unsigned long long wait_ticks;
};

// The wave-streamed SOR of one clamped grid: phi = `iters` lexicographic SOR sweeps from zero.  FP64 sweeps
// the potential itself; mixed sweeps an FP32 correction per tile of SW sweeps on an FP64 base whose residual
// it evaluates in FP64 (tile-local iterative refinement).
class SorWs {
   public:
    cc_status init(int Nx, int Ny, int iters, double dx, double dy, double G, double omega, int precision);
    void destroy();
    ~SorWs() { destroy(); }
    // phi (grid layout, mean not removed) of rho; rhobar: device scalar, the mean of rho.
    cc_status solve(const double* rho, const double* rhobar, double* phi, hipStream_t s, Profiler* prof);
    cc_status error(hipStream_t s);  // a producer wait timed out (results are wrong); syncs s
    bool ready() const { return D != nullptr; }

    int Nx = 0, Ny = 0, iters = 0, SW = 4, NB = 0, pitch = 0, K = 0, ntiles = 0, grid = 0;
    bool mixed = false;
    double kx = 0, ky = 0, om = 0, c4piG = 0, inv_denom = 0;
    void *D = nullptr, *C = nullptr, *rec = nullptr;
    int* state = nullptr;
    size_t rec_stride = 0;
    unsigned long long wait_ticks = 0;
};

// Whether a clamped grid's solve uses the wave-streamed tiles (CC_SOR_WS=1; the LDS tiles otherwise).
bool sor_ws_enabled();

}  // namespace cc
