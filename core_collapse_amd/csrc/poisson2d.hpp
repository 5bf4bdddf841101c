// poisson2d.hpp -- 2D Poisson gravity by SOR on gfx950.
//
// Restates compute_poisson_gravity_periodic (mhd2d_resistive_hlld_ct_gravity.cpp:178-237):
// rhs = 4 pi G (rho - mean rho); phi0 = 0; `iters` lexicographic Gauss-Seidel SOR sweeps;
// g = -central grad(phi).  The per-sweep mean removal of the reference (M2:219-222) commutes
// with the sweep (the stencil weights sum to the diagonal), so removing it once at the end
// gives the same iterate and g does not depend on it at all.
//
// Three schedules reproduce the lexicographic iterate exactly (up to rounding):
//  * clamped ("outflow") boundaries on both axes: a time-skewed, LDS-tiled wavefront; or (CC_SOR_WS=1)
//    wave-streamed tiles (sorws.hpp), one wave of 64 rows x SW sweeps streaming over the columns with its
//    neighbours in registers.  Cell
//    (i,j) of sweep s runs at wavefront step i+j+2s, i.e. every step updates one colour of a
//    red-black checkerboard inside a moving band; tiles of 64 x WJ cells x S sweeps are
//    parallelograms in (i+s, j+s, s) and run level by level (one launch per tile level).
//  * any periodic axis: one workgroup, anti-diagonal by anti-diagonal; the wrap-around
//    neighbours allow sweeps to overlap only at a skew of max(Nx, Ny) steps per sweep (against
//    Nx + Ny - 1 without overlap).
// Mixed precision (tiled schedule): each tile evaluates the FP64 residual of its region once,
// then runs its S sweeps on an FP32 correction in LDS and adds it back in FP64 (tile-local
// iterative refinement).
#pragma once
#include <vector>

#include "common.hpp"
#include "profiler.hpp"
#include "sorws.hpp"

namespace cc {

constexpr int kSorTileW = 64;  // tile width along i of the LDS-tiled schedules (one wave64: lane = column)

// Tile shape of the tiled schedule: S sweeps per tile (16 waves x S/16), WJ rows; FP64 keeps the
// source term in LDS (rhs_lds) or loads it from global memory one step ahead.
struct SorTile {
    int S = 32, WJ = 64;    // sweeps per tile; rows (rectangular) or k values (sheared) per tile
    bool rhs_lds = true;
    bool sheared = false;   // boxes in (sweep, i + j + 2 sweep, i + sweep) instead of (sweep, j + sweep, i + sweep)
    bool fp64 = false;      // a mixed-precision solve running the FP64 kernels (full FP64 sweeps, the reference's
                            // arithmetic): the autotuner takes them where they are faster than the mixed ones
};
// The kernels a solve of precision `mixed` runs with tile t: the mixed (FP32 correction) kernels or the FP64 ones.
inline bool sor_mixed_kernels(const SorTile& t, bool mixed) { return mixed && !t.fp64; }

struct SorArgs {
    int Nx, Ny, nsw;
    int px, py;
    double idx2, idy2, inv_denom, omega, c4piG;
    float kxf, kyf, omegaf;  // FP32 correction-sweep coefficients: idx2/denom, idy2/denom, omega
};

// The compiled tile shapes (S, WJ[, source term in LDS]); FP64 with the source term in LDS needs
// 16 B per box cell, which fits 160 KB of LDS only at WJ = 64.
#define CC_SOR_TILES_FP64(X) X(16, 64, true) X(32, 64, true) X(32, 64, false) X(16, 128, false) X(32, 128, false)
#define CC_SOR_TILES_MIXED(X) X(16, 64) X(32, 64) X(16, 128) X(32, 128)
// The compiled sheared tile shapes (S, K); the source term is in LDS.
#define CC_SOR_SHEARED_FP64(X) X(16, 64) X(16, 32) X(32, 32) X(16, 80)
#define CC_SOR_SHEARED_MIXED(X) X(16, 64) X(16, 48) X(16, 32) X(16, 128) X(32, 64) X(32, 32) X(16, 80)

// One grid of a batched tiled SOR launch: its arguments and arrays.
struct SorBatchEntry {
    SorArgs a;
    double* phi;
    const double* rho;
    const double* rhobar;
};

class Poisson2D {
   public:
    cc_status init(int Nx, int Ny, bool px, bool py, double dx, double dy, double G, int iters, double omega,
                   int precision);
    // The fastest compiled tile shape for this grid (timed once per grid in the process, see cc_sor_autotune).
    cc_status autotuned_tile(SorTile& out);
    void destroy();
    // g = -grad phi(rho); rho, gx, gy are device arrays of Nx*Ny.  solve = prologue + sweeps + epilogue;
    // a batch of tiled grids runs the three separately, with the sweeps of all grids merged per tile level.
    cc_status solve(const double* rho, double* gx, double* gy, hipStream_t s, Profiler* prof);
    cc_status prologue(const double* rho, hipStream_t s, Profiler* prof);  // mean of rho, phi = 0
    cc_status sweeps(const double* rho, hipStream_t s, Profiler* prof);
    cc_status epilogue(double* gx, double* gy, hipStream_t s, Profiler* prof);  // g = -central grad phi
    const std::vector<int>& schedule_host() const { return sched_host_; }
    // The dataflow schedule hit its bounded wait (a fault in the schedule; results are wrong).  Syncs s.
    cc_status flow_error(hipStream_t s);
    // phi with its mean removed (after solve), device array.
    cc_status phi_mean_removed(double* out_dev, hipStream_t s);
    bool tiled() const { return tiled_; }
    int levels() const { return (int)level_off_.size() - 1; }
    int tiles_total() const { return level_off_.empty() ? 0 : level_off_.back(); }
    ~Poisson2D() { destroy(); }

    int Nx = 0, Ny = 0;
    bool px = true, py = true, tiled_ = false;
    double dx = 1, dy = 1, G = 1, omega = 1.8;
    int iters = 0, precision = 0;
    double* phi = nullptr;      // Nx*Ny
    double* partial = nullptr;  // reduction partials
    double* rhobar = nullptr;   // device scalar(s): [0] mean rho, [1] mean phi
    int* sched = nullptr;       // (TS,TJ,TI) per tile, grouped by level
    SorTile tile_{};            // tile shape of the tiled schedule
    std::vector<int> level_off_;
    std::vector<int> sched_host_;  // the schedule's (TS, TJ, TI) triples, grouped by level
    bool flow_ = false;            // tiled sweeps as one dataflow launch (else one launch per level)
    int* dep_off = nullptr;        // dataflow dependencies (CSR over the tiles of sched)
    int* dep_idx = nullptr;
    int* flow_state = nullptr;     // [0] tile counter, [1] error flag, [2 ..] done flags
    SorArgs args{};
    SorWs* ws_ = nullptr;          // the wave-streamed solve of a clamped grid (else the LDS tiles)
    bool wave_streamed() const { return ws_ != nullptr; }
};

// Launch one level of a batched tiled SOR: nt tiles given as (instance, TS, TJ, TI), every instance
// with tile shape t.
cc_status launch_tile_batch(const SorTile& t, bool mixed, int nt, hipStream_t s, const SorBatchEntry* tab,
                            const int* tiles4);

// Whether a tiled solve of ntiles tiles runs as one dataflow launch (else one launch per level).
bool sor_flow_enabled(int ntiles);
// Dependencies of the dataflow schedule of `flat` (level order): CSR offsets and tile indices.
void sor_flow_deps(const std::vector<int>& flat, std::vector<int>& off, std::vector<int>& idx);
// The dataflow launch over a merged schedule (instance, TS, TJ|TK, TI per tile, level order) with its waits
// (CSR over merged indices) and a state of ntiles + 2 ints, which it clears first.
cc_status launch_flow_batch(const SorTile& t, bool mixed, int ntiles, hipStream_t s, const SorBatchEntry* tab,
                            const int* tiles4, const int* dep_off, const int* dep_idx, int* state);
// Reports (and clears) a timed-out wait recorded in a dataflow state; synchronises s.
cc_status sor_flow_error(const int* state, hipStream_t s);

// The tile levels of the tiled schedule for a tile shape (flat (TS, TJ, TI) triples, level offsets).
void sor_levels_w(int W, int WJ, int S, int Nx, int Ny, int iters, std::vector<int>& flat, std::vector<int>& level_off);
void sor_levels_sheared(int W, int K, int S, int Nx, int Ny, int iters, std::vector<int>& flat, std::vector<int>& level_off);
void sor_levels(const SorTile& t, int Nx, int Ny, int iters, std::vector<int>& flat, std::vector<int>& level_off);
bool sor_autotune_enabled();

// Deterministic device mean of n doubles into out (two-pass fixed-order reduction).
cc_status device_mean(const double* x, size_t n, double* partial, double* out, hipStream_t s);

}  // namespace cc
