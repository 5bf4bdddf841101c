// sorws.hip -- lexicographic SOR on clamped grids with wave-streamed tiles (gfx950).
//
// The same iterate as the lexicographic sweep of compute_poisson_gravity_periodic (M2:207-223): cell (c, r)
// of sweep s reads W and S of sweep s and E, N and itself of sweep s - 1, so all cells with one value of
// c + r + 2 s are independent.  A tile is one wave: 64 rows (lanes) x SW sweeps (registers) streaming over all
// columns.  Lane l of tile (b, q) runs sweep s = q SW + m on row r = 64 b + l - m, and at local step T on
// column c = T - l - m.  Then every value a lane needs at step T comes from step T - 1 or T - 2 of itself or
// of lane l - 1 (a DPP shift), so a step is SW independent cell updates per lane with no LDS and no barrier:
//   W  = own value of sweep m at T-1           S    = lane l-1, sweep m, T-1
//   N  = own value of sweep m-1 at T-1         E    = lane l-1, sweep m-1, T-1
//   self (previous sweep) = lane l-1, sweep m-1, T-2;  source term = lane l-1's of sweep m-1 at T-2.
// Sweep m = 0 reads E, N, self and the source from memory; lane 0 reads its lane -1 values from the
// record that lane 63 of the tile below, (b-1, q), writes at every step (its values of all SW sweeps).  A
// record is loaded kWsRP steps before its first use, spread over the lanes (lane j holds its element j), into
// a ring of kWsRS registers; a DPP row shift brings element j to lane 0 where the step needs it.
// The final values of the last sweep go to memory.  The arrays are kept along anti-diagonals, D[k][r] =
// cell (k - r, r), so every lane of a step touches one diagonal and all global accesses are coalesced.
//
// Tiles run in one persistent launch: waves take tiles in (q, b) order from a counter.  Tile (b, q) waits for
// its producers (b-1, q) (the records), (b, q-1) and (b+1, q-1) (the sweep-(s0-1) values of its rows) to be
// ahead of it on the global time axis c + r + 2 s; every tile publishes its progress every 12 steps.
// Everything one wave writes and another reads (iterate, records, progress) goes through device-coherent
// (sc1) loads and stores, which bypass the caches that are not coherent across the chip: a publication is
// "all of this wave's memory operations complete (s_waitcnt), then the progress store", and a consumer issues
// its data loads only after its progress load has returned a value past what it needs.  No agent-scope
// fence (an L2 write-back or invalidate on this chip) is executed in the sweep.  A model of this kernel (its registers, loads, stores, waits
// and publications, in a random interleaving of tiles, with in-flight stores and loads checked for races)
// reproduces the lexicographic sweep bit for bit in tests/test_sor_schedule.py.
//
// Clamped boundaries (M2:196-198): a missing neighbour is the cell's own previous-sweep value.  A lane whose
// column has not started holds the E input (the clamp value of W at c = 0); a lane past the last column holds
// its last value (the clamp value of E at c = Nx - 1 for the next sweep); rows 0 and Ny - 1 select self for S
// and N.  Interior tiles run the middle columns without any of these selects.
//
// Mixed precision (R = float): tile (b, q) sweeps a correction d = phi - phi_b in FP32, where phi_b is the
// potential after sweep s0 - 1 (FP64, in memory).  Sweep 0 evaluates the FP64 residual of phi_b at its cell,
// res = src + kx (E + W) + ky (N + S) - phi_b with all neighbours from phi_b, and runs on d with res as its
// source term (d of E, N and the cell itself is zero); sweeps m >= 1 are the plain FP32 sweep on d and res;
// the last sweep stores phi_b + d in FP64.  This is the FP64 sweep written as phi = phi_b + d, so each tile
// is one step of iterative refinement, with FP32 arithmetic in every sweep of the tile.
#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <type_traits>

#include "sorws.hpp"

namespace cc {

// ------------------------------------------------------------------ device helpers
// lane l <- lane l - 1 (DPP wave_shr:1); lane 0 <- `old`
__device__ __forceinline__ float ws_shr(float old, float v) {
    return __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(__builtin_bit_cast(int, old),
                                                                 __builtin_bit_cast(int, v), 0x138, 0xf, 0xf, false));
}
__device__ __forceinline__ double ws_shr(double old, double v) {
    const long long o = __builtin_bit_cast(long long, old), u = __builtin_bit_cast(long long, v);
    const int lo = __builtin_amdgcn_update_dpp((int)(o & 0xffffffffll), (int)(u & 0xffffffffll), 0x138, 0xf, 0xf, false);
    const int hi = __builtin_amdgcn_update_dpp((int)(o >> 32), (int)(u >> 32), 0x138, 0xf, 0xf, false);
    return __builtin_bit_cast(double, ((long long)hi << 32) | (unsigned)lo);
}

// 16-byte chunks of a record: lane 63's SW values and SW source terms, and in the mixed form its FP64 base
// value (two floats) and two floats of padding
template <typename R>
using WsVec = R __attribute__((ext_vector_type(16 / sizeof(R))));
template <typename R, int SW>
constexpr int ws_recn() { return 2 * SW + (std::is_same<R, float>::value ? 4 : 0); }
constexpr int kWsRS = 12;           // record ring slots (= steps per unrolled group)
constexpr int kWsRP = kWsRS - 3;    // steps a record is loaded before its first use

// Lane 0 <- lane j of v (DPP row_shl:j within the first row of 16 lanes; readlane beyond it).  Only lane 0
// of the result is used; j is a constant after unrolling, and the switch folds to one instruction.
__device__ __forceinline__ int ws_lane0(int v, int j) {
#define CC_WS_ROWSHL(J) \
    case J:             \
        return __builtin_amdgcn_update_dpp(0, v, 0x100 + J, 0xf, 0xf, false);
    switch (j) {
        case 0:
            return v;
        CC_WS_ROWSHL(1) CC_WS_ROWSHL(2) CC_WS_ROWSHL(3) CC_WS_ROWSHL(4) CC_WS_ROWSHL(5) CC_WS_ROWSHL(6)
        CC_WS_ROWSHL(7) CC_WS_ROWSHL(8) CC_WS_ROWSHL(9) CC_WS_ROWSHL(10) CC_WS_ROWSHL(11) CC_WS_ROWSHL(12)
        CC_WS_ROWSHL(13) CC_WS_ROWSHL(14) CC_WS_ROWSHL(15)
        default:
            return __builtin_amdgcn_readlane(v, j);
    }
#undef CC_WS_ROWSHL
}
__device__ __forceinline__ float ws_elem(float v, int j) {
    return __builtin_bit_cast(float, ws_lane0(__builtin_bit_cast(int, v), j));
}
__device__ __forceinline__ double ws_elem(double v, int j) {
    const long long u = __builtin_bit_cast(long long, v);
    const int lo = ws_lane0((int)(u & 0xffffffffll), j), hi = ws_lane0((int)(u >> 32), j);
    return __builtin_bit_cast(double, ((long long)hi << 32) | (unsigned)lo);
}
// The FP64 base value of a mixed-form record (floats j and j + 1), in lane 0.
__device__ __forceinline__ double ws_elem_base(float v, int j) {
    const int lo = ws_lane0(__builtin_bit_cast(int, v), j), hi = ws_lane0(__builtin_bit_cast(int, v), j + 1);
    return __builtin_bit_cast(double, ((long long)hi << 32) | (unsigned)lo);
}

// Device-coherent scalar accesses of values that other waves read or wrote.
template <typename T>
__device__ __forceinline__ T ws_ld(const T* p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
template <typename T>
__device__ __forceinline__ void ws_st(T* p, T v) {
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
// Records: 16-byte device-coherent buffer accesses (cache policy sc1) through a resource over one tile's
// records (the raw-buffer descriptor word 3 of gfx9).
constexpr int kWsSc1 = 16;
__device__ __forceinline__ __amdgpu_buffer_rsrc_t ws_rsrc(const void* base, unsigned bytes) {
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(base), 0, (int)bytes, 0x00020000);
}

// Registers of one tile.  Ring indices are compile-time in the unrolled step loop.
template <typename R, int SW>
struct WsRegs {
    R X[3][SW];                   // value of sweep m computed at step T (slot T % 3)
    R Sr[3][SW];                  // source term of sweep m at step T
    R RD[kWsRS];                  // records of tile (b-1, q), its local step X in slot X % RS, element j in lane j
    double E[kWsPD], N[kWsPD], C[kWsPD];  // sweep-0 inputs of steps T .. T + PD - 1 (slot T % PD)
    double Bw[kWsPD];             // mixed: base of the cell the last sweep stores
    double Eprev, Eprev2;         // E inputs of the two previous steps (self and, mixed, W of the base)
};

struct WsTile {
    int b, q, nm, gbase, tile;
    int pB, pL, pD;      // producer tiles (b-1, q), (b, q-1), (b+1, q-1); -1: none
    int needB, needLD;   // producer progress (global time) already seen
};

// Waits (all lanes) until tile p has published progress >= need (global time); returns the progress seen.
// Bounded by wall-clock ticks: on a timeout the error flag is raised and the wave goes on; once the flag is
// up no wave waits any more (the results are wrong, the launch drains).  The data loads that follow are
// issued after the progress load has returned (the exit test depends on it); the fence keeps the compiler
// from moving them above it.
__device__ __forceinline__ int ws_wait(const WsArgs& a, int p, int need) {
    int v = ws_ld(&a.state[2 + p]);
    if (v < need) {
        const unsigned long long t0 = wall_clock64();
        while (v < need) {
            if (wall_clock64() - t0 > a.wait_ticks || ws_ld(&a.state[1]) != 0) {
                ws_st(&a.state[1], 1);
                break;
            }
            __builtin_amdgcn_s_sleep(2);
            v = ws_ld(&a.state[2 + p]);
        }
    }
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
    return v;
}

// All lanes: every load and store this wave has issued is complete (its device-coherent stores have
// reached the point of coherence) before lane 0 stores the progress value.
__device__ __forceinline__ void ws_publish(const WsArgs& a, int tile, int g, int lane) {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
    __builtin_amdgcn_s_waitcnt(0);
    if (lane == 0) ws_st(&a.state[2 + tile], g);
}

// The sweep-0 loads of step T into ring slot PH % PD: E = cell (c+1, r), N = (c, r+1), the source term, and
// for the mixed form the base of the cell the last sweep stores, (c-nm+1, r-nm+1) (the lane's own cell where
// that is outside the grid: always in bounds).
template <typename R, int SW, int PH>
__device__ __forceinline__ void ws_prefetch(const WsArgs& a, const WsTile& t, WsRegs<R, SW>& g, int T, int lane) {
    constexpr int e = PH % kWsPD;
    const double* D = static_cast<const double*>(a.D);
    const double* Cs = static_cast<const double*>(a.C);
    const int r0 = 64 * t.b + lane;
    const long k = T + 64 * t.b;
    const long P = a.pitch;
    g.E[e] = ws_ld(&D[(k + 1) * P + r0]);
    g.N[e] = ws_ld(&D[(k + 1) * P + r0 + 1]);
    g.C[e] = Cs[k * P + r0];
    if constexpr (std::is_same<R, float>::value) {
        const long kw = k - 2 * (t.nm - 1);
        const int rw = r0 - (t.nm - 1);
        g.Bw[e] = ws_ld(&D[(rw >= 0 && kw >= 0) ? kw * P + rw : k * P + r0]);
    }
}

// One local step T (T % 12 == PH).  GEN: with the column / row selects (ramps, edge tiles, partial sweep block).
// The lane -1 values come from the records of the producer's steps T + 63 (slot r1) and T + 62 (slot r2).
template <typename R, int SW, int PH, bool GEN>
__device__ __forceinline__ void ws_step(const WsArgs& a, const WsTile& t, WsRegs<R, SW>& g, int T, int lane,
                                        R* __restrict__ srec) {
    constexpr bool MIX = std::is_same<R, float>::value;
    constexpr int cur = PH % 3, p1 = (PH + 2) % 3, p2 = (PH + 1) % 3, e = PH % kWsPD;
    constexpr int r1 = (PH + 63) % kWsRS, r2 = (PH + 62) % kWsRS;
    const R kx = (R)a.kx, ky = (R)a.ky, om = (R)a.om;
    double* D = static_cast<double*>(a.D);
    const int r0 = 64 * t.b + lane;
#pragma unroll
    for (int m = 0; m < SW; ++m) {
        const int c = T - lane - m, r = r0 - m;
        R self, E, N, src;
        if (m == 0) {
            if constexpr (MIX) {
                // S of the base: lane l-1's base of step T-1 (its E input of step T-2); lane 0: the record's
                double sb = g.Eprev, eb = g.E[e], nb = g.N[e], wb = g.Eprev2;
                double s_b = ws_shr(ws_elem_base(g.RD[r1], 2 * SW), g.Eprev2);
                if constexpr (GEN) {
                    if (c == a.Nx - 1) eb = sb;
                    if (c == 0) wb = sb;
                    if (r == 0) s_b = sb;
                    if (r == a.Ny - 1) nb = sb;
                }
                src = (R)(((g.C[e] + a.kx * (eb + wb)) + a.ky * (nb + s_b)) - sb);
                self = E = N = R(0);
            } else {
                self = g.Eprev;
                E = g.E[e];
                N = g.N[e];
                src = g.C[e];
            }
        } else {
            self = ws_shr(ws_elem(g.RD[r2], m - 1), g.X[p2][m - 1]);
            E = ws_shr(ws_elem(g.RD[r1], m - 1), g.X[p1][m - 1]);
            N = g.X[p1][m - 1];
            src = ws_shr(ws_elem(g.RD[r2], SW + m - 1), g.Sr[p2][m - 1]);
        }
        R S = ws_shr(ws_elem(g.RD[r1], m), g.X[p1][m]);
        const R W = g.X[p1][m];
        if constexpr (GEN) {
            if (m == 0 && c == a.Nx - 1) E = self;
            if (r == 0) S = self;
            if (r == a.Ny - 1) N = self;
        }
        const R xn = self + om * (((src + kx * (E + W)) + ky * (N + S)) - self);
        bool store = m == SW - 1;
        if constexpr (GEN) {
            const bool active = c >= 0 && c < a.Nx && r >= 0 && r < a.Ny && m < t.nm;
            g.X[cur][m] = active ? xn : (c < 0 ? E : W);
            store = active && m == t.nm - 1;
        } else {
            g.X[cur][m] = xn;
        }
        g.Sr[cur][m] = src;
        if (store) {
            const size_t q = (size_t)(c + r) * a.pitch + r;
            if constexpr (MIX)
                ws_st(&D[q], g.Bw[e] + (double)xn);
            else
                ws_st(&D[q], (double)xn);
        }
    }
    // the record of this step (lane 63's values) for the tile above
    if (lane == 63) {
        constexpr int RN = ws_recn<R, SW>(), L = 16 / sizeof(R);
        R v[RN];
#pragma unroll
        for (int m = 0; m < SW; ++m) {
            v[m] = g.X[cur][m];
            v[SW + m] = g.Sr[cur][m];
        }
        if constexpr (RN > 2 * SW) {
            const long long bits = __builtin_bit_cast(long long, g.Eprev);
            v[2 * SW] = __builtin_bit_cast(R, (int)(bits & 0xffffffffll));
            v[2 * SW + 1] = __builtin_bit_cast(R, (int)(bits >> 32));
            v[2 * SW + 2] = v[2 * SW + 3] = R(0);
        }
        const __amdgpu_buffer_rsrc_t rs = ws_rsrc(srec, a.rec_bytes);
#pragma unroll
        for (int j = 0; j < RN / L; ++j) {
            WsVec<R> w;
#pragma unroll
            for (int i = 0; i < L; ++i) w[i] = v[j * L + i];
            __builtin_amdgcn_raw_buffer_store_b128(__builtin_bit_cast(WsVec<int>, w), rs,
                                                   (T * RN + j * L) * (int)sizeof(R), 0, kWsSc1);
        }
    }
    g.Eprev2 = g.Eprev;
    g.Eprev = g.E[e];
}

// The producer tile's record of its local step Tp, spread over the lanes (lane j < RN: element j).
template <typename R, int SW>
__device__ __forceinline__ R ws_load_record(const WsArgs& a, const R* rin, int Tp, int lane) {
    constexpr int RN = ws_recn<R, SW>();
    const __amdgpu_buffer_rsrc_t rs = ws_rsrc(rin, a.rec_bytes);
    const int off = (Tp * RN + (lane < RN ? lane : 0)) * (int)sizeof(R);
    if constexpr (sizeof(R) == 8)
        return __builtin_bit_cast(R, __builtin_amdgcn_raw_buffer_load_b64(rs, off, 0, kWsSc1));
    else
        return __builtin_bit_cast(R, __builtin_amdgcn_raw_buffer_load_b32(rs, off, 0, kWsSc1));
}

// Before step T: the record of the producer's step T + 64 + RP (first used at step T + RP + 1, at global time
// gbase + T + RP) into its ring slot.
template <typename R, int SW, int PH>
__device__ __forceinline__ void ws_pre(const WsArgs& a, WsTile& t, WsRegs<R, SW>& g, int T, const R* rin, int lane) {
    if (t.pB < 0) return;
    const int need = t.gbase + T + kWsRP + 1;
    if (t.needB < need) t.needB = ws_wait(a, t.pB, need);
    g.RD[(PH + 64 + kWsRP) % kWsRS] = ws_load_record<R, SW>(a, rin, T + 64 + kWsRP, lane);
}

// After step T: the sweep-0 inputs of step T + PD into the slot step T has used; producers (b, q-1) and
// (b+1, q-1) must have passed global time gbase + T + PD.
template <typename R, int SW, int PH>
__device__ __forceinline__ void ws_post(const WsArgs& a, WsTile& t, WsRegs<R, SW>& g, int T, int lane) {
    const int need = t.gbase + T + kWsPD + 1;
    if (t.needLD < need) {
        int v = 1 << 30;
        if (t.pL >= 0) v = min(v, ws_wait(a, t.pL, need));
        if (t.pD >= 0) v = min(v, ws_wait(a, t.pD, need));
        t.needLD = v;
    }
    ws_prefetch<R, SW, PH>(a, t, g, T + kWsPD, lane);
}

template <typename R, int SW, bool GEN>
__device__ __forceinline__ void ws_run(const WsArgs& a, WsTile& t, WsRegs<R, SW>& g, int T0, int T1, int lane,
                                       R* srec, const R* rin) {
    // T0 and T1 are multiples of 12; progress is published after every group of 12 steps
    for (int T = T0; T < T1; T += 12) {
#define CC_WS_STEP(J)                                   \
    ws_pre<R, SW, J>(a, t, g, T + J, rin, lane);        \
    ws_step<R, SW, J, GEN>(a, t, g, T + J, lane, srec); \
    ws_post<R, SW, J>(a, t, g, T + J, lane);
        CC_WS_STEP(0) CC_WS_STEP(1) CC_WS_STEP(2) CC_WS_STEP(3) CC_WS_STEP(4) CC_WS_STEP(5)
        CC_WS_STEP(6) CC_WS_STEP(7) CC_WS_STEP(8) CC_WS_STEP(9) CC_WS_STEP(10) CC_WS_STEP(11)
#undef CC_WS_STEP
        ws_publish(a, t.tile, t.gbase + T + 12, lane);
    }
}

template <typename R, int SW>
__global__ __launch_bounds__(64) void k_sor_ws(WsArgs a) {
    const int lane = threadIdx.x;
    __shared__ int s_tile;
    for (;;) {
        // The next tile, claimed by lane 0 and handed to the wave through LDS between barriers.  (Handing it
        // over with readfirstlane after a lane-0 branch lets the compiler put the branch on the loop's back
        // edge, so that lanes 1-63 re-enter the loop without lane 0 and read a stale tile.)
        __syncthreads();
        if (lane == 0) s_tile = atomicAdd(&a.state[0], 1);
        __syncthreads();
        const int tile = s_tile;
        if (tile >= a.ntiles) break;
        WsTile t;
        t.tile = tile;
        t.q = tile / a.NB;
        t.b = tile - t.q * a.NB;
        t.nm = min(SW, a.nsw - t.q * SW);
        t.gbase = 64 * t.b + 2 * SW * t.q;
        t.pB = t.b > 0 ? tile - 1 : -1;
        t.pL = t.q > 0 ? tile - a.NB : -1;
        t.pD = (t.q > 0 && t.b + 1 < a.NB) ? tile - a.NB + 1 : -1;
        t.needB = t.needLD = -(1 << 30);
        // rows 64 b - SW + 1 .. 64 b + 63 of the tile hold row 0 or Ny - 1 (or rows outside), or a partial block
        const bool edge = 64 * t.b - SW + 1 <= 0 || 64 * t.b + 63 >= a.Ny - 1 || t.nm < SW;
        // the last active step is (Nx - 1) + 63 + (nm - 1); whole groups of 6 steps
        const int Tend = (a.Nx + 63 + t.nm + 11) / 12 * 12;
        R* srec = static_cast<R*>(a.rec) + (size_t)tile * a.rec_stride;
        const R* rin = static_cast<const R*>(a.rec) + (size_t)(t.pB >= 0 ? t.pB : 0) * a.rec_stride;
        WsRegs<R, SW> g;
#pragma unroll
        for (int j = 0; j < 3; ++j)
#pragma unroll
            for (int m = 0; m < SW; ++m) g.X[j][m] = g.Sr[j][m] = R(0);
#pragma unroll
        for (int j = 0; j < kWsRS; ++j) g.RD[j] = R(0);
        // sweep-0 inputs of steps 0 .. PD - 1; E inputs of steps -1 and -2 (the cell itself, and the cell
        // west of it, which only lanes outside the grid use); lane -1 records of steps -2 and -1
        {
            int v = 1 << 30;
            if (t.pL >= 0) v = min(v, ws_wait(a, t.pL, t.gbase + kWsPD));
            if (t.pD >= 0) v = min(v, ws_wait(a, t.pD, t.gbase + kWsPD));
            t.needLD = v;
        }
        ws_prefetch<R, SW, 0>(a, t, g, 0, lane);
        ws_prefetch<R, SW, 1>(a, t, g, 1, lane);
        ws_prefetch<R, SW, 2>(a, t, g, 2, lane);
        ws_prefetch<R, SW, 3>(a, t, g, 3, lane);
        ws_prefetch<R, SW, 4>(a, t, g, 4, lane);
        ws_prefetch<R, SW, 5>(a, t, g, 5, lane);
        g.Eprev = ws_ld(&static_cast<const double*>(a.D)[(size_t)(64 * t.b) * a.pitch + 64 * t.b + lane]);
        g.Eprev2 = 0.0;
        // W of sweep 0 at column 0 (lane 0, step 0): the value a lane holds before its column starts, the
        // E input of step -1 (the correction d is zero there)
        if constexpr (!std::is_same<R, float>::value) g.X[2][0] = g.Eprev;
        if (t.pB >= 0) {  // the producer's steps 62 .. 63 + RP
            t.needB = ws_wait(a, t.pB, t.gbase + kWsRP);
#pragma unroll
            for (int X = 62; X < 64 + kWsRP; ++X) g.RD[X % kWsRS] = ws_load_record<R, SW>(a, rin, X, lane);
        }
        // ramp-up, middle, ramp-down: from step Ta every lane and sweep is past column 0, before step Tb
        // no lane has reached column Nx - 1
        const int Ta = min(Tend, (SW + 63 + 11) / 12 * 12);
        const int Tb = max(Ta, (a.Nx - 1) / 12 * 12);
        if (edge || Tb <= Ta) {
            ws_run<R, SW, true>(a, t, g, 0, Tend, lane, srec, rin);
        } else {
            ws_run<R, SW, true>(a, t, g, 0, Ta, lane, srec, rin);
            ws_run<R, SW, false>(a, t, g, Ta, Tb, lane, srec, rin);
            ws_run<R, SW, true>(a, t, g, Tb, Tend, lane, srec, rin);
        }
        ws_publish(a, tile, 1 << 30, lane);
    }
}

// ------------------------------------------------------------------ layout kernels
// src = -4 pi G (rho - mean rho) / denom into the diagonal layout, and the iterate = 0 (also outside the grid).
__global__ __launch_bounds__(256) void k_ws_prepare(int Nx, int Ny, int pitch, int K, double c4piG, double inv_denom,
                                                    const double* __restrict__ rho, const double* __restrict__ rhobar,
                                                    double* __restrict__ C, double* __restrict__ D) {
    const int r = blockIdx.x * 256 + threadIdx.x, k = blockIdx.y;
    if (r >= pitch || k >= K) return;
    const int c = k - r;
    double v = 0.0;
    if (r < Ny && c >= 0 && c < Nx) v = -(c4piG * (rho[(size_t)r * Nx + c] - rhobar[0])) * inv_denom;
    C[(size_t)k * pitch + r] = v;
    D[(size_t)k * pitch + r] = 0.0;
}

// The iterate in diagonal layout -> phi in the grid layout, through a 32 x 32 LDS tile (both sides coalesced).
__global__ __launch_bounds__(256) void k_ws_finish(int Nx, int Ny, int pitch, const double* __restrict__ D,
                                                   double* __restrict__ phi) {
    __shared__ double s[32][33];
    const int c0 = blockIdx.x * 32, r0 = blockIdx.y * 32;
    // walk the tile's diagonals: consecutive threads read consecutive rows of one diagonal
    for (int idx = threadIdx.x; idx < 63 * 32; idx += 256) {
        const int kk = idx / 32, rr = idx % 32, cc = kk - rr;
        const int c = c0 + cc, r = r0 + rr;
        if (cc < 0 || cc >= 32 || c >= Nx || r >= Ny) continue;
        s[rr][cc] = D[(size_t)(c + r) * pitch + r];
    }
    __syncthreads();
    for (int idx = threadIdx.x; idx < 32 * 32; idx += 256) {
        const int rr = idx / 32, cc = idx % 32;
        const int c = c0 + cc, r = r0 + rr;
        if (c < Nx && r < Ny) phi[(size_t)r * Nx + c] = s[rr][cc];
    }
}

__global__ void k_ws_reset(int* __restrict__ state, int n) {
    for (int q = blockIdx.x * blockDim.x + threadIdx.x; q < n; q += gridDim.x * blockDim.x)
        if (q != 1) state[q] = 0;
}

// ------------------------------------------------------------------ host side
bool sor_ws_enabled() {
    const char* e = std::getenv("CC_SOR_WS");
    return e && e[0] == '1';
}

cc_status SorWs::init(int Nx_, int Ny_, int iters_, double dx, double dy, double G, double omega, int precision) {
    destroy();
    Nx = Nx_; Ny = Ny_; iters = iters_;
    mixed = precision == CC_MIXED;
    SW = mixed ? 8 : 4;
    if (const char* e = std::getenv("CC_SOR_WS_SW")) SW = std::atoi(e);
    if (SW != 4 && SW != 8) {
        set_error(CC_ERR_INVALID_ARG, "SOR wave-streamed tiles: sweeps per tile must be 4 or 8");
        return CC_ERR_INVALID_ARG;
    }
    if (Nx < 1 || Ny < 1 || iters < 1) {
        set_error(CC_ERR_INVALID_ARG, "SOR wave-streamed tiles: empty grid or no sweeps");
        return CC_ERR_INVALID_ARG;
    }
    NB = (Ny + SW - 1 + 63) / 64;
    pitch = 64 * NB + 2;
    K = Nx + 64 * NB + SW + 16;
    ntiles = NB * ((iters + SW - 1) / SW);
    const double idx2 = 1.0 / (dx * dx), idy2 = 1.0 / (dy * dy), denom = 2.0 * (idx2 + idy2);
    inv_denom = 1.0 / denom;
    kx = idx2 * inv_denom;
    ky = idy2 * inv_denom;
    om = omega;
    c4piG = 4.0 * M_PI * G;
    // records: local steps < Tend <= Nx + 63 + SW + 11 are written, steps up to Tend + 63 + 9 are read
    rec_stride = (size_t)(Nx + SW + 160) * (2 * SW + (mixed ? 4 : 0));
    const size_t rsz = mixed ? sizeof(float) : sizeof(double);
    const size_t ND = (size_t)K * pitch;
    if ((double)K * pitch > 2.0e9 || K > 65535) {
        set_error(CC_ERR_INVALID_ARG, "SOR wave-streamed tiles: grid too large");
        return CC_ERR_INVALID_ARG;
    }
    auto al = [&](void** p, size_t bytes) {
        if (hipMalloc(p, bytes) != hipSuccess) return false;
        return hipMemset(*p, 0, bytes) == hipSuccess;
    };
    if (!al(&D, ND * sizeof(double)) || !al(&C, ND * sizeof(double)) || !al(&rec, (size_t)ntiles * rec_stride * rsz) ||
        !al((void**)&state, ((size_t)ntiles + 2) * sizeof(int))) {
        set_error(CC_ERR_OOM, "SOR wave-streamed tiles: device allocation failed");
        return CC_ERR_OOM;
    }
    int dev = 0, khz = 0, ncu = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, dev) != hipSuccess || khz <= 0)
        khz = 100000;
    if (hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || ncu <= 0) ncu = 256;
    wait_ticks = (unsigned long long)khz * 100ull;  // 100 ms
    int per_cu = 8;
    if (const char* e = std::getenv("CC_SOR_WS_WAVES")) per_cu = std::max(1, std::atoi(e));
    grid = std::min(ntiles, ncu * per_cu);
    return CC_OK;
}

void SorWs::destroy() {
    for (void* p : {D, C, rec, (void*)state})
        if (p) hipFree(p);
    D = C = rec = nullptr;
    state = nullptr;
}

cc_status SorWs::solve(const double* rho, const double* rhobar, double* phi, hipStream_t s, Profiler* prof) {
    const dim3 dg((pitch + 255) / 256, K);
    hipLaunchKernelGGL(k_ws_prepare, dg, dim3(256), 0, s, Nx, Ny, pitch, K, c4piG, inv_denom, rho, rhobar,
                       static_cast<double*>(C), static_cast<double*>(D));
    CC_CHECK_LAUNCH();
    hipLaunchKernelGGL(k_ws_reset, dim3(std::min(64, (ntiles + 2 + 255) / 256)), dim3(256), 0, s, state, ntiles + 2);
    CC_CHECK_LAUNCH();
    WsArgs a;
    a.Nx = Nx; a.Ny = Ny; a.nsw = iters; a.NB = NB; a.pitch = pitch; a.ntiles = ntiles;
    a.kx = kx; a.ky = ky; a.om = om;
    a.D = D; a.C = C; a.rec = rec; a.rec_stride = rec_stride;
    a.rec_bytes = (unsigned)(rec_stride * (mixed ? sizeof(float) : sizeof(double)));
    a.state = state;
    a.wait_ticks = wait_ticks;
    {
        ProfScope ps(prof, s, "sor");
        if (mixed && SW == 8)
            hipLaunchKernelGGL((k_sor_ws<float, 8>), dim3(grid), dim3(64), 0, s, a);
        else if (mixed)
            hipLaunchKernelGGL((k_sor_ws<float, 4>), dim3(grid), dim3(64), 0, s, a);
        else if (SW == 8)
            hipLaunchKernelGGL((k_sor_ws<double, 8>), dim3(grid), dim3(64), 0, s, a);
        else
            hipLaunchKernelGGL((k_sor_ws<double, 4>), dim3(grid), dim3(64), 0, s, a);
        CC_CHECK_LAUNCH();
    }
    hipLaunchKernelGGL(k_ws_finish, dim3((Nx + 31) / 32, (Ny + 31) / 32), dim3(256), 0, s, Nx, Ny, pitch,
                       static_cast<const double*>(D), phi);
    CC_CHECK_LAUNCH();
    return CC_OK;
}

cc_status SorWs::error(hipStream_t s) {
    int e = 0;
    CC_HIP(hipMemcpyAsync(&e, state + 1, sizeof(int), hipMemcpyDeviceToHost, s));
    CC_HIP(hipStreamSynchronize(s));
    if (e != 0) {
        CC_HIP(hipMemsetAsync(state + 1, 0, sizeof(int), s));
        CC_HIP(hipStreamSynchronize(s));
        set_error(CC_ERR_HIP, "SOR wave-streamed tiles: a wait for a producer tile timed out");
        return CC_ERR_HIP;
    }
    return CC_OK;
}

}  // namespace cc
