// mhd3d.hip -- gfx950 solver step for 3D resistive MHD with HLLD + constrained transport
// (restating mhd3d_resistive_hlld_ct.cpp, cited as M3:<line>), with the reference's hybrid
// cold + thermal + bounce EOS or an ideal gas.
//
// One forward-Euler stage (advance_stage, M3:694-1030) is:
//   k_face3d    x, y and (mixed) z faces, one face per thread: PLM-MC + HLLD -> 5 face fluxes, 2 face EMFs
//               (from the primitives in the flux precision, left by the previous stage's k_joule3d or,
//               at the start of a call, by k_prim3d)
//   k_update3d  (mixed) flux difference + floors (old faces), FP64               -> U*
//   k_zupd3d    (FP64) z faces + flux difference + floors in one column march    -> U*
//   k_ect3d     edge EMFs (1/4 sum of face EMFs + eta * curl of cell-centred B0) and the face-field
//               induction, in one z-marching pass over LDS tiles              -> B1
//   k_joule3d   Joule heating from cell-centred B1 differences + floors, marching along z per column;
//               the stage-2 instance also forms the Heun average, its floors and the next step's
//               signal-speed maximum; both store the primitives of their result for the next stage.
// k_edge3d + k_ct3d (edge arrays in memory) replace k_ect3d when CC_M3_FUSED=0.
// Mixed precision: primitives recovered in FP64 and stored in FP32; PLM and HLLD in FP32 (faces near
// the Alfven degeneracy, where the reference's star states are ill-conditioned, or near the guard's
// threshold are queued and solved again in FP64 from the conserved state by k_face3d_fp64); face fluxes
// and EMFs stored in FP32; flux differences, the conserved update, EMF averages and CT in FP64.
#include <algorithm>
#include <cstdlib>
#include <cmath>
#include <type_traits>
#include <vector>

#include "profiler.hpp"
#include "diag.hpp"
#include "riemann.hpp"
#include "stepcommon.hpp"

namespace cc {

struct Geom3D {
    int Nx, Ny, Nz, px, py, pz;
    double dx, dy, dz, floor_rho, floor_P, eta, cfl;
    float tau;          // mixed precision: relative Alfven-degeneracy margin for the FP64 fallback
    int ct_comp;        // compensated face update (FP32 low words of the face field)
    double hlld_guard;  // HLLD degeneracy guard (0: the reference's 1e-14 only)
};

// FP64 EOS for pressure recovery and floors: hybrid (M3:114-212) or ideal.
struct Eos3D {
    int ideal;
    double gamma;  // ideal
    HybridEOS<double> hy;
};

struct State3D {
    double *rho, *mx, *my, *mz, *E;  // Nx*Ny*Nz, (k*Ny+j)*Nx+i
    double *Bx, *By, *Bz;            // (Nx+1)Ny Nz, Nx(Ny+1)Nz, Nx Ny(Nz+1)
    float *Bx_lo, *By_lo, *Bz_lo;    // rounding errors of the faces (compensated face update only)
};

struct Edges {
    double *Ex, *Ey, *Ez;  // Nx(Ny+1)(Nz+1), (Nx+1)Ny(Nz+1), (Nx+1)(Ny+1)Nz
};

// ------------------------------------------------------------------ indexing (M3:250-308)
__device__ __forceinline__ size_t fx_(const Geom3D& g, int i, int j, int k) {
    return ((size_t)k * g.Ny + j) * (size_t)(g.Nx + 1) + i;
}
__device__ __forceinline__ size_t fy_(const Geom3D& g, int i, int j, int k) {
    return ((size_t)k * (g.Ny + 1) + j) * (size_t)g.Nx + i;
}
__device__ __forceinline__ size_t fz_(const Geom3D& g, int i, int j, int k) {
    return ((size_t)k * g.Ny + j) * (size_t)g.Nx + i;
}
// face-index maps: wrap modulo n+1 (periodic) or clamp to [0, n]
__device__ __forceinline__ int fmap(int i, int n, bool per) { return per ? imod(i, n + 1) : iclamp(i, 0, n); }

__device__ __forceinline__ double get_bx(const Geom3D& g, const double* B, int i, int j, int k) {
    return B[fx_(g, fmap(i, g.Nx, g.px), bmap(j, g.Ny, g.py), bmap(k, g.Nz, g.pz))];
}
__device__ __forceinline__ double get_by(const Geom3D& g, const double* B, int i, int j, int k) {
    return B[fy_(g, bmap(i, g.Nx, g.px), fmap(j, g.Ny, g.py), bmap(k, g.Nz, g.pz))];
}
__device__ __forceinline__ double get_bz(const Geom3D& g, const double* B, int i, int j, int k) {
    return B[fz_(g, bmap(i, g.Nx, g.px), bmap(j, g.Ny, g.py), fmap(k, g.Nz, g.pz))];
}
// cell-centred field from faces with the reference's index handling (M3:280-296)
__device__ __forceinline__ double bxc(const Geom3D& g, const double* B, int i, int j, int k) {
    return 0.5 * (get_bx(g, B, i, j, k) + get_bx(g, B, i + 1, j, k));
}
__device__ __forceinline__ double byc(const Geom3D& g, const double* B, int i, int j, int k) {
    return 0.5 * (get_by(g, B, i, j, k) + get_by(g, B, i, j + 1, k));
}
__device__ __forceinline__ double bzc(const Geom3D& g, const double* B, int i, int j, int k) {
    return 0.5 * (get_bz(g, B, i, j, k) + get_bz(g, B, i, j, k + 1));
}

// ------------------------------------------------------------------ FP64 EOS: floors, recovery
__device__ __forceinline__ double hy_pressure(const HybridEOS<double>& h, double rho, double e_spec) {  // M3:146-152
    rho = smax(rho, h.floor_rho);
    const auto c = h.cold(rho);
    const double eth = smax(0.0, e_spec - c.ec);
    const double Pth = (h.gamma_th - 1.0) * rho * eth;
    return smax(c.Pc + Pth, h.floor_P);
}

__device__ __forceinline__ void floors3d(const Eos3D& e, double& rho, double& mx, double& my, double& mz, double& E,
                                         double Bx, double By, double Bz) {
    const HybridEOS<double>& h = e.hy;
    rho = smax(rho, h.floor_rho);
    const double u = mx / rho, v = my / rho, w = mz / rho;
    const double kin = 0.5 * rho * (u * u + v * v + w * w);
    const double mag = 0.5 * (Bx * Bx + By * By + Bz * Bz);
    if (e.ideal) {  // M2:114-124
        const double P = (e.gamma - 1.0) * (E - kin - mag);
        if (P < h.floor_P) E = kin + mag + h.floor_P / (e.gamma - 1.0);
        return;
    }
    double eint_tot = E - kin - mag;  // M3:165-183
    if (eint_tot < 0.0) eint_tot = 0.0;
    const double e_spec = eint_tot / rho;
    double P = hy_pressure(h, rho, e_spec);
    if (P < h.floor_P) P = h.floor_P;
    const auto c = h.cold(rho);
    const double Pth = smax(0.0, P - c.Pc);
    const double eth = (h.gamma_th > 1.0) ? (Pth / ((h.gamma_th - 1.0) * rho)) : 0.0;
    E = kin + mag + rho * (c.ec + eth);
}

__device__ __forceinline__ Prim<double> cons_to_prim3d(const Eos3D& e, double rho, double mx, double my, double mz,
                                                       double E, double Bx, double By, double Bz) {
    const HybridEOS<double>& h = e.hy;
    rho = smax(rho, h.floor_rho);
    const double u = mx / rho, v = my / rho, w = mz / rho;
    const double kin = 0.5 * rho * (u * u + v * v + w * w);
    const double mag = 0.5 * (Bx * Bx + By * By + Bz * Bz);
    double P;
    if (e.ideal) {  // M2:126-135
        P = (e.gamma - 1.0) * (E - kin - mag);
        if (P < h.floor_P) P = h.floor_P;
    } else {  // M3:185-197
        double eint_tot = E - kin - mag;
        if (eint_tot < 0.0) eint_tot = 0.0;
        const double e_spec = eint_tot / rho;
        P = hy_pressure(h, rho, e_spec);
        if (P < h.floor_P) P = h.floor_P;
    }
    return Prim<double>{rho, u, v, w, P, Bx, By, Bz};
}

__device__ __forceinline__ Prim<double> prim_at(const Geom3D& g, const Eos3D& e, const State3D& U, int i, int j,
                                                int k) {  // get_prim_cc, M3:298-308
    i = bmap(i, g.Nx, g.px);
    j = bmap(j, g.Ny, g.py);
    k = bmap(k, g.Nz, g.pz);
    const size_t q = ((size_t)k * g.Ny + j) * g.Nx + i;
    return cons_to_prim3d(e, U.rho[q], U.mx[q], U.my[q], U.mz[q], U.E[q], bxc(g, U.Bx, i, j, k),
                          byc(g, U.By, i, j, k), bzc(g, U.Bz, i, j, k));
}

__device__ __forceinline__ double signal_speed3d(const Eos3D& e, const Prim<double>& W) {  // M3:633-636
    double a2;
    if (e.ideal) {
        a2 = e.gamma * W.P / smax(W.rho, e.hy.floor_rho);
    } else {
        a2 = e.hy.sound_speed2(W, e.hy.aux(W));
    }
    auto fs = [&](double Bn) {
        const double b2 = W.Bx * W.Bx + W.By * W.By + W.Bz * W.Bz;
        const double rho = smax(W.rho, 1e-300);
        const double vA2 = b2 / rho, vAn2 = (Bn * Bn) / rho;
        double disc = (a2 + vA2) * (a2 + vA2) - 4.0 * a2 * vAn2;
        disc = smax(0.0, disc);
        return sqrt(smax(0.0, 0.5 * ((a2 + vA2) + sqrt(disc))));
    };
    return smax(smax(fabs(W.u) + fs(W.Bx), fabs(W.v) + fs(W.By)), fabs(W.w) + fs(W.Bz));
}

// ------------------------------------------------------------------ flux path: primitives, faces, update
// Primitives of every cell (FP64 recovery, M3:298-308), stored in the flux precision R as 8 SoA arrays.
template <typename R>
struct PrimArrays {
    R* c[8];  // rho, u, v, w, P, Bx, By, Bz
};
// Face fluxes of one axis in the reference's component order (rho, mx, my, mz, E).
template <typename R>
struct FaceFlux {
    R* f[5];
};
template <typename R>
struct EmfArr {
    R *Ey_x, *Ez_x, *Ez_y, *Ex_y, *Ex_z, *Ey_z;
};

// The cell kernels take the EOS kind as a template argument (IDEAL), so the other EOS's code and
// registers drop out of the instance that runs.
template <bool IDEAL>
__device__ __forceinline__ Eos3D eos_of(Eos3D e) {
    e.ideal = IDEAL ? 1 : 0;
    return e;
}

template <typename R, bool IDEAL>
__global__ __launch_bounds__(256) void k_prim3d(Geom3D g, Eos3D e_in, State3D U0, PrimArrays<R> W) {
    const Eos3D e = eos_of<IDEAL>(e_in);
    const unsigned qq = blockIdx.x * 256u + threadIdx.x;  // cells < 2^31 (checked at create)
    if (qq >= (unsigned)g.Nx * (unsigned)g.Ny * (unsigned)g.Nz) return;
    const size_t q = qq;
    const int i = (int)(qq % (unsigned)g.Nx);
    const unsigned r = qq / (unsigned)g.Nx;
    const int j = (int)(r % (unsigned)g.Ny), k = (int)(r / (unsigned)g.Ny);
    const Prim<double> P = cons_to_prim3d(e, U0.rho[q], U0.mx[q], U0.my[q], U0.mz[q], U0.E[q], bxc(g, U0.Bx, i, j, k),
                                          byc(g, U0.By, i, j, k), bzc(g, U0.Bz, i, j, k));
    W.c[0][q] = (R)P.rho; W.c[1][q] = (R)P.u; W.c[2][q] = (R)P.v; W.c[3][q] = (R)P.w;
    W.c[4][q] = (R)P.P; W.c[5][q] = (R)P.Bx; W.c[6][q] = (R)P.By; W.c[7][q] = (R)P.Bz;
}

template <typename R>
__device__ __forceinline__ Prim<R> load_prim(const PrimArrays<R>& W, size_t q) {
    return Prim<R>{W.c[0][q], W.c[1][q], W.c[2][q], W.c[3][q], W.c[4][q], W.c[5][q], W.c[6][q], W.c[7][q]};
}

// The face solver of the flux path on FP64 face states (for the batched entry point): FP64 HLLD, or
// the mixed path's FP32 HLLD on the rounded states with its FP64 fallback on the FP64 states, the result
// rounded to FP32 as the mixed step stores its fluxes.
template <bool MIXED, typename EOS64, typename EOS32>
__device__ __forceinline__ void face_hlld3d(const EOS64& e64, const EOS32& e32, const Prim<double>& WL,
                                            const Prim<double>& WR, double Bn, double F[7], float tau, double guard) {
    if constexpr (MIXED) {
        hlld_x_mixed(e32, e64, WL, WR, Bn, F, tau, guard);
#pragma unroll
        for (int c = 0; c < 7; ++c) F[c] = (float)F[c];
    } else {
        hlld_x<double>(e64, WL, WR, Bn, fabs(Bn) < 1e-12, F, nullptr, 0.0, guard);
    }
}

// PLM-MC on the four cells f-2 .. f+1 of face (i, j, k) of axis AX (boundary-mapped) and the face
// states in the axis frame of hlld_x (y: the (u,v)/(Bx,By) swap of M3:577-595, z: the (u,v,w)->(w,u,v)
// map of M3:604-622).  cell_prim(i, j, k) gives the primitives of an in-range cell.
template <int AX, typename R, typename CellPrim>
__device__ __forceinline__ void face_states3d_of(const Geom3D& g, const CellPrim& cell_prim, double Bn, int i, int j,
                                                 int k, Prim<R>& L, Prim<R>& Rr) {
    const int f = AX == 0 ? i : (AX == 1 ? j : k);
    auto cell = [&](int o) {  // the cell o steps from the face along the axis
        int ii = i, jj = j, kk = k;
        if (AX == 0) ii = bmap(f + o, g.Nx, g.px);
        if (AX == 1) jj = bmap(f + o, g.Ny, g.py);
        if (AX == 2) kk = bmap(f + o, g.Nz, g.pz);
        return cell_prim(ii, jj, kk);
    };
    const Prim<R> w0 = cell(-2), w1 = cell(-1), w2 = cell(0), w3 = cell(1);
    const R frho = (R)g.floor_rho, fP = (R)g.floor_P;
    Prim<R> a, b, WL, WR;
    plm_mc<R>(w0, w1, w2, AX, frho, fP, a, WL);  // right state of cell f-1
    plm_mc<R>(w1, w2, w3, AX, frho, fP, WR, b);  // left state of cell f
    L = WL;
    Rr = WR;
    if (AX == 1) {
        L = Prim<R>{WL.rho, WL.v, WL.u, WL.w, WL.P, (R)Bn, WL.Bx, WL.Bz};
        Rr = Prim<R>{WR.rho, WR.v, WR.u, WR.w, WR.P, (R)Bn, WR.Bx, WR.Bz};
    } else if (AX == 2) {
        L = Prim<R>{WL.rho, WL.w, WL.u, WL.v, WL.P, (R)Bn, WL.Bx, WL.By};
        Rr = Prim<R>{WR.rho, WR.w, WR.u, WR.v, WR.P, (R)Bn, WR.Bx, WR.By};
    }
}

// Face states from the stored primitives of the stage.
template <int AX, typename R>
__device__ __forceinline__ void face_states3d(const Geom3D& g, const PrimArrays<R>& W, double Bn, int i, int j, int k,
                                              Prim<R>& L, Prim<R>& Rr) {
    face_states3d_of<AX, R>(
        g, [&](int ii, int jj, int kk) { return load_prim(W, ((size_t)kk * g.Ny + jj) * g.Nx + ii); }, Bn, i, j, k, L,
        Rr);
}

// The mixed path's FP64 solve of a face its FP32 HLLD flagged: the four cells' primitives recovered in FP64
// from the conserved state U of the stage, PLM-MC and HLLD in FP64, i.e. the FP64 path's flux of that face.
template <int AX, typename EOS64>
__device__ __forceinline__ void face_hlld_fp64_from_cons(const Geom3D& g, const Eos3D& e, const EOS64& e64,
                                                         const State3D& U, double Bn, int i, int j, int k,
                                                         double o[7]) {
    Prim<double> L, Rr;
    face_states3d_of<AX, double>(
        g, [&](int ii, int jj, int kk) { return prim_at(g, e, U, ii, jj, kk); }, Bn, i, j, k, L, Rr);
    hlld_x<double>(e64, L, Rr, Bn, fabs(Bn) < 1e-12, o, nullptr, 0.0, g.hlld_guard);
}

// The five fluxes of a face of axis AX in the conserved order (rho, mx, my, mz, E) from hlld_x's output in
// the axis frame.
template <int AX, typename R>
__device__ __forceinline__ void face_fluxes3d(const R o[7], R f[5]) {
    f[0] = o[0];
    f[4] = o[4];
    if (AX == 0) { f[1] = o[1]; f[2] = o[2]; f[3] = o[3]; }
    else if (AX == 1) { f[1] = o[2]; f[2] = o[1]; f[3] = o[3]; }
    else { f[1] = o[2]; f[2] = o[3]; f[3] = o[1]; }
}

// The two face EMFs of face t (Ey_x = F(Bz), Ez_x = -F(By); Ez_y = G(Bx), Ex_y = -G(Bz); Ex_z = H(By),
// Ey_z = -H(Bx)).
template <int AX, typename R>
__device__ __forceinline__ void face_emfs3d(EmfArr<R>& ef, size_t t, const R o[7]) {
    if (AX == 0) {
        ef.Ey_x[t] = o[6];
        ef.Ez_x[t] = -o[5];
    } else if (AX == 1) {
        ef.Ez_y[t] = o[5];
        ef.Ex_y[t] = -o[6];
    } else {
        ef.Ex_z[t] = o[6];
        ef.Ey_z[t] = -o[5];
    }
}

// The five fluxes in reference order and the two face EMFs of face t.
template <int AX, typename R>
__device__ __forceinline__ void face_store3d(FaceFlux<R>& F, EmfArr<R>& ef, size_t t, const R o[7]) {
    R f[5];
    face_fluxes3d<AX, R>(o, f);
#pragma unroll
    for (int c = 0; c < 5; ++c) F.f[c][t] = f[c];
    face_emfs3d<AX, R>(ef, t, o);
}

// Faces whose FP32 HLLD flagged the Alfven degeneracy, per axis, for the FP64 recomputation pass, and (z faces +
// update in one pass) the cells with such a z face, whose update waits for the FP64 flux.  Queue q (faces of axis
// 0, 1, 2; cells) is split into kQSub sub-queues: sub-queue k takes the indices [k span[q], (k+1) span[q]) into
// the slots of the same range, so its capacity is exact, and has its own counter on its own cache line.  One
// counter serialises: in a turbulence run ~70 % of the z-face waves can hold a flagged face, and their returning
// atomics on one address took the z-face kernel from 0.27 to 1.08 ms at 256^3.
constexpr int kQSub = 256;    // sub-queues per queue
constexpr int kQStride = 32;  // ints between counters (128 bytes)
struct DegQueue {
    int* count;    // counter of sub-queue k of queue q at count[(q * kQSub + k) * kQStride]
    int* face[3];  // face indices per axis (capacity: the axis' face count)
    int* cell;     // cell indices (capacity: the cell count)
    int span[4];   // indices per sub-queue: ceil(n / kQSub)
};

// Appends v to queue q for the lanes of the wave where `take` holds.  The lanes whose v falls in the first
// taking lane's sub-queue share one atomic (their slots contiguous in lane order); the others (a wave straddling
// a sub-queue boundary) take one each.  Every active lane of the wave must call it (uniform control flow).
__device__ __forceinline__ void queue_push(const DegQueue& dq, int q, bool take, int v) {
    const unsigned long long mask = __ballot(take);
    if (mask == 0ull) return;
    int* arr = q < 3 ? dq.face[q] : dq.cell;
    const int span = dq.span[q];
    const int k = v / span;
    const int lane = (int)__lane_id();
    const int leader = __ffsll((long long)mask) - 1;
    const int kl = __shfl(k, leader, 64);
    const bool same = take && k == kl;
    const unsigned long long ms = __ballot(same);
    int base = 0;
    if (lane == leader) base = atomicAdd(&dq.count[(q * kQSub + kl) * kQStride], __popcll(ms));
    base = __shfl(base, leader, 64);
    if (same) arr[kl * span + base + __popcll(ms & ((1ull << lane) - 1ull))] = v;
    else if (take) arr[k * span + atomicAdd(&dq.count[(q * kQSub + k) * kQStride], 1)] = v;
}

// Visits the entries of queue q with the blocks of a grid-stride launch: block b takes sub-queue b mod kQSub,
// the blocks with that remainder share its entries.
template <typename F>
__device__ __forceinline__ void queue_for_each(const DegQueue& dq, int q, F&& f) {
    const int* arr = q < 3 ? dq.face[q] : dq.cell;
    const int parts = max(1, (int)gridDim.x / kQSub);
    for (int k = (int)blockIdx.x % kQSub; k < kQSub; k += (int)min(gridDim.x, (unsigned)kQSub)) {
        const int n = dq.count[(q * kQSub + k) * kQStride];
        for (int e = ((int)blockIdx.x / kQSub) * (int)blockDim.x + (int)threadIdx.x; e < n; e += parts * (int)blockDim.x)
            f(arr[k * dq.span[q] + e]);
    }
}

// One face of axis AX per thread (x: M3:727-750, y: M3:752-775, z: M3:777-800): PLM-MC, HLLD, the
// fluxes and EMFs.  The mixed form runs the FP32 HLLD only and queues the faces it flags (near the
// Alfven degeneracy, or with the degeneracy guard on, near the guard's threshold); k_face3d_fp64 then
// recomputes those in FP64 from the conserved state U (the FP64 solver's registers would otherwise halve
// this kernel's occupancy).  Without the guard a field along z (the turbulence problem's) keeps most z
// faces near the degeneracy: INLINE z faces then recompute inline (a queue of most faces would cost more
// than the occupancy gains).
template <int AX, bool MIXED, typename EOS64, typename EOS32, bool INLINE = false>
__global__ __launch_bounds__(256) void k_face3d(Geom3D g, EOS64 e64, EOS32 e32,
                                                PrimArrays<typename std::conditional<MIXED, float, double>::type> W,
                                                const double* __restrict__ Bface,
                                                FaceFlux<typename std::conditional<MIXED, float, double>::type> F,
                                                EmfArr<typename std::conditional<MIXED, float, double>::type> ef,
                                                DegQueue dq, Eos3D e, State3D U) {
    using R = typename std::conditional<MIXED, float, double>::type;
    const int Nx = g.Nx, Ny = g.Ny, Nz = g.Nz;
    const int nx = Nx + (AX == 0), ny = Ny + (AX == 1);
    int i, j, k;
    if (AX == 0) {  // consecutive threads take consecutive faces of the (Nx+1) x Ny x Nz array
        const unsigned tt = blockIdx.x * 256u + threadIdx.x;
        if (tt >= (unsigned)nx * (unsigned)ny * (unsigned)Nz) return;
        i = (int)(tt % (unsigned)nx);
        const unsigned r = tt / (unsigned)nx;
        j = (int)(r % (unsigned)ny);
        k = (int)(r / (unsigned)ny);
    } else {  // a block is 64 consecutive i x 4 consecutive faces along AX, so its waves share cells in L1
        i = blockIdx.x * 64 + (threadIdx.x & 63);
        const int fa = blockIdx.z * 4 + (threadIdx.x >> 6);
        if (AX == 1) { j = fa; k = blockIdx.y; } else { k = fa; j = blockIdx.y; }
        if (i >= Nx || fa > (AX == 1 ? Ny : Nz)) return;
    }
    const size_t t = ((size_t)k * ny + j) * nx + i;
    const double Bn = Bface[t];
    Prim<R> L, Rr;
    face_states3d<AX, R>(g, W, Bn, i, j, k, L, Rr);
    R o[7];
    if constexpr (MIXED && !INLINE) {
        bool deg = false;
        hlld_x<float>(e32, L, Rr, (float)Bn, fabs(Bn) < 1e-12, o, &deg, g.tau, (float)g.hlld_guard);
        queue_push(dq, AX, deg, (int)t);
    } else if constexpr (MIXED) {  // recomputed inline
        bool deg = false;
        hlld_x<float>(e32, L, Rr, (float)Bn, fabs(Bn) < 1e-12, o, &deg, g.tau, (float)g.hlld_guard);
        if (deg) {
            double o64[7];
            face_hlld_fp64_from_cons<AX>(g, e, e64, U, Bn, i, j, k, o64);
#pragma unroll
            for (int c = 0; c < 7; ++c) o[c] = (float)o64[c];
        }
    } else {
        hlld_x<double>(e64, L, Rr, Bn, fabs(Bn) < 1e-12, o, nullptr, 0.0, g.hlld_guard);
    }
    face_store3d<AX, R>(F, ef, t, o);
}

template <int AX, typename EOS64>
__device__ __forceinline__ void face_fp64_one(const Geom3D& g, const Eos3D& e, const EOS64& e64, const State3D& U,
                                              const double* __restrict__ Bface, FaceFlux<float>& F, EmfArr<float>& ef,
                                              int tf) {
    const int nx = g.Nx + (AX == 0), ny = g.Ny + (AX == 1);
    const int i = tf % nx, r = tf / nx, j = r % ny, k = r / ny;
    double o64[7];
    face_hlld_fp64_from_cons<AX>(g, e, e64, U, Bface[tf], i, j, k, o64);
    float o[7];
#pragma unroll
    for (int c = 0; c < 7; ++c) o[c] = (float)o64[c];
    face_store3d<AX, float>(F, ef, (size_t)tf, o);
}

// The FP64 faces of the mixed path: the queued faces of the axes in the mask `axes` (bit a: axis a), each
// solved from the stage's conserved state U (the FP64 path's flux of the face, rounded to FP32).
template <typename EOS64>
__global__ __launch_bounds__(256) void k_face3d_fp64(Geom3D g, Eos3D e, EOS64 e64, State3D U, FaceFlux<float> F0,
                                                     FaceFlux<float> F1, FaceFlux<float> F2, EmfArr<float> ef,
                                                     DegQueue dq, int axes) {
    if (axes & 1) queue_for_each(dq, 0, [&](int tf) { face_fp64_one<0>(g, e, e64, U, U.Bx, F0, ef, tf); });
    if (axes & 2) queue_for_each(dq, 1, [&](int tf) { face_fp64_one<1>(g, e, e64, U, U.By, F1, ef, tf); });
    if (axes & 4) queue_for_each(dq, 2, [&](int tf) { face_fp64_one<2>(g, e, e64, U, U.Bz, F2, ef, tf); });
}

__global__ void k_zero_counts(int* __restrict__ c) {  // the 4 x kQSub counters
    for (int k = (int)threadIdx.x; k < 4 * kQSub; k += (int)blockDim.x) c[k * kQStride] = 0;
}

// U0 + dt * (flux difference) and the floors with the old faces' cell-centred field (M3:802-834), in FP64.
// flux(ax, side, c) is component c (conserved order) of the flux through the cell's lower (side 0) or
// upper (side 1) face of axis ax.
template <typename FluxAt>
__device__ __forceinline__ void cell_update3d(const Geom3D& g, const Eos3D& e, double dt, const double u0[5],
                                              const FluxAt& flux, double bx, double by, double bz, double u[5]) {
    const double invdx = 1.0 / g.dx, invdy = 1.0 / g.dy, invdz = 1.0 / g.dz;
    double d[5];
#pragma unroll
    for (int c = 0; c < 5; c++)
        d[c] = -(flux(0, 1, c) - flux(0, 0, c)) * invdx - (flux(1, 1, c) - flux(1, 0, c)) * invdy -
               (flux(2, 1, c) - flux(2, 0, c)) * invdz;
    double rho = u0[0] + dt * d[0], mx = u0[1] + dt * d[1], my = u0[2] + dt * d[2];
    double mz = u0[3] + dt * d[3], E = u0[4] + dt * d[4];
    floors3d(e, rho, mx, my, mz, E, bx, by, bz);
    u[0] = rho; u[1] = mx; u[2] = my; u[3] = mz; u[4] = E;
}

// The update of cell (i, j, k) from the stored face fluxes.
template <typename R>
__device__ __forceinline__ void update_cell_at3d(const Geom3D& g, const Eos3D& e, const State3D& U0,
                                                 const FaceFlux<R>& FX, const FaceFlux<R>& FY, const FaceFlux<R>& FZ,
                                                 double dt, int i, int j, int k, State3D& Us) {
    const size_t q = ((size_t)k * g.Ny + j) * g.Nx + i;
    const size_t xL = fx_(g, i, j, k), yD = fy_(g, i, j, k), zB = fz_(g, i, j, k);
    const size_t xo[2] = {xL, xL + 1}, yo[2] = {yD, yD + g.Nx}, zo[2] = {zB, zB + (size_t)g.Nx * g.Ny};
    const double u0[5] = {U0.rho[q], U0.mx[q], U0.my[q], U0.mz[q], U0.E[q]};
    double u[5];
    cell_update3d(
        g, e, dt, u0,
        [&](int ax, int side, int c) {
            return ax == 0 ? (double)FX.f[c][xo[side]] : ax == 1 ? (double)FY.f[c][yo[side]] : (double)FZ.f[c][zo[side]];
        },
        bxc(g, U0.Bx, i, j, k), byc(g, U0.By, i, j, k), bzc(g, U0.Bz, i, j, k), u);
    Us.rho[q] = u[0]; Us.mx[q] = u[1]; Us.my[q] = u[2]; Us.mz[q] = u[3]; Us.E[q] = u[4];
}

// Flux difference, conserved update and floors with the old faces (M3:802-834), in FP64.
template <typename R, bool IDEAL>
__global__ __launch_bounds__(256) void k_update3d(Geom3D g, Eos3D e_in, State3D U0, FaceFlux<R> FX, FaceFlux<R> FY,
                                                  FaceFlux<R> FZ, const Scalars* __restrict__ sc, State3D Us) {
    const Eos3D e = eos_of<IDEAL>(e_in);
    const unsigned qq = blockIdx.x * 256u + threadIdx.x;  // cells < 2^31 (checked at create)
    if (qq >= (unsigned)g.Nx * (unsigned)g.Ny * (unsigned)g.Nz) return;
    const int i = (int)(qq % (unsigned)g.Nx);
    const unsigned r = qq / (unsigned)g.Nx;
    const int j = (int)(r % (unsigned)g.Ny), k = (int)(r / (unsigned)g.Ny);
    update_cell_at3d<R>(g, e, U0, FX, FY, FZ, sc->dt, i, j, k, Us);
}

// ------------------------------------------------------------------ z faces + update in one pass
// One face of axis AX from PLM states in the flux precision R: the axis-frame rotation of face_states3d, the
// HLLD (FP32: also whether the FP64 solver must redo the face), the fluxes in conserved order.
template <int AX, typename R, typename EOSR>
__device__ __forceinline__ bool solve_face3d(const Geom3D& g, const EOSR& eos, const Prim<R>& WL, const Prim<R>& WR,
                                             double Bn, R o[7], R f[5]) {
    Prim<R> L = WL, Rr = WR;
    if (AX == 1) {
        L = Prim<R>{WL.rho, WL.v, WL.u, WL.w, WL.P, (R)Bn, WL.Bx, WL.Bz};
        Rr = Prim<R>{WR.rho, WR.v, WR.u, WR.w, WR.P, (R)Bn, WR.Bx, WR.Bz};
    } else if (AX == 2) {
        L = Prim<R>{WL.rho, WL.w, WL.u, WL.v, WL.P, (R)Bn, WL.Bx, WL.By};
        Rr = Prim<R>{WR.rho, WR.w, WR.u, WR.v, WR.P, (R)Bn, WR.Bx, WR.By};
    }
    bool deg = false;
    if constexpr (std::is_same<R, float>::value)
        hlld_x<float>(eos, L, Rr, (float)Bn, fabs(Bn) < 1e-12, o, &deg, g.tau, (float)g.hlld_guard);
    else
        hlld_x<double>(eos, L, Rr, Bn, fabs(Bn) < 1e-12, o, nullptr, 0.0, g.hlld_guard);
    face_fluxes3d<AX, R>(o, f);
    return deg;
}

// The z faces and the update of a stage in one column march (z faces: M3:777-800, update: M3:802-834), in
// place of k_face3d<2> + k_update3d.  One thread per (i, j) column marches over a chunk of planes: at plane s
// it forms PLM-MC of cell s+1 from the primitives of cells s, s+1 (carried) and s+2 (loaded one plane ahead),
// solves z face s+1 between the carried plus state of cell s and the minus state of cell s+1, and updates
// cell s from its two z-face fluxes (registers) and the x and y face fluxes the face kernels stored (final:
// in the mixed path the FP64 redo of flagged x and y faces runs before).  The z fluxes never reach memory.
// The march starts two planes early, so that a chunk's bottom face (also the top face of the chunk below)
// comes from the same code as every other z face.  Mixed: the chunk that owns a z face the FP32 HLLD flags
// queues it; a cell with such a face stores its two z-face fluxes and is queued for k_update3d_cells, which
// redoes it after k_face3d_fp64 has stored the FP64 flux.
template <bool MIXED, typename EOSR, bool IDEAL>
__global__ __launch_bounds__(256) void k_zupd3d(Geom3D g, Eos3D e_in, EOSR eos,
                                                PrimArrays<typename std::conditional<MIXED, float, double>::type> W,
                                                State3D U0, const Scalars* __restrict__ sc, State3D Us,
                                                FaceFlux<typename std::conditional<MIXED, float, double>::type> FX,
                                                FaceFlux<typename std::conditional<MIXED, float, double>::type> FY,
                                                FaceFlux<typename std::conditional<MIXED, float, double>::type> FZ,
                                                EmfArr<typename std::conditional<MIXED, float, double>::type> ef,
                                                DegQueue dq, int zchunk) {
    using R = typename std::conditional<MIXED, float, double>::type;
    const Eos3D e = eos_of<IDEAL>(e_in);
    const int Nx = g.Nx, Ny = g.Ny, Nz = g.Nz;
    const int i = (int)blockIdx.x * 64 + (int)(threadIdx.x & 63), j = (int)blockIdx.y * 4 + (int)(threadIdx.x >> 6);
    if (i >= Nx || j >= Ny) return;  // no barriers below
    const int c0 = (int)blockIdx.z * zchunk, c1 = min(Nz, c0 + zchunk);
    const size_t NXY = (size_t)Nx * Ny, NX1Y = (size_t)(Nx + 1) * Ny, NXY1 = (size_t)Nx * (Ny + 1);
    const size_t coff = (size_t)j * Nx + i, xoff = (size_t)j * (Nx + 1) + i;
    const R frho = (R)g.floor_rho, fP = (R)g.floor_P;
    const double dt = sc->dt;
    auto prim = [&](int k) { return load_prim(W, (size_t)bmap(k, Nz, g.pz) * NXY + coff); };

    Prim<R> wp{};   // plus state of cell s
    R fzb[5] = {};  // flux of z face s
    bool dzb = false;
    double bzb = 0.0;
    for (int s = c0 - 2; s < c1; ++s) {
        // cells s and s+1 were loaded by the previous plane's step (cache hits): reloading them costs less
        // than carrying three primitive sets through the solver
        Prim<R> mn, wpn;
        plm_mc<R>(prim(s), prim(s + 1), prim(s + 2), 2, frho, fP, mn, wpn);  // minus and plus states of cell s+1
        if (s >= c0 - 1) {
            const size_t zf = (size_t)(s + 1) * NXY;
            const double bzt = U0.Bz[zf + coff];
            R o[7], f[5];
            const bool d = solve_face3d<2, R>(g, eos, wp, mn, bzt, o, f);
            const bool own = s + 1 < c1 || s + 1 == Nz;  // z faces c0 .. c1-1 (and Nz) belong to this chunk
            const bool defer = MIXED && s >= c0 && (dzb || d);
            if constexpr (MIXED) queue_push(dq, 3, defer, (int)((size_t)s * NXY + coff));
            if (s >= c0) {  // cell s
                const size_t zc = (size_t)s * NXY, q = zc + coff;
                if (defer) {
#pragma unroll
                    for (int k = 0; k < 5; ++k) {
                        FZ.f[k][q] = fzb[k];
                        FZ.f[k][zf + coff] = f[k];
                    }
                } else {
                    const size_t xL = (size_t)s * NX1Y + xoff, yD = (size_t)s * NXY1 + coff;
                    const double u0[5] = {U0.rho[q], U0.mx[q], U0.my[q], U0.mz[q], U0.E[q]};
                    double u[5];
                    cell_update3d(
                        g, e, dt, u0,
                        [&](int ax, int side, int k) {
                            return ax == 0 ? (double)FX.f[k][xL + side]
                                           : ax == 1 ? (double)FY.f[k][yD + (side ? (size_t)Nx : 0)]
                                                     : (double)(side ? f[k] : fzb[k]);
                        },
                        0.5 * (U0.Bx[xL] + U0.Bx[xL + 1]), 0.5 * (U0.By[yD] + U0.By[yD + Nx]), 0.5 * (bzb + bzt), u);
                    Us.rho[q] = u[0]; Us.mx[q] = u[1]; Us.my[q] = u[2]; Us.mz[q] = u[3]; Us.E[q] = u[4];
                }
            }
            if (own) face_emfs3d<2, R>(ef, zf + coff, o);  // after the update: before its loads it holds registers
            if constexpr (MIXED) queue_push(dq, 2, own && d, (int)(zf + coff));
#pragma unroll
            for (int k = 0; k < 5; ++k) fzb[k] = f[k];
            dzb = d;
            bzb = bzt;
        }
        wp = wpn;
    }
}

// The update of the cells k_zupd3d queued (those with a z face the FP32 HLLD flagged), after k_face3d_fp64 has
// stored the FP64 fluxes of the flagged faces beside the FP32 ones the cell wrote.
template <bool IDEAL>
__global__ __launch_bounds__(256) void k_update3d_cells(Geom3D g, Eos3D e_in, State3D U0, FaceFlux<float> FX,
                                                        FaceFlux<float> FY, FaceFlux<float> FZ,
                                                        const Scalars* __restrict__ sc, State3D Us, DegQueue dq) {
    const Eos3D e = eos_of<IDEAL>(e_in);
    queue_for_each(dq, 3, [&](int c) {
        const unsigned qq = (unsigned)c;
        const int i = (int)(qq % (unsigned)g.Nx);
        const unsigned r = qq / (unsigned)g.Nx;
        const int j = (int)(r % (unsigned)g.Ny), k = (int)(r / (unsigned)g.Ny);
        update_cell_at3d<float>(g, e, U0, FX, FY, FZ, sc->dt, i, j, k, Us);
    });
}

// ------------------------------------------------------------------ edge EMFs (M3:836-939)
// Face EMFs at boundary-mapped indices (the reference's wrap of face and cell indices).
template <typename R>
struct EmfAt {
    const Geom3D& g;
    const EmfArr<R>& ef;
    __device__ __forceinline__ double exy(int i, int j, int k) const { return (double)ef.Ex_y[fy_(g, bmap(i, g.Nx, g.px), fmap(j, g.Ny, g.py), bmap(k, g.Nz, g.pz))]; }
    __device__ __forceinline__ double ezx(int i, int j, int k) const { return (double)ef.Ez_x[fx_(g, fmap(i, g.Nx, g.px), bmap(j, g.Ny, g.py), bmap(k, g.Nz, g.pz))]; }
    __device__ __forceinline__ double eyx(int i, int j, int k) const { return (double)ef.Ey_x[fx_(g, fmap(i, g.Nx, g.px), bmap(j, g.Ny, g.py), bmap(k, g.Nz, g.pz))]; }
    __device__ __forceinline__ double exz(int i, int j, int k) const { return (double)ef.Ex_z[fz_(g, bmap(i, g.Nx, g.px), bmap(j, g.Ny, g.py), fmap(k, g.Nz, g.pz))]; }
    __device__ __forceinline__ double eyz(int i, int j, int k) const { return (double)ef.Ey_z[fz_(g, bmap(i, g.Nx, g.px), bmap(j, g.Ny, g.py), fmap(k, g.Nz, g.pz))]; }
    __device__ __forceinline__ double ezy(int i, int j, int k) const { return (double)ef.Ez_y[fy_(g, bmap(i, g.Nx, g.px), fmap(j, g.Ny, g.py), bmap(k, g.Nz, g.pz))]; }
};

// The three edge EMFs: 1/4 of the four adjacent face EMFs + eta * curl of the cell-centred B0, with the
// reference's index handling (M3:836-939).  The resistive parts take the cell-centred differences as
// arguments so that callers holding the cell-centred field can share it.
__device__ __forceinline__ double edge_sum(double ideal, double eta, double d1, double d2) {
    double res = 0.0;
    if (eta > 0.0) res = eta * (d1 - d2);
    return ideal + res;
}
template <typename R>
__device__ __forceinline__ double edge_ex(const Geom3D& g, const EmfArr<R>& ef, const State3D& B0, int i, int jf, int kf) {
    const EmfAt<R> m{g, ef};
    const double ex_ideal = 0.25 * (m.exy(i, jf, kf - 1) + m.exy(i, jf, kf) + m.exz(i, jf - 1, kf) + m.exz(i, jf, kf));
    double dBzdy = 0.0, dBydz = 0.0;
    if (g.eta > 0.0) {
        dBzdy = (bzc(g, B0.Bz, i, jf, kf) - bzc(g, B0.Bz, i, jf - 1, kf)) / g.dy;
        dBydz = (byc(g, B0.By, i, jf, kf) - byc(g, B0.By, i, jf, kf - 1)) / g.dz;
    }
    return edge_sum(ex_ideal, g.eta, dBzdy, dBydz);
}
template <typename R>
__device__ __forceinline__ double edge_ey(const Geom3D& g, const EmfArr<R>& ef, const State3D& B0, int i, int j, int kf) {
    const EmfAt<R> m{g, ef};
    const double ey_ideal = 0.25 * (m.eyx(i, j, kf - 1) + m.eyx(i, j, kf) + m.eyz(i - 1, j, kf) + m.eyz(i, j, kf));
    double dBxdz = 0.0, dBzdx = 0.0;
    if (g.eta > 0.0) {
        dBxdz = (bxc(g, B0.Bx, i, j, kf) - bxc(g, B0.Bx, i, j, kf - 1)) / g.dz;
        dBzdx = (bzc(g, B0.Bz, i, j, kf) - bzc(g, B0.Bz, i - 1, j, kf)) / g.dx;
    }
    return edge_sum(ey_ideal, g.eta, dBxdz, dBzdx);
}
template <typename R>
__device__ __forceinline__ double edge_ez(const Geom3D& g, const EmfArr<R>& ef, const State3D& B0, int i, int jf, int k) {
    const EmfAt<R> m{g, ef};
    const double ez_ideal = 0.25 * (m.ezx(i, jf - 1, k) + m.ezx(i, jf, k) + m.ezy(i - 1, jf, k) + m.ezy(i, jf, k));
    double dBydx = 0.0, dBxdy = 0.0;
    if (g.eta > 0.0) {
        dBydx = (byc(g, B0.By, i, jf, k) - byc(g, B0.By, i - 1, jf, k)) / g.dx;
        dBxdy = (bxc(g, B0.Bx, i, jf, k) - bxc(g, B0.Bx, i, jf - 1, k)) / g.dy;
    }
    return edge_sum(ez_ideal, g.eta, dBydx, dBxdy);
}

template <typename R>
__global__ __launch_bounds__(256) void k_edge3d(Geom3D g, EmfArr<R> ef, State3D B0, Edges ed) {
    const int ia = blockIdx.x * 64 + (threadIdx.x & 63), ja = blockIdx.y, ka = blockIdx.z * 4 + (threadIdx.x >> 6);
    const int Nx = g.Nx, Ny = g.Ny, Nz = g.Nz;
    if (ia > Nx || ka > Nz) return;
    if (ia < Nx) ed.Ex[((size_t)ka * (Ny + 1) + ja) * Nx + ia] = edge_ex(g, ef, B0, ia, ja, ka);          // Ex(i, jf, kf)
    if (ja < Ny) ed.Ey[((size_t)ka * Ny + ja) * (Nx + 1) + ia] = edge_ey(g, ef, B0, ia, ja, ka);          // Ey(if, j, kf)
    if (ka < Nz) ed.Ez[((size_t)ka * (Ny + 1) + ja) * (Nx + 1) + ia] = edge_ez(g, ef, B0, ia, ja, ka);    // Ez(if, jf, k)
}

// ------------------------------------------------------------------ CT (M3:941-1015)
// Face-field increments from the edge EMFs around the face.
__device__ __forceinline__ double ct_dbx(const Geom3D& g, double dt, double ez_j, double ez_j1, double ey_k, double ey_k1) {
    return -(dt / g.dy) * (ez_j1 - ez_j) + (dt / g.dz) * (ey_k1 - ey_k);
}
__device__ __forceinline__ double ct_dby(const Geom3D& g, double dt, double ez_i, double ez_i1, double ex_k, double ex_k1) {
    return +(dt / g.dx) * (ez_i1 - ez_i) - (dt / g.dz) * (ex_k1 - ex_k);
}
__device__ __forceinline__ double ct_dbz(const Geom3D& g, double dt, double ey_i, double ey_i1, double ex_j, double ex_j1) {
    return -(dt / g.dx) * (ey_i1 - ey_i) + (dt / g.dy) * (ex_j1 - ex_j);
}
// B1 = B0 + increment: the reference's sum, or (compensated) the exact sum carried as hi + FP32 lo.
__device__ __forceinline__ void ct_new(bool comp, double b0, float b0lo, double d, double& hi, float& lo) {
    if (comp) {
        ct_add(b0, b0lo, d, hi, lo);
    } else {
        hi = b0 + d;
        lo = 0.f;
    }
}

__global__ __launch_bounds__(256) void k_ct3d(Geom3D g, Edges ed, State3D B0, State3D B1, const Scalars* __restrict__ sc) {
    const int ia = blockIdx.x * 64 + (threadIdx.x & 63), ja = blockIdx.y, ka = blockIdx.z * 4 + (threadIdx.x >> 6);
    const int Nx = g.Nx, Ny = g.Ny, Nz = g.Nz;
    if (ia > Nx || ka > Nz) return;
    const double dt = sc->dt;
    auto EX = [&](int i, int jf, int kf) { return ed.Ex[((size_t)kf * (Ny + 1) + jf) * Nx + i]; };
    auto EY = [&](int i, int j, int kf) { return ed.Ey[((size_t)kf * Ny + j) * (Nx + 1) + i]; };
    auto EZ = [&](int i, int jf, int k) { return ed.Ez[((size_t)k * (Ny + 1) + jf) * (Nx + 1) + i]; };
    // face indices j+1 <= Ny and k+1 <= Nz need no wrapping (imod(n, n+1) = n, clamp(n, 0, n) = n)
    auto put = [&](double* B1v, float* B1lo, const double* B0v, const float* B0lo, size_t f, double d) {
        double hi;
        float lo;
        ct_new(g.ct_comp, B0v[f], g.ct_comp ? B0lo[f] : 0.f, d, hi, lo);
        B1v[f] = hi;
        if (g.ct_comp) B1lo[f] = lo;
    };
    if (ja < Ny && ka < Nz) {
        const int i = ia, j = ja, k = ka;
        put(B1.Bx, B1.Bx_lo, B0.Bx, B0.Bx_lo, fx_(g, i, j, k), ct_dbx(g, dt, EZ(i, j, k), EZ(i, j + 1, k), EY(i, j, k), EY(i, j, k + 1)));
    }
    if (ia < Nx && ka < Nz) {
        const int i = ia, j = ja, k = ka;
        put(B1.By, B1.By_lo, B0.By, B0.By_lo, fy_(g, i, j, k), ct_dby(g, dt, EZ(i, j, k), EZ(i + 1, j, k), EX(i, j, k), EX(i, j, k + 1)));
    }
    if (ia < Nx && ja < Ny) {
        const int i = ia, j = ja, k = ka;
        put(B1.Bz, B1.Bz_lo, B0.Bz, B0.Bz_lo, fz_(g, i, j, k), ct_dbz(g, dt, EY(i, j, k), EY(i + 1, j, k), EX(i, j, k), EX(i, j + 1, k)));
    }
}

// ------------------------------------------------------------------ edge EMFs + CT in one pass
// The edge EMFs and CT (M3:836-1015) without the edge arrays.  A block owns a TX x TY column of cells and
// marches along z over a chunk of planes; position (xs, ys) in [0, TX+2) x [0, TY+2) holds (a, b) =
// (i0-1+xs, j0-1+ys).  Per plane s of the march:
//   (1) every position evaluates the cell-centred B0 (bxc, byc, bzc) at raw (a, b, s+1) with the
//       reference's index maps and shares it through LDS;
//   (2) positions xs, ys >= 1 evaluate Ex, Ey at z-face s+1 and Ez at plane s with the reference's
//       expressions: the -x / -y cell-centred values come from LDS, those of plane s from registers;
//   (3) positions owning faces (the tile's, and those of index Nx / Ny / Nz on the grid's high sides)
//       form Bx, By of plane s and Bz of z-face s+1 from their own edges and those of the +x / +y
//       neighbour (LDS), and write them.
// Every edge is evaluated once per tile plus a one-position halo, and neither the edges nor the face
// EMFs of a plane are read from memory twice.
template <int TX, int TY>
struct EctShape {
    static constexpr int PX = TX + 2, PY = TY + 2, NP = PX * PY, NT = NP <= 512 ? 512 : 1024;
    static_assert(NP <= 1024, "tile positions exceed the block");
};

template <typename R, int TX, int TY, int MINW = 1>
__global__ __launch_bounds__((EctShape<TX, TY>::NT), MINW) void k_ect3d(Geom3D g, EmfArr<R> ef, State3D B0, State3D B1,
                                                                const Scalars* __restrict__ sc, int zchunk) {
    constexpr int PX = EctShape<TX, TY>::PX, NP = EctShape<TX, TY>::NP;
    __shared__ double s_cbx[NP], s_cby[NP], s_cbz[NP];  // bxc, byc of plane s; bzc of plane s+1
    __shared__ double s_ez[NP], s_ey[NP], s_ex[NP];     // Ez of plane s; Ey, Ex of z-face s+1
    const int Nx = g.Nx, Ny = g.Ny, Nz = g.Nz;
    const bool comp = g.ct_comp != 0;
    const int ntx = (Nx + TX - 1) / TX;
    const int i0 = (int)(blockIdx.x % (unsigned)ntx) * TX, j0 = (int)(blockIdx.x / (unsigned)ntx) * TY;
    const int c0 = (int)blockIdx.y * zchunk, c1 = min(Nz, c0 + zchunk);
    const bool act = threadIdx.x < (unsigned)NP;
    const int p = act ? (int)threadIdx.x : 0;
    const int xs = p % PX, ys = p / PX;
    const int a = i0 - 1 + xs, b = j0 - 1 + ys;
    const bool has_edge = act && xs >= 1 && ys >= 1;
    const bool own_bx = act && ys >= 1 && ys <= TY && b < Ny && xs >= 1 && a <= Nx && (xs <= TX || a == Nx);
    const bool own_by = act && xs >= 1 && xs <= TX && a < Nx && ys >= 1 && b <= Ny && (ys <= TY || b == Ny);
    const bool own_bz = act && xs >= 1 && xs <= TX && a < Nx && ys >= 1 && ys <= TY && b < Ny;
    const double dt = sc->dt, eta = g.eta;
    // Boundary-mapped indices split into an x/y part per position (fixed along the march) and a z part per
    // plane (uniform over the block): fx_ = (k Ny + j)(Nx+1) + i with i face-mapped and j, k cell-mapped,
    // fy_ = (k (Ny+1) + j) Nx + i with j face-mapped, fz_ = (k Ny + j) Nx + i with k face-mapped.
    const int NX1 = Nx + 1, NY1 = Ny + 1;
    const int fa = fmap(a, Nx, g.px), fa1 = fmap(a + 1, Nx, g.px), ca = bmap(a, Nx, g.px), cam = bmap(a - 1, Nx, g.px);
    const int fb = fmap(b, Ny, g.py), fb1 = fmap(b + 1, Ny, g.py), cb = bmap(b, Ny, g.py), cbm = bmap(b - 1, Ny, g.py);
    const int ox_ab = cb * NX1 + fa, ox_a1b = cb * NX1 + fa1, ox_abm = cbm * NX1 + fa;
    const int oy_ab = fb * Nx + ca, oy_ab1 = fb1 * Nx + ca, oy_amb = fb * Nx + cam;
    const int oz_ab = cb * Nx + ca, oz_abm = cbm * Nx + ca, oz_amb = cb * Nx + cam;
    auto zx = [&](int k) { return (size_t)bmap(k, Nz, g.pz) * ((size_t)Ny * NX1); };
    auto zy = [&](int k) { return (size_t)bmap(k, Nz, g.pz) * ((size_t)NY1 * Nx); };
    auto zz = [&](int kf) { return (size_t)fmap(kf, Nz, g.pz) * ((size_t)Ny * Nx); };

    // Global values of one march step s: B0 faces of plane s+1 (the cell-centred field), the face EMFs
    // Ex_y, Ey_x of plane s+1, Ex_z, Ey_z of z-face s+1, Ez_x, Ez_y of plane s, and the low words of the
    // faces this position updates.  Loaded one step ahead.
    struct StepIn {
        double bxa, bxb, bya, byb, bza, bzb;
        R exy, eyx, exzm, exz, eyzm, eyz, ezxm, ezx, ezym, ezy;  // face EMFs in the flux precision
        float lbx, lby, lbz;
    };
    auto load_step = [&](int s, StepIn& L) {
        const int pn = s + 1;
        const size_t xp = zx(pn), yp = zy(pn), zp = zz(pn), zp1 = zz(pn + 1), xs0 = zx(s), ys0 = zy(s);
        L.bxa = B0.Bx[xp + ox_ab]; L.bxb = B0.Bx[xp + ox_a1b];
        L.bya = B0.By[yp + oy_ab]; L.byb = B0.By[yp + oy_ab1];
        L.bza = B0.Bz[zp + oz_ab]; L.bzb = B0.Bz[zp1 + oz_ab];
        L.exy = ef.Ex_y[yp + oy_ab]; L.eyx = ef.Ey_x[xp + ox_ab];
        L.exzm = ef.Ex_z[zp + oz_abm]; L.exz = ef.Ex_z[zp + oz_ab];
        L.eyzm = ef.Ey_z[zp + oz_amb]; L.eyz = ef.Ey_z[zp + oz_ab];
        L.ezxm = ef.Ez_x[xs0 + ox_abm]; L.ezx = ef.Ez_x[xs0 + ox_ab];
        L.ezym = ef.Ez_y[ys0 + oy_amb]; L.ezy = ef.Ez_y[ys0 + oy_ab];
        L.lbx = L.lby = L.lbz = 0.f;
        if (comp) {  // own faces are in range: the maps are the identity there
            if (own_bx && s >= c0) L.lbx = B0.Bx_lo[xs0 + ox_ab];
            if (own_by && s >= c0) L.lby = B0.By_lo[ys0 + oy_ab];
            if (own_bz) L.lbz = B0.Bz_lo[zp + oz_ab];
        }
    };

    // carried along z: cell-centred B0 and the face values of plane s, the face EMFs Ex_y, Ey_x of plane
    // s, the edges Ex, Ey of z-face s
    double cbx_o, cby_o, bxa_o, bya_o, exy_o, eyx_o;
    {
        const int q = c0 - 1;
        const size_t xq = zx(q), yq = zy(q);
        bxa_o = B0.Bx[xq + ox_ab];
        bya_o = B0.By[yq + oy_ab];
        cbx_o = 0.5 * (bxa_o + B0.Bx[xq + ox_a1b]);
        cby_o = 0.5 * (bya_o + B0.By[yq + oy_ab1]);
        exy_o = (double)ef.Ex_y[yq + oy_ab];
        eyx_o = (double)ef.Ey_x[xq + ox_ab];
    }
    double ex_c = 0.0, ey_c = 0.0;
    StepIn cur;
    load_step(c0 - 1, cur);

    for (int s = c0 - 1; s < c1; ++s) {
        const int pn = s + 1;  // z-face and cell plane s+1 (<= c1 <= Nz)
        // (1) cell-centred B0 of plane s+1
        const double cbx_n = 0.5 * (cur.bxa + cur.bxb), cby_n = 0.5 * (cur.bya + cur.byb),
                     cbz_n = 0.5 * (cur.bza + cur.bzb);
        if (act) {
            s_cbx[p] = cbx_o;
            s_cby[p] = cby_o;
            s_cbz[p] = cbz_n;
        }
        __syncthreads();
        StepIn nxt;
        if (s + 1 < c1) load_step(s + 1, nxt);
        // (2) edges: Ex, Ey of z-face s+1, Ez of plane s
        double ex_n = 0.0, ey_n = 0.0, ez = 0.0;
        if (has_edge) {
            const double exi = 0.25 * (exy_o + (double)cur.exy + (double)cur.exzm + (double)cur.exz);
            const double eyi = 0.25 * (eyx_o + (double)cur.eyx + (double)cur.eyzm + (double)cur.eyz);
            double dBzdy = 0.0, dBydz = 0.0, dBxdz = 0.0, dBzdx = 0.0;
            if (eta > 0.0) {
                dBzdy = (cbz_n - s_cbz[p - PX]) / g.dy;
                dBydz = (cby_n - cby_o) / g.dz;
                dBxdz = (cbx_n - cbx_o) / g.dz;
                dBzdx = (cbz_n - s_cbz[p - 1]) / g.dx;
            }
            ex_n = edge_sum(exi, eta, dBzdy, dBydz);
            ey_n = edge_sum(eyi, eta, dBxdz, dBzdx);
            if (s >= c0) {
                const double ezi = 0.25 * ((double)cur.ezxm + (double)cur.ezx + (double)cur.ezym + (double)cur.ezy);
                double dBydx = 0.0, dBxdy = 0.0;
                if (eta > 0.0) {
                    dBydx = (cby_o - s_cby[p - 1]) / g.dx;
                    dBxdy = (cbx_o - s_cbx[p - PX]) / g.dy;
                }
                ez = edge_sum(ezi, eta, dBydx, dBxdy);
            }
        }
        if (act) {
            s_ez[p] = ez;
            s_ey[p] = ey_n;
            s_ex[p] = ex_n;
        }
        __syncthreads();
        // (3) the face field: Bx, By of plane s, Bz of z-face s+1 (z-face c1 belongs to the next chunk)
        if (s >= c0) {
            if (own_bx) {
                const size_t f = zx(s) + ox_ab;
                double hi;
                float lo;
                ct_new(comp, bxa_o, cur.lbx, ct_dbx(g, dt, ez, s_ez[p + PX], ey_c, ey_n), hi, lo);
                B1.Bx[f] = hi;
                if (comp) B1.Bx_lo[f] = lo;
            }
            if (own_by) {
                const size_t f = zy(s) + oy_ab;
                double hi;
                float lo;
                ct_new(comp, bya_o, cur.lby, ct_dby(g, dt, ez, s_ez[p + 1], ex_c, ex_n), hi, lo);
                B1.By[f] = hi;
                if (comp) B1.By_lo[f] = lo;
            }
        }
        if (own_bz && (pn < c1 || pn == Nz)) {
            const size_t f = zz(pn) + oz_ab;
            double hi;
            float lo;
            ct_new(comp, cur.bza, cur.lbz, ct_dbz(g, dt, ey_n, s_ey[p + 1], ex_n, s_ex[p + PX]), hi, lo);
            B1.Bz[f] = hi;
            if (comp) B1.Bz_lo[f] = lo;
        }
        cbx_o = cbx_n; cby_o = cby_n; bxa_o = cur.bxa; bya_o = cur.bya;
        exy_o = (double)cur.exy; eyx_o = (double)cur.eyx; ex_c = ex_n; ey_c = ey_n;
        cur = nxt;
    }
}

// ------------------------------------------------------------------ Joule + floors (+ Heun)
// The cell kernels that end a stage or reduce over cells march along z: one thread per (i, j) column of a
// chunk of zchunk planes, 64 i x 4 j per block.  Their per-block maxima therefore number
// col_blocks = ceil(Nx/64) ceil(Ny/4) ceil(Nz/zchunk) for every writer (speed, Joule, divB).
struct ColGrid {
    int i, j, c0, c1;
    __device__ ColGrid(const Geom3D& g, int zchunk)
        : i((int)blockIdx.x * 64 + (int)(threadIdx.x & 63)), j((int)blockIdx.y * 4 + (int)(threadIdx.x >> 6)),
          c0((int)blockIdx.z * zchunk), c1(min(g.Nz, (int)blockIdx.z * zchunk + zchunk)) {}
    __device__ bool valid(const Geom3D& g) const { return i < g.Nx && j < g.Ny; }
};

// Joule heating from the central differences of the cell-centred B1 (add_joule_heating, M3:652-692),
// floors (M3:1019-1029) and in stage 2 the Heun average with its floors (M3:1042-1065) and the next
// step's signal-speed maximum.  Along the column the cell-centred B1 of planes k-1, k, k+1 and the z faces
// k, k+1 of the column and of its four x/y neighbours stay in registers; the boundary-mapped face indices
// are split into a per-thread x/y part and a per-plane z part, so a plane costs 17 face loads and no index
// maps.  Same expressions as the reference, in the same order.
// The primitives of the result (cons_to_prim on its face-averaged B, M3:298-308) are stored in the flux
// precision for the next stage's face kernels, which k_prim3d would otherwise recompute from the result.
// The stage-2 instance holds more live values (Heun average, compensated faces): it runs at 2 waves per SIMD
// without spilling rather than at 3 with.
template <typename R, bool FINAL, bool IDEAL>
__global__ __launch_bounds__(256, FINAL ? 2 : 4) void k_joule3d(Geom3D g, Eos3D e_in, State3D Us, State3D B1,
                                                 const Scalars* __restrict__ sc, State3D Un, State3D Out,
                                                 double* __restrict__ spd, int zchunk, PrimArrays<R> W) {
    const Eos3D e = eos_of<IDEAL>(e_in);
    const ColGrid cg(g, zchunk);
    const int i = cg.i, j = cg.j, Nx = g.Nx, Ny = g.Ny, Nz = g.Nz;
    const bool comp = g.ct_comp != 0;
    double speed = 0.0;
    if (cg.valid(g)) {
        const double dt = sc->dt;
        const int NX1 = Nx + 1, NY1 = Ny + 1;
        const int im = bmap(i - 1, Nx, g.px), ip = bmap(i + 1, Nx, g.px), jm = bmap(j - 1, Ny, g.py), jp = bmap(j + 1, Ny, g.py);
        // x faces (fx_, j cell-mapped): rows j, j-1, j+1 at faces i and i+1; y faces (fy_, i cell-mapped):
        // columns i, i-1, i+1 at faces j and j+1; z faces (fz_): (i, j) and its four x/y neighbours
        const int ox0 = j * NX1 + i, oxm = jm * NX1 + i, oxp = jp * NX1 + i;
        const int oy0 = j * Nx + i, oym = j * Nx + im, oyp = j * Nx + ip;
        const int oz0 = j * Nx + i, ozjm = jm * Nx + i, ozjp = jp * Nx + i, ozim = j * Nx + im, ozip = j * Nx + ip;
        const size_t sx = (size_t)Ny * NX1, sy = (size_t)NY1 * Nx, sz = (size_t)Ny * Nx;
        auto bxc_at = [&](int k) {
            const size_t o = (size_t)bmap(k, Nz, g.pz) * sx + ox0;
            return 0.5 * (B1.Bx[o] + B1.Bx[o + 1]);
        };
        auto byc_at = [&](int k) {
            const size_t o = (size_t)bmap(k, Nz, g.pz) * sy + oy0;
            return 0.5 * (B1.By[o] + B1.By[o + Nx]);
        };
        const int c0 = cg.c0, c1 = cg.c1;
        double bxc_m = bxc_at(c0 - 1), byc_m = byc_at(c0 - 1), bxc_0 = bxc_at(c0), byc_0 = byc_at(c0);
        size_t zf = (size_t)c0 * sz;  // z face k (face indices k, k+1 <= Nz need no map)
        double bz0 = B1.Bz[zf + oz0], bzjm = B1.Bz[zf + ozjm], bzjp = B1.Bz[zf + ozjp], bzim = B1.Bz[zf + ozim],
               bzip = B1.Bz[zf + ozip];
        for (int k = c0; k < c1; ++k) {
            const size_t zf1 = zf + sz;
            const double bxc_p = bxc_at(k + 1), byc_p = byc_at(k + 1);
            const double bz1 = B1.Bz[zf1 + oz0], bz1jm = B1.Bz[zf1 + ozjm], bz1jp = B1.Bz[zf1 + ozjp],
                         bz1im = B1.Bz[zf1 + ozim], bz1ip = B1.Bz[zf1 + ozip];
            const size_t xk = (size_t)k * sx, yk = (size_t)k * sy, q = zf + oz0;  // q = ((k Ny + j) Nx + i)
            double rho = Us.rho[q], mx = Us.mx[q], my = Us.my[q], mz = Us.mz[q], E = Us.E[q];
            if (g.eta > 0.0) {  // add_joule_heating, M3:652-692
                const double dBzdy = (0.5 * (bzjp + bz1jp) - 0.5 * (bzjm + bz1jm)) / (2 * g.dy);
                const double dBydz = (byc_p - byc_m) / (2 * g.dz);
                const double dBxdz = (bxc_p - bxc_m) / (2 * g.dz);
                const double dBzdx = (0.5 * (bzip + bz1ip) - 0.5 * (bzim + bz1im)) / (2 * g.dx);
                const double dBydx = (0.5 * (B1.By[yk + oyp] + B1.By[yk + oyp + Nx]) -
                                      0.5 * (B1.By[yk + oym] + B1.By[yk + oym + Nx])) / (2 * g.dx);
                const double dBxdy = (0.5 * (B1.Bx[xk + oxp] + B1.Bx[xk + oxp + 1]) -
                                      0.5 * (B1.Bx[xk + oxm] + B1.Bx[xk + oxm + 1])) / (2 * g.dy);
                const double Jx = dBzdy - dBydz, Jy = dBxdz - dBzdx, Jz = dBydx - dBxdy;
                const double J2 = Jx * Jx + Jy * Jy + Jz * Jz;
                E += dt * g.eta * J2;
            }
            const size_t fL = xk + ox0, fR = fL + 1, fD = yk + oy0, fU = fD + Nx, fB = q, fF = zf1 + oz0;
            const double bxL = B1.Bx[fL], bxR = B1.Bx[fR], byD = B1.By[fD], byU = B1.By[fU], bzB = bz0, bzF = bz1;
            floors3d(e, rho, mx, my, mz, E, 0.5 * (bxL + bxR), 0.5 * (byD + byU), 0.5 * (bzB + bzF));  // M3:1019-1029
            Prim<double> P;
            if (!FINAL) {
                Out.rho[q] = rho; Out.mx[q] = mx; Out.my[q] = my; Out.mz[q] = mz; Out.E[q] = E;
                P = cons_to_prim3d(e, rho, mx, my, mz, E, 0.5 * (bxL + bxR), 0.5 * (byD + byU), 0.5 * (bzB + bzF));
            } else {  // Heun average and floors, M3:1042-1065
                rho = 0.5 * (Un.rho[q] + rho);
                mx = 0.5 * (Un.mx[q] + mx);
                my = 0.5 * (Un.my[q] + my);
                mz = 0.5 * (Un.mz[q] + mz);
                E = 0.5 * (Un.E[q] + E);
                double nxL, nxR, nyD, nyU, nzB, nzF;
                float lxL = 0.f, lxR = 0.f, lyD = 0.f, lyU = 0.f, lzB = 0.f, lzF = 0.f;
                if (comp) {  // 1/2 (B^n + B^2) of the compensated fields
                    ct_avg(Un.Bx[fL], Un.Bx_lo[fL], bxL, B1.Bx_lo[fL], nxL, lxL);
                    ct_avg(Un.Bx[fR], Un.Bx_lo[fR], bxR, B1.Bx_lo[fR], nxR, lxR);
                    ct_avg(Un.By[fD], Un.By_lo[fD], byD, B1.By_lo[fD], nyD, lyD);
                    ct_avg(Un.By[fU], Un.By_lo[fU], byU, B1.By_lo[fU], nyU, lyU);
                    ct_avg(Un.Bz[fB], Un.Bz_lo[fB], bzB, B1.Bz_lo[fB], nzB, lzB);
                    ct_avg(Un.Bz[fF], Un.Bz_lo[fF], bzF, B1.Bz_lo[fF], nzF, lzF);
                } else {
                    nxL = 0.5 * (Un.Bx[fL] + bxL); nxR = 0.5 * (Un.Bx[fR] + bxR);
                    nyD = 0.5 * (Un.By[fD] + byD); nyU = 0.5 * (Un.By[fU] + byU);
                    nzB = 0.5 * (Un.Bz[fB] + bzB); nzF = 0.5 * (Un.Bz[fF] + bzF);
                }
                const double bx = 0.5 * (nxL + nxR), by = 0.5 * (nyD + nyU), bz = 0.5 * (nzB + nzF);
                floors3d(e, rho, mx, my, mz, E, bx, by, bz);
                Out.rho[q] = rho; Out.mx[q] = mx; Out.my[q] = my; Out.mz[q] = mz; Out.E[q] = E;
                Out.Bx[fL] = nxL;
                if (i == Nx - 1) Out.Bx[fR] = nxR;
                Out.By[fD] = nyD;
                if (j == Ny - 1) Out.By[fU] = nyU;
                Out.Bz[fB] = nzB;
                if (k == Nz - 1) Out.Bz[fF] = nzF;
                if (comp) {
                    Out.Bx_lo[fL] = lxL;
                    if (i == Nx - 1) Out.Bx_lo[fR] = lxR;
                    Out.By_lo[fD] = lyD;
                    if (j == Ny - 1) Out.By_lo[fU] = lyU;
                    Out.Bz_lo[fB] = lzB;
                    if (k == Nz - 1) Out.Bz_lo[fF] = lzF;
                }
                P = cons_to_prim3d(e, rho, mx, my, mz, E, bx, by, bz);
                speed = smax(speed, signal_speed3d(e, P));
            }
            W.c[0][q] = (R)P.rho; W.c[1][q] = (R)P.u; W.c[2][q] = (R)P.v; W.c[3][q] = (R)P.w;
            W.c[4][q] = (R)P.P; W.c[5][q] = (R)P.Bx; W.c[6][q] = (R)P.By; W.c[7][q] = (R)P.Bz;
            bxc_m = bxc_0; bxc_0 = bxc_p; byc_m = byc_0; byc_0 = byc_p;
            bz0 = bz1; bzjm = bz1jm; bzjp = bz1jp; bzim = bz1im; bzip = bz1ip;
            zf = zf1;
        }
    }
    if (FINAL) block_max_to(speed, spd);
}

template <bool IDEAL>
__global__ __launch_bounds__(256) void k_speed_max3d(Geom3D g, Eos3D e_in, State3D U, double* __restrict__ spd,
                                                     int zchunk) {
    const Eos3D e = eos_of<IDEAL>(e_in);
    const ColGrid cg(g, zchunk);
    double speed = 0.0;
    if (cg.valid(g))
        for (int k = cg.c0; k < cg.c1; ++k) speed = smax(speed, signal_speed3d(e, prim_at(g, e, U, cg.i, cg.j, k)));
    block_max_to(speed, spd);
}

// First level of the dt reduction: nspd per-block maxima -> kDtPart3 values.
constexpr int kDtPart3 = 256;
__global__ __launch_bounds__(256) void k_partmax3d(const double* __restrict__ part, int n, double* __restrict__ out) {
    double m = 0.0;
    for (int k = blockIdx.x * 256 + threadIdx.x; k < n; k += gridDim.x * 256) m = smax(m, part[k]);
    block_max_to(m, out);
}

__global__ __launch_bounds__(256) void k_dt_finalize3d(Geom3D g, Scalars* sc, double* dts, const double* spd, int nspd) {
    const double maxs = block_reduce_max(spd, nspd, 1e-14);
    if (threadIdx.x == 0) finalize_dt(sc, maxs, smin(smin(g.dx, g.dy), g.dz), g.cfl, g.eta, dts);
}

__global__ __launch_bounds__(256) void k_divb_max3d(Geom3D g, State3D U, double* __restrict__ spd, int zchunk) {
    const ColGrid cg(g, zchunk);
    const int i = cg.i, j = cg.j;
    double v = 0.0;
    if (cg.valid(g))
        for (int k = cg.c0; k < cg.c1; ++k)  // divB_cc, M3:1070-1078 (face i+1 <= Nx etc. need no wrapping)
            v = smax(v, fabs((U.Bx[fx_(g, i + 1, j, k)] - U.Bx[fx_(g, i, j, k)]) / g.dx +
                             (U.By[fy_(g, i, j + 1, k)] - U.By[fy_(g, i, j, k)]) / g.dy +
                             (U.Bz[fz_(g, i, j, k + 1)] - U.Bz[fz_(g, i, j, k)]) / g.dz));
    block_max_to(v, spd);
}

// Snapshot frame (the store block of M3:1293-1325): cons_to_prim on face-averaged B, sound speed of
// the EOS, Mach (floor 1e-12), cell divB.  No FP contraction.
__global__ __launch_bounds__(256) void k_frame3d(Geom3D g, Eos3D e, State3D U, double* __restrict__ out) {
#pragma clang fp contract(off)
    const int i = blockIdx.x * blockDim.x + threadIdx.x, j = blockIdx.y, k = blockIdx.z;
    if (i >= g.Nx) return;
    const size_t N = (size_t)g.Nx * g.Ny * g.Nz, q = ((size_t)k * g.Ny + j) * g.Nx + i;
    const double bxL = U.Bx[fx_(g, i, j, k)], bxR = U.Bx[fx_(g, i + 1, j, k)];
    const double byD = U.By[fy_(g, i, j, k)], byU = U.By[fy_(g, i, j + 1, k)];
    const double bzB = U.Bz[fz_(g, i, j, k)], bzF = U.Bz[fz_(g, i, j, k + 1)];
    const double bx = 0.5 * (bxL + bxR), by = 0.5 * (byD + byU), bz = 0.5 * (bzB + bzF);
    const Prim<double> W = cons_to_prim3d(e, U.rho[q], U.mx[q], U.my[q], U.mz[q], U.E[q], bx, by, bz);
    double cs2;
    if (e.ideal) {
        cs2 = e.gamma * W.P / smax(W.rho, e.hy.floor_rho);
    } else {
        cs2 = e.hy.sound_speed2(W, e.hy.aux(W));  // M3:206-212
    }
    const double cs = sqrt(cs2);
    const double vabs = sqrt(W.u * W.u + W.v * W.v + W.w * W.w);
    const double mach = vabs / smax(cs, 1e-12);
    const double divb = (bxR - bxL) / g.dx + (byU - byD) / g.dy + (bzF - bzB) / g.dz;  // M3:1070-1078
    const double f[11] = {W.rho, W.P, mach, W.u, W.v, W.w, vabs, W.Bx, W.By, W.Bz, divb};
#pragma unroll
    for (int c = 0; c < 11; c++) out[(size_t)c * N + q] = f[c];
}

__global__ __launch_bounds__(256) void k_reduce_max_to(const double* spd, int nspd, double* out) {
    const double m = block_reduce_max(spd, nspd, 0.0);
    if (threadIdx.x == 0) out[0] = m;
}

template <bool MIXED, typename EOS64, typename EOS32>
__global__ void k_hlld3d_batch(EOS64 e64, EOS32 e32, float tau, double guard, int n, const double* __restrict__ WLa,
                               const double* __restrict__ WRa, const double* __restrict__ Bna, double* __restrict__ out) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    const double* a = WLa + 8 * (size_t)t;
    const double* b = WRa + 8 * (size_t)t;
    const Prim<double> WL{a[0], a[1], a[2], a[3], a[4], a[5], a[6], a[7]};
    const Prim<double> WR{b[0], b[1], b[2], b[3], b[4], b[5], b[6], b[7]};
    double F[7];
    face_hlld3d<MIXED>(e64, e32, WL, WR, Bna[t], F, tau, guard);
    for (int c = 0; c < 7; c++) out[7 * (size_t)t + c] = F[c];
}

template <typename R>
HybridEOS<R> make_hybrid(const cc_mhd3d_params* p) {
    HybridEOS<R> h;
    const double K2 = p->K1 * std::pow(p->rho_nuc, p->gamma1 - p->gamma2);
    h.gamma1 = (R)p->gamma1; h.gamma2 = (R)p->gamma2; h.K1 = (R)p->K1; h.K2 = (R)K2; h.rho_nuc = (R)p->rho_nuc;
    h.gamma_th = (R)p->gamma_th; h.rho_bounce = (R)p->rho_bounce; h.gamma_bounce = (R)p->gamma_bounce;
    h.Kbounce = (R)p->Kbounce;
    h.ec1_n = (R)(p->K1 * std::pow(p->rho_nuc, p->gamma1 - 1.0) / (p->gamma1 - 1.0));
    h.ec2_n = (R)(K2 * std::pow(p->rho_nuc, p->gamma2 - 1.0) / (p->gamma2 - 1.0));
    h.rb_g = (R)std::pow(p->rho_bounce, p->gamma_bounce);
    h.rb_g1 = (R)std::pow(p->rho_bounce, p->gamma_bounce - 1.0);
    h.floor_rho = (R)p->floor_rho;
    h.floor_P = (R)p->floor_P;
    return h;
}

}  // namespace cc

// The EOS in the flux precision: FP32 for the mixed path, FP64 otherwise.
template <bool MIXED, typename EOS64, typename EOS32>
const typename std::conditional<MIXED, EOS32, EOS64>::type& select_eos(const EOS64& e64, const EOS32& e32) {
    if constexpr (MIXED) return e32;
    else return e64;
}

// ==================================================================== host handle
using namespace cc;

static constexpr int kDtLog3 = 1 << 16;

struct cc_mhd3d {
    cc_mhd3d_params p{};
    Geom3D g{};
    Eos3D e{};
    hipStream_t stream = nullptr;
    State3D S[3]{};
    int cur = 0;
    State3D Us{};   // cells after the flux update
    State3D Bt{};   // stage-2 face field before averaging
    void* prim[8] = {};     // primitives, flux precision
    void* flx[3][5] = {};   // face fluxes per axis, flux precision
    void* emf[6] = {};      // face EMFs Ey_x, Ez_x, Ez_y, Ex_y, Ex_z, Ey_z, flux precision
    DegQueue dq{};          // mixed: faces queued for the FP64 HLLD
    Edges ed{};
    Scalars* sc = nullptr;
    double* dts_dev = nullptr;
    double* spd = nullptr;
    double* spd2 = nullptr;  // kDtPart3 second-level maxima
    double* red = nullptr;
    int nspd = 0;
    // edge EMFs + CT in one pass (k_ect3d, tile shape ect_shape); k_edge3d + k_ct3d otherwise
    bool fused = true;
    int ect_shape = -1;  // -1: 3 for the mixed path (fits two blocks per CU), 1 for FP64 (spills at that bound)
    int zchunk = 32;
    // z faces + update in one column march (k_zupd3d, z chunks of zuchunk planes): 1 always, 0 never, -1
    // (default) for FP64 only.  In the mixed path every z face the FP32 HLLD flags defers its two cells to a
    // second pass, which costs more than the saved z-flux traffic (256^3 turbulence: 5.5-6.1 against 5.5-5.7 ms
    // per step); in FP64 the pass saves 0.3 ms per stage.
    int zupd = -1;
    int zuchunk = 32;
    bool use_zupd() const { return zupd > 0 || (zupd < 0 && p.precision != CC_MIXED); }
    Profiler prof;
    FrameBuffer frame;
    FieldStats stats;
    std::vector<void*> allocs;
    hipGraphExec_t graph[3] = {nullptr, nullptr, nullptr};

    size_t N() const { return (size_t)p.Nx * p.Ny * p.Nz; }
    // The snapshot frame of the current state into the frame buffer.
    cc_status make_frame() {
        CC_TRY(frame.ensure((size_t)CC_MHD3D_FRAME_FIELDS * N()));
        hipLaunchKernelGGL(k_frame3d, dim3((p.Nx + 255) / 256, p.Ny, p.Nz), dim3(256), 0, stream, g, e, S[cur],
                           frame.d);
        CC_CHECK_LAUNCH();
        return CC_OK;
    }
    size_t NFx() const { return (size_t)(p.Nx + 1) * p.Ny * p.Nz; }
    // grid of the column-marching cell kernels (Joule, speed, divB): 64 i x 4 j columns per block, z chunks
    int jchunk = 16;
    dim3 col_grid() const { return dim3((p.Nx + 63) / 64, (p.Ny + 3) / 4, (p.Nz + jchunk - 1) / jchunk); }
    size_t NFy() const { return (size_t)p.Nx * (p.Ny + 1) * p.Nz; }
    size_t NFz() const { return (size_t)p.Nx * p.Ny * (p.Nz + 1); }
    Profiler* pr() { return prof.on ? &prof : nullptr; }

    cc_status alloc(double** ptr, size_t n) {
        if (hipMalloc((void**)ptr, n * sizeof(double)) != hipSuccess) {
            set_error(CC_ERR_OOM, "cc_mhd3d: device allocation failed");
            return CC_ERR_OOM;
        }
        allocs.push_back(*ptr);
        CC_HIP(hipMemsetAsync(*ptr, 0, n * sizeof(double), stream));
        return CC_OK;
    }
    cc_status alloc_cells(State3D& s) {
        CC_TRY(alloc(&s.rho, N())); CC_TRY(alloc(&s.mx, N())); CC_TRY(alloc(&s.my, N()));
        CC_TRY(alloc(&s.mz, N())); CC_TRY(alloc(&s.E, N()));
        return CC_OK;
    }
    cc_status alloc_faces(State3D& s) {
        CC_TRY(alloc(&s.Bx, NFx())); CC_TRY(alloc(&s.By, NFy())); CC_TRY(alloc(&s.Bz, NFz()));
        if (g.ct_comp) {
            CC_TRY(alloc_bytes((void**)&s.Bx_lo, NFx() * sizeof(float)));
            CC_TRY(alloc_bytes((void**)&s.By_lo, NFy() * sizeof(float)));
            CC_TRY(alloc_bytes((void**)&s.Bz_lo, NFz() * sizeof(float)));
        }
        return CC_OK;
    }

    cc_status alloc_bytes(void** ptr, size_t bytes) {
        if (hipMalloc(ptr, bytes) != hipSuccess) {
            set_error(CC_ERR_OOM, "cc_mhd3d: device allocation failed");
            return CC_ERR_OOM;
        }
        allocs.push_back(*ptr);
        CC_HIP(hipMemsetAsync(*ptr, 0, bytes, stream));
        return CC_OK;
    }
    cc_status alloc_flux_path() {
        const size_t s = p.precision == CC_MIXED ? sizeof(float) : sizeof(double);
        const size_t nf[3] = {NFx(), NFy(), NFz()};
        for (auto& a : prim) CC_TRY(alloc_bytes(&a, N() * s));
        for (int ax = 0; ax < 3; ++ax)
            for (auto& a : flx[ax]) CC_TRY(alloc_bytes(&a, nf[ax] * s));
        const size_t ne[6] = {NFx(), NFx(), NFy(), NFy(), NFz(), NFz()};
        for (int c = 0; c < 6; ++c) CC_TRY(alloc_bytes(&emf[c], ne[c] * s));
        if (p.precision == CC_MIXED) {
            CC_TRY(alloc_bytes((void**)&dq.count, (size_t)4 * kQSub * kQStride * sizeof(int)));
            for (int ax = 0; ax < 3; ++ax) {
                CC_TRY(alloc_bytes((void**)&dq.face[ax], nf[ax] * sizeof(int)));
                dq.span[ax] = (int)((nf[ax] + kQSub - 1) / kQSub);
            }
            CC_TRY(alloc_bytes((void**)&dq.cell, N() * sizeof(int)));
            dq.span[3] = (int)((N() + kQSub - 1) / kQSub);
        }
        return CC_OK;
    }
    template <typename R>
    PrimArrays<R> W() const {
        PrimArrays<R> w;
        for (int c = 0; c < 8; ++c) w.c[c] = static_cast<R*>(prim[c]);
        return w;
    }
    template <typename R>
    FaceFlux<R> FF(int ax) const {
        FaceFlux<R> f;
        for (int c = 0; c < 5; ++c) f.f[c] = static_cast<R*>(flx[ax][c]);
        return f;
    }
    template <typename R>
    EmfArr<R> EF() const {
        return EmfArr<R>{static_cast<R*>(emf[0]), static_cast<R*>(emf[1]), static_cast<R*>(emf[2]),
                         static_cast<R*>(emf[3]), static_cast<R*>(emf[4]), static_cast<R*>(emf[5])};
    }
    static unsigned blocks(size_t n) { return (unsigned)((n + 255) / 256); }

    template <bool MIXED, typename EOS64, typename EOS32>
    cc_status launch_flux(const EOS64& e64, const EOS32& e32, const State3D& in) {
        using R = typename std::conditional<MIXED, float, double>::type;
        if (MIXED) {
            hipLaunchKernelGGL(k_zero_counts, dim3(1), dim3(256), 0, stream, dq.count);
            CC_CHECK_LAUNCH();
        }
        if (use_zupd()) {  // x and y faces; z faces + update in one column march
            {
                ProfScope ps(pr(), stream, "face");
                hipLaunchKernelGGL((k_face3d<0, MIXED, EOS64, EOS32>), dim3(blocks(NFx())), dim3(256), 0, stream, g, e64,
                                   e32, W<R>(), in.Bx, FF<R>(0), EF<R>(), dq, e, in);
                CC_CHECK_LAUNCH();
            }
            {
                ProfScope ps(pr(), stream, "face");
                hipLaunchKernelGGL((k_face3d<1, MIXED, EOS64, EOS32>), dim3((p.Nx + 63) / 64, p.Nz, (p.Ny + 1 + 3) / 4),
                                   dim3(256), 0, stream, g, e64, e32, W<R>(), in.By, FF<R>(1), EF<R>(), dq, e, in);
                CC_CHECK_LAUNCH();
            }
            if constexpr (MIXED) {
                ProfScope ps(pr(), stream, "face_fp64");
                hipLaunchKernelGGL((k_face3d_fp64<EOS64>), dim3(2048), dim3(256), 0, stream, g, e, e64, in,
                                   FF<float>(0), FF<float>(1), FF<float>(2), EF<float>(), dq, 3);
                CC_CHECK_LAUNCH();
            }
            {
                ProfScope ps(pr(), stream, "zupd");
                const dim3 grid((p.Nx + 63) / 64, (p.Ny + 3) / 4, (p.Nz + zuchunk - 1) / zuchunk);
                using EOSR = typename std::conditional<MIXED, EOS32, EOS64>::type;
                const EOSR& er = select_eos<MIXED>(e64, e32);
                hipLaunchKernelGGL((e.ideal ? k_zupd3d<MIXED, EOSR, true> : k_zupd3d<MIXED, EOSR, false>), grid,
                                   dim3(256), 0, stream, g, e, er, W<R>(), in, sc, Us, FF<R>(0), FF<R>(1), FF<R>(2),
                                   EF<R>(), dq, zuchunk);
                CC_CHECK_LAUNCH();
            }
            if constexpr (MIXED) {
                ProfScope ps(pr(), stream, "face_fp64");
                hipLaunchKernelGGL((k_face3d_fp64<EOS64>), dim3(2048), dim3(256), 0, stream, g, e, e64, in,
                                   FF<float>(0), FF<float>(1), FF<float>(2), EF<float>(), dq, 4);
                hipLaunchKernelGGL((e.ideal ? k_update3d_cells<true> : k_update3d_cells<false>), dim3(1024),
                                   dim3(256), 0, stream, g, e, in, FF<float>(0), FF<float>(1), FF<float>(2), sc, Us,
                                   dq);
                CC_CHECK_LAUNCH();
            }
            return CC_OK;
        }
        {
            ProfScope ps(pr(), stream, "face");
            hipLaunchKernelGGL((k_face3d<0, MIXED, EOS64, EOS32>), dim3(blocks(NFx())), dim3(256), 0, stream, g, e64, e32,
                               W<R>(), in.Bx, FF<R>(0), EF<R>(), dq, e, in);
            CC_CHECK_LAUNCH();
        }
        {
            ProfScope ps(pr(), stream, "face");
            hipLaunchKernelGGL((k_face3d<1, MIXED, EOS64, EOS32>), dim3((p.Nx + 63) / 64, p.Nz, (p.Ny + 1 + 3) / 4),
                               dim3(256), 0, stream, g, e64, e32,
                               W<R>(), in.By, FF<R>(1), EF<R>(), dq, e, in);
            CC_CHECK_LAUNCH();
        }
        {
            ProfScope ps(pr(), stream, "face");
            if (MIXED && p.hlld_guard <= 0.0)
                hipLaunchKernelGGL((k_face3d<2, MIXED, EOS64, EOS32, true>), dim3((p.Nx + 63) / 64, p.Ny, (p.Nz + 1 + 3) / 4),
                                   dim3(256), 0, stream, g, e64, e32, W<R>(), in.Bz, FF<R>(2), EF<R>(), dq, e, in);
            else
                hipLaunchKernelGGL((k_face3d<2, MIXED, EOS64, EOS32>), dim3((p.Nx + 63) / 64, p.Ny, (p.Nz + 1 + 3) / 4),
                                   dim3(256), 0, stream, g, e64, e32, W<R>(), in.Bz, FF<R>(2), EF<R>(), dq, e, in);
            CC_CHECK_LAUNCH();
        }
        if constexpr (MIXED) {
            ProfScope ps(pr(), stream, "face_fp64");
            hipLaunchKernelGGL((k_face3d_fp64<EOS64>), dim3(2048), dim3(256), 0, stream, g, e, e64, in, FF<float>(0),
                               FF<float>(1), FF<float>(2), EF<float>(), dq, 7);
            CC_CHECK_LAUNCH();
        }
        {
            ProfScope ps(pr(), stream, "update");
            hipLaunchKernelGGL((e.ideal ? k_update3d<R, true> : k_update3d<R, false>), dim3(blocks(N())), dim3(256), 0, stream, g, e, in, FF<R>(0), FF<R>(1),
                               FF<R>(2), sc, Us);
            CC_CHECK_LAUNCH();
        }
        return CC_OK;
    }

    cc_status flux(const State3D& in) {
        const bool mixed = p.precision == CC_MIXED;
        if (e.ideal) {
            const IdealEOS3D<double> e64{p.gamma - 1.0, p.gamma, p.floor_rho};
            const IdealEOS3D<float> e32{(float)(p.gamma - 1.0), (float)p.gamma, (float)p.floor_rho};
            return mixed ? launch_flux<true>(e64, e32, in) : launch_flux<false>(e64, e32, in);
        }
        const HybridEOS<float> h32 = make_hybrid<float>(&p);
        return mixed ? launch_flux<true>(e.hy, h32, in) : launch_flux<false>(e.hy, h32, in);
    }

    // One stage in -> out; FINAL averages with hn (Heun) and writes the averaged faces to out.
    template <bool FINAL>
    cc_status stage(const State3D& in, const State3D& out, const State3D& hn) {
        const int Nx = p.Nx, Ny = p.Ny, Nz = p.Nz;
        CC_TRY(flux(in));
        const State3D& bdst = FINAL ? Bt : out;
        if (fused) {
            ProfScope ps(pr(), stream, "ect");
            CC_TRY(p.precision == CC_MIXED ? launch_ect<float>(in, bdst) : launch_ect<double>(in, bdst));
        } else {
            ProfScope ps(pr(), stream, "edge");
            if (p.precision == CC_MIXED)
                hipLaunchKernelGGL(k_edge3d<float>, dim3((Nx + 1 + 63) / 64, Ny + 1, (Nz + 1 + 3) / 4), dim3(256), 0, stream, g,
                                   EF<float>(), in, ed);
            else
                hipLaunchKernelGGL(k_edge3d<double>, dim3((Nx + 1 + 63) / 64, Ny + 1, (Nz + 1 + 3) / 4), dim3(256), 0, stream, g,
                                   EF<double>(), in, ed);
            CC_CHECK_LAUNCH();
        }
        if (!fused) {
            ProfScope ps(pr(), stream, "ct");
            hipLaunchKernelGGL(k_ct3d, dim3((Nx + 1 + 63) / 64, Ny + 1, (Nz + 1 + 3) / 4), dim3(256), 0, stream, g, ed, in, bdst, sc);
            CC_CHECK_LAUNCH();
        }
        {
            ProfScope ps(pr(), stream, "joule");
            if (p.precision == CC_MIXED)
                hipLaunchKernelGGL((e.ideal ? k_joule3d<float, FINAL, true> : k_joule3d<float, FINAL, false>), col_grid(), dim3(256), 0,
                                   stream, g, e, Us, bdst, sc, hn, out, spd, jchunk, W<float>());
            else
                hipLaunchKernelGGL((e.ideal ? k_joule3d<double, FINAL, true> : k_joule3d<double, FINAL, false>), col_grid(), dim3(256), 0,
                                   stream, g, e, Us, bdst, sc, hn, out, spd, jchunk, W<double>());
            CC_CHECK_LAUNCH();
        }
        return CC_OK;
    }

    template <typename R, int TX, int TY, int MINW = 1>
    cc_status launch_ect_shape(const State3D& in, const State3D& B1) {
        const dim3 grid((unsigned)(((p.Nx + TX - 1) / TX) * ((p.Ny + TY - 1) / TY)),
                        (unsigned)((p.Nz + zchunk - 1) / zchunk));
        hipLaunchKernelGGL((k_ect3d<R, TX, TY, MINW>), grid, dim3(EctShape<TX, TY>::NT), 0, stream, g, EF<R>(), in, B1, sc,
                           zchunk);
        CC_CHECK_LAUNCH();
        return CC_OK;
    }
    template <typename R>
    cc_status launch_ect(const State3D& in, const State3D& B1) {
        const int shape = ect_shape >= 0 ? ect_shape : (std::is_same<R, float>::value ? 3 : 1);
        switch (shape) {
            case 0: return launch_ect_shape<R, 32, 28>(in, B1);
            case 2: return launch_ect_shape<R, 64, 5>(in, B1);
            case 3: return launch_ect_shape<R, 32, 12, 4>(in, B1);
            default: return launch_ect_shape<R, 32, 12>(in, B1);
        }
    }

    cc_status heun(int a, int b, int c, bool driver, double* dts) {
        if (driver) {
            ProfScope ps(pr(), stream, "dt");
            hipLaunchKernelGGL(k_partmax3d, dim3(kDtPart3), dim3(256), 0, stream, spd, nspd, spd2);
            hipLaunchKernelGGL(k_dt_finalize3d, dim3(1), dim3(256), 0, stream, g, sc, dts, spd2, kDtPart3);
            CC_CHECK_LAUNCH();
        }
        CC_TRY(stage<false>(S[a], S[b], S[a]));
        CC_TRY(stage<true>(S[b], S[c], S[a]));
        return CC_OK;
    }

    // Primitives of s in the flux precision for the first stage's face kernels; every stage's Joule pass
    // leaves those of its result for the next stage, so a step sequence needs this once at its start.
    cc_status launch_prim(const State3D& s) {
        ProfScope ps(pr(), stream, "prim");
        const bool mixed = p.precision == CC_MIXED;
        if (mixed && e.ideal) hipLaunchKernelGGL((k_prim3d<float, true>), dim3(blocks(N())), dim3(256), 0, stream, g, e, s, W<float>());
        else if (mixed) hipLaunchKernelGGL((k_prim3d<float, false>), dim3(blocks(N())), dim3(256), 0, stream, g, e, s, W<float>());
        else if (e.ideal) hipLaunchKernelGGL((k_prim3d<double, true>), dim3(blocks(N())), dim3(256), 0, stream, g, e, s, W<double>());
        else hipLaunchKernelGGL((k_prim3d<double, false>), dim3(blocks(N())), dim3(256), 0, stream, g, e, s, W<double>());
        CC_CHECK_LAUNCH();
        return CC_OK;
    }

    cc_status speed_max(const State3D& s) {
        hipLaunchKernelGGL((e.ideal ? k_speed_max3d<true> : k_speed_max3d<false>), col_grid(), dim3(256), 0, stream, g, e, s, spd, jchunk);
        CC_CHECK_LAUNCH();
        return CC_OK;
    }

    ~cc_mhd3d() {
        for (auto& ge : graph)
            if (ge) hipGraphExecDestroy(ge);
        for (void* a : allocs) hipFree(a);
        if (sc) hipFree(sc);
        if (stream) hipStreamDestroy(stream);
    }
};

extern "C" {

cc_status cc_mhd3d_create(const cc_mhd3d_params* p, cc_mhd3d** out) {
    if (!out || !p) { set_error(CC_ERR_INVALID_ARG, "null argument"); return CC_ERR_INVALID_ARG; }
    *out = nullptr;
    if (p->Nx < 2 || p->Ny < 2 || p->Nz < 2 || p->Nx > 4096 || p->Ny > 4096 || p->Nz > 4096 ||
        (size_t)(p->Nx + 1) * (p->Ny + 1) * (p->Nz + 1) >= (size_t)1 << 31) {
        set_error(CC_ERR_INVALID_ARG, "grid must be within [2, 4096]^3 with (Nx+1)(Ny+1)(Nz+1) < 2^31");
        return CC_ERR_INVALID_ARG;
    }
    if (p->precision != CC_FP64 && p->precision != CC_MIXED) { set_error(CC_ERR_INVALID_ARG, "bad precision"); return CC_ERR_INVALID_ARG; }
    if (p->eos != 0 && p->eos != 1) { set_error(CC_ERR_INVALID_ARG, "eos must be 0 (hybrid) or 1 (ideal)"); return CC_ERR_INVALID_ARG; }
    if (p->eos == 1 && !(p->gamma > 1.0)) { set_error(CC_ERR_INVALID_ARG, "ideal EOS needs gamma > 1"); return CC_ERR_INVALID_ARG; }
    if (!(p->hlld_guard >= 0.0 && p->hlld_guard < 1.0)) { set_error(CC_ERR_INVALID_ARG, "hlld_guard must be in [0, 1)"); return CC_ERR_INVALID_ARG; }
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) { set_error(CC_ERR_NO_DEVICE, "no HIP device"); return CC_ERR_NO_DEVICE; }
    cc_mhd3d* h = new cc_mhd3d();
    h->p = *p;
    Geom3D& g = h->g;
    g.Nx = p->Nx; g.Ny = p->Ny; g.Nz = p->Nz; g.px = p->periodic_x != 0; g.py = p->periodic_y != 0; g.pz = p->periodic_z != 0;
    g.dx = (p->x1 - p->x0) / p->Nx; g.dy = (p->y1 - p->y0) / p->Ny; g.dz = (p->z1 - p->z0) / p->Nz;
    g.floor_rho = p->floor_rho; g.floor_P = p->floor_P; g.eta = p->eta; g.cfl = p->cfl;
    g.tau = p->degen_tau > 0.0 ? (float)p->degen_tau : default_degen_tau(p->hlld_guard);
    g.ct_comp = p->ct_compensated != 0;
    g.hlld_guard = p->hlld_guard;
    h->e.ideal = p->eos == 1;
    h->e.gamma = p->gamma;
    h->e.hy = make_hybrid<double>(p);
    if (hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking) != hipSuccess) {
        set_error(CC_ERR_HIP, "hipStreamCreate failed");
        delete h;
        return CC_ERR_HIP;
    }
    cc_status st = CC_OK;
    for (int k = 0; k < 3 && st == CC_OK; k++) {
        st = h->alloc_cells(h->S[k]);
        if (st == CC_OK) st = h->alloc_faces(h->S[k]);
    }
    if (st == CC_OK) st = h->alloc_cells(h->Us);
    if (st == CC_OK) st = h->alloc_faces(h->Bt);
    if (st == CC_OK) st = h->alloc_flux_path();
    if (const char* v = std::getenv("CC_M3_FUSED")) h->fused = std::atoi(v) != 0;
    if (const char* v = std::getenv("CC_M3_ECT_SHAPE")) h->ect_shape = std::atoi(v);
    if (const char* v = std::getenv("CC_M3_ZCHUNK")) h->zchunk = std::max(1, std::atoi(v));
    if (const char* v = std::getenv("CC_M3_JCHUNK")) h->jchunk = std::max(1, std::atoi(v));
    if (const char* v = std::getenv("CC_M3_ZUPD")) h->zupd = std::atoi(v);
    if (const char* v = std::getenv("CC_M3_ZUCHUNK")) h->zuchunk = std::max(1, std::atoi(v));
    if (st == CC_OK && !h->fused) st = h->alloc(&h->ed.Ex, (size_t)p->Nx * (p->Ny + 1) * (p->Nz + 1));
    if (st == CC_OK && !h->fused) st = h->alloc(&h->ed.Ey, (size_t)(p->Nx + 1) * p->Ny * (p->Nz + 1));
    if (st == CC_OK && !h->fused) st = h->alloc(&h->ed.Ez, (size_t)(p->Nx + 1) * (p->Ny + 1) * p->Nz);
    if (st == CC_OK) st = h->alloc(&h->dts_dev, kDtLog3);
    {
        const dim3 cg = h->col_grid();  // the per-block maxima of the cell kernels (speed, Joule, divB)
        h->nspd = (int)(cg.x * cg.y * cg.z);
    }
    if (st == CC_OK) st = h->alloc(&h->spd, (size_t)h->nspd);
    if (st == CC_OK) st = h->alloc(&h->spd2, (size_t)kDtPart3);
    if (st == CC_OK) st = h->alloc(&h->red, 4);
    if (st == CC_OK && hipMalloc((void**)&h->sc, sizeof(Scalars)) != hipSuccess) {
        set_error(CC_ERR_OOM, "scalar allocation failed");
        st = CC_ERR_OOM;
    }
    if (st == CC_OK) {
        hipLaunchKernelGGL(k_set_scalars, dim3(1), dim3(1), 0, h->stream, h->sc, 0.0, 0.0, 1e300, 1);
        if (hipStreamSynchronize(h->stream) != hipSuccess) { set_error(CC_ERR_HIP, "sync failed"); st = CC_ERR_HIP; }
    }
    if (st != CC_OK) { delete h; return st; }
    *out = h;
    return CC_OK;
}

cc_status cc_mhd3d_destroy(cc_mhd3d* h) {
    if (!h) return CC_OK;
    hipStreamSynchronize(h->stream);
    delete h;
    return CC_OK;
}

#define CHECK_H3()                                                  \
    do {                                                            \
        if (!h) { set_error(CC_ERR_INVALID_ARG, "null handle"); return CC_ERR_INVALID_ARG; } \
    } while (0)

cc_status cc_mhd3d_upload(cc_mhd3d* h, const double* rho, const double* mx, const double* my, const double* mz,
                          const double* E, const double* Bx_f, const double* By_f, const double* Bz_f) {
    CHECK_H3();
    if (!rho || !mx || !my || !mz || !E || !Bx_f || !By_f || !Bz_f) { set_error(CC_ERR_INVALID_ARG, "null state pointer"); return CC_ERR_INVALID_ARG; }
    State3D& s = h->S[h->cur];
    const size_t n = h->N() * 8;
    CC_HIP(hipMemcpyAsync(s.rho, rho, n, hipMemcpyHostToDevice, h->stream));
    CC_HIP(hipMemcpyAsync(s.mx, mx, n, hipMemcpyHostToDevice, h->stream));
    CC_HIP(hipMemcpyAsync(s.my, my, n, hipMemcpyHostToDevice, h->stream));
    CC_HIP(hipMemcpyAsync(s.mz, mz, n, hipMemcpyHostToDevice, h->stream));
    CC_HIP(hipMemcpyAsync(s.E, E, n, hipMemcpyHostToDevice, h->stream));
    CC_HIP(hipMemcpyAsync(s.Bx, Bx_f, h->NFx() * 8, hipMemcpyHostToDevice, h->stream));
    CC_HIP(hipMemcpyAsync(s.By, By_f, h->NFy() * 8, hipMemcpyHostToDevice, h->stream));
    CC_HIP(hipMemcpyAsync(s.Bz, Bz_f, h->NFz() * 8, hipMemcpyHostToDevice, h->stream));
    if (h->g.ct_comp) {  // an uploaded face field is exact as given
        CC_HIP(hipMemsetAsync(s.Bx_lo, 0, h->NFx() * sizeof(float), h->stream));
        CC_HIP(hipMemsetAsync(s.By_lo, 0, h->NFy() * sizeof(float), h->stream));
        CC_HIP(hipMemsetAsync(s.Bz_lo, 0, h->NFz() * sizeof(float), h->stream));
    }
    CC_HIP(hipStreamSynchronize(h->stream));
    return CC_OK;
}

cc_status cc_mhd3d_download(cc_mhd3d* h, double* rho, double* mx, double* my, double* mz, double* E, double* Bx_f,
                            double* By_f, double* Bz_f) {
    CHECK_H3();
    State3D& s = h->S[h->cur];
    const size_t n = h->N() * 8;
    struct { double* d; const double* s; size_t b; } cp[8] = {
        {rho, s.rho, n}, {mx, s.mx, n}, {my, s.my, n}, {mz, s.mz, n}, {E, s.E, n},
        {Bx_f, s.Bx, h->NFx() * 8}, {By_f, s.By, h->NFy() * 8}, {Bz_f, s.Bz, h->NFz() * 8}};
    for (auto& c : cp)
        if (c.d) CC_HIP(hipMemcpyAsync(c.d, c.s, c.b, hipMemcpyDeviceToHost, h->stream));
    CC_HIP(hipStreamSynchronize(h->stream));
    return CC_OK;
}

cc_status cc_mhd3d_compute_dt(cc_mhd3d* h, double* dt) {
    CHECK_H3();
    if (!dt) { set_error(CC_ERR_INVALID_ARG, "null dt"); return CC_ERR_INVALID_ARG; }
    hipLaunchKernelGGL(k_set_scalars, dim3(1), dim3(1), 0, h->stream, h->sc, 0.0, 0.0, 1e300, 1);
    CC_TRY(h->speed_max(h->S[h->cur]));
    hipLaunchKernelGGL(k_dt_finalize3d, dim3(1), dim3(256), 0, h->stream, h->g, h->sc, (double*)nullptr, h->spd, h->nspd);
    CC_CHECK_LAUNCH();
    Scalars s;
    CC_HIP(hipMemcpyAsync(&s, h->sc, sizeof(Scalars), hipMemcpyDeviceToHost, h->stream));
    CC_HIP(hipStreamSynchronize(h->stream));
    *dt = s.dt;
    return CC_OK;
}

cc_status cc_mhd3d_stage(cc_mhd3d* h, double dt) {
    CHECK_H3();
    hipLaunchKernelGGL(k_set_scalars, dim3(1), dim3(1), 0, h->stream, h->sc, dt, 0.0, 1e300, 0);
    const int a = h->cur, b = (a + 1) % 3;
    CC_TRY(h->launch_prim(h->S[a]));
    CC_TRY(h->stage<false>(h->S[a], h->S[b], h->S[a]));
    h->cur = b;
    return CC_OK;
}

cc_status cc_mhd3d_rk2_step(cc_mhd3d* h, double dt) {
    CHECK_H3();
    if (!std::isfinite(dt) || dt < 0.0) { set_error(CC_ERR_BAD_DT, "dt must be finite and >= 0"); return CC_ERR_BAD_DT; }
    hipLaunchKernelGGL(k_set_scalars, dim3(1), dim3(1), 0, h->stream, h->sc, dt, 0.0, 1e300, 0);
    const int a = h->cur, b = (a + 1) % 3, c = (a + 2) % 3;
    CC_TRY(h->launch_prim(h->S[a]));
    CC_TRY(h->heun(a, b, c, false, nullptr));
    h->cur = c;
    return CC_OK;
}

cc_status cc_mhd3d_run(cc_mhd3d* h, int nsteps, double* t_inout, double t_end, double* dts) {
    CHECK_H3();
    if (nsteps < 0 || !t_inout) { set_error(CC_ERR_INVALID_ARG, "nsteps < 0 or null t"); return CC_ERR_INVALID_ARG; }
    if (nsteps == 0) return CC_OK;
    if (nsteps > kDtLog3) {
        double t = *t_inout;
        for (int done = 0; done < nsteps;) {
            const int n = std::min(kDtLog3, nsteps - done);
            CC_TRY(cc_mhd3d_run(h, n, &t, t_end, dts ? dts + done : nullptr));
            done += n;
        }
        *t_inout = t;
        return CC_OK;
    }
    hipLaunchKernelGGL(k_set_scalars, dim3(1), dim3(1), 0, h->stream, h->sc, 0.0, *t_inout, t_end, 1);
    CC_TRY(h->speed_max(h->S[h->cur]));
    CC_TRY(h->launch_prim(h->S[h->cur]));
    const bool use_graph = !h->prof.on;
    for (int s = 0; s < nsteps; ++s) {
        const int a = h->cur, b = (a + 1) % 3, c = (a + 2) % 3;
        if (use_graph) {
            if (!h->graph[a]) {
                hipGraph_t gr;
                CC_HIP(hipStreamBeginCapture(h->stream, hipStreamCaptureModeThreadLocal));
                cc_status st = h->heun(a, b, c, true, h->dts_dev);
                hipError_t e = hipStreamEndCapture(h->stream, &gr);
                if (st != CC_OK) return st;
                if (e != hipSuccess) { set_error(CC_ERR_HIP, "graph capture failed"); return CC_ERR_HIP; }
                CC_HIP(hipGraphInstantiate(&h->graph[a], gr, nullptr, nullptr, 0));
                hipGraphDestroy(gr);
            }
            CC_HIP(hipGraphLaunch(h->graph[a], h->stream));
        } else {
            CC_TRY(h->heun(a, b, c, true, h->dts_dev));
        }
        h->cur = c;
    }
    Scalars s;
    CC_HIP(hipMemcpyAsync(&s, h->sc, sizeof(Scalars), hipMemcpyDeviceToHost, h->stream));
    if (dts) CC_HIP(hipMemcpyAsync(dts, h->dts_dev, (size_t)nsteps * 8, hipMemcpyDeviceToHost, h->stream));
    CC_HIP(hipStreamSynchronize(h->stream));
    *t_inout = s.t;
    if (!std::isfinite(s.dt) || s.dt < 0.0) { set_error(CC_ERR_BAD_DT, "non-finite or negative dt on device"); return CC_ERR_BAD_DT; }
    return CC_OK;
}

cc_status cc_mhd3d_divb_max(cc_mhd3d* h, double* out) {
    CHECK_H3();
    if (!out) { set_error(CC_ERR_INVALID_ARG, "null output"); return CC_ERR_INVALID_ARG; }
    hipLaunchKernelGGL(k_divb_max3d, h->col_grid(), dim3(256), 0, h->stream, h->g,
                       h->S[h->cur], h->spd, h->jchunk);
    hipLaunchKernelGGL(k_reduce_max_to, dim3(1), dim3(256), 0, h->stream, h->spd, h->nspd, h->red);
    CC_CHECK_LAUNCH();
    CC_HIP(hipMemcpyAsync(out, h->red, 8, hipMemcpyDeviceToHost, h->stream));
    CC_HIP(hipStreamSynchronize(h->stream));
    return CC_OK;
}

cc_status cc_mhd3d_frame(cc_mhd3d* h, double* out) {
    CHECK_H3();
    if (!out) { set_error(CC_ERR_INVALID_ARG, "null output"); return CC_ERR_INVALID_ARG; }
    CC_TRY(h->make_frame());
    CC_HIP(hipMemcpyAsync(out, h->frame.d, sizeof(double) * CC_MHD3D_FRAME_FIELDS * h->N(), hipMemcpyDeviceToHost,
                          h->stream));
    CC_HIP(hipStreamSynchronize(h->stream));
    return CC_OK;
}

cc_status cc_mhd3d_totals(cc_mhd3d* h, double* out6) {
    CHECK_H3();
    if (!out6) { set_error(CC_ERR_INVALID_ARG, "null output"); return CC_ERR_INVALID_ARG; }
    const State3D& s = h->S[h->cur];
    FieldPtrs a{{s.rho, s.mx, s.my, s.mz, s.E}};
    double r[15];
    CC_TRY(h->stats.run(a, nullptr, 5, h->N(), h->stream, r));
    const double dV = h->g.dx * h->g.dy * h->g.dz;
    for (int f = 0; f < 5; f++) out6[f] = r[3 * f] * dV;
    return cc_mhd3d_divb_max(h, out6 + 5);
}

cc_status cc_mhd3d_compare(cc_mhd3d* a, cc_mhd3d* b, double* out) {
    if (!a || !b || !out) { set_error(CC_ERR_INVALID_ARG, "null handle or output"); return CC_ERR_INVALID_ARG; }
    if (a->p.Nx != b->p.Nx || a->p.Ny != b->p.Ny || a->p.Nz != b->p.Nz) {
        set_error(CC_ERR_INVALID_ARG, "cc_mhd3d_compare: grids differ");
        return CC_ERR_INVALID_ARG;
    }
    CC_TRY(a->make_frame());
    CC_TRY(b->make_frame());
    CC_HIP(hipStreamSynchronize(b->stream));
    return compare_frames(a->stats, a->frame.d, b->frame.d, CC_MHD3D_FRAME_FIELDS, a->N(), a->stream, out);
}

cc_status cc_mhd3d_sync(cc_mhd3d* h) {
    CHECK_H3();
    CC_HIP(hipStreamSynchronize(h->stream));
    return CC_OK;
}

cc_status cc_mhd3d_set_profiling(cc_mhd3d* h, int on) {
    CHECK_H3();
    h->prof.on = on != 0;
    return CC_OK;
}

cc_status cc_mhd3d_kernel_time(cc_mhd3d* h, const char* kernel, int* count, double* total_ms) {
    CHECK_H3();
    if (!kernel || !count || !total_ms) { set_error(CC_ERR_INVALID_ARG, "null argument"); return CC_ERR_INVALID_ARG; }
    CC_HIP(hipStreamSynchronize(h->stream));
    h->prof.get_total(kernel, count, total_ms);
    return CC_OK;
}

cc_status cc_mhd3d_reset_profiling(cc_mhd3d* h) {
    CHECK_H3();
    CC_HIP(hipStreamSynchronize(h->stream));
    h->prof.reset();
    return CC_OK;
}

cc_status cc_hlld3d_batch(const cc_mhd3d_params* p, int n, const double* WL, const double* WR, const double* Bn,
                          double* out) {
    if (!p || n < 0 || (n > 0 && (!WL || !WR || !Bn || !out))) { set_error(CC_ERR_INVALID_ARG, "invalid batch arguments"); return CC_ERR_INVALID_ARG; }
    if (n == 0) return CC_OK;
    double *dL, *dR, *dB, *dO;
    const size_t b8 = (size_t)n * 8 * 8;
    if (hipMalloc(&dL, b8) != hipSuccess) { set_error(CC_ERR_OOM, "batch alloc"); return CC_ERR_OOM; }
    hipMalloc(&dR, b8);
    hipMalloc(&dB, (size_t)n * 8);
    hipMalloc(&dO, (size_t)n * 7 * 8);
    hipMemcpy(dL, WL, b8, hipMemcpyHostToDevice);
    hipMemcpy(dR, WR, b8, hipMemcpyHostToDevice);
    hipMemcpy(dB, Bn, (size_t)n * 8, hipMemcpyHostToDevice);
    const dim3 grid((n + 255) / 256);
    const bool mixed = p->precision == CC_MIXED;
    const float tau = p->degen_tau > 0.0 ? (float)p->degen_tau : default_degen_tau(p->hlld_guard);
    if (p->eos == 1) {
        const IdealEOS3D<double> e64{p->gamma - 1.0, p->gamma, p->floor_rho};
        const IdealEOS3D<float> e32{(float)(p->gamma - 1.0), (float)p->gamma, (float)p->floor_rho};
        if (mixed)
            hipLaunchKernelGGL((k_hlld3d_batch<true, IdealEOS3D<double>, IdealEOS3D<float>>), grid, dim3(256), 0, 0, e64, e32, tau, p->hlld_guard, n, dL, dR, dB, dO);
        else
            hipLaunchKernelGGL((k_hlld3d_batch<false, IdealEOS3D<double>, IdealEOS3D<float>>), grid, dim3(256), 0, 0, e64, e32, tau, p->hlld_guard, n, dL, dR, dB, dO);
    } else {
        const HybridEOS<double> e64 = make_hybrid<double>(p);
        const HybridEOS<float> e32 = make_hybrid<float>(p);
        if (mixed)
            hipLaunchKernelGGL((k_hlld3d_batch<true, HybridEOS<double>, HybridEOS<float>>), grid, dim3(256), 0, 0, e64, e32, tau, p->hlld_guard, n, dL, dR, dB, dO);
        else
            hipLaunchKernelGGL((k_hlld3d_batch<false, HybridEOS<double>, HybridEOS<float>>), grid, dim3(256), 0, 0, e64, e32, tau, p->hlld_guard, n, dL, dR, dB, dO);
    }
    hipError_t e = hipGetLastError();
    if (e == hipSuccess) e = hipMemcpy(out, dO, (size_t)n * 7 * 8, hipMemcpyDeviceToHost);
    hipFree(dL); hipFree(dR); hipFree(dB); hipFree(dO);
    if (e != hipSuccess) { set_error(CC_ERR_HIP, hipGetErrorString(e)); return CC_ERR_HIP; }
    return CC_OK;
}

}  // extern "C"
