// euler3d.hip -- gfx950 solver step of the 3D Euler HLLC + FFT self-gravity problem (the periodic
// uniform-sphere collapse): the 2D Euler step of euler2d_gravity_hllc.cpp (EH:406-550) extended to
// three axes, with the gravity of poisson3d.hpp.
//
// One stage (rhs + update) is five launches:
//   k_e3_prim       floored primitives (rho, u, v, w, P) of every cell (EH:415-425), FP64 or FP32
//   k_e3_face<AX>   one thread per face along AX (N faces per axis, periodic): 4-point MUSCL-MC
//                   face states (EH:437-455) and the EH HLLC with two transverse velocities
//   k_e3_update     flux divergence + rho g source (EH:505-524), g differenced from the potential on
//                   the fly, and the Heun stage with floors (EH:529-550); the last stage also leaves
//                   the block maxima of the signal speed
// Heun: dt (device) -> g(U) -> U1 = U + dt rhs(U) -> g(U1) -> U = U/2 + (U1 + dt rhs(U1))/2, the
// final update in place (each thread reads U only at its own cell).  Mixed precision keeps the
// primitives, reconstruction, HLLC and face fluxes in FP32; state, update and gravity source FP64.
// This file is compiled without FP contraction, so the FP64 path repeats the oracle's rounding.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "diag.hpp"
#include "hllc.hpp"
#include "poisson3d.hpp"
#include "profiler.hpp"
#include "riemann.hpp"
#include "stepcommon.hpp"

namespace cc {

struct GeomE3 {
    int Nx, Ny, Nz, lx, ly;
    double dx, dy, dz, gamma, gm1, floor_rho, floor_P, cfl;
    int gravity_on;
};

struct StateE3 {
    double *rho, *mx, *my, *mz, *E;
};

template <typename R>
struct PrimE3 {
    R* w[5];  // rho, u, v, w, P
};

template <typename R>
struct FluxE3 {
    R* f[3][5];  // per axis the 5 fluxes through the lower face of every cell
};

// EXACT = false (the mixed path's update): the velocities multiply by 1/rho instead of three quotients.
template <bool EXACT = true>
__device__ __forceinline__ void floors_e3(const GeomE3& g, double& rho, double& mx, double& my, double& mz,
                                          double& E) {
    rho = smax(rho, g.floor_rho);
    const double ir = EXACT ? 0.0 : 1.0 / rho;
    const double u = EXACT ? mx / rho : mx * ir, v = EXACT ? my / rho : my * ir, w = EXACT ? mz / rho : mz * ir;
    const double kin = 0.5 * rho * ((u * u + v * v) + w * w);
    const double P = g.gm1 * (E - kin);
    if (P < g.floor_P) E = kin + g.floor_P / g.gm1;
}

template <bool EXACT = true>
__device__ __forceinline__ void prim_e3(const GeomE3& g, double rho, double mx, double my, double mz, double E,
                                        double& u, double& v, double& w, double& P, double& a) {
    rho = smax(rho, g.floor_rho);
    const double ir = EXACT ? 0.0 : 1.0 / rho;
    u = EXACT ? mx / rho : mx * ir;
    v = EXACT ? my / rho : my * ir;
    w = EXACT ? mz / rho : mz * ir;
    const double kin = 0.5 * rho * ((u * u + v * v) + w * w);
    P = g.gm1 * (E - kin);
    if (P < g.floor_P) P = g.floor_P;
    a = sqrt(EXACT ? g.gamma * P / rho : g.gamma * P * ir);
}

template <bool EXACT = true>
__device__ __forceinline__ double speed_e3(const GeomE3& g, double rho, double mx, double my, double mz, double E) {
    double u, v, w, P, a;
    prim_e3<EXACT>(g, rho, mx, my, mz, E, u, v, w, P, a);
    return smax(smax(fabs(u) + a, fabs(v) + a), fabs(w) + a);
}

template <typename R>
__global__ __launch_bounds__(256) void k_e3_prim(GeomE3 g, StateE3 U, PrimE3<R> W, size_t n) {
    const size_t q = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (q >= n) return;
    double rho = U.rho[q], mx = U.mx[q], my = U.my[q], mz = U.mz[q], E = U.E[q];
    floors_e3(g, rho, mx, my, mz, E);
    double u, v, w, P, a;
    prim_e3(g, rho, mx, my, mz, E, u, v, w, P, a);
    W.w[0][q] = (R)smax(rho, g.floor_rho);
    W.w[1][q] = (R)u;
    W.w[2][q] = (R)v;
    W.w[3][q] = (R)w;
    W.w[4][q] = (R)P;
}

// The flux through the face between cells q - e_AX and q from primitive c of the cells q - 2e, q - e, q,
// q + e (cell(o, c), o = 0..3): 4-point MUSCL-MC states (EH:437-455) and the EH HLLC (EH:125-219).
template <typename R, int AX, typename CellW>
__device__ __forceinline__ void face_e3(const GeomE3& g, const CellW& cell, R out[5]) {
    R L[5], Rr[5];
#pragma unroll
    for (int c = 0; c < 5; ++c) {
        const R wm = cell(0, c), w0 = cell(1, c), wp = cell(2, c), wpp = cell(3, c);
        const R sL = mc_limiter<R>(w0 - wm, wp - w0);
        const R sR = mc_limiter<R>(wp - w0, wpp - wp);
        L[c] = w0 + R(0.5) * sL;
        Rr[c] = wp - R(0.5) * sR;
    }
    const R fr = (R)g.floor_rho, fP = (R)g.floor_P;
    L[0] = smax(L[0], fr); L[4] = smax(L[4], fP);
    Rr[0] = smax(Rr[0], fr); Rr[4] = smax(Rr[4], fP);
    hllc_e3<R, AX>((R)g.gamma, fr, L, Rr, out);
}

// Face between cells q - e_AX and q: states from cells q - 2e, q - e, q, q + e (periodic).
// x: consecutive threads take consecutive faces; y, z: a block is 64 consecutive i x 4 consecutive faces
// along the axis (its waves then share the stencil cells in L1).
template <typename R, int AX>
__global__ __launch_bounds__(256) void k_e3_face(GeomE3 g, PrimE3<R> W, FluxE3<R> F, size_t n) {
    const int Nx = g.Nx, Ny = g.Ny, Nz = g.Nz;
    int i, j, k;
    size_t q;
    if (AX == 0) {
        q = (size_t)blockIdx.x * 256 + threadIdx.x;
        if (q >= n) return;
        i = (int)(q & (size_t)(Nx - 1));
        const size_t line = q >> g.lx;
        j = (int)(line & (size_t)(Ny - 1));
        k = (int)(line >> g.ly);
    } else {
        i = blockIdx.x * 64 + (threadIdx.x & 63);
        const int fa = blockIdx.z * 4 + (threadIdx.x >> 6);
        if (AX == 1) { j = fa; k = blockIdx.y; } else { k = fa; j = blockIdx.y; }
        if (i >= Nx || fa >= (AX == 1 ? Ny : Nz)) return;
        q = ((size_t)k * Ny + j) * Nx + i;
    }
    size_t qm2, qm1, qp1;
    if (AX == 0) {
        const size_t r = q - i;
        qm2 = r + ((i - 2) & (Nx - 1)); qm1 = r + ((i - 1) & (Nx - 1)); qp1 = r + ((i + 1) & (Nx - 1));
    } else if (AX == 1) {
        const size_t r = q - (size_t)j * Nx;
        qm2 = r + (size_t)((j - 2) & (Ny - 1)) * Nx; qm1 = r + (size_t)((j - 1) & (Ny - 1)) * Nx;
        qp1 = r + (size_t)((j + 1) & (Ny - 1)) * Nx;
    } else {
        const size_t sxy = (size_t)Nx * Ny, r = q - (size_t)k * sxy;
        qm2 = r + (size_t)((k - 2) & (Nz - 1)) * sxy; qm1 = r + (size_t)((k - 1) & (Nz - 1)) * sxy;
        qp1 = r + (size_t)((k + 1) & (Nz - 1)) * sxy;
    }
    R out[5];
    face_e3<R, AX>(
        g, [&](int o, int c) {
            const R* a = W.w[c];
            return o == 0 ? a[qm2] : o == 1 ? a[qm1] : o == 2 ? a[q] : a[qp1];
        },
        out);
#pragma unroll
    for (int c = 0; c < 5; ++c) F.f[AX][c][q] = out[c];
}

// MODE 0: out = rhs(Uin).  MODE 1: out = Uin + dt rhs, floors.  MODE 2: out = Uold/2 + (Uin + dt
// rhs)/2, floors, and the block maxima of the new state's signal speed.
// The gravity g = -central grad phi is differenced here from the potential (the expression of the
// solver's own gradient kernel), so g is never stored.
// The update of cell q (MODE as below) from its six face fluxes (flux(ax, side, c): component c through the
// lower (side 0) or upper (side 1) face of axis ax) and the potential at its six neighbours (phi(ax, side));
// returns the new state's signal speed in MODE 2.  EXACT (the FP64 path): the reference's quotients, so the
// step repeats the oracle's rounding; otherwise (mixed, whose fluxes carry FP32 rounding) the flux
// differences, the potential gradient and the velocities multiply by reciprocals, 20 fewer FP64 divisions.
template <int MODE, bool EXACT, typename FluxAt, typename PhiAt>
__device__ __forceinline__ double update_e3(const GeomE3& g, const FluxAt& flux, const PhiAt& phi, const StateE3& Uin,
                                            const StateE3& Uold, const StateE3& out, const Scalars* __restrict__ sc,
                                            size_t q) {
    double d[5];
    const double idx = EXACT ? 0.0 : 1.0 / g.dx, idy = EXACT ? 0.0 : 1.0 / g.dy, idz = EXACT ? 0.0 : 1.0 / g.dz;
#pragma unroll
    for (int c = 0; c < 5; ++c) {
        const double xL = flux(0, 0, c), xR = flux(0, 1, c);
        const double yD = flux(1, 0, c), yU = flux(1, 1, c);
        const double zD = flux(2, 0, c), zU = flux(2, 1, c);
        d[c] = EXACT ? -(xR - xL) / g.dx - (yU - yD) / g.dy - (zU - zD) / g.dz
                     : -(xR - xL) * idx - (yU - yD) * idy - (zU - zD) * idz;
    }
    const double r0 = Uin.rho[q], m0 = Uin.mx[q], n0 = Uin.my[q], o0 = Uin.mz[q], e0 = Uin.E[q];
    if (g.gravity_on) {
        const double rho = smax(r0, g.floor_rho);
        double u, v, w, ax, ay, az;
        if (EXACT) {
            u = m0 / rho; v = n0 / rho; w = o0 / rho;
            ax = -(phi(0, 1) - phi(0, 0)) / (2.0 * g.dx);
            ay = -(phi(1, 1) - phi(1, 0)) / (2.0 * g.dy);
            az = -(phi(2, 1) - phi(2, 0)) / (2.0 * g.dz);
        } else {
            const double ir = 1.0 / rho;
            u = m0 * ir; v = n0 * ir; w = o0 * ir;
            ax = -(phi(0, 1) - phi(0, 0)) * (0.5 * idx);
            ay = -(phi(1, 1) - phi(1, 0)) * (0.5 * idy);
            az = -(phi(2, 1) - phi(2, 0)) * (0.5 * idz);
        }
        d[1] += rho * ax;
        d[2] += rho * ay;
        d[3] += rho * az;
        d[4] += rho * ((u * ax + v * ay) + w * az);
    }
    double speed = 0.0;
    if (MODE == 0) {
        out.rho[q] = d[0]; out.mx[q] = d[1]; out.my[q] = d[2]; out.mz[q] = d[3]; out.E[q] = d[4];
    } else {
        const double dt = sc->dt;
        double rho, mx, my, mz, E;
        if (MODE == 1) {
            rho = r0 + dt * d[0]; mx = m0 + dt * d[1]; my = n0 + dt * d[2]; mz = o0 + dt * d[3]; E = e0 + dt * d[4];
        } else {
            rho = 0.5 * Uold.rho[q] + 0.5 * (r0 + dt * d[0]);
            mx = 0.5 * Uold.mx[q] + 0.5 * (m0 + dt * d[1]);
            my = 0.5 * Uold.my[q] + 0.5 * (n0 + dt * d[2]);
            mz = 0.5 * Uold.mz[q] + 0.5 * (o0 + dt * d[3]);
            E = 0.5 * Uold.E[q] + 0.5 * (e0 + dt * d[4]);
        }
        floors_e3<EXACT>(g, rho, mx, my, mz, E);
        out.rho[q] = rho; out.mx[q] = mx; out.my[q] = my; out.mz[q] = mz; out.E[q] = E;
        if (MODE == 2) speed = speed_e3<EXACT>(g, rho, mx, my, mz, E);
    }
    return speed;
}

// MODE 0: out = rhs(Uin).  MODE 1: out = Uin + dt rhs, floors.  MODE 2: out = Uold/2 + (Uin + dt
// rhs)/2, floors, and the block maxima of the new state's signal speed.
// The gravity g = -central grad phi is differenced here from the potential (the expression of the
// solver's own gradient kernel), so g is never stored.
template <int MODE, typename R>
__global__ __launch_bounds__(256) void k_e3_update(GeomE3 g, FluxE3<R> F, StateE3 Uin, StateE3 Uold, StateE3 out,
                                                   const double* __restrict__ phi, const Scalars* __restrict__ sc,
                                                   double* __restrict__ spd, size_t n) {
    const size_t q = (size_t)blockIdx.x * 256 + threadIdx.x;
    double speed = 0.0;
    if (q < n) {
        const int Nx = g.Nx, Ny = g.Ny, Nz = g.Nz;
        const int i = (int)(q & (size_t)(Nx - 1));
        const size_t line = q >> g.lx;
        const int j = (int)(line & (size_t)(Ny - 1)), k = (int)(line >> g.ly);
        const size_t sxy = (size_t)Nx * Ny;
        const size_t qn[3][2] = {{q - i + ((i - 1) & (Nx - 1)), q - i + ((i + 1) & (Nx - 1))},
                                 {q - (size_t)j * Nx + (size_t)((j - 1) & (Ny - 1)) * Nx,
                                  q - (size_t)j * Nx + (size_t)((j + 1) & (Ny - 1)) * Nx},
                                 {q - (size_t)k * sxy + (size_t)((k - 1) & (Nz - 1)) * sxy,
                                  q - (size_t)k * sxy + (size_t)((k + 1) & (Nz - 1)) * sxy}};
        speed = update_e3<MODE, std::is_same<R, double>::value>(
            g, [&](int ax, int side, int c) { return (double)F.f[ax][c][side ? qn[ax][1] : q]; },
            [&](int ax, int side) { return phi[qn[ax][side]]; }, Uin, Uold, out, sc, q);
    }
    if (MODE == 2) block_max_to(speed, spd);
}

// ------------------------------------------------------------------ faces + update in one pass
// A block owns a kE3TX x kE3TY column of cells and marches along z over a chunk of planes.  Per plane s its
// threads solve the (TX+1) TY x faces and TX (TY+1) y faces of plane s from the plane's primitives in LDS (a
// 2-cell halo, loaded one plane ahead into registers), and each column thread the z face between cells s and
// s+1 from the primitives of cells s-1 .. s+2 (three carried, one loaded a plane ahead); the x and y fluxes go
// through LDS, the z fluxes stay in registers, and the thread updates cell s (update_e3, the split update's
// arithmetic).  The march starts one plane early so that a chunk's lowest z face comes from the same code as
// the others; faces on a tile boundary are solved by both tiles, with the same arithmetic.  Replaces the face
// kernels and k_e3_update: no flux arrays.  The speed maxima of MODE 2 are one partial per block (nblk = the
// grid's block count, which k_e3_speed also uses).
constexpr int kE3TX = 32, kE3TY = 7;
template <int MODE, typename R>
__global__ __launch_bounds__(256, (std::is_same<R, float>::value ? 3 : 2)) void k_e3_fused(GeomE3 g, PrimE3<R> W, StateE3 Uin, StateE3 Uold, StateE3 out,
                                                  const double* __restrict__ phi, const Scalars* __restrict__ sc,
                                                  double* __restrict__ spd, int zchunk) {
    constexpr int TX = kE3TX, TY = kE3TY, HX = TX + 4, HY = TY + 4, NP = HX * HY;
    constexpr int NXF = (TX + 1) * TY, NYF = TX * (TY + 1), NC = TX * TY;
    static_assert(NXF <= 256 && NYF <= 256 && NC <= 256 && NP <= 512, "tile does not fit the block");
    __shared__ R s_w[5][NP];                      // primitives of plane s
    __shared__ R s_fx[5][NXF], s_fy[5][NYF];      // x and y face fluxes of plane s
    const int Nx = g.Nx, Ny = g.Ny, Nz = g.Nz;
    const int t = threadIdx.x;
    const int ntx = (Nx + TX - 1) / TX;
    const int i0 = (int)(blockIdx.x % (unsigned)ntx) * TX, j0 = (int)(blockIdx.x / (unsigned)ntx) * TY;
    const int c0 = (int)blockIdx.y * zchunk, c1 = min(Nz, c0 + zchunk);
    const size_t NXY = (size_t)Nx * Ny;
    // plane positions this thread stages (t and t + 256; the 2 x 2 corners are not read)
    size_t pofs[2];
    bool pact[2];
#pragma unroll
    for (int r = 0; r < 2; ++r) {
        const int pos = t + 256 * r;
        const int hx = pos % HX, hy = pos / HX;
        pact[r] = pos < NP && !((hx < 2 || hx >= TX + 2) && (hy < 2 || hy >= TY + 2));
        pofs[r] = (size_t)((j0 - 2 + hy) & (Ny - 1)) * Nx + (size_t)((i0 - 2 + hx) & (Nx - 1));
    }
    R pv[2][5];
    auto load_plane = [&](int s) {
        const size_t zb = (size_t)(s & (Nz - 1)) * NXY;
#pragma unroll
        for (int r = 0; r < 2; ++r)
            if (pact[r]) {
#pragma unroll
                for (int c = 0; c < 5; ++c) pv[r][c] = W.w[c][zb + pofs[r]];
            }
    };
    auto store_plane = [&]() {
#pragma unroll
        for (int r = 0; r < 2; ++r)
            if (pact[r]) {
#pragma unroll
                for (int c = 0; c < 5; ++c) s_w[c][t + 256 * r] = pv[r][c];
            }
    };
    // x-face slot: the lower x face of cell (xi, xj); y-face slot: the lower y face of cell (yi, yj)
    const int xs = t % (TX + 1), ys = t / (TX + 1), xi = i0 + xs, xj = j0 + ys;
    const bool xact = t < NXF && xi <= Nx && xj < Ny;
    const int gx = t % TX, gy = t / TX, yi = i0 + gx, yj = j0 + gy;
    const bool yact = t < NYF && yi < Nx && yj <= Ny;
    // column: cell (ci, cj) of plane s
    const int ci = yi, cj = yj;
    const bool cact = t < NC && ci < Nx && cj < Ny;
    const size_t coff = (size_t)cj * Nx + ci;
    auto colw = [&](int k, R v[5]) {
        const size_t q = (size_t)(k & (Nz - 1)) * NXY + coff;
#pragma unroll
        for (int c = 0; c < 5; ++c) v[c] = W.w[c][q];
    };
    // the column's cells s-1, s, s+1 (carried), s+2 (loaded a plane ahead), at the first step s = c0-1
    R pa[5] = {}, pb[5] = {}, pc[5] = {}, pn[5] = {};
    R fzb[5] = {}, fzt[5] = {};
    if (cact) {
        colw(c0 - 2, pa);
        colw(c0 - 1, pb);
        colw(c0, pc);
        colw(c0 + 1, pn);
    }
    load_plane(c0);
    store_plane();
    if (c0 + 1 < c1) load_plane(c0 + 1);
    __syncthreads();
    double speed = 0.0;
    for (int s = c0 - 1; s < c1; ++s) {
        // (1) faces: x and y of plane s (LDS), the z face between the column's cells s and s+1
        if (s >= c0 && xact) {
            const int pb0 = (ys + 2) * HX + xs;
            R f[5];
            face_e3<R, 0>(g, [&](int o, int c) { return s_w[c][pb0 + o]; }, f);
#pragma unroll
            for (int c = 0; c < 5; ++c) s_fx[c][t] = f[c];
        }
        if (s >= c0 && yact) {
            const int pb0 = gy * HX + gx + 2;
            R f[5];
            face_e3<R, 1>(g, [&](int o, int c) { return s_w[c][pb0 + o * HX]; }, f);
#pragma unroll
            for (int c = 0; c < 5; ++c) s_fy[c][t] = f[c];
        }
        if (cact)
            face_e3<R, 2>(g, [&](int o, int c) { return o == 0 ? pa[c] : o == 1 ? pb[c] : o == 2 ? pc[c] : pn[c]; },
                          fzt);
        __syncthreads();
        // (2) the update of cell s
        if (s >= c0 && cact) {
            const int sx = gy * (TX + 1) + gx, sy = gy * TX + gx;
            const size_t q = (size_t)s * NXY + coff;
            const size_t qn[3][2] = {{q - ci + ((ci - 1) & (Nx - 1)), q - ci + ((ci + 1) & (Nx - 1))},
                                     {q - (size_t)cj * Nx + (size_t)((cj - 1) & (Ny - 1)) * Nx,
                                      q - (size_t)cj * Nx + (size_t)((cj + 1) & (Ny - 1)) * Nx},
                                     {q - (size_t)s * NXY + (size_t)((s - 1) & (Nz - 1)) * NXY,
                                      q - (size_t)s * NXY + (size_t)((s + 1) & (Nz - 1)) * NXY}};
            const double sp = update_e3<MODE, std::is_same<R, double>::value>(
                g,
                [&](int ax, int side, int c) {
                    return ax == 0 ? (double)s_fx[c][sx + side]
                                   : ax == 1 ? (double)s_fy[c][sy + side * TX] : (double)(side ? fzt[c] : fzb[c]);
                },
                [&](int ax, int side) { return phi[qn[ax][side]]; }, Uin, Uold, out, sc, q);
            speed = smax(speed, sp);
        }
        if (cact) {  // the z face between s and s+1 becomes the lower face of cell s+1; shift the column
#pragma unroll
            for (int c = 0; c < 5; ++c) {
                fzb[c] = fzt[c];
                pa[c] = pb[c];
                pb[c] = pc[c];
                pc[c] = pn[c];
            }
            if (s + 1 < c1) colw(s + 3, pn);
        }
        // (3) plane s+1 into LDS and the next plane's loads
        if (s >= c0 && s + 1 < c1) {
            store_plane();
            if (s + 2 < c1) load_plane(s + 2);
        }
        __syncthreads();
    }
    if (MODE == 2) block_max_to(speed, spd);
}

// The signal-speed maxima of U, one partial per block over a grid-stride loop (the grid is the block count of
// the kernel that writes the partials during a run).
__global__ __launch_bounds__(256) void k_e3_speed(GeomE3 g, StateE3 U, double* __restrict__ spd, size_t n) {
    double s = 0.0;
    for (size_t q = (size_t)blockIdx.x * 256 + threadIdx.x; q < n; q += (size_t)gridDim.x * 256)
        s = smax(s, speed_e3(g, U.rho[q], U.mx[q], U.my[q], U.mz[q], U.E[q]));
    block_max_to(s, spd);
}

// First level of the speed reduction: the per-block maxima of a launch over all cells -> kDtPart values.
constexpr int kDtPart = 256;
__global__ __launch_bounds__(256) void k_e3_partmax(const double* __restrict__ part, int n, double* __restrict__ out) {
    double m = 0.0;
    for (int k = blockIdx.x * 256 + threadIdx.x; k < n; k += gridDim.x * 256) m = smax(m, part[k]);
    block_max_to(m, out);
}

// dt = cfl min(dx, dy, dz) / max speed, clamped to t_end, t advanced, dt logged
__global__ __launch_bounds__(256) void k_e3_dt(GeomE3 g, const double* __restrict__ spd, int nspd, Scalars* sc,
                                               double* __restrict__ dts) {
    const double maxs = block_reduce_max(spd, nspd, 1e-30);
    if (threadIdx.x == 0) finalize_dt(sc, maxs, smin(smin(g.dx, g.dy), g.dz), g.cfl, 0.0, dts);
}

__global__ __launch_bounds__(256) void k_e3_dt_value(GeomE3 g, const double* __restrict__ spd, int nspd,
                                                     double* __restrict__ out) {
    const double maxs = block_reduce_max(spd, nspd, 1e-30);
    if (threadIdx.x == 0) out[0] = g.cfl * smin(smin(g.dx, g.dy), g.dz) / maxs;
}

static constexpr int kDtLogE3 = 1 << 16;

}  // namespace cc

using namespace cc;

struct cc_euler3d {
    cc_euler3d_params p{};
    GeomE3 g{};
    hipStream_t stream = nullptr;
    StateE3 S{}, S1{};
    PrimE3<double> W64{};
    PrimE3<float> W32{};
    FluxE3<double> F64{};
    FluxE3<float> F32{};
    double *gx = nullptr, *gy = nullptr, *gz = nullptr;
    Scalars* sc = nullptr;
    double* dts_dev = nullptr;
    double* spd = nullptr;
    double* dtv = nullptr;
    double* spd2 = nullptr;  // kDtPart second-level maxima
    int nspd = 0;   // speed partials: the block count of the stage-2 kernel that writes them
    // faces + update in one z-marching pass (k_e3_fused, z chunks of zchunk planes) instead of the face kernels
    // and k_e3_update (CC_E3_FUSED=0)
    bool fused = true;
    int zchunk = 32;
    dim3 fused_grid() const {
        return dim3((unsigned)(((p.Nx + kE3TX - 1) / kE3TX) * ((p.Ny + kE3TY - 1) / kE3TY)),
                    (unsigned)((p.Nz + zchunk - 1) / zchunk));
    }
    Poisson3D pois;
    Profiler prof;
    FieldStats stats;
    std::vector<void*> allocs;
    hipGraphExec_t graph = nullptr;

    size_t N() const { return (size_t)p.Nx * p.Ny * p.Nz; }
    unsigned blocks() const { return (unsigned)((N() + 255) / 256); }
    Profiler* pr() { return prof.on ? &prof : nullptr; }

    template <typename T>
    cc_status alloc(T** ptr, size_t n) {
        if (hipMalloc((void**)ptr, n * sizeof(T)) != hipSuccess) {
            set_error(CC_ERR_OOM, "cc_euler3d: device allocation failed");
            return CC_ERR_OOM;
        }
        allocs.push_back(*ptr);
        CC_HIP(hipMemsetAsync(*ptr, 0, n * sizeof(T), stream));
        return CC_OK;
    }
    cc_status alloc_state(StateE3& s) {
        CC_TRY(alloc(&s.rho, N())); CC_TRY(alloc(&s.mx, N())); CC_TRY(alloc(&s.my, N()));
        CC_TRY(alloc(&s.mz, N())); CC_TRY(alloc(&s.E, N()));
        return CC_OK;
    }

    // phi of rho (the stage update differences it); with arrays, also g into gx, gy, gz
    cc_status gravity_of(const double* rho, bool arrays = false) {
        if (!p.gravity_on) return CC_OK;
        return arrays ? pois.solve(rho, gx, gy, gz, stream, pr()) : pois.solve(rho, nullptr, nullptr, nullptr, stream, pr());
    }

    // rhs of `in` and the stage update (MODE as in k_e3_update)
    template <int MODE>
    cc_status stage(const StateE3& in, const StateE3& old, const StateE3& out) {
        const size_t n = N();
        const bool mixed = p.precision == CC_MIXED;
        {
            ProfScope ps(pr(), stream, "prim");
            if (mixed)
                hipLaunchKernelGGL(k_e3_prim<float>, dim3(blocks()), dim3(256), 0, stream, g, in, W32, n);
            else
                hipLaunchKernelGGL(k_e3_prim<double>, dim3(blocks()), dim3(256), 0, stream, g, in, W64, n);
            CC_CHECK_LAUNCH();
        }
        if (fused) {
            ProfScope ps(pr(), stream, "fused");
            if (mixed)
                hipLaunchKernelGGL((k_e3_fused<MODE, float>), fused_grid(), dim3(256), 0, stream, g, W32, in, old, out,
                                   pois.phi, sc, spd, zchunk);
            else
                hipLaunchKernelGGL((k_e3_fused<MODE, double>), fused_grid(), dim3(256), 0, stream, g, W64, in, old, out,
                                   pois.phi, sc, spd, zchunk);
            CC_CHECK_LAUNCH();
            return CC_OK;
        }
        {
            ProfScope ps(pr(), stream, "face");
            const dim3 gy((p.Nx + 63) / 64, p.Nz, (p.Ny + 3) / 4), gz((p.Nx + 63) / 64, p.Ny, (p.Nz + 3) / 4);
            if (mixed) {
                hipLaunchKernelGGL((k_e3_face<float, 0>), dim3(blocks()), dim3(256), 0, stream, g, W32, F32, n);
                hipLaunchKernelGGL((k_e3_face<float, 1>), gy, dim3(256), 0, stream, g, W32, F32, n);
                hipLaunchKernelGGL((k_e3_face<float, 2>), gz, dim3(256), 0, stream, g, W32, F32, n);
            } else {
                hipLaunchKernelGGL((k_e3_face<double, 0>), dim3(blocks()), dim3(256), 0, stream, g, W64, F64, n);
                hipLaunchKernelGGL((k_e3_face<double, 1>), gy, dim3(256), 0, stream, g, W64, F64, n);
                hipLaunchKernelGGL((k_e3_face<double, 2>), gz, dim3(256), 0, stream, g, W64, F64, n);
            }
            CC_CHECK_LAUNCH();
        }
        {
            ProfScope ps(pr(), stream, "update");
            if (mixed)
                hipLaunchKernelGGL((k_e3_update<MODE, float>), dim3(blocks()), dim3(256), 0, stream, g, F32, in, old, out,
                                   pois.phi, sc, spd, n);
            else
                hipLaunchKernelGGL((k_e3_update<MODE, double>), dim3(blocks()), dim3(256), 0, stream, g, F64, in, old,
                                   out, pois.phi, sc, spd, n);
            CC_CHECK_LAUNCH();
        }
        return CC_OK;
    }

    // One Heun step of S (with driver: dt from the last speed maxima first).
    cc_status heun(bool driver) {
        if (driver) {
            ProfScope ps(pr(), stream, "dt");
            hipLaunchKernelGGL(k_e3_partmax, dim3(kDtPart), dim3(256), 0, stream, spd, nspd, spd2);
            hipLaunchKernelGGL(k_e3_dt, dim3(1), dim3(256), 0, stream, g, spd2, kDtPart, sc, dts_dev);
            CC_CHECK_LAUNCH();
        }
        CC_TRY(gravity_of(S.rho));
        CC_TRY(stage<1>(S, S, S1));
        CC_TRY(gravity_of(S1.rho));
        CC_TRY(stage<2>(S1, S, S));
        return CC_OK;
    }

    cc_status speed_max() {
        ProfScope ps(pr(), stream, "dt");
        hipLaunchKernelGGL(k_e3_speed, dim3(nspd), dim3(256), 0, stream, g, S, spd, N());
        CC_CHECK_LAUNCH();
        return CC_OK;
    }

    ~cc_euler3d() {
        if (graph) hipGraphExecDestroy(graph);
        for (void* a : allocs) hipFree(a);
        pois.destroy();
        if (stream) hipStreamDestroy(stream);
    }
};

#define CHECK_E3()                                                                         \
    do {                                                                                   \
        if (!h) { set_error(CC_ERR_INVALID_ARG, "null handle"); return CC_ERR_INVALID_ARG; } \
    } while (0)

namespace {
bool pow2_8_1024(int n) { return n >= 8 && n <= 1024 && (n & (n - 1)) == 0; }
int log2i(int n) {
    int l = 0;
    while ((1 << l) < n) ++l;
    return l;
}
}  // namespace

extern "C" {

cc_status cc_euler3d_create(const cc_euler3d_params* p, cc_euler3d** out) {
    if (!p || !out) { set_error(CC_ERR_INVALID_ARG, "null argument"); return CC_ERR_INVALID_ARG; }
    *out = nullptr;
    if (!pow2_8_1024(p->Nx) || !pow2_8_1024(p->Ny) || !pow2_8_1024(p->Nz)) {
        set_error(CC_ERR_NON_POW2, "cc_euler3d: Nx, Ny, Nz must be powers of two in [8, 1024] (periodic FFT gravity)");
        return CC_ERR_NON_POW2;
    }
    if (!(p->x1 > p->x0) || !(p->y1 > p->y0) || !(p->z1 > p->z0) || !(p->gamma > 1.0) || !(p->cfl > 0.0)) {
        set_error(CC_ERR_INVALID_ARG, "cc_euler3d: invalid domain, gamma or cfl");
        return CC_ERR_INVALID_ARG;
    }
    if (p->precision != CC_FP64 && p->precision != CC_MIXED) { set_error(CC_ERR_INVALID_ARG, "bad precision"); return CC_ERR_INVALID_ARG; }
    if (p->poisson_max_iter < 0 || p->poisson_max_iter > 64 || !(p->poisson_tol >= 0.0)) {
        set_error(CC_ERR_INVALID_ARG, "cc_euler3d: poisson_max_iter must be in [0, 64] and poisson_tol >= 0");
        return CC_ERR_INVALID_ARG;
    }
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) { set_error(CC_ERR_NO_DEVICE, "no HIP device"); return CC_ERR_NO_DEVICE; }
    cc_euler3d* h = new cc_euler3d();
    h->p = *p;
    GeomE3& g = h->g;
    g.Nx = p->Nx; g.Ny = p->Ny; g.Nz = p->Nz; g.lx = log2i(p->Nx); g.ly = log2i(p->Ny);
    g.dx = (p->x1 - p->x0) / p->Nx; g.dy = (p->y1 - p->y0) / p->Ny; g.dz = (p->z1 - p->z0) / p->Nz;
    g.gamma = p->gamma; g.gm1 = p->gamma - 1.0; g.floor_rho = p->floor_rho; g.floor_P = p->floor_P; g.cfl = p->cfl;
    g.gravity_on = p->gravity_on != 0;
    if (hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking) != hipSuccess) {
        set_error(CC_ERR_HIP, "hipStreamCreate failed");
        delete h;
        return CC_ERR_HIP;
    }
    const size_t n = h->N();
    cc_status st = h->alloc_state(h->S);
    if (st == CC_OK) st = h->alloc_state(h->S1);
    if (p->precision == CC_MIXED) {
        for (int c = 0; c < 5 && st == CC_OK; ++c) st = h->alloc(&h->W32.w[c], n);
        for (int a = 0; a < 3; ++a)
            for (int c = 0; c < 5 && st == CC_OK; ++c) st = h->alloc(&h->F32.f[a][c], n);
    } else {
        for (int c = 0; c < 5 && st == CC_OK; ++c) st = h->alloc(&h->W64.w[c], n);
        for (int a = 0; a < 3; ++a)
            for (int c = 0; c < 5 && st == CC_OK; ++c) st = h->alloc(&h->F64.f[a][c], n);
    }
    if (st == CC_OK) st = h->alloc(&h->gx, n);
    if (st == CC_OK) st = h->alloc(&h->gy, n);
    if (st == CC_OK) st = h->alloc(&h->gz, n);
    if (st == CC_OK) st = h->alloc(&h->dts_dev, kDtLogE3);
    if (const char* v = std::getenv("CC_E3_FUSED")) h->fused = std::atoi(v) != 0;
    if (const char* v = std::getenv("CC_E3_ZCHUNK")) h->zchunk = std::max(1, std::atoi(v));
    h->nspd = h->fused ? (int)(h->fused_grid().x * h->fused_grid().y) : (int)h->blocks();
    if (st == CC_OK) st = h->alloc(&h->spd, (size_t)h->nspd);
    if (st == CC_OK) st = h->alloc(&h->dtv, 1);
    if (st == CC_OK) st = h->alloc(&h->spd2, kDtPart);
    if (st == CC_OK) st = h->alloc(&h->sc, 1);
    // The gravity solve of a mixed handle: one FP32-FFT correction gains ~7 digits of the residual, so a
    // tolerance below 1e-6 takes two of them, which cost more than the one FP64-FFT correction that reaches
    // it (512^3: 2 x 2.8 against 3.9 ms, the solves 7.1 against 5.7 ms); such a handle runs the FP64 FFT.
    // CC_E3_FFT=fp32 / fp64 overrides the choice.
    int grav_prec = p->precision;
    if (p->precision == CC_MIXED && p->poisson_tol < 1e-6) grav_prec = CC_FP64;
    if (const char* v = std::getenv("CC_E3_FFT"))
        grav_prec = p->precision == CC_MIXED && std::string(v) == "fp32" ? CC_MIXED : CC_FP64;
    if (st == CC_OK && p->gravity_on)
        st = h->pois.init(p->Nx, p->Ny, p->Nz, g.dx, g.dy, g.dz, p->G, p->floor_rho, p->poisson_tol, p->poisson_max_iter,
                          grav_prec, h->stream);
    if (st == CC_OK) {
        hipLaunchKernelGGL(k_set_scalars, dim3(1), dim3(1), 0, h->stream, h->sc, 0.0, 0.0, 1e300, 1);
        if (hipStreamSynchronize(h->stream) != hipSuccess) { set_error(CC_ERR_HIP, "sync failed"); st = CC_ERR_HIP; }
    }
    if (st != CC_OK) { delete h; return st; }
    *out = h;
    return CC_OK;
}

cc_status cc_euler3d_destroy(cc_euler3d* h) {
    if (!h) return CC_OK;
    hipStreamSynchronize(h->stream);
    delete h;
    return CC_OK;
}

cc_status cc_euler3d_upload(cc_euler3d* h, const double* rho, const double* mx, const double* my, const double* mz,
                            const double* E) {
    CHECK_E3();
    if (!rho || !mx || !my || !mz || !E) { set_error(CC_ERR_INVALID_ARG, "null field"); return CC_ERR_INVALID_ARG; }
    const size_t b = h->N() * sizeof(double);
    CC_HIP(hipMemcpyAsync(h->S.rho, rho, b, hipMemcpyHostToDevice, h->stream));
    CC_HIP(hipMemcpyAsync(h->S.mx, mx, b, hipMemcpyHostToDevice, h->stream));
    CC_HIP(hipMemcpyAsync(h->S.my, my, b, hipMemcpyHostToDevice, h->stream));
    CC_HIP(hipMemcpyAsync(h->S.mz, mz, b, hipMemcpyHostToDevice, h->stream));
    CC_HIP(hipMemcpyAsync(h->S.E, E, b, hipMemcpyHostToDevice, h->stream));
    CC_HIP(hipStreamSynchronize(h->stream));
    return CC_OK;
}

static cc_status download_state(cc_euler3d* h, const StateE3& s, double* rho, double* mx, double* my, double* mz,
                                double* E) {
    const size_t b = h->N() * sizeof(double);
    if (rho) CC_HIP(hipMemcpyAsync(rho, s.rho, b, hipMemcpyDeviceToHost, h->stream));
    if (mx) CC_HIP(hipMemcpyAsync(mx, s.mx, b, hipMemcpyDeviceToHost, h->stream));
    if (my) CC_HIP(hipMemcpyAsync(my, s.my, b, hipMemcpyDeviceToHost, h->stream));
    if (mz) CC_HIP(hipMemcpyAsync(mz, s.mz, b, hipMemcpyDeviceToHost, h->stream));
    if (E) CC_HIP(hipMemcpyAsync(E, s.E, b, hipMemcpyDeviceToHost, h->stream));
    CC_HIP(hipStreamSynchronize(h->stream));
    return CC_OK;
}

cc_status cc_euler3d_download(cc_euler3d* h, double* rho, double* mx, double* my, double* mz, double* E) {
    CHECK_E3();
    return download_state(h, h->S, rho, mx, my, mz, E);
}

cc_status cc_euler3d_gravity(cc_euler3d* h, double* gx, double* gy, double* gz, double* phi) {
    CHECK_E3();
    CC_TRY(h->gravity_of(h->S.rho, true));
    const size_t b = h->N() * sizeof(double);
    if (!h->p.gravity_on) {
        CC_HIP(hipStreamSynchronize(h->stream));
        for (double* o : {gx, gy, gz, phi})
            if (o) std::memset(o, 0, b);
        return CC_OK;
    }
    if (gx) CC_HIP(hipMemcpyAsync(gx, h->gx, b, hipMemcpyDeviceToHost, h->stream));
    if (gy) CC_HIP(hipMemcpyAsync(gy, h->gy, b, hipMemcpyDeviceToHost, h->stream));
    if (gz) CC_HIP(hipMemcpyAsync(gz, h->gz, b, hipMemcpyDeviceToHost, h->stream));
    if (phi) CC_HIP(hipMemcpyAsync(phi, h->pois.phi, b, hipMemcpyDeviceToHost, h->stream));
    CC_HIP(hipStreamSynchronize(h->stream));
    return CC_OK;
}

cc_status cc_euler3d_poisson_stats(cc_euler3d* h, int* solves, int* iters_last, int* iters_max, int* iters_total,
                                   double* resid_last, double* resid_worst, int* unconverged) {
    CHECK_E3();
    Poisson3DStatus s{};
    if (h->p.gravity_on) {
        CC_HIP(hipMemcpyAsync(&s, h->pois.st, sizeof(s), hipMemcpyDeviceToHost, h->stream));
        CC_HIP(hipStreamSynchronize(h->stream));
    }
    if (solves) *solves = s.solves;
    if (iters_last) *iters_last = s.iters_last;
    if (iters_max) *iters_max = s.iters_max;
    if (iters_total) *iters_total = s.iters_total;
    if (resid_last) *resid_last = s.resid;
    if (resid_worst) *resid_worst = s.resid_worst;
    if (unconverged) *unconverged = s.unconverged;
    return CC_OK;
}

cc_status cc_euler3d_rhs(cc_euler3d* h, double* drho, double* dmx, double* dmy, double* dmz, double* dE) {
    CHECK_E3();
    CC_TRY(h->gravity_of(h->S.rho));
    CC_TRY(h->stage<0>(h->S, h->S, h->S1));
    return download_state(h, h->S1, drho, dmx, dmy, dmz, dE);
}

cc_status cc_euler3d_compute_dt(cc_euler3d* h, double* dt) {
    CHECK_E3();
    if (!dt) { set_error(CC_ERR_INVALID_ARG, "null dt"); return CC_ERR_INVALID_ARG; }
    CC_TRY(h->speed_max());
    hipLaunchKernelGGL(k_e3_partmax, dim3(kDtPart), dim3(256), 0, h->stream, h->spd, h->nspd, h->spd2);
    hipLaunchKernelGGL(k_e3_dt_value, dim3(1), dim3(256), 0, h->stream, h->g, h->spd2, kDtPart, h->dtv);
    CC_CHECK_LAUNCH();
    CC_HIP(hipMemcpyAsync(dt, h->dtv, sizeof(double), hipMemcpyDeviceToHost, h->stream));
    CC_HIP(hipStreamSynchronize(h->stream));
    return CC_OK;
}

cc_status cc_euler3d_rk2_step(cc_euler3d* h, double dt) {
    CHECK_E3();
    hipLaunchKernelGGL(k_set_scalars, dim3(1), dim3(1), 0, h->stream, h->sc, dt, 0.0, 1e300, 0);
    CC_TRY(h->heun(false));
    CC_HIP(hipStreamSynchronize(h->stream));
    return CC_OK;
}

cc_status cc_euler3d_run(cc_euler3d* h, int nsteps, double* t_inout, double t_end, double* dts) {
    CHECK_E3();
    if (nsteps < 0 || !t_inout) { set_error(CC_ERR_INVALID_ARG, "nsteps < 0 or null t"); return CC_ERR_INVALID_ARG; }
    if (nsteps == 0) return CC_OK;
    if (nsteps > kDtLogE3) {
        double t = *t_inout;
        for (int done = 0; done < nsteps;) {
            const int n = std::min(kDtLogE3, nsteps - done);
            CC_TRY(cc_euler3d_run(h, n, &t, t_end, dts ? dts + done : nullptr));
            done += n;
        }
        *t_inout = t;
        return CC_OK;
    }
    hipLaunchKernelGGL(k_set_scalars, dim3(1), dim3(1), 0, h->stream, h->sc, 0.0, *t_inout, t_end, 1);
    CC_TRY(h->speed_max());
    const bool use_graph = !h->prof.on;
    for (int s = 0; s < nsteps; ++s) {
        if (use_graph) {
            if (!h->graph) {
                hipGraph_t gr;
                CC_HIP(hipStreamBeginCapture(h->stream, hipStreamCaptureModeThreadLocal));
                cc_status st = h->heun(true);
                hipError_t e = hipStreamEndCapture(h->stream, &gr);
                if (st != CC_OK) return st;
                if (e != hipSuccess) { set_error(CC_ERR_HIP, "graph capture failed"); return CC_ERR_HIP; }
                e = hipGraphInstantiate(&h->graph, gr, nullptr, nullptr, 0);
                hipGraphDestroy(gr);
                if (e != hipSuccess) { set_error(CC_ERR_HIP, "graph instantiation failed"); return CC_ERR_HIP; }
            }
            CC_HIP(hipGraphLaunch(h->graph, h->stream));
        } else {
            CC_TRY(h->heun(true));
        }
    }
    Scalars s;
    CC_HIP(hipMemcpyAsync(&s, h->sc, sizeof(Scalars), hipMemcpyDeviceToHost, h->stream));
    if (dts) CC_HIP(hipMemcpyAsync(dts, h->dts_dev, (size_t)nsteps * sizeof(double), hipMemcpyDeviceToHost, h->stream));
    CC_HIP(hipStreamSynchronize(h->stream));
    *t_inout = s.t;
    if (!std::isfinite(s.dt) || s.dt < 0.0) {
        set_error(CC_ERR_BAD_DT, "non-finite or negative dt on device");
        return CC_ERR_BAD_DT;
    }
    return CC_OK;
}

cc_status cc_euler3d_totals(cc_euler3d* h, double* out5) {
    CHECK_E3();
    if (!out5) { set_error(CC_ERR_INVALID_ARG, "null output"); return CC_ERR_INVALID_ARG; }
    FieldPtrs a{{h->S.rho, h->S.mx, h->S.my, h->S.mz, h->S.E}};
    double r[15];
    CC_TRY(h->stats.run(a, nullptr, 5, h->N(), h->stream, r));
    const double dV = h->g.dx * h->g.dy * h->g.dz;
    for (int f = 0; f < 5; f++) out5[f] = r[3 * f] * dV;
    return CC_OK;
}

cc_status cc_euler3d_sync(cc_euler3d* h) {
    CHECK_E3();
    CC_HIP(hipStreamSynchronize(h->stream));
    return CC_OK;
}

cc_status cc_euler3d_set_profiling(cc_euler3d* h, int on) {
    CHECK_E3();
    h->prof.on = on != 0;
    return CC_OK;
}

cc_status cc_euler3d_kernel_time(cc_euler3d* h, const char* kernel, int* count, double* total_ms) {
    CHECK_E3();
    if (!kernel || !count || !total_ms) { set_error(CC_ERR_INVALID_ARG, "null argument"); return CC_ERR_INVALID_ARG; }
    CC_HIP(hipStreamSynchronize(h->stream));
    h->prof.get_total(kernel, count, total_ms);
    return CC_OK;
}

cc_status cc_euler3d_reset_profiling(cc_euler3d* h) {
    CHECK_E3();
    CC_HIP(hipStreamSynchronize(h->stream));
    h->prof.reset();
    return CC_OK;
}

}  // extern "C"
