// mhd2d.hip -- gfx950 solver step for 2D resistive MHD with HLLD + constrained transport + SOR
// self-gravity (restating mhd2d_resistive_hlld_ct_gravity.cpp, cited as M2:<line>).
//
// One forward-Euler stage (advance_stage, M2:690-834) is two kernels:
//   k_flux2d      PLM-MC reconstruction + HLLD on x- and y-faces of a 32x8 tile from primitives
//                 staged once per cell in LDS, flux difference + floors (old faces B0)  -> U*,
//                 face EMFs Ez_x, Ez_y
//   k_ct_src2d    corner Ez = 1/4 sum of face EMFs + eta Jz(B0) of a 64x4 tile into LDS, CT face
//                 update, gravity source, floors (B1), Joule heating, Bz diffusion, floors (B1); the
//                 stage-2 instance also forms the Heun average, its floors and the next step's CFL
//                 signal-speed maximum (wave64 shuffles, one partial per block).
// Heun (rk2_step, M2:838-869) = stage(U^n, g) -> U1; g1 = Poisson(U1.rho); stage(U1, g1) + average.
// In the mixed-precision form the HLLD runs in FP32 (recomputed in FP64 for faces near the Alfven
// degeneracy) on FP64 primitives and PLM states; flux differences, conserved updates, EMF averages
// and CT stay FP64.
#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "diag.hpp"
#include "poisson2d.hpp"
#include "riemann.hpp"
#include "stepcommon.hpp"

namespace cc {

struct Geom2D {
    int Nx, Ny, px, py;
    double dx, dy, gamma, gm1, floor_rho, floor_P, eta, cfl;
    int gravity_on;
    int ct_comp;        // compensated face update (FP32 low words of the face field)
    double hlld_guard;  // HLLD degeneracy guard (0: the reference's 1e-14 only)
};

struct State2D {
    double *rho, *mx, *my, *mz, *E, *Bz;  // Nx*Ny
    double *Bx, *By;                      // (Nx+1)*Ny, Nx*(Ny+1)
    float *Bx_lo, *By_lo;                 // rounding errors of Bx, By (compensated face update only)
};

// ------------------------------------------------------------------ cell helpers (FP64)
__device__ __forceinline__ void enforce_floors2d(const Geom2D& g, double& rho, double& mx, double& my, double& mz,
                                                 double& E, double Bx, double By, double Bz) {  // M2:114-124
    rho = smax(rho, g.floor_rho);
    const double u = mx / rho, v = my / rho, w = mz / rho;
    const double kin = 0.5 * rho * (u * u + v * v + w * w);
    const double mag = 0.5 * (Bx * Bx + By * By + Bz * Bz);
    const double P = g.gm1 * (E - kin - mag);
    if (P < g.floor_P) {
        const double eint = g.floor_P / g.gm1;
        E = kin + mag + eint;
    }
}

__device__ __forceinline__ Prim<double> cons_to_prim2d(const Geom2D& g, double rho, double mx, double my, double mz,
                                                       double E, double Bx, double By, double Bz) {  // M2:126-135
    rho = smax(rho, g.floor_rho);
    const double u = mx / rho, v = my / rho, w = mz / rho;
    const double kin = 0.5 * rho * (u * u + v * v + w * w);
    const double mag = 0.5 * (Bx * Bx + By * By + Bz * Bz);
    double P = g.gm1 * (E - kin - mag);
    if (P < g.floor_P) P = g.floor_P;
    return Prim<double>{rho, u, v, w, P, Bx, By, Bz};
}

__device__ __forceinline__ double signal_speed2d(const Geom2D& g, const Prim<double>& W) {  // M2:585-587
    IdealEOS2D<double> eos{g.gm1, g.gamma, g.floor_rho};
    const NoAux ax{};
    const double sx = fabs(W.u) + fast_speed<double>(eos, W, W.Bx, ax);
    const double sy = fabs(W.v) + fast_speed<double>(eos, W, W.By, ax);
    return smax(sx, sy);
}

// ------------------------------------------------------------------ flux + FV kernel
constexpr int TX2 = 32, TY2 = 8, RX2 = TX2 + 4, RY2 = TY2 + 4;

// The HLLD of one face from FP64 face states: FP64, or (MIXED) FP32 with FP64 fallback near the
// Alfven degeneracy.
template <bool MIXED>
__device__ __forceinline__ void face_hlld2d(const Geom2D& g, const Prim<double>& WL, const Prim<double>& WR, double Bn,
                                            double F[7]) {
    const IdealEOS2D<double> e64{g.gm1, g.gamma, g.floor_rho};
    if (MIXED) {
        const IdealEOS2D<float> e32{(float)g.gm1, (float)g.gamma, (float)g.floor_rho};
        hlld_x_mixed(e32, e64, WL, WR, Bn, F, default_degen_tau(g.hlld_guard), g.hlld_guard);
    } else {
        hlld_x<double>(e64, WL, WR, Bn, fabs(Bn) < 1e-12, F, nullptr, 0.0, g.hlld_guard);
    }
}

__device__ __forceinline__ Prim<double> lds_prim2(const double (*sW)[RY2][RX2], int ly, int lx) {
    return Prim<double>{sW[0][ly][lx], sW[1][ly][lx], sW[2][ly][lx], sW[3][ly][lx],
                        sW[4][ly][lx], sW[5][ly][lx], sW[6][ly][lx], sW[7][ly][lx]};
}

template <bool MIXED>
__global__ __launch_bounds__(256) void k_flux2d(Geom2D g, State2D U0, const Scalars* __restrict__ sc, State2D Us,
                                                double* __restrict__ ezx, double* __restrict__ ezy) {
    constexpr int TX = TX2, TY = TY2, RX = RX2, RY = RY2;
    __shared__ double sW[8][RY][RX];
    __shared__ double sFx[6][TY][TX + 1];
    __shared__ double sFy[6][TY + 1][TX];
    const int tid = threadIdx.x;
    const int i0 = blockIdx.x * TX, j0 = blockIdx.y * TY;
    const int Nx = g.Nx, Ny = g.Ny;

    // 1. primitives of the tile + 2-cell halo, boundary-mapped as get_prim_cc does (M2:278-286)
    for (int c = tid; c < RX * RY; c += 256) {
        const int lx = c % RX, ly = c / RX;
        const int i = bmap(i0 - 2 + lx, Nx, g.px), j = bmap(j0 - 2 + ly, Ny, g.py);
        const size_t q = (size_t)j * Nx + i;
        const double bx = 0.5 * (U0.Bx[(size_t)j * (Nx + 1) + i] + U0.Bx[(size_t)j * (Nx + 1) + i + 1]);
        const double by = 0.5 * (U0.By[(size_t)j * Nx + i] + U0.By[(size_t)(j + 1) * Nx + i]);
        const Prim<double> W = cons_to_prim2d(g, U0.rho[q], U0.mx[q], U0.my[q], U0.mz[q], U0.E[q], bx, by, U0.Bz[q]);
        sW[0][ly][lx] = W.rho; sW[1][ly][lx] = W.u; sW[2][ly][lx] = W.v; sW[3][ly][lx] = W.w;
        sW[4][ly][lx] = W.P; sW[5][ly][lx] = W.Bx; sW[6][ly][lx] = W.By; sW[7][ly][lx] = W.Bz;
    }
    __syncthreads();
    const double frho = g.floor_rho, fP = g.floor_P;

    // 2. x-faces i0..i0+TX of rows j0..j0+TY-1 (M2:706-725)
    for (int f = tid; f < (TX + 1) * TY; f += 256) {
        const int fx = f % (TX + 1), fy = f / (TX + 1);
        const int i = i0 + fx, j = j0 + fy;
        if (i > Nx || j >= Ny) continue;
        const int ly = fy + 2;
        Prim<double> a, b, WL, WR;
        plm_mc<double>(lds_prim2(sW, ly, fx), lds_prim2(sW, ly, fx + 1), lds_prim2(sW, ly, fx + 2), 0, frho, fP, a,
                       WL);  // right state of cell i-1
        plm_mc<double>(lds_prim2(sW, ly, fx + 1), lds_prim2(sW, ly, fx + 2), lds_prim2(sW, ly, fx + 3), 0, frho, fP,
                       WR, b);  // left state of cell i
        double F[7];
        face_hlld2d<MIXED>(g, WL, WR, U0.Bx[(size_t)j * (Nx + 1) + i], F);
        sFx[0][fy][fx] = F[0]; sFx[1][fy][fx] = F[1]; sFx[2][fy][fx] = F[2];
        sFx[3][fy][fx] = F[3]; sFx[4][fy][fx] = F[4]; sFx[5][fy][fx] = F[6];
        if (fx < TX || i == Nx) ezx[(size_t)j * (Nx + 1) + i] = -F[5];
    }
    // 3. y-faces j0..j0+TY of columns i0..i0+TX-1 (M2:728-747), by the (u,v)/(Bx,By) swap of M2:540-561
    for (int f = tid; f < TX * (TY + 1); f += 256) {
        const int fx = f % TX, fy = f / TX;
        const int i = i0 + fx, j = j0 + fy;
        if (i >= Nx || j > Ny) continue;
        const int lx = fx + 2;
        Prim<double> a, b, WL, WR;
        plm_mc<double>(lds_prim2(sW, fy, lx), lds_prim2(sW, fy + 1, lx), lds_prim2(sW, fy + 2, lx), 1, frho, fP, a, WL);
        plm_mc<double>(lds_prim2(sW, fy + 1, lx), lds_prim2(sW, fy + 2, lx), lds_prim2(sW, fy + 3, lx), 1, frho, fP,
                       WR, b);
        const double Bn = U0.By[(size_t)j * Nx + i];
        const Prim<double> aL{WL.rho, WL.v, WL.u, WL.w, WL.P, Bn, WL.Bx, WL.Bz};
        const Prim<double> aR{WR.rho, WR.v, WR.u, WR.w, WR.P, Bn, WR.Bx, WR.Bz};
        double F[7];
        face_hlld2d<MIXED>(g, aL, aR, Bn, F);
        sFy[0][fy][fx] = F[0]; sFy[1][fy][fx] = F[2]; sFy[2][fy][fx] = F[1];
        sFy[3][fy][fx] = F[3]; sFy[4][fy][fx] = F[4]; sFy[5][fy][fx] = F[6];
        if (fy < TY || j == Ny) ezy[(size_t)j * Nx + i] = F[5];
    }
    __syncthreads();

    // 4. flux difference and floors with the old faces (M2:751-779)
    const int tx = tid % TX, ty = tid / TX;
    const int i = i0 + tx, j = j0 + ty;
    if (i >= Nx || j >= Ny) return;
    const size_t q = (size_t)j * Nx + i;
    const double dt = sc->dt, invdx = 1.0 / g.dx, invdy = 1.0 / g.dy;
    double d[6];
#pragma unroll
    for (int k = 0; k < 6; k++)
        d[k] = -(sFx[k][ty][tx + 1] - sFx[k][ty][tx]) * invdx - (sFy[k][ty + 1][tx] - sFy[k][ty][tx]) * invdy;
    double rho = U0.rho[q] + dt * d[0];
    double mx = U0.mx[q] + dt * d[1];
    double my = U0.my[q] + dt * d[2];
    double mz = U0.mz[q] + dt * d[3];
    double E = U0.E[q] + dt * d[4];
    const double Bz = U0.Bz[q] + dt * d[5];
    const double bx = 0.5 * (U0.Bx[(size_t)j * (Nx + 1) + i] + U0.Bx[(size_t)j * (Nx + 1) + i + 1]);
    const double by = 0.5 * (U0.By[(size_t)j * Nx + i] + U0.By[(size_t)(j + 1) * Nx + i]);
    enforce_floors2d(g, rho, mx, my, mz, E, bx, by, Bz);
    Us.rho[q] = rho; Us.mx[q] = mx; Us.my[q] = my; Us.mz[q] = mz; Us.E[q] = E; Us.Bz[q] = Bz;
}

// ------------------------------------------------------------------ corner EMF (M2:781-804)
// Corner (ie, je), ie in [0, Nx], je in [0, Ny]: 1/4 of the four face EMFs + eta Jz of B0 (Jz_edge,
// M2:565-575), with the reference's index maps.
__device__ __forceinline__ double corner_ez(const Geom2D& g, const double* __restrict__ ezx, const double* __restrict__ ezy,
                                           const double* __restrict__ Bx, const double* __restrict__ By, int ie, int je) {
    const int Nx = g.Nx, Ny = g.Ny;
    const int jm = bmap(je - 1, Ny, g.py), jc = bmap(je, Ny, g.py);
    const int im = bmap(ie - 1, Nx, g.px), ic = bmap(ie, Nx, g.px);
    const double ez_ideal = 0.25 * (ezx[(size_t)jm * (Nx + 1) + ie] + ezx[(size_t)jc * (Nx + 1) + ie] +
                                    ezy[(size_t)je * Nx + im] + ezy[(size_t)je * Nx + ic]);
    double ez_res = 0.0;
    if (g.eta > 0.0) {
        const double dBydx = (By[(size_t)je * Nx + ic] - By[(size_t)je * Nx + im]) / g.dx;
        const double dBxdy = (Bx[(size_t)jc * (Nx + 1) + ie] - Bx[(size_t)jm * (Nx + 1) + ie]) / g.dy;
        ez_res = g.eta * (dBydx - dBxdy);
    }
    return ez_ideal + ez_res;
}

// The corner EMFs a 64 x 4 cell tile reads: corners (i0 .. i0+65, j0 .. j0+5) in LDS (the CT of the
// tile's faces and the Joule term's B1 of the +x / +y neighbour faces), any other one (across a periodic
// boundary) evaluated where it is read.
struct CornerTile {
    static constexpr int TX = 64, TY = 4, CX = TX + 2, CY = TY + 2;
    const Geom2D& g;
    const double* ezx;
    const double* ezy;
    const double* Bx;
    const double* By;
    const double* lds;
    int i0, j0;
    __device__ double operator()(int ie, int je) const {
        const int x = ie - i0, y = je - j0;
        if (x >= 0 && x < CX && y >= 0 && y < CY && ie <= g.Nx && je <= g.Ny) return lds[y * CX + x];
        return corner_ez(g, ezx, ezy, Bx, By, ie, je);
    }
};

// ------------------------------------------------------------------ CT + sources (+ Heun) kernel
// Increments of x-face (fi, fj) and y-face (fi, fj) from the corner EMFs (M2:807-830).
template <typename EZ>
__device__ __forceinline__ double ct_dbx(const Geom2D& g, const EZ& ez, int fi, int fj, double dt) {
    return (-(dt / g.dy)) * (ez(fi, fj + 1) - ez(fi, fj));
}
template <typename EZ>
__device__ __forceinline__ double ct_dby(const Geom2D& g, const EZ& ez, int fi, int fj, double dt) {
    return (dt / g.dx) * (ez(fi + 1, fj) - ez(fi, fj));
}
// B1 of a face: the reference's B0 + increment, or its compensated form (hi word; lo into *lo).
__device__ __forceinline__ double ct_face(const Geom2D& g, const double* B0, const float* B0lo, size_t f, double d,
                                          float* lo = nullptr) {
    if (!g.ct_comp) return B0[f] + d;
    double h;
    float l;
    ct_add(B0[f], B0lo[f], d, h, l);
    if (lo) *lo = l;
    return h;
}
__device__ __forceinline__ size_t fxi(const Geom2D& g, int fi, int fj) { return (size_t)fj * (g.Nx + 1) + fi; }
__device__ __forceinline__ size_t fyi(const Geom2D& g, int fi, int fj) { return (size_t)fj * g.Nx + fi; }

// One kernel for corner EMFs, CT and the cell sources of a stage: a 64 x 4 tile evaluates the corner EMFs
// it reads into LDS (1.5 per cell), then updates its faces and cells.
template <bool FINAL>
__global__ __launch_bounds__(256) void k_ct_src2d(Geom2D g, State2D Us, const double* __restrict__ Bx0,
                                                  const double* __restrict__ By0, const float* __restrict__ Bx0lo,
                                                  const float* __restrict__ By0lo, const double* __restrict__ ezx,
                                                  const double* __restrict__ ezy,
                                                  const double* __restrict__ gx, const double* __restrict__ gy,
                                                  Scalars* __restrict__ sc, State2D Un, State2D Out,
                                                  double* __restrict__ spd) {
    constexpr int CX = CornerTile::CX, CY = CornerTile::CY;
    __shared__ double s_ez[CX * CY];
    const int i0 = blockIdx.x * CornerTile::TX, j0 = blockIdx.y * CornerTile::TY;
    const int i = i0 + (int)(threadIdx.x & 63), j = j0 + (int)(threadIdx.x >> 6);
    const int Nx = g.Nx, Ny = g.Ny;
    for (int t = threadIdx.x; t < CX * CY; t += 256) {
        const int ie = i0 + t % CX, je = j0 + t / CX;
        if (ie <= Nx && je <= Ny) s_ez[t] = corner_ez(g, ezx, ezy, Bx0, By0, ie, je);
    }
    __syncthreads();
    const CornerTile ez{g, ezx, ezy, Bx0, By0, s_ez, i0, j0};
    const bool active = i < Nx && j < Ny;
    double speed = 0.0;
    if (active) {
        const size_t q = (size_t)j * Nx + i;
        const double dt = sc->dt;
        const double dxL = ct_dbx(g, ez, i, j, dt), dxR = ct_dbx(g, ez, i + 1, j, dt);
        const double dyD = ct_dby(g, ez, i, j, dt), dyU = ct_dby(g, ez, i, j + 1, dt);
        float lxL = 0.f, lxR = 0.f, lyD = 0.f, lyU = 0.f;
        const double bxL = ct_face(g, Bx0, Bx0lo, fxi(g, i, j), dxL, &lxL);
        const double bxR = ct_face(g, Bx0, Bx0lo, fxi(g, i + 1, j), dxR, &lxR);
        const double byD = ct_face(g, By0, By0lo, fyi(g, i, j), dyD, &lyD);
        const double byU = ct_face(g, By0, By0lo, fyi(g, i, j + 1), dyU, &lyU);
        const double bx1 = 0.5 * (bxL + bxR), by1 = 0.5 * (byD + byU);
        double rho = Us.rho[q], mx = Us.mx[q], my = Us.my[q], mz = Us.mz[q], E = Us.E[q], Bz = Us.Bz[q];
        if (g.gravity_on) {  // add_gravity_sources, M2:602-629
            const double gxq = gx[q], gyq = gy[q];
            const double u = mx / smax(rho, 1e-300), v = my / smax(rho, 1e-300);
            mx += dt * rho * gxq;
            my += dt * rho * gyq;
            E += dt * rho * (u * gxq + v * gyq);
            enforce_floors2d(g, rho, mx, my, mz, E, bx1, by1, Bz);
        }
        if (g.eta > 0.0) {  // add_joule_heating_and_Bz_resistive, M2:631-686
            const int iE = bmap(i + 1, Nx, g.px), iW = bmap(i - 1, Nx, g.px);
            const int jN = bmap(j + 1, Ny, g.py), jS = bmap(j - 1, Ny, g.py);
            const double BzE = Us.Bz[(size_t)j * Nx + iE], BzW = Us.Bz[(size_t)j * Nx + iW];
            const double BzN = Us.Bz[(size_t)jN * Nx + i], BzS = Us.Bz[(size_t)jS * Nx + i];
            const double dBzdy = (BzN - BzS) / (2 * g.dy);
            const double dBzdx = (BzE - BzW) / (2 * g.dx);
            const double dBydx = (ct_face(g, By0, By0lo, fyi(g, iE, j), ct_dby(g, ez, iE, j, dt)) - byD) / g.dx;
            const double dBxdy = (ct_face(g, Bx0, Bx0lo, fxi(g, i, jN), ct_dbx(g, ez, i, jN, dt)) - bxL) / g.dy;
            const double Jx = dBzdy, Jy = -dBzdx, Jz = dBydx - dBxdy;
            const double J2 = Jx * Jx + Jy * Jy + Jz * Jz;
            E += dt * g.eta * J2;
            const double idx2 = 1.0 / (g.dx * g.dx), idy2 = 1.0 / (g.dy * g.dy);
            const double lap = (BzE - 2 * Bz + BzW) * idx2 + (BzN - 2 * Bz + BzS) * idy2;
            Bz = Bz + dt * g.eta * lap;
            enforce_floors2d(g, rho, mx, my, mz, E, bx1, by1, Bz);
        }
        if (!FINAL) {
            Out.rho[q] = rho; Out.mx[q] = mx; Out.my[q] = my; Out.mz[q] = mz; Out.E[q] = E; Out.Bz[q] = Bz;
            Out.Bx[(size_t)j * (Nx + 1) + i] = bxL;
            if (i == Nx - 1) Out.Bx[(size_t)j * (Nx + 1) + Nx] = bxR;
            Out.By[(size_t)j * Nx + i] = byD;
            if (j == Ny - 1) Out.By[(size_t)Ny * Nx + i] = byU;
            if (g.ct_comp) {
                Out.Bx_lo[(size_t)j * (Nx + 1) + i] = lxL;
                if (i == Nx - 1) Out.Bx_lo[(size_t)j * (Nx + 1) + Nx] = lxR;
                Out.By_lo[(size_t)j * Nx + i] = lyD;
                if (j == Ny - 1) Out.By_lo[(size_t)Ny * Nx + i] = lyU;
            }
        } else {  // Heun average and floors, M2:848-868
            rho = 0.5 * (Un.rho[q] + rho);
            mx = 0.5 * (Un.mx[q] + mx);
            my = 0.5 * (Un.my[q] + my);
            mz = 0.5 * (Un.mz[q] + mz);
            E = 0.5 * (Un.E[q] + E);
            Bz = 0.5 * (Un.Bz[q] + Bz);
            double nbxL, nbxR, nbyD, nbyU;
            float nlxL = 0.f, nlxR = 0.f, nlyD = 0.f, nlyU = 0.f;
            if (g.ct_comp) {  // 1/2 (B^n + B^2) of the compensated fields
                const size_t fL = fxi(g, i, j), fR = fxi(g, i + 1, j), fD = fyi(g, i, j), fU = fyi(g, i, j + 1);
                ct_avg(Un.Bx[fL], Un.Bx_lo[fL], bxL, lxL, nbxL, nlxL);
                ct_avg(Un.Bx[fR], Un.Bx_lo[fR], bxR, lxR, nbxR, nlxR);
                ct_avg(Un.By[fD], Un.By_lo[fD], byD, lyD, nbyD, nlyD);
                ct_avg(Un.By[fU], Un.By_lo[fU], byU, lyU, nbyU, nlyU);
            } else {
                nbxL = 0.5 * (Un.Bx[(size_t)j * (Nx + 1) + i] + bxL);
                nbxR = 0.5 * (Un.Bx[(size_t)j * (Nx + 1) + i + 1] + bxR);
                nbyD = 0.5 * (Un.By[(size_t)j * Nx + i] + byD);
                nbyU = 0.5 * (Un.By[(size_t)(j + 1) * Nx + i] + byU);
            }
            const double nbx = 0.5 * (nbxL + nbxR), nby = 0.5 * (nbyD + nbyU);
            enforce_floors2d(g, rho, mx, my, mz, E, nbx, nby, Bz);
            Out.rho[q] = rho; Out.mx[q] = mx; Out.my[q] = my; Out.mz[q] = mz; Out.E[q] = E; Out.Bz[q] = Bz;
            Out.Bx[(size_t)j * (Nx + 1) + i] = nbxL;
            if (i == Nx - 1) Out.Bx[(size_t)j * (Nx + 1) + Nx] = nbxR;
            Out.By[(size_t)j * Nx + i] = nbyD;
            if (j == Ny - 1) Out.By[(size_t)Ny * Nx + i] = nbyU;
            if (g.ct_comp) {
                Out.Bx_lo[(size_t)j * (Nx + 1) + i] = nlxL;
                if (i == Nx - 1) Out.Bx_lo[(size_t)j * (Nx + 1) + Nx] = nlxR;
                Out.By_lo[(size_t)j * Nx + i] = nlyD;
                if (j == Ny - 1) Out.By_lo[(size_t)Ny * Nx + i] = nlyU;
            }
            // next step's compute_dt (M2:579-598) on the final state
            speed = signal_speed2d(g, cons_to_prim2d(g, rho, mx, my, mz, E, nbx, nby, Bz));
        }
    }
    if (FINAL) block_max_to(speed, spd);
}

// ------------------------------------------------------------------ dt, diagnostics, small kernels
// (the grid of k_ct_src2d, 64 x 4 cells per block, so both leave the same number of per-block maxima)
__global__ __launch_bounds__(256) void k_speed_max2d(Geom2D g, State2D U, double* __restrict__ spd) {
    const int i = blockIdx.x * 64 + (int)(threadIdx.x & 63), j = blockIdx.y * 4 + (int)(threadIdx.x >> 6);
    double speed = 0.0;
    if (i < g.Nx && j < g.Ny) {
        const size_t q = (size_t)j * g.Nx + i;
        const double bx = 0.5 * (U.Bx[(size_t)j * (g.Nx + 1) + i] + U.Bx[(size_t)j * (g.Nx + 1) + i + 1]);
        const double by = 0.5 * (U.By[(size_t)j * g.Nx + i] + U.By[(size_t)(j + 1) * g.Nx + i]);
        speed = signal_speed2d(g, cons_to_prim2d(g, U.rho[q], U.mx[q], U.my[q], U.mz[q], U.E[q], bx, by, U.Bz[q]));
    }
    block_max_to(speed, spd);
}


// dt from the signal-speed maximum over the per-block partial maxima of the last speed pass,
// seeded with 1e-14 as at M2:581.
__global__ __launch_bounds__(256) void k_dt_finalize(Geom2D g, Scalars* sc, double* dts,
                                                     const double* __restrict__ spd, int nspd) {
    const double maxs = block_reduce_max(spd, nspd, 1e-14);
    if (threadIdx.x == 0) finalize_dt(sc, maxs, smin(g.dx, g.dy), g.cfl, g.eta, dts);
}

__global__ void k_set_scalars(Scalars* sc, double dt, double t, double t_end, int reset_step) {
    sc->dt = dt;
    sc->t = t;
    sc->t_end = t_end;
    if (reset_step) sc->step = 0;
    sc->divb = 0.0;
}

__global__ __launch_bounds__(256) void k_divb_max2d(Geom2D g, State2D U, Scalars* __restrict__ sc) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x, j = blockIdx.y;
    double v = 0.0;
    if (i < g.Nx) {  // divB_cc, M2:873-879 (faces i+1 <= Nx and j+1 <= Ny need no wrapping)
        const double BxR = U.Bx[(size_t)j * (g.Nx + 1) + i + 1], BxL = U.Bx[(size_t)j * (g.Nx + 1) + i];
        const double ByU = U.By[(size_t)(j + 1) * g.Nx + i], ByD = U.By[(size_t)j * g.Nx + i];
        v = fabs((BxR - BxL) / g.dx + (ByU - ByD) / g.dy);
    }
    v = wave_max(v);
    if ((threadIdx.x & 63) == 0 && v > 0.0) atomic_max_pos((unsigned long long*)&sc->divb, v);
}

// Snapshot frame (the store lambda of M2:1065-1106): cons_to_prim on face-averaged B, sound speed,
// Mach, cell divB.  No FP contraction, so the frame rounds as the reference's does.
__global__ __launch_bounds__(256) void k_frame2d(Geom2D g, State2D U, double* __restrict__ out) {
#pragma clang fp contract(off)
    const int i = blockIdx.x * blockDim.x + threadIdx.x, j = blockIdx.y;
    if (i >= g.Nx) return;
    const size_t N = (size_t)g.Nx * g.Ny, q = (size_t)j * g.Nx + i, fl = (size_t)j * (g.Nx + 1) + i;
    const double bxL = U.Bx[fl], bxR = U.Bx[fl + 1], byD = U.By[q], byU = U.By[q + g.Nx];
    const double bx = 0.5 * (bxL + bxR), by = 0.5 * (byD + byU), bz = U.Bz[q];
    const double rho = smax(U.rho[q], g.floor_rho);  // cons_to_prim, M2:126-135
    const double u = U.mx[q] / rho, v = U.my[q] / rho, w = U.mz[q] / rho;
    const double kin = 0.5 * rho * (u * u + v * v + w * w);
    const double mag = 0.5 * (bx * bx + by * by + bz * bz);
    double P = g.gm1 * (U.E[q] - kin - mag);
    if (P < g.floor_P) P = g.floor_P;
    const double cs = sqrt(smax(0.0, g.gamma * P / smax(rho, g.floor_rho)));  // M2:148-150, M2:1092
    const double vabs = sqrt(u * u + v * v + w * w);
    const double mach = vabs / smax(cs, 1e-14);
    const double divb = (bxR - bxL) / g.dx + (byU - byD) / g.dy;  // divB_cc, M2:873-879
    const double f[10] = {rho, P, mach, u, v, vabs, bx, by, bz, divb};
#pragma unroll
    for (int k = 0; k < 10; k++) out[(size_t)k * N + q] = f[k];
}

// Batched HLLD for tests (axis 0: hlld_x M2:372; axis 1: hlld_y M2:540), through the same face
// routine as the flux kernel.
template <bool MIXED>
__global__ void k_hlld_batch(int axis, int n, const double* __restrict__ WLa, const double* __restrict__ WRa,
                             const double* __restrict__ Bna, double gamma, double floor_rho, double* __restrict__ out) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    const double* a = WLa + 8 * (size_t)k;
    const double* b = WRa + 8 * (size_t)k;
    const double Bn = Bna[k];
    Geom2D g{};
    g.gamma = gamma; g.gm1 = gamma - 1.0; g.floor_rho = floor_rho;
    Prim<double> WL{a[0], a[1], a[2], a[3], a[4], a[5], a[6], a[7]};
    Prim<double> WR{b[0], b[1], b[2], b[3], b[4], b[5], b[6], b[7]};
    double F[7];
    double* o = out + 8 * (size_t)k;
    if (axis == 0) {
        face_hlld2d<MIXED>(g, WL, WR, Bn, F);
        for (int c = 0; c < 7; c++) o[c] = F[c];
        o[7] = -F[5];
    } else {
        const Prim<double> aL{WL.rho, WL.v, WL.u, WL.w, WL.P, Bn, WL.Bx, WL.Bz};
        const Prim<double> aR{WR.rho, WR.v, WR.u, WR.w, WR.P, Bn, WR.Bx, WR.Bz};
        face_hlld2d<MIXED>(g, aL, aR, Bn, F);
        o[0] = F[0]; o[1] = F[2]; o[2] = F[1]; o[3] = F[3]; o[4] = F[4]; o[5] = F[5]; o[6] = F[6]; o[7] = F[5];
    }
}

}  // namespace cc

// ==================================================================== host handle
using namespace cc;

static constexpr int kDtLog = 1 << 16;

struct cc_mhd2d {
    cc_mhd2d_params p{};
    Geom2D g{};
    hipStream_t stream = nullptr;
    State2D S[3]{};  // state sets; cur/aux roles rotate
    int cur = 0;
    State2D Us{};    // post-flux cells (only the 6 cell arrays are used)
    double *ezx = nullptr, *ezy = nullptr;
    double *gx = nullptr, *gy = nullptr, *g1x = nullptr, *g1y = nullptr;
    Scalars* sc = nullptr;
    double* dts_dev = nullptr;  // device log of per-step dt (kDtLog entries)
    double* spd = nullptr;      // per-block signal-speed maxima of the last speed pass
    int nspd = 0;
    // grid of the CT + sources kernel and of the signal-speed kernel: 64 x 4 cells per 256-thread block
    dim3 ct_grid() const { return dim3((unsigned)((p.Nx + 63) / 64), (unsigned)((p.Ny + 3) / 4)); }
    Poisson2D pois;
    Profiler prof;
    FrameBuffer frame;
    FieldStats stats;
    std::vector<void*> allocs;
    hipGraphExec_t graph[3] = {nullptr, nullptr, nullptr};  // step graph starting from state set k
    bool graph_ok[3] = {false, false, false};

    size_t N() const { return (size_t)p.Nx * p.Ny; }
    size_t NFx() const { return (size_t)(p.Nx + 1) * p.Ny; }
    size_t NFy() const { return (size_t)p.Nx * (p.Ny + 1); }
    Profiler* pr() { return prof.on ? &prof : nullptr; }

    cc_status alloc(double** ptr, size_t n) {
        if (hipMalloc((void**)ptr, n * sizeof(double)) != hipSuccess) {
            set_error(CC_ERR_OOM, "cc_mhd2d: device allocation failed");
            return CC_ERR_OOM;
        }
        allocs.push_back(*ptr);
        CC_HIP(hipMemsetAsync(*ptr, 0, n * sizeof(double), stream));
        return CC_OK;
    }
    cc_status alloc_f(float** ptr, size_t n) {
        if (hipMalloc((void**)ptr, n * sizeof(float)) != hipSuccess) {
            set_error(CC_ERR_OOM, "cc_mhd2d: device allocation failed");
            return CC_ERR_OOM;
        }
        allocs.push_back(*ptr);
        CC_HIP(hipMemsetAsync(*ptr, 0, n * sizeof(float), stream));
        return CC_OK;
    }
    cc_status alloc_state(State2D& s, bool lo) {
        CC_TRY(alloc(&s.rho, N())); CC_TRY(alloc(&s.mx, N())); CC_TRY(alloc(&s.my, N()));
        CC_TRY(alloc(&s.mz, N())); CC_TRY(alloc(&s.E, N())); CC_TRY(alloc(&s.Bz, N()));
        CC_TRY(alloc(&s.Bx, NFx())); CC_TRY(alloc(&s.By, NFy()));
        if (lo) {
            CC_TRY(alloc_f(&s.Bx_lo, NFx()));
            CC_TRY(alloc_f(&s.By_lo, NFy()));
        }
        return CC_OK;
    }

    // One forward-Euler stage: in -> out with gravity (gxp, gyp); FINAL adds the Heun average with `hn`.
    template <bool FINAL>
    cc_status stage(const State2D& in, const State2D& out, const double* gxp, const double* gyp, const State2D& hn) {
        const int Nx = p.Nx, Ny = p.Ny;
        {
            ProfScope ps(pr(), stream, "flux");
            dim3 grid((Nx + 31) / 32, (Ny + 7) / 8);
            if (p.precision == CC_MIXED)
                hipLaunchKernelGGL(k_flux2d<true>, grid, dim3(256), 0, stream, g, in, sc, Us, ezx, ezy);
            else
                hipLaunchKernelGGL(k_flux2d<false>, grid, dim3(256), 0, stream, g, in, sc, Us, ezx, ezy);
            CC_CHECK_LAUNCH();
        }
        {
            ProfScope ps(pr(), stream, "ct");
            hipLaunchKernelGGL(k_ct_src2d<FINAL>, ct_grid(), dim3(256), 0, stream, g, Us, in.Bx, in.By,
                               in.Bx_lo, in.By_lo, ezx, ezy, gxp, gyp, sc, hn, out, spd);
            CC_CHECK_LAUNCH();
        }
        return CC_OK;
    }

    cc_status gravity_of(const double* rho, double* gxp, double* gyp) {
        if (!p.gravity_on) return CC_OK;  // g stays zero (M2:185)
        return pois.solve(rho, gxp, gyp, stream, pr());
    }

    // The Heun step from state set a into set c using set b for U1 (rk2_step, M2:838-869).
    // With `driver`, gravity and dt come first as in the loop of M2:1112-1114.
    cc_status heun(int a, int b, int c, bool driver, double* dts) {
        if (driver) {
            CC_TRY(gravity_of(S[a].rho, gx, gy));
            ProfScope ps(pr(), stream, "dt");
            hipLaunchKernelGGL(k_dt_finalize, dim3(1), dim3(256), 0, stream, g, sc, dts, spd, nspd);
            CC_CHECK_LAUNCH();
        }
        CC_TRY(stage<false>(S[a], S[b], gx, gy, S[a]));
        CC_TRY(gravity_of(S[b].rho, g1x, g1y));
        CC_TRY(stage<true>(S[b], S[c], g1x, g1y, S[a]));
        return CC_OK;
    }

    // The snapshot frame of the current state into the frame buffer.
    cc_status make_frame() {
        CC_TRY(frame.ensure((size_t)CC_MHD2D_FRAME_FIELDS * N()));
        hipLaunchKernelGGL(k_frame2d, dim3((p.Nx + 255) / 256, p.Ny), dim3(256), 0, stream, g, S[cur], frame.d);
        CC_CHECK_LAUNCH();
        return CC_OK;
    }

    cc_status speed_max(const State2D& s) {
        ProfScope ps(pr(), stream, "dt");
        hipLaunchKernelGGL(k_speed_max2d, ct_grid(), dim3(256), 0, stream, g, s, spd);
        CC_CHECK_LAUNCH();
        return CC_OK;
    }

    ~cc_mhd2d() {
        for (auto& ge : graph)
            if (ge) hipGraphExecDestroy(ge);
        for (void* a : allocs) hipFree(a);
        if (sc) hipFree(sc);
        pois.destroy();
        if (stream) hipStreamDestroy(stream);
    }
};

namespace {
cc_status check_params2d(const cc_mhd2d_params* p) {
    if (!p) { set_error(CC_ERR_INVALID_ARG, "null params"); return CC_ERR_INVALID_ARG; }
    if (p->Nx < 2 || p->Ny < 2 || p->Nx > 65536 || p->Ny > 65536) {
        set_error(CC_ERR_INVALID_ARG, "Nx, Ny must be in [2, 65536]");
        return CC_ERR_INVALID_ARG;
    }
    if (!(p->x1 > p->x0) || !(p->y1 > p->y0) || !(p->gamma > 1.0) || p->sor_iters < 0) {
        set_error(CC_ERR_INVALID_ARG, "invalid domain, gamma or sor_iters");
        return CC_ERR_INVALID_ARG;
    }
    if (p->precision != CC_FP64 && p->precision != CC_MIXED) {
        set_error(CC_ERR_INVALID_ARG, "precision must be CC_FP64 or CC_MIXED");
        return CC_ERR_INVALID_ARG;
    }
    if (!(p->hlld_guard >= 0.0 && p->hlld_guard < 1.0)) {
        set_error(CC_ERR_INVALID_ARG, "hlld_guard must be in [0, 1)");
        return CC_ERR_INVALID_ARG;
    }
    return CC_OK;
}
}  // namespace

extern "C" {

cc_status cc_mhd2d_create(const cc_mhd2d_params* p, cc_mhd2d** out) {
    if (!out) { set_error(CC_ERR_INVALID_ARG, "null output handle"); return CC_ERR_INVALID_ARG; }
    *out = nullptr;
    CC_TRY(check_params2d(p));
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) {
        set_error(CC_ERR_NO_DEVICE, "no HIP device visible");
        return CC_ERR_NO_DEVICE;
    }
    cc_mhd2d* h = new cc_mhd2d();
    h->p = *p;
    Geom2D& g = h->g;
    g.Nx = p->Nx; g.Ny = p->Ny; g.px = p->periodic_x != 0; g.py = p->periodic_y != 0;
    g.dx = (p->x1 - p->x0) / p->Nx; g.dy = (p->y1 - p->y0) / p->Ny;
    g.gamma = p->gamma; g.gm1 = p->gamma - 1.0; g.floor_rho = p->floor_rho; g.floor_P = p->floor_P;
    g.eta = p->eta; g.cfl = p->cfl; g.gravity_on = p->gravity_on != 0;
    g.ct_comp = p->ct_compensated != 0; g.hlld_guard = p->hlld_guard;
    cc_status st = CC_OK;
    if (hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking) != hipSuccess) {
        set_error(CC_ERR_HIP, "hipStreamCreate failed");
        delete h;
        return CC_ERR_HIP;
    }
    for (int k = 0; k < 3 && st == CC_OK; k++) st = h->alloc_state(h->S[k], g.ct_comp);
    if (st == CC_OK) st = h->alloc_state(h->Us, false);
    if (st == CC_OK) st = h->alloc(&h->ezx, h->NFx());
    if (st == CC_OK) st = h->alloc(&h->ezy, h->NFy());
    if (st == CC_OK) st = h->alloc(&h->gx, h->N());
    if (st == CC_OK) st = h->alloc(&h->gy, h->N());
    if (st == CC_OK) st = h->alloc(&h->g1x, h->N());
    if (st == CC_OK) st = h->alloc(&h->g1y, h->N());
    if (st == CC_OK) st = h->alloc(&h->dts_dev, kDtLog);
    h->nspd = (int)(h->ct_grid().x * h->ct_grid().y);
    if (st == CC_OK) st = h->alloc(&h->spd, (size_t)h->nspd);
    if (st == CC_OK && hipMalloc((void**)&h->sc, sizeof(Scalars)) != hipSuccess) {
        set_error(CC_ERR_OOM, "scalar allocation failed");
        st = CC_ERR_OOM;
    }
    if (st == CC_OK && p->gravity_on)
        st = h->pois.init(p->Nx, p->Ny, g.px, g.py, g.dx, g.dy, p->G, p->sor_iters, p->sor_omega, p->precision);
    if (st == CC_OK) {
        hipLaunchKernelGGL(k_set_scalars, dim3(1), dim3(1), 0, h->stream, h->sc, 0.0, 0.0, 1e300, 1);
        if (hipStreamSynchronize(h->stream) != hipSuccess) {
            set_error(CC_ERR_HIP, "stream sync failed in create");
            st = CC_ERR_HIP;
        }
    }
    if (st != CC_OK) {
        delete h;
        return st;
    }
    *out = h;
    return CC_OK;
}

cc_status cc_mhd2d_destroy(cc_mhd2d* h) {
    if (!h) return CC_OK;
    hipStreamSynchronize(h->stream);
    delete h;
    return CC_OK;
}

#define CHECK_H()                                                   \
    do {                                                            \
        if (!h) {                                                   \
            set_error(CC_ERR_INVALID_ARG, "null handle");           \
            return CC_ERR_INVALID_ARG;                              \
        }                                                           \
    } while (0)

cc_status cc_mhd2d_upload(cc_mhd2d* h, const double* rho, const double* mx, const double* my, const double* mz,
                          const double* E, const double* Bz, const double* Bx_f, const double* By_f) {
    CHECK_H();
    if (!rho || !mx || !my || !mz || !E || !Bz || !Bx_f || !By_f) {
        set_error(CC_ERR_INVALID_ARG, "null state pointer");
        return CC_ERR_INVALID_ARG;
    }
    State2D& s = h->S[h->cur];
    const size_t n = h->N() * sizeof(double);
    CC_HIP(hipMemcpyAsync(s.rho, rho, n, hipMemcpyHostToDevice, h->stream));
    CC_HIP(hipMemcpyAsync(s.mx, mx, n, hipMemcpyHostToDevice, h->stream));
    CC_HIP(hipMemcpyAsync(s.my, my, n, hipMemcpyHostToDevice, h->stream));
    CC_HIP(hipMemcpyAsync(s.mz, mz, n, hipMemcpyHostToDevice, h->stream));
    CC_HIP(hipMemcpyAsync(s.E, E, n, hipMemcpyHostToDevice, h->stream));
    CC_HIP(hipMemcpyAsync(s.Bz, Bz, n, hipMemcpyHostToDevice, h->stream));
    CC_HIP(hipMemcpyAsync(s.Bx, Bx_f, h->NFx() * sizeof(double), hipMemcpyHostToDevice, h->stream));
    CC_HIP(hipMemcpyAsync(s.By, By_f, h->NFy() * sizeof(double), hipMemcpyHostToDevice, h->stream));
    if (h->g.ct_comp) {  // an uploaded face field is exact as given
        CC_HIP(hipMemsetAsync(s.Bx_lo, 0, h->NFx() * sizeof(float), h->stream));
        CC_HIP(hipMemsetAsync(s.By_lo, 0, h->NFy() * sizeof(float), h->stream));
    }
    CC_HIP(hipStreamSynchronize(h->stream));
    return CC_OK;
}

cc_status cc_mhd2d_download(cc_mhd2d* h, double* rho, double* mx, double* my, double* mz, double* E, double* Bz,
                            double* Bx_f, double* By_f) {
    CHECK_H();
    State2D& s = h->S[h->cur];
    const size_t n = h->N() * sizeof(double);
    struct { double* dst; const double* src; size_t bytes; } cp[8] = {
        {rho, s.rho, n}, {mx, s.mx, n}, {my, s.my, n}, {mz, s.mz, n}, {E, s.E, n}, {Bz, s.Bz, n},
        {Bx_f, s.Bx, h->NFx() * sizeof(double)}, {By_f, s.By, h->NFy() * sizeof(double)}};
    for (auto& c : cp)
        if (c.dst) CC_HIP(hipMemcpyAsync(c.dst, c.src, c.bytes, hipMemcpyDeviceToHost, h->stream));
    CC_HIP(hipStreamSynchronize(h->stream));
    return CC_OK;
}

cc_status cc_mhd2d_gravity(cc_mhd2d* h) {
    CHECK_H();
    CC_TRY(h->gravity_of(h->S[h->cur].rho, h->gx, h->gy));
    return CC_OK;
}

cc_status cc_mhd2d_get_gravity(cc_mhd2d* h, double* gx, double* gy) {
    CHECK_H();
    if (gx) CC_HIP(hipMemcpyAsync(gx, h->gx, h->N() * sizeof(double), hipMemcpyDeviceToHost, h->stream));
    if (gy) CC_HIP(hipMemcpyAsync(gy, h->gy, h->N() * sizeof(double), hipMemcpyDeviceToHost, h->stream));
    CC_HIP(hipStreamSynchronize(h->stream));
    return CC_OK;
}

cc_status cc_mhd2d_set_gravity(cc_mhd2d* h, const double* gx, const double* gy) {
    CHECK_H();
    if (!gx || !gy) { set_error(CC_ERR_INVALID_ARG, "null gravity pointer"); return CC_ERR_INVALID_ARG; }
    CC_HIP(hipMemcpyAsync(h->gx, gx, h->N() * sizeof(double), hipMemcpyHostToDevice, h->stream));
    CC_HIP(hipMemcpyAsync(h->gy, gy, h->N() * sizeof(double), hipMemcpyHostToDevice, h->stream));
    CC_HIP(hipStreamSynchronize(h->stream));
    return CC_OK;
}

cc_status cc_mhd2d_compute_dt(cc_mhd2d* h, double* dt) {
    CHECK_H();
    if (!dt) { set_error(CC_ERR_INVALID_ARG, "null dt"); return CC_ERR_INVALID_ARG; }
    hipLaunchKernelGGL(k_set_scalars, dim3(1), dim3(1), 0, h->stream, h->sc, 0.0, 0.0, 1e300, 1);
    CC_TRY(h->speed_max(h->S[h->cur]));
    hipLaunchKernelGGL(k_dt_finalize, dim3(1), dim3(256), 0, h->stream, h->g, h->sc, (double*)nullptr, h->spd, h->nspd);
    CC_CHECK_LAUNCH();
    Scalars s;
    CC_HIP(hipMemcpyAsync(&s, h->sc, sizeof(Scalars), hipMemcpyDeviceToHost, h->stream));
    CC_HIP(hipStreamSynchronize(h->stream));
    *dt = s.dt;
    return CC_OK;
}

cc_status cc_mhd2d_stage(cc_mhd2d* h, double dt) {
    CHECK_H();
    hipLaunchKernelGGL(k_set_scalars, dim3(1), dim3(1), 0, h->stream, h->sc, dt, 0.0, 1e300, 0);
    const int a = h->cur, b = (h->cur + 1) % 3;
    CC_TRY(h->stage<false>(h->S[a], h->S[b], h->gx, h->gy, h->S[a]));
    h->cur = b;
    return CC_OK;
}

cc_status cc_mhd2d_rk2_step(cc_mhd2d* h, double dt) {
    CHECK_H();
    if (!std::isfinite(dt) || dt < 0.0) { set_error(CC_ERR_BAD_DT, "dt must be finite and >= 0"); return CC_ERR_BAD_DT; }
    hipLaunchKernelGGL(k_set_scalars, dim3(1), dim3(1), 0, h->stream, h->sc, dt, 0.0, 1e300, 0);
    const int a = h->cur, b = (a + 1) % 3, c = (a + 2) % 3;
    CC_TRY(h->heun(a, b, c, false, nullptr));
    h->cur = c;
    return CC_OK;
}

cc_status cc_mhd2d_run(cc_mhd2d* h, int nsteps, double* t_inout, double t_end, double* dts) {
    CHECK_H();
    if (nsteps < 0 || !t_inout) { set_error(CC_ERR_INVALID_ARG, "nsteps < 0 or null t"); return CC_ERR_INVALID_ARG; }
    if (nsteps == 0) return CC_OK;
    if (nsteps > kDtLog) {  // the device dt log holds kDtLog steps: run in chunks
        double t = *t_inout;
        for (int done = 0; done < nsteps;) {
            const int n = std::min(kDtLog, nsteps - done);
            CC_TRY(cc_mhd2d_run(h, n, &t, t_end, dts ? dts + done : nullptr));
            done += n;
        }
        *t_inout = t;
        return CC_OK;
    }
    hipLaunchKernelGGL(k_set_scalars, dim3(1), dim3(1), 0, h->stream, h->sc, 0.0, *t_inout, t_end, 1);
    CC_TRY(h->speed_max(h->S[h->cur]));
    const bool use_graph = !h->prof.on;
    for (int s = 0; s < nsteps; ++s) {
        const int a = h->cur, b = (a + 1) % 3, c = (a + 2) % 3;
        if (use_graph) {
            if (!h->graph_ok[a]) {
                hipGraph_t gr;
                CC_HIP(hipStreamBeginCapture(h->stream, hipStreamCaptureModeThreadLocal));
                cc_status st = h->heun(a, b, c, true, h->dts_dev);
                hipError_t e = hipStreamEndCapture(h->stream, &gr);
                if (st != CC_OK) return st;
                if (e != hipSuccess) { set_error(CC_ERR_HIP, "graph capture failed"); return CC_ERR_HIP; }
                CC_HIP(hipGraphInstantiate(&h->graph[a], gr, nullptr, nullptr, 0));
                hipGraphDestroy(gr);
                h->graph_ok[a] = true;
            }
            CC_HIP(hipGraphLaunch(h->graph[a], h->stream));
        } else {
            CC_TRY(h->heun(a, b, c, true, h->dts_dev));
        }
        h->cur = c;
    }
    Scalars s;
    CC_HIP(hipMemcpyAsync(&s, h->sc, sizeof(Scalars), hipMemcpyDeviceToHost, h->stream));
    if (dts) CC_HIP(hipMemcpyAsync(dts, h->dts_dev, (size_t)nsteps * sizeof(double), hipMemcpyDeviceToHost, h->stream));
    CC_HIP(hipStreamSynchronize(h->stream));
    *t_inout = s.t;
    CC_TRY(h->pois.flow_error(h->stream));
    if (!std::isfinite(s.dt) || s.dt < 0.0) {
        set_error(CC_ERR_BAD_DT, "non-finite or negative dt on device");
        return CC_ERR_BAD_DT;
    }
    return CC_OK;
}

cc_status cc_mhd2d_divb_max(cc_mhd2d* h, double* out) {
    CHECK_H();
    if (!out) { set_error(CC_ERR_INVALID_ARG, "null output"); return CC_ERR_INVALID_ARG; }
    hipLaunchKernelGGL(k_set_scalars, dim3(1), dim3(1), 0, h->stream, h->sc, 0.0, 0.0, 1e300, 0);
    hipLaunchKernelGGL(k_divb_max2d, dim3((h->p.Nx + 255) / 256, h->p.Ny), dim3(256), 0, h->stream, h->g,
                       h->S[h->cur], h->sc);
    CC_CHECK_LAUNCH();
    Scalars s;
    CC_HIP(hipMemcpyAsync(&s, h->sc, sizeof(Scalars), hipMemcpyDeviceToHost, h->stream));
    CC_HIP(hipStreamSynchronize(h->stream));
    *out = s.divb;
    return CC_OK;
}

cc_status cc_mhd2d_frame(cc_mhd2d* h, double* out) {
    CHECK_H();
    if (!out) { set_error(CC_ERR_INVALID_ARG, "null output"); return CC_ERR_INVALID_ARG; }
    CC_TRY(h->make_frame());
    CC_HIP(hipMemcpyAsync(out, h->frame.d, sizeof(double) * CC_MHD2D_FRAME_FIELDS * h->N(), hipMemcpyDeviceToHost,
                          h->stream));
    CC_HIP(hipStreamSynchronize(h->stream));
    return CC_OK;
}

cc_status cc_mhd2d_totals(cc_mhd2d* h, double* out6) {
    CHECK_H();
    if (!out6) { set_error(CC_ERR_INVALID_ARG, "null output"); return CC_ERR_INVALID_ARG; }
    const State2D& s = h->S[h->cur];
    FieldPtrs a{{s.rho, s.mx, s.my, s.mz, s.E}};
    double r[15];
    CC_TRY(h->stats.run(a, nullptr, 5, h->N(), h->stream, r));
    const double dA = h->g.dx * h->g.dy;
    for (int f = 0; f < 5; f++) out6[f] = r[3 * f] * dA;
    return cc_mhd2d_divb_max(h, out6 + 5);
}

cc_status cc_mhd2d_compare(cc_mhd2d* a, cc_mhd2d* b, double* out) {
    if (!a || !b || !out) { set_error(CC_ERR_INVALID_ARG, "null handle or output"); return CC_ERR_INVALID_ARG; }
    if (a->p.Nx != b->p.Nx || a->p.Ny != b->p.Ny) {
        set_error(CC_ERR_INVALID_ARG, "cc_mhd2d_compare: grids differ");
        return CC_ERR_INVALID_ARG;
    }
    CC_TRY(a->make_frame());
    CC_TRY(b->make_frame());
    CC_HIP(hipStreamSynchronize(b->stream));
    return compare_frames(a->stats, a->frame.d, b->frame.d, CC_MHD2D_FRAME_FIELDS, a->N(), a->stream, out);
}

cc_status cc_mhd2d_sync(cc_mhd2d* h) {
    CHECK_H();
    CC_HIP(hipStreamSynchronize(h->stream));
    return CC_OK;
}

cc_status cc_mhd2d_set_profiling(cc_mhd2d* h, int on) {
    CHECK_H();
    h->prof.on = on != 0;
    return CC_OK;
}

cc_status cc_mhd2d_kernel_time(cc_mhd2d* h, const char* kernel, int* count, double* total_ms) {
    CHECK_H();
    if (!kernel || !count || !total_ms) { set_error(CC_ERR_INVALID_ARG, "null argument"); return CC_ERR_INVALID_ARG; }
    CC_HIP(hipStreamSynchronize(h->stream));
    h->prof.get_total(kernel, count, total_ms);
    return CC_OK;
}

cc_status cc_mhd2d_reset_profiling(cc_mhd2d* h) {
    CHECK_H();
    CC_HIP(hipStreamSynchronize(h->stream));
    h->prof.reset();
    return CC_OK;
}

}  // extern "C"

// ==================================================================== batches of instances
struct cc_mhd2d_batch {
    std::vector<cc_mhd2d*> h;
    hipStream_t root = nullptr;
    hipEvent_t fork = nullptr;
    std::vector<hipEvent_t> join;
    hipGraphExec_t graph[3] = {nullptr, nullptr, nullptr};  // one step of every instance, by rotation phase
    // Merged SOR (every instance with gravity, clamped boundaries, one precision and tile shape): the
    // tile levels of all instances run as one launch per level on the root stream.
    bool merged = false;
    bool mixed = false;
    SorTile tile{};
    int* tiles4 = nullptr;           // (instance, TS, TJ, TI) per tile, grouped by level
    std::vector<int> level_off;      // level offsets, in tiles
    SorBatchEntry* tab[3][2] = {};   // per rotation phase and solve (U^n, U^1): each instance's SOR arguments
    // The merged solve as one dataflow launch (unless CC_SOR_LEVELS): per tile the tiles it waits for,
    // as CSR over merged indices, and the launch state (counter, error flag, done flags).
    bool flow = false;
    int ntiles = 0;
    int *dep_off = nullptr, *dep_idx = nullptr, *fstate = nullptr;

    // Fork from the root stream to every instance stream, run fn(instance) there, join back.
    template <typename F>
    cc_status fan(F&& fn) {
        if (hipEventRecord(fork, root) != hipSuccess) return CC_ERR_HIP;
        for (size_t i = 0; i < h.size(); i++) {
            if (hipStreamWaitEvent(h[i]->stream, fork, 0) != hipSuccess) return CC_ERR_HIP;
            CC_TRY(fn(h[i]));
            if (hipEventRecord(join[i], h[i]->stream) != hipSuccess || hipStreamWaitEvent(root, join[i], 0) != hipSuccess)
                return CC_ERR_HIP;
        }
        return CC_OK;
    }
    cc_status merged_sweeps(int a, int q) {
        if (flow) return launch_flow_batch(tile, sor_mixed_kernels(tile, mixed), ntiles, root, tab[a][q], tiles4, dep_off, dep_idx, fstate);
        for (size_t l = 0; l + 1 < level_off.size(); ++l)
            CC_TRY(launch_tile_batch(tile, sor_mixed_kernels(tile, mixed), level_off[l + 1] - level_off[l], root, tab[a][q], tiles4 + 4 * level_off[l]));
        return CC_OK;
    }
    // Device copies of a merged schedule: tiles, and with waits the CSR arrays and a launch state.
    struct MergedDev {
        int *t4 = nullptr, *doff = nullptr, *didx = nullptr, *st = nullptr;
        void free_all() {
            for (int* p : {t4, doff, didx, st})
                if (p) hipFree(p);
            t4 = doff = didx = st = nullptr;
        }
    };
    static cc_status upload_merged(const std::vector<int>& flat, const std::vector<int>* doff,
                                   const std::vector<int>* didx, MergedDev& d) {
        const size_t nt = flat.size() / 4;
        bool ok = hipMalloc(&d.t4, flat.size() * sizeof(int)) == hipSuccess &&
                  hipMemcpy(d.t4, flat.data(), flat.size() * sizeof(int), hipMemcpyHostToDevice) == hipSuccess;
        if (ok && doff) {
            ok = hipMalloc(&d.doff, doff->size() * sizeof(int)) == hipSuccess &&
                 hipMalloc(&d.didx, std::max<size_t>(didx->size(), 1) * sizeof(int)) == hipSuccess &&
                 hipMalloc(&d.st, (nt + 2) * sizeof(int)) == hipSuccess &&
                 hipMemcpy(d.doff, doff->data(), doff->size() * sizeof(int), hipMemcpyHostToDevice) == hipSuccess &&
                 (didx->empty() || hipMemcpy(d.didx, didx->data(), didx->size() * sizeof(int), hipMemcpyHostToDevice) == hipSuccess) &&
                 hipMemset(d.st, 0, (nt + 2) * sizeof(int)) == hipSuccess;
        }
        if (!ok) {
            d.free_all();
            set_error(CC_ERR_OOM, "batch schedule allocation failed");
            return CC_ERR_OOM;
        }
        return CC_OK;
    }
    // Can the instances share merged SOR launches?  If so build the merged schedule and the tables.
    cc_status setup_merged() {
        const cc_mhd2d* h0 = h[0];
        for (const cc_mhd2d* x : h) {
            const Poisson2D& P = x->pois;
            // the merged launches run one LDS tile shape of their own; the instances' shapes may differ.
            // Wave-streamed solves stay per instance (each a branch of the batch graph).
            if (!x->p.gravity_on || !P.tiled() || P.iters <= 0 || x->p.precision != h0->p.precision ||
                P.wave_streamed())
                return CC_OK;  // not mergeable: the instances keep their own SOR launches
        }
        mixed = h0->p.precision == CC_MIXED;
        for (int a = 0; a < 3; ++a)
            for (int q = 0; q < 2; ++q) {
                std::vector<SorBatchEntry> e(h.size());
                for (size_t i = 0; i < h.size(); i++) {
                    Poisson2D& P = h[i]->pois;
                    e[i] = SorBatchEntry{P.args, P.phi, h[i]->S[q == 0 ? a : (a + 1) % 3].rho, P.rhobar};
                }
                if (hipMalloc(&tab[a][q], e.size() * sizeof(SorBatchEntry)) != hipSuccess) { set_error(CC_ERR_OOM, "batch table allocation failed"); return CC_ERR_OOM; }
                CC_HIP(hipMemcpy(tab[a][q], e.data(), e.size() * sizeof(SorBatchEntry), hipMemcpyHostToDevice));
            }
        // the tile shape: the instances' own, or with autotuning the compiled shape whose merged solve is fastest
        std::vector<SorTile> cands{h0->pois.tile_};
        if (sor_autotune_enabled()) {
#define CC_BCAND64(S_, WJ_, R_) cands.push_back(SorTile{S_, WJ_, R_, false, mixed});
#define CC_BCAND32(S_, WJ_) if (mixed) cands.push_back(SorTile{S_, WJ_, true});
#define CC_BSCAND64(S_, K_) cands.push_back(SorTile{S_, K_, true, true, mixed});
#define CC_BSCAND32(S_, K_) if (mixed) cands.push_back(SorTile{S_, K_, true, true});
            CC_SOR_TILES_FP64(CC_BCAND64)
            CC_SOR_TILES_MIXED(CC_BCAND32)
            CC_SOR_SHEARED_FP64(CC_BSCAND64)
            CC_SOR_SHEARED_MIXED(CC_BSCAND32)
#undef CC_BCAND64
#undef CC_BCAND32
#undef CC_BSCAND64
#undef CC_BSCAND32
        }
        float best = 3.0e38f;
        MergedDev keep;
        for (size_t ci = 0; ci < cands.size(); ++ci) {
            std::vector<int> flat, off, doff, didx;
            build_merged(cands[ci], flat, off, &doff, &didx);
            const int nt = (int)flat.size() / 4;
            const bool can_flow = sor_flow_enabled(nt);
            MergedDev d;
            CC_TRY(upload_merged(flat, can_flow ? &doff : nullptr, can_flow ? &didx : nullptr, d));
            // with several candidates, each shape is timed in both launch forms: the levels of a merged
            // schedule are already full, so one launch per level can beat the dataflow launch
            float ms_form[2] = {0.f, 3.0e38f};  // [0] one launch per level, [1] one dataflow launch
            for (int fl = 0; fl < (can_flow ? 2 : 1) && cands.size() > 1; ++fl) {
                hipEvent_t e0, e1;
                hipEventCreate(&e0);
                hipEventCreate(&e1);
                cc_status st = CC_OK;
                for (int rep = 0; rep < 2 && st == CC_OK; ++rep) {  // the first run pays first-launch costs; phi is scratch
                    hipEventRecord(e0, root);
                    if (fl) {
                        st = launch_flow_batch(cands[ci], sor_mixed_kernels(cands[ci], mixed), nt, root, tab[0][0], d.t4, d.doff, d.didx, d.st);
                    } else {
                        for (size_t l = 0; l + 1 < off.size() && st == CC_OK; ++l)
                            st = launch_tile_batch(cands[ci], sor_mixed_kernels(cands[ci], mixed), off[l + 1] - off[l], root, tab[0][0], d.t4 + 4 * off[l]);
                    }
                    hipEventRecord(e1, root);
                    hipEventSynchronize(e1);
                    hipEventElapsedTime(&ms_form[fl], e0, e1);
                }
                hipEventDestroy(e0);
                hipEventDestroy(e1);
                if (st == CC_OK && fl) st = sor_flow_error(d.st, root);
                if (st != CC_OK) { d.free_all(); keep.free_all(); return st; }
                if (std::getenv("CC_SOR_AUTOTUNE_LOG"))
                    std::fprintf(stderr, "sor autotune batch of %zu %s: S %d, %s %d%s%s, %s: %.3f ms\n", h.size(),
                                 mixed ? "mixed" : "fp64", cands[ci].S, cands[ci].sheared ? "K" : "WJ", cands[ci].WJ,
                                 cands[ci].sheared ? " sheared" : (cands[ci].rhs_lds ? "" : " rhs-global"),
                                 cands[ci].fp64 ? " fp64-kernels" : "", fl ? "dataflow" : "levels", ms_form[fl]);
            }
            // a single candidate (no autotuning) runs as the single-grid solves do: dataflow unless CC_SOR_LEVELS
            const bool use_flow = cands.size() > 1 ? ms_form[1] < ms_form[0] : can_flow;
            const float ms = ms_form[use_flow ? 1 : 0];
            if (ci == 0 || ms < best) {
                best = ms;
                keep.free_all();
                keep = d;
                level_off = off;
                tile = cands[ci];
                flow = use_flow;
                ntiles = nt;
            } else {
                d.free_all();
            }
        }
        tiles4 = keep.t4;
        dep_off = keep.doff;
        dep_idx = keep.didx;
        fstate = keep.st;
        merged = true;
        return CC_OK;
    }
    // The merged schedule for tile shape t: level L holds level L of every instance, (instance, TS, TJ, TI).
    // The merged schedule of tile shape t: level L of every instance, then level L + 1; with doff/didx also
    // each tile's waits (the instance's own, sor_flow_deps) as merged indices.
    void build_merged(const SorTile& t, std::vector<int>& flat, std::vector<int>& off, std::vector<int>* doff = nullptr,
                      std::vector<int>* didx = nullptr) {
        std::vector<std::vector<int>> fl(h.size()), lo(h.size()), pos(h.size());
        for (size_t i = 0; i < h.size(); i++) {
            sor_levels(t, h[i]->pois.Nx, h[i]->pois.Ny, h[i]->pois.iters, fl[i], lo[i]);
            pos[i].assign(fl[i].size() / 3, -1);
        }
        std::vector<std::pair<int, int>> who;  // (instance, local tile) of each merged tile
        flat.clear();
        off.assign(1, 0);
        for (int L = 0;; ++L) {
            bool any = false;
            for (size_t i = 0; i < h.size(); i++) {
                if (L + 1 >= (int)lo[i].size()) continue;
                any = true;
                for (int k = lo[i][L]; k < lo[i][L + 1]; ++k) {
                    pos[i][k] = (int)flat.size() / 4;
                    who.emplace_back((int)i, k);
                    flat.push_back((int)i);
                    flat.push_back(fl[i][3 * k]);
                    flat.push_back(fl[i][3 * k + 1]);
                    flat.push_back(fl[i][3 * k + 2]);
                }
            }
            if (!any) break;
            off.push_back((int)flat.size() / 4);
        }
        if (doff && didx) {
            std::vector<std::vector<int>> loff(h.size()), lidx(h.size());
            for (size_t i = 0; i < h.size(); i++) sor_flow_deps(fl[i], loff[i], lidx[i]);
            doff->assign(1, 0);
            didx->clear();
            for (const auto& w : who) {
                const int i = w.first, k = w.second;
                for (int q = loff[i][k]; q < loff[i][k + 1]; ++q) didx->push_back(pos[i][lidx[i][q]]);
                doff->push_back((int)didx->size());
            }
        }
    }
    // One Heun step of every instance (heun(a, b, c, true) of each) with the SOR sweeps merged:
    // per instance mean + phi = 0 | merged sweeps | g, dt, stage 1, mean + phi = 0 | merged sweeps | g, stage 2.
    cc_status capture_merged_body(int a) {
        const int b = (a + 1) % 3, c = (a + 2) % 3;
        CC_TRY(fan([&](cc_mhd2d* x) { return x->pois.prologue(x->S[a].rho, x->stream, nullptr); }));
        CC_TRY(merged_sweeps(a, 0));
        CC_TRY(fan([&](cc_mhd2d* x) {
            CC_TRY(x->pois.epilogue(x->gx, x->gy, x->stream, nullptr));
            hipLaunchKernelGGL(k_dt_finalize, dim3(1), dim3(256), 0, x->stream, x->g, x->sc, x->dts_dev, x->spd, x->nspd);
            CC_CHECK_LAUNCH();
            CC_TRY(x->stage<false>(x->S[a], x->S[b], x->gx, x->gy, x->S[a]));
            return x->pois.prologue(x->S[b].rho, x->stream, nullptr);
        }));
        CC_TRY(merged_sweeps(a, 1));
        return fan([&](cc_mhd2d* x) {
            CC_TRY(x->pois.epilogue(x->g1x, x->g1y, x->stream, nullptr));
            return x->stage<true>(x->S[b], x->S[c], x->g1x, x->g1y, x->S[a]);
        });
    }

    // One step of every instance as parallel branches of one graph: the root stream forks to the
    // instances' streams and joins them back.
    cc_status capture(int a) {
        const int b = (a + 1) % 3, c = (a + 2) % 3;
        CC_HIP(hipStreamBeginCapture(root, hipStreamCaptureModeGlobal));
        cc_status st = merged ? capture_merged_body(a)
                              : fan([&](cc_mhd2d* x) { return x->heun(a, b, c, true, x->dts_dev); });
        hipGraph_t gr = nullptr;
        const hipError_t e = hipStreamEndCapture(root, &gr);
        if (st != CC_OK || e != hipSuccess) {
            if (gr) hipGraphDestroy(gr);
            if (st == CC_OK) set_error(CC_ERR_HIP, "batch graph capture failed");
            return st == CC_OK ? CC_ERR_HIP : st;
        }
        const hipError_t ei = hipGraphInstantiate(&graph[a], gr, nullptr, nullptr, 0);
        hipGraphDestroy(gr);
        if (ei != hipSuccess) {
            set_error(CC_ERR_HIP, "batch graph instantiation failed");
            return CC_ERR_HIP;
        }
        return CC_OK;
    }
    ~cc_mhd2d_batch() {
        for (auto& g : graph)
            if (g) hipGraphExecDestroy(g);
        if (tiles4) hipFree(tiles4);
        for (int* p : {dep_off, dep_idx, fstate})
            if (p) hipFree(p);
        for (auto& ta : tab)
            for (auto* t : ta)
                if (t) hipFree(t);
        for (auto e : join)
            if (e) hipEventDestroy(e);
        if (fork) hipEventDestroy(fork);
        if (root) hipStreamDestroy(root);
    }
};

extern "C" {

cc_status cc_mhd2d_batch_create(cc_mhd2d* const* handles, int n, cc_mhd2d_batch** out) {
    if (!handles || n < 1 || !out) { set_error(CC_ERR_INVALID_ARG, "cc_mhd2d_batch_create: bad arguments"); return CC_ERR_INVALID_ARG; }
    *out = nullptr;
    for (int i = 0; i < n; i++)
        if (!handles[i]) { set_error(CC_ERR_INVALID_ARG, "cc_mhd2d_batch_create: null handle"); return CC_ERR_INVALID_ARG; }
    cc_mhd2d_batch* b = new cc_mhd2d_batch();
    b->h.assign(handles, handles + n);
    b->join.assign((size_t)n, nullptr);
    bool ok = hipStreamCreateWithFlags(&b->root, hipStreamNonBlocking) == hipSuccess &&
              hipEventCreateWithFlags(&b->fork, hipEventDisableTiming) == hipSuccess;
    for (int i = 0; i < n && ok; i++) ok = hipEventCreateWithFlags(&b->join[i], hipEventDisableTiming) == hipSuccess;
    if (!ok) {
        delete b;
        set_error(CC_ERR_HIP, "cc_mhd2d_batch_create: stream / event creation failed");
        return CC_ERR_HIP;
    }
    for (int i = 0; i < n; i++)  // the instances' set-up copies (schedules, scalars) are done before any batch launch
        if (hipStreamSynchronize(handles[i]->stream) != hipSuccess) { delete b; set_error(CC_ERR_HIP, "sync failed"); return CC_ERR_HIP; }
    if (std::getenv("CC_BATCH_SEPARATE_SOR") == nullptr) {
        const cc_status st = b->setup_merged();
        if (st != CC_OK) { delete b; return st; }
    }
    *out = b;
    return CC_OK;
}

cc_status cc_mhd2d_batch_info(cc_mhd2d_batch* b, int* merged, int* S, int* WJ, int* sheared, int* fp64_kernels) {
    if (!b || !merged || !S || !WJ) { set_error(CC_ERR_INVALID_ARG, "cc_mhd2d_batch_info: null argument"); return CC_ERR_INVALID_ARG; }
    *merged = b->merged ? 1 : 0;
    *S = b->merged ? b->tile.S : 0;
    *WJ = b->merged ? b->tile.WJ : 0;
    if (sheared) *sheared = b->merged && b->tile.sheared ? 1 : 0;
    if (fp64_kernels) *fp64_kernels = b->merged && b->tile.fp64 ? 1 : 0;
    return CC_OK;
}

cc_status cc_mhd2d_batch_destroy(cc_mhd2d_batch* b) {
    if (!b) return CC_OK;
    hipStreamSynchronize(b->root);
    delete b;
    return CC_OK;
}

cc_status cc_mhd2d_batch_run(cc_mhd2d_batch* b, int nsteps, double* t_inout, const double* t_end) {
    if (!b || nsteps < 0 || !t_inout || !t_end) { set_error(CC_ERR_INVALID_ARG, "cc_mhd2d_batch_run: bad arguments"); return CC_ERR_INVALID_ARG; }
    if (nsteps > kDtLog) { set_error(CC_ERR_INVALID_ARG, "cc_mhd2d_batch_run: at most 65536 steps per call"); return CC_ERR_INVALID_ARG; }
    const int a0 = b->h[0]->cur;
    for (auto* h : b->h) {
        if (h->cur != a0) { set_error(CC_ERR_INVALID_ARG, "cc_mhd2d_batch_run: instances at different rotation phases"); return CC_ERR_INVALID_ARG; }
        if (h->prof.on) { set_error(CC_ERR_INVALID_ARG, "cc_mhd2d_batch_run: profiling is on for an instance (the batch graph cannot time kernels)"); return CC_ERR_INVALID_ARG; }
    }
    if (nsteps == 0) return CC_OK;
    // per-instance start: t, t_end, the signal speeds of the current state (as cc_mhd2d_run)
    for (size_t i = 0; i < b->h.size(); i++) {
        cc_mhd2d* h = b->h[i];
        CC_HIP(hipStreamSynchronize(h->stream));
        hipLaunchKernelGGL(k_set_scalars, dim3(1), dim3(1), 0, h->stream, h->sc, 0.0, t_inout[i], t_end[i], 1);
        CC_TRY(h->speed_max(h->S[h->cur]));
        CC_HIP(hipEventRecord(b->join[i], h->stream));
        CC_HIP(hipStreamWaitEvent(b->root, b->join[i], 0));
    }
    int cur = a0;
    for (int s = 0; s < nsteps; ++s) {
        if (!b->graph[cur]) CC_TRY(b->capture(cur));
        CC_HIP(hipGraphLaunch(b->graph[cur], b->root));
        cur = (cur + 2) % 3;
    }
    CC_HIP(hipStreamSynchronize(b->root));
    cc_status st = CC_OK;
    if (b->merged && b->flow && sor_flow_error(b->fstate, b->root) != CC_OK) st = CC_ERR_HIP;
    for (size_t i = 0; i < b->h.size(); i++) {
        cc_mhd2d* h = b->h[i];
        h->cur = cur;
        Scalars sc;
        CC_HIP(hipMemcpy(&sc, h->sc, sizeof(Scalars), hipMemcpyDeviceToHost));
        t_inout[i] = sc.t;
        // the instances' own dataflow SOR runs only when the batch does not merge their SOR
        if (!b->merged && h->pois.flow_error(h->stream) != CC_OK) st = CC_ERR_HIP;
        if (!std::isfinite(sc.dt) || sc.dt < 0.0) {
            set_error(CC_ERR_BAD_DT, "non-finite or negative dt on device in batch instance " + std::to_string(i));
            st = CC_ERR_BAD_DT;
        }
    }
    return st;
}

cc_status cc_hlld2d_batch(int axis, int n, const double* WL, const double* WR, const double* Bn, double gamma,
                          double floor_rho, int precision, double* out) {
    if (n < 0 || (axis != 0 && axis != 1) || (n > 0 && (!WL || !WR || !Bn || !out))) {
        set_error(CC_ERR_INVALID_ARG, "invalid batch arguments");
        return CC_ERR_INVALID_ARG;
    }
    if (n == 0) return CC_OK;
    double *dL, *dR, *dB, *dO;
    const size_t b8 = (size_t)n * 8 * sizeof(double);
    if (hipMalloc(&dL, b8) != hipSuccess) { set_error(CC_ERR_OOM, "batch alloc"); return CC_ERR_OOM; }
    hipMalloc(&dR, b8);
    hipMalloc(&dB, (size_t)n * sizeof(double));
    hipMalloc(&dO, b8);
    hipMemcpy(dL, WL, b8, hipMemcpyHostToDevice);
    hipMemcpy(dR, WR, b8, hipMemcpyHostToDevice);
    hipMemcpy(dB, Bn, (size_t)n * sizeof(double), hipMemcpyHostToDevice);
    if (precision == CC_MIXED)
        hipLaunchKernelGGL(k_hlld_batch<true>, dim3((n + 255) / 256), dim3(256), 0, 0, axis, n, dL, dR, dB, gamma,
                           floor_rho, dO);
    else
        hipLaunchKernelGGL(k_hlld_batch<false>, dim3((n + 255) / 256), dim3(256), 0, 0, axis, n, dL, dR, dB, gamma,
                           floor_rho, dO);
    hipError_t e = hipGetLastError();
    if (e == hipSuccess) e = hipMemcpy(out, dO, b8, hipMemcpyDeviceToHost);
    hipFree(dL); hipFree(dR); hipFree(dB); hipFree(dO);
    if (e != hipSuccess) { set_error(CC_ERR_HIP, hipGetErrorString(e)); return CC_ERR_HIP; }
    return CC_OK;
}

cc_status cc_poisson2d_sor(int Nx, int Ny, int periodic_x, int periodic_y, double dx, double dy, double G,
                           int sor_iters, double omega, int precision, const double* rho, double* gx, double* gy,
                           double* phi) {
    if (Nx < 2 || Ny < 2 || !rho || sor_iters < 0 || !(dx > 0) || !(dy > 0)) {
        set_error(CC_ERR_INVALID_ARG, "invalid Poisson arguments");
        return CC_ERR_INVALID_ARG;
    }
    Poisson2D P;
    CC_TRY(P.init(Nx, Ny, !!periodic_x, !!periodic_y, dx, dy, G, sor_iters, omega, precision));
    const size_t N = (size_t)Nx * Ny;
    double *drho, *dgx, *dgy;
    if (hipMalloc(&drho, N * 8) != hipSuccess) { set_error(CC_ERR_OOM, "alloc"); return CC_ERR_OOM; }
    hipMalloc(&dgx, N * 8);
    hipMalloc(&dgy, N * 8);
    hipMemcpy(drho, rho, N * 8, hipMemcpyHostToDevice);
    cc_status st = P.solve(drho, dgx, dgy, 0, nullptr);
    if (st == CC_OK && phi) st = P.phi_mean_removed(P.phi, 0);
    hipError_t e = hipDeviceSynchronize();
    if (st == CC_OK && e == hipSuccess) st = P.flow_error(0);
    if (st == CC_OK && e == hipSuccess) {
        if (gx) hipMemcpy(gx, dgx, N * 8, hipMemcpyDeviceToHost);
        if (gy) hipMemcpy(gy, dgy, N * 8, hipMemcpyDeviceToHost);
        if (phi) hipMemcpy(phi, P.phi, N * 8, hipMemcpyDeviceToHost);
    }
    hipFree(drho); hipFree(dgx); hipFree(dgy);
    if (st != CC_OK) return st;
    if (e != hipSuccess) { set_error(CC_ERR_HIP, hipGetErrorString(e)); return CC_ERR_HIP; }
    return CC_OK;
}

}  // extern "C"
