// euler2d.hip -- gfx950 solver step of the 2D Euler HLLC + SOR self-gravity code
// (euler2d_gravity_hllc.cpp, cited as EH:<line>).
//
// One right-hand side (rhs, EH:406-527) is two launches:
//   k_eh_sor     periodic Poisson solve by lexicographic SOR until the in-sweep residual is below
//                tol (EH:288-386), on ONE workgroup: a sweep is a sequence of anti-diagonals
//                i + j = d (one barrier each), which is the lexicographic dependency order including
//                the wrap-around reads; then the mean is removed and the residual tested, as the
//                reference does after every sweep; finally g = -central grad(phi).  phi lives in LDS
//                when the grid fits (<= 8192 cells), else in global memory.
//   k_eh_stage   32x8 cell tile: floored primitives of the tile + 2-cell halo in LDS, 4-point MUSCL-MC
//                face states and HLLC on x- and y-faces into LDS, flux divergence + rho g source, and
//                the Heun stage update with floors (rk2_step, EH:529-550) -- or dU itself (MODE 0).
// Heun = rhs(U) -> U1 = U + dt k1 (floors); rhs(U1) -> U = U/2 + (U1 + dt k2)/2 (floors), written in
// place (each thread reads U^n only at its own cell).  Mixed precision runs the HLLC in FP32 on the
// FP64 face states; everything else stays FP64.
// This file is compiled without FP contraction so that the FP64 path keeps the reference's rounding.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "diag.hpp"
#include "hllc.hpp"
#include "profiler.hpp"
#include "riemann.hpp"
#include "stepcommon.hpp"

namespace cc {

struct GeomEH {
    int Nx, Ny, px, py;
    double dx, dy, gamma, gm1, floor_rho, floor_P, cfl;
    int gravity_on;
};

struct StateEH {
    double *rho, *mx, *my, *E;
};

struct SorEH {
    int Nx, Ny, max_iter;
    double idx2, idy2, denom, omega, tol, c4piG, floor_rho, dx, dy;
};

// ------------------------------------------------------------------ cell helpers
__device__ __forceinline__ void floors_eh(const GeomEH& g, double& rho, double& mx, double& my, double& E) {
    rho = smax(rho, g.floor_rho);  // EH:62-72
    const double u = mx / rho, v = my / rho;
    const double kin = 0.5 * rho * (u * u + v * v);
    const double P = g.gm1 * (E - kin);
    if (P < g.floor_P) {
        const double eint = g.floor_P / g.gm1;
        E = kin + eint;
    }
}

__device__ __forceinline__ void prim_eh(const GeomEH& g, double rho, double mx, double my, double E, double& u,
                                        double& v, double& P, double& a) {  // cons_to_prim_full EH:74-83
    rho = smax(rho, g.floor_rho);
    u = mx / rho;
    v = my / rho;
    const double kin = 0.5 * rho * (u * u + v * v);
    P = g.gm1 * (E - kin);
    if (P < g.floor_P) P = g.floor_P;
    a = sqrt(g.gamma * P / rho);
}

__device__ __forceinline__ double speed_eh(const GeomEH& g, double rho, double mx, double my, double E) {
    double u, v, P, a;
    prim_eh(g, rho, mx, my, E, u, v, P, a);
    return smax(fabs(u) + a, fabs(v) + a);  // EH:399-400
}

// ------------------------------------------------------------------ block reductions (any blockDim <= 1024)
__device__ __forceinline__ void block_sum_max(double& s, double& m, double* red_s, double* red_m) {
    s = wave_sum(s);
    m = wave_max(m);
    const int w = threadIdx.x >> 6, nw = (blockDim.x + 63) >> 6;
    if ((threadIdx.x & 63) == 0) { red_s[w] = s; red_m[w] = m; }
    __syncthreads();
    if (threadIdx.x == 0) {
        double ts = 0.0, tm = 0.0;
        for (int k = 0; k < nw; ++k) { ts += red_s[k]; tm = smax(tm, red_m[k]); }
        red_s[16] = ts;
        red_m[16] = tm;
    }
    __syncthreads();
    s = red_s[16];
    m = red_m[16];
    __syncthreads();
}

// ------------------------------------------------------------------ SOR to tolerance (EH:288-386)
// LDS: phi lives in LDS (N doubles of dynamic LDS); RLDS: the source term 4 pi G (rho - rhobar) too
// (2N doubles), so a diagonal step waits on LDS only, never on a global load.
template <bool LDS, bool RLDS>
__global__ __launch_bounds__(1024) void k_eh_sor(SorEH a, const double* __restrict__ rho, double* __restrict__ gphi,
                                                 double* __restrict__ gx, double* __restrict__ gy,
                                                 int* __restrict__ sweeps) {
    extern __shared__ double sphi[];
    __shared__ double red_s[17], red_m[17];
    double* phi = LDS ? sphi : gphi;
    const int Nx = a.Nx, Ny = a.Ny, B = blockDim.x, tid = threadIdx.x;
    const int N = Nx * Ny;
    double* rsrc = sphi + N;  // RLDS only
    // rhobar of the floored density (EH:306-315) and phi0 = 0 (EH:325)
    double s = 0.0, m = 0.0;
    for (int k = tid; k < N; k += B) {
        s += smax(rho[k], a.floor_rho);
        phi[k] = 0.0;
    }
    block_sum_max(s, m, red_s, red_m);
    const double rhobar = s / (double)N;
    if (RLDS) {
        for (int k = tid; k < N; k += B) rsrc[k] = a.c4piG * (smax(rho[k], a.floor_rho) - rhobar);
        __syncthreads();
    }
    int it = 0;
    while (it < a.max_iter) {
        double mres = 0.0;
        for (int d = 0; d < Nx + Ny - 1; ++d) {
            for (int j = tid; j < Ny; j += B) {
                const int i = d - j;
                if (i < 0 || i >= Nx) continue;
                const int k = j * Nx + i;
                const int row = j * Nx;
                const double phiE = phi[row + (i + 1 == Nx ? 0 : i + 1)];
                const double phiW = phi[row + (i == 0 ? Nx - 1 : i - 1)];
                const double phiN = phi[(j + 1 == Ny ? 0 : j + 1) * Nx + i];
                const double phiS = phi[(j == 0 ? Ny - 1 : j - 1) * Nx + i];
                const double r = RLDS ? rsrc[k] : a.c4piG * (smax(rho[k], a.floor_rho) - rhobar);
                const double phi_new = ((phiE + phiW) * a.idx2 + (phiN + phiS) * a.idy2 - r) / a.denom;
                double pk = phi[k];
                const double delta = phi_new - pk;
                pk += a.omega * delta;
                phi[k] = pk;
                const double lap = (phiE - 2.0 * pk + phiW) * a.idx2 + (phiN - 2.0 * pk + phiS) * a.idy2;
                mres = smax(mres, fabs(lap - r));
            }
            __syncthreads();
        }
        // mean removal after every sweep (EH:361-366)
        double ps = 0.0;
        for (int k = tid; k < N; k += B) ps += phi[k];
        block_sum_max(ps, mres, red_s, red_m);
        const double phibar = ps / (double)N;
        for (int k = tid; k < N; k += B) phi[k] -= phibar;
        __syncthreads();
        ++it;
        if (mres < a.tol) break;
    }
    // g = -grad(phi), central, periodic (EH:370-383)
    for (int k = tid; k < N; k += B) {
        const int j = k / Nx, i = k - j * Nx;
        const double phiE = phi[j * Nx + (i + 1 == Nx ? 0 : i + 1)], phiW = phi[j * Nx + (i == 0 ? Nx - 1 : i - 1)];
        const double phiN = phi[(j + 1 == Ny ? 0 : j + 1) * Nx + i], phiS = phi[(j == 0 ? Ny - 1 : j - 1) * Nx + i];
        gx[k] = -(phiE - phiW) / (2.0 * a.dx);
        gy[k] = -(phiN - phiS) / (2.0 * a.dy);
        if (LDS) gphi[k] = phi[k];
    }
    if (tid == 0) sweeps[0] = it;
}

// ------------------------------------------------------------------ rhs + stage kernel (EH:406-550)
constexpr int TXE = 32, TYE = 8, RXE = TXE + 4, RYE = TYE + 4;

// MUSCL face states from four cells along the face normal (EH:437-451): component c of the left
// state is w0 + s_L/2, of the right state wp - s_R/2, then the rho / P floors of EH:453-454.
__device__ __forceinline__ void muscl_eh(const GeomEH& g, const double (*sW)[RYE][RXE], int y0, int x0, int dy,
                                         int dx, Hydro4<double>& L, Hydro4<double>& R, bool yaxis) {
    double l[4], r[4];
#pragma unroll
    for (int c = 0; c < 4; ++c) {
        const double wm = sW[c][y0][x0], w0 = sW[c][y0 + dy][x0 + dx];
        const double wp = sW[c][y0 + 2 * dy][x0 + 2 * dx], wpp = sW[c][y0 + 3 * dy][x0 + 3 * dx];
        const double sL = mc_limiter<double>(w0 - wm, wp - w0);
        const double sR = mc_limiter<double>(wp - w0, wpp - wp);
        l[c] = w0 + 0.5 * sL;
        r[c] = wp - 0.5 * sR;
    }
    // (rho, u, v, P) -> (rho, un, ut, P)
    L = Hydro4<double>{smax(l[0], g.floor_rho), yaxis ? l[2] : l[1], yaxis ? l[1] : l[2], smax(l[3], g.floor_P)};
    R = Hydro4<double>{smax(r[0], g.floor_rho), yaxis ? r[2] : r[1], yaxis ? r[1] : r[2], smax(r[3], g.floor_P)};
}

template <bool MIXED, bool YAXIS>
__device__ __forceinline__ void face_hllc_eh(const GeomEH& g, const Hydro4<double>& L, const Hydro4<double>& R,
                                             double F[4]) {
    if (MIXED) {
        const Hydro4<float> l{(float)L.rho, (float)L.un, (float)L.ut, (float)L.P};
        const Hydro4<float> r{(float)R.rho, (float)R.un, (float)R.ut, (float)R.P};
        float f[4];
        hllc_eh<float, YAXIS>((float)g.gamma, (float)g.floor_rho, l, r, f);
        F[0] = f[0]; F[1] = f[1]; F[2] = f[2]; F[3] = f[3];
    } else {
        hllc_eh<double, YAXIS>(g.gamma, g.floor_rho, L, R, F);
    }
}

// MODE 0: dU = rhs(Uin) -> out.  MODE 1: out = Uin + dt rhs(Uin), floors.  MODE 2: out = Uold/2 +
// (Uin + dt rhs(Uin))/2, floors, and the block maxima of the new state's signal speed (EH:399-400).
template <int MODE, bool MIXED>
__global__ __launch_bounds__(256) void k_eh_stage(GeomEH g, StateEH Uin, StateEH Uold, StateEH out,
                                                  const double* __restrict__ gx, const double* __restrict__ gy,
                                                  double dt, double* __restrict__ spd) {
    __shared__ double sW[4][RYE][RXE];
    __shared__ double sFx[4][TYE][TXE + 1];
    __shared__ double sFy[4][TYE + 1][TXE];
    const int tid = threadIdx.x;
    const int i0 = blockIdx.x * TXE, j0 = blockIdx.y * TYE;
    const int Nx = g.Nx, Ny = g.Ny;

    // 1. floored primitives of the tile + halo (EH:415-425)
    for (int c = tid; c < RXE * RYE; c += 256) {
        const int lx = c % RXE, ly = c / RXE;
        const int i = bmap(i0 - 2 + lx, Nx, g.px), j = bmap(j0 - 2 + ly, Ny, g.py);
        const size_t q = (size_t)j * Nx + i;
        double rho = Uin.rho[q], mx = Uin.mx[q], my = Uin.my[q], E = Uin.E[q];
        floors_eh(g, rho, mx, my, E);
        double u, v, P, a;
        prim_eh(g, rho, mx, my, E, u, v, P, a);
        sW[0][ly][lx] = smax(rho, g.floor_rho); sW[1][ly][lx] = u; sW[2][ly][lx] = v; sW[3][ly][lx] = P;
    }
    __syncthreads();
    // 2. x-faces i0..i0+TX of rows j0..j0+TY-1 (EH:427-464)
    for (int f = tid; f < (TXE + 1) * TYE; f += 256) {
        const int fx = f % (TXE + 1), fy = f / (TXE + 1);
        if (i0 + fx > Nx || j0 + fy >= Ny) continue;
        Hydro4<double> L, R;
        muscl_eh(g, sW, fy + 2, fx, 0, 1, L, R, false);
        double F[4];
        face_hllc_eh<MIXED, false>(g, L, R, F);
        sFx[0][fy][fx] = F[0]; sFx[1][fy][fx] = F[1]; sFx[2][fy][fx] = F[2]; sFx[3][fy][fx] = F[3];
    }
    // 3. y-faces j0..j0+TY of columns i0..i0+TX-1 (EH:466-503); normal momentum is my
    for (int f = tid; f < TXE * (TYE + 1); f += 256) {
        const int fx = f % TXE, fy = f / TXE;
        if (i0 + fx >= Nx || j0 + fy > Ny) continue;
        Hydro4<double> L, R;
        muscl_eh(g, sW, fy, fx + 2, 1, 0, L, R, true);
        double F[4];
        face_hllc_eh<MIXED, true>(g, L, R, F);
        sFy[0][fy][fx] = F[0]; sFy[1][fy][fx] = F[2]; sFy[2][fy][fx] = F[1]; sFy[3][fy][fx] = F[3];
    }
    __syncthreads();
    // 4. flux divergence + gravity source (EH:505-524), then the stage update
    const int tx = tid % TXE, ty = tid / TXE;
    const int i = i0 + tx, j = j0 + ty;
    double speed = 0.0;
    if (i < Nx && j < Ny) {
        const size_t q = (size_t)j * Nx + i;
        double d[4];
#pragma unroll
        for (int c = 0; c < 4; ++c)
            d[c] = -(sFx[c][ty][tx + 1] - sFx[c][ty][tx]) / g.dx - (sFy[c][ty + 1][tx] - sFy[c][ty][tx]) / g.dy;
        const double r0 = Uin.rho[q], m0 = Uin.mx[q], n0 = Uin.my[q], e0 = Uin.E[q];
        if (g.gravity_on) {
            const double rho = smax(r0, g.floor_rho);
            const double u = m0 / rho, v = n0 / rho;
            d[1] += rho * gx[q];
            d[2] += rho * gy[q];
            d[3] += rho * (u * gx[q] + v * gy[q]);
        }
        if (MODE == 0) {
            out.rho[q] = d[0]; out.mx[q] = d[1]; out.my[q] = d[2]; out.E[q] = d[3];
        } else {
            double rho, mx, my, E;
            if (MODE == 1) {  // EH:534-541
                rho = r0 + dt * d[0]; mx = m0 + dt * d[1]; my = n0 + dt * d[2]; E = e0 + dt * d[3];
            } else {  // EH:544-549
                rho = 0.5 * Uold.rho[q] + 0.5 * (r0 + dt * d[0]);
                mx = 0.5 * Uold.mx[q] + 0.5 * (m0 + dt * d[1]);
                my = 0.5 * Uold.my[q] + 0.5 * (n0 + dt * d[2]);
                E = 0.5 * Uold.E[q] + 0.5 * (e0 + dt * d[3]);
            }
            floors_eh(g, rho, mx, my, E);
            out.rho[q] = rho; out.mx[q] = mx; out.my[q] = my; out.E[q] = E;
            if (MODE == 2) speed = speed_eh(g, rho, mx, my, E);
        }
    }
    if (MODE == 2) block_max_to(speed, spd);
}

__global__ __launch_bounds__(256) void k_eh_speed(GeomEH g, StateEH U, double* __restrict__ spd) {
    const int i = blockIdx.x * 256 + threadIdx.x, j = blockIdx.y;
    double s = 0.0;
    if (i < g.Nx) {
        const size_t q = (size_t)j * g.Nx + i;
        s = speed_eh(g, U.rho[q], U.mx[q], U.my[q], U.E[q]);
    }
    block_max_to(s, spd);
}

// dt = cfl min(dx, dy) / max(1e-30, max speed) (EH:390-404)
__global__ __launch_bounds__(256) void k_eh_dt(GeomEH g, const double* __restrict__ spd, int nspd,
                                               double* __restrict__ dt_out) {
    const double maxs = block_reduce_max(spd, nspd, 1e-30);
    if (threadIdx.x == 0) dt_out[0] = g.cfl * smin(g.dx, g.dy) / maxs;
}

template <bool MIXED>
__global__ void k_hllc_batch(int variant, int axis, int n, const double* __restrict__ WLa,
                             const double* __restrict__ WRa, double gamma, double floor_rho, double* __restrict__ out) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    const double* a = WLa + 4 * (size_t)k;
    const double* b = WRa + 4 * (size_t)k;
    const bool y = axis == 1;
    // (rho, u, v, P) -> (rho, un, ut, P)
    const Hydro4<double> L{a[0], y ? a[2] : a[1], y ? a[1] : a[2], a[3]};
    const Hydro4<double> R{b[0], y ? b[2] : b[1], y ? b[1] : b[2], b[3]};
    double F[4];
    if (MIXED) {
        const Hydro4<float> l{(float)L.rho, (float)L.un, (float)L.ut, (float)L.P};
        const Hydro4<float> r{(float)R.rho, (float)R.un, (float)R.ut, (float)R.P};
        float f[4];
        if (variant == 0) {
            if (y) hllc_eh<float, true>((float)gamma, (float)floor_rho, l, r, f);
            else hllc_eh<float, false>((float)gamma, (float)floor_rho, l, r, f);
        } else {
            hllc_ns<float>((float)gamma, (float)floor_rho, l, r, f);
        }
        for (int c = 0; c < 4; ++c) F[c] = f[c];
    } else if (variant == 0) {
        if (y) hllc_eh<double, true>(gamma, floor_rho, L, R, F);
        else hllc_eh<double, false>(gamma, floor_rho, L, R, F);
    } else {
        hllc_ns<double>(gamma, floor_rho, L, R, F);
    }
    double* o = out + 4 * (size_t)k;
    o[0] = F[0];
    o[1] = y ? F[2] : F[1];  // f_mx
    o[2] = y ? F[1] : F[2];  // f_my
    o[3] = F[3];
}

// Snapshot frame (the store lambda of EH:661-691): enforce_floors on a copy (EH:62-72), then
// cons_to_prim_full (EH:74-83); Mach = |v| / a for a > 0, else 0.
__global__ __launch_bounds__(256) void k_eh_frame(GeomEH g, StateEH U, double* __restrict__ out) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x, j = blockIdx.y;
    if (i >= g.Nx) return;
    const size_t N = (size_t)g.Nx * g.Ny, q = (size_t)j * g.Nx + i;
    double rho = smax(U.rho[q], g.floor_rho);
    const double mx = U.mx[q], my = U.my[q];
    double E = U.E[q];
    {
        const double u = mx / rho, v = my / rho;
        const double kin = 0.5 * rho * (u * u + v * v);
        const double P = g.gm1 * (E - kin);
        if (P < g.floor_P) E = kin + g.floor_P / g.gm1;
    }
    const double r = smax(rho, g.floor_rho);
    const double u = mx / r, v = my / r;
    const double kin = 0.5 * r * (u * u + v * v);
    double P = g.gm1 * (E - kin);
    if (P < g.floor_P) P = g.floor_P;
    const double a = sqrt(g.gamma * P / r);
    const double vmag = sqrt(u * u + v * v);
    out[q] = rho;
    out[N + q] = P;
    out[2 * N + q] = (a > 0.0) ? (vmag / a) : 0.0;
}

}  // namespace cc

using namespace cc;

// ------------------------------------------------------------------ host handle
struct cc_euler2d {
    cc_euler2d_params p{};
    GeomEH g{};
    SorEH sa{};
    hipStream_t stream = nullptr;
    StateEH U{}, U1{}, T{};  // U^n (updated in place), U1, scratch for dU / downloads
    double *gx = nullptr, *gy = nullptr, *phi = nullptr, *spd = nullptr, *dtd = nullptr;
    int* sweeps = nullptr;
    int nspd = 0, sor_threads = 64, last_sweeps = 0;
    bool sor_lds = false;
    long long total_sweeps = 0;
    int solves = 0;
    Profiler prof;
    FrameBuffer frame;
    FieldStats stats;
    std::vector<void*> allocs;

    size_t N() const { return (size_t)p.Nx * p.Ny; }
    Profiler* pr() { return prof.on ? &prof : nullptr; }
    cc_status make_frame() {
        CC_TRY(frame.ensure((size_t)CC_EULER2D_FRAME_FIELDS * N()));
        hipLaunchKernelGGL(k_eh_frame, dim3((p.Nx + 255) / 256, p.Ny), dim3(256), 0, stream, g, U, frame.d);
        CC_CHECK_LAUNCH();
        return CC_OK;
    }

    cc_status alloc(void** ptr, size_t bytes) {
        if (hipMalloc(ptr, bytes) != hipSuccess) {
            set_error(CC_ERR_OOM, "cc_euler2d: device allocation failed");
            return CC_ERR_OOM;
        }
        allocs.push_back(*ptr);
        CC_HIP(hipMemsetAsync(*ptr, 0, bytes, stream));
        return CC_OK;
    }
    cc_status alloc_state(StateEH& s) {
        CC_TRY(alloc((void**)&s.rho, N() * 8)); CC_TRY(alloc((void**)&s.mx, N() * 8));
        CC_TRY(alloc((void**)&s.my, N() * 8)); CC_TRY(alloc((void**)&s.E, N() * 8));
        return CC_OK;
    }

    // g of the density of state s (compute_poisson_gravity_periodic, EH:288-386)
    cc_status gravity(const StateEH& s) {
        if (!p.gravity_on) return CC_OK;  // g stays zero (EH:293)
        ProfScope ps(pr(), stream, "sor");
        const bool rlds = N() <= 4096;  // phi and the source term in 64 KB of LDS
        const size_t lds = sor_lds ? N() * 8 * (rlds ? 2 : 1) : 0;
        if (sor_lds && rlds)
            hipLaunchKernelGGL((k_eh_sor<true, true>), dim3(1), dim3(sor_threads), lds, stream, sa, s.rho, phi, gx, gy,
                               sweeps);
        else if (sor_lds)
            hipLaunchKernelGGL((k_eh_sor<true, false>), dim3(1), dim3(sor_threads), lds, stream, sa, s.rho, phi, gx,
                               gy, sweeps);
        else
            hipLaunchKernelGGL((k_eh_sor<false, false>), dim3(1), dim3(sor_threads), 0, stream, sa, s.rho, phi, gx, gy,
                               sweeps);
        CC_CHECK_LAUNCH();
        return CC_OK;
    }
    template <int MODE>
    cc_status stage(const StateEH& in, const StateEH& old, const StateEH& out, double dt) {
        ProfScope ps(pr(), stream, "flux");
        dim3 grid((p.Nx + TXE - 1) / TXE, (p.Ny + TYE - 1) / TYE);
        if (p.precision == CC_MIXED)
            hipLaunchKernelGGL((k_eh_stage<MODE, true>), grid, dim3(256), 0, stream, g, in, old, out, gx, gy, dt, spd);
        else
            hipLaunchKernelGGL((k_eh_stage<MODE, false>), grid, dim3(256), 0, stream, g, in, old, out, gx, gy, dt, spd);
        CC_CHECK_LAUNCH();
        return CC_OK;
    }
    cc_status read_sweeps() {
        if (!p.gravity_on) return CC_OK;
        CC_HIP(hipMemcpyAsync(&last_sweeps, sweeps, sizeof(int), hipMemcpyDeviceToHost, stream));
        CC_HIP(hipStreamSynchronize(stream));
        total_sweeps += last_sweeps;
        solves += 1;
        return CC_OK;
    }
    // rk2_step EH:529-550
    cc_status rk2(double dt) {
        CC_TRY(gravity(U));
        CC_TRY(stage<1>(U, U, U1, dt));
        CC_TRY(read_sweeps());
        CC_TRY(gravity(U1));
        CC_TRY(stage<2>(U1, U, U, dt));
        CC_TRY(read_sweeps());
        return CC_OK;
    }
    cc_status compute_dt(double* dt) {
        {
            ProfScope ps(pr(), stream, "dt");
            hipLaunchKernelGGL(k_eh_speed, dim3((p.Nx + 255) / 256, p.Ny), dim3(256), 0, stream, g, U, spd);
            CC_CHECK_LAUNCH();
            hipLaunchKernelGGL(k_eh_dt, dim3(1), dim3(256), 0, stream, g, spd, nspd, dtd);
            CC_CHECK_LAUNCH();
        }
        CC_HIP(hipMemcpyAsync(dt, dtd, sizeof(double), hipMemcpyDeviceToHost, stream));
        CC_HIP(hipStreamSynchronize(stream));
        return CC_OK;
    }
    ~cc_euler2d() {
        for (void* a : allocs) hipFree(a);
        if (stream) hipStreamDestroy(stream);
    }
};

#define CHECK_H()                                         \
    do {                                                  \
        if (!h) {                                         \
            set_error(CC_ERR_INVALID_ARG, "null handle"); \
            return CC_ERR_INVALID_ARG;                    \
        }                                                 \
    } while (0)

extern "C" {

cc_status cc_euler2d_create(const cc_euler2d_params* p, cc_euler2d** out) {
    if (!out || !p) { set_error(CC_ERR_INVALID_ARG, "null params or output handle"); return CC_ERR_INVALID_ARG; }
    *out = nullptr;
    if (p->Nx < 2 || p->Ny < 2 || p->Nx > 32768 || p->Ny > 32768 || !(p->x1 > p->x0) || !(p->y1 > p->y0) ||
        !(p->gamma > 1.0) || p->max_iter < 0 || !(p->omega > 0.0 && p->omega < 2.0)) {
        set_error(CC_ERR_INVALID_ARG, "cc_euler2d: Nx, Ny in [2, 32768], x1 > x0, y1 > y0, gamma > 1, max_iter >= 0, "
                                      "0 < omega < 2 required");
        return CC_ERR_INVALID_ARG;
    }
    if (p->precision != CC_FP64 && p->precision != CC_MIXED) {
        set_error(CC_ERR_INVALID_ARG, "precision must be CC_FP64 or CC_MIXED");
        return CC_ERR_INVALID_ARG;
    }
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) {
        set_error(CC_ERR_NO_DEVICE, "no HIP device visible");
        return CC_ERR_NO_DEVICE;
    }
    cc_euler2d* h = new cc_euler2d();
    h->p = *p;
    GeomEH& g = h->g;
    g.Nx = p->Nx; g.Ny = p->Ny; g.px = p->periodic_x != 0; g.py = p->periodic_y != 0;
    g.dx = (p->x1 - p->x0) / p->Nx; g.dy = (p->y1 - p->y0) / p->Ny;
    g.gamma = p->gamma; g.gm1 = p->gamma - 1.0; g.floor_rho = p->floor_rho; g.floor_P = p->floor_P;
    g.cfl = p->cfl; g.gravity_on = p->gravity_on != 0;
    SorEH& a = h->sa;
    a.Nx = p->Nx; a.Ny = p->Ny; a.max_iter = p->max_iter;
    a.idx2 = 1.0 / (g.dx * g.dx); a.idy2 = 1.0 / (g.dy * g.dy);
    a.denom = 2.0 * (a.idx2 + a.idy2);
    a.omega = p->omega; a.tol = p->tol;
    const double FOURPI = 4.0 * M_PI;
    a.c4piG = FOURPI * p->G;  // rhs = (4 pi G)(rho - rhobar), EH:320-323
    a.floor_rho = p->floor_rho; a.dx = g.dx; a.dy = g.dy;
    h->sor_lds = (size_t)p->Nx * p->Ny <= 8192;
    h->sor_threads = std::min(1024, ((p->Ny + 63) / 64) * 64);
    h->nspd = ((p->Nx + 255) / 256) * p->Ny;
    const int nblk = ((p->Nx + TXE - 1) / TXE) * ((p->Ny + TYE - 1) / TYE);
    cc_status st = CC_OK;
    if (hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking) != hipSuccess) {
        set_error(CC_ERR_HIP, "hipStreamCreate failed");
        delete h;
        return CC_ERR_HIP;
    }
    if (st == CC_OK) st = h->alloc_state(h->U);
    if (st == CC_OK) st = h->alloc_state(h->U1);
    if (st == CC_OK) st = h->alloc_state(h->T);
    if (st == CC_OK) st = h->alloc((void**)&h->gx, h->N() * 8);
    if (st == CC_OK) st = h->alloc((void**)&h->gy, h->N() * 8);
    if (st == CC_OK) st = h->alloc((void**)&h->phi, h->N() * 8);
    if (st == CC_OK) st = h->alloc((void**)&h->spd, (size_t)std::max(h->nspd, nblk) * 8);
    if (st == CC_OK) st = h->alloc((void**)&h->dtd, 8);
    if (st == CC_OK) st = h->alloc((void**)&h->sweeps, 8);
    if (st == CC_OK && hipStreamSynchronize(h->stream) != hipSuccess) {
        set_error(CC_ERR_HIP, "stream sync failed in create");
        st = CC_ERR_HIP;
    }
    if (st != CC_OK) {
        delete h;
        return st;
    }
    *out = h;
    return CC_OK;
}

cc_status cc_euler2d_destroy(cc_euler2d* h) {
    if (!h) return CC_OK;
    hipStreamSynchronize(h->stream);
    delete h;
    return CC_OK;
}

cc_status cc_euler2d_upload(cc_euler2d* h, const double* rho, const double* mx, const double* my, const double* E) {
    CHECK_H();
    if (!rho || !mx || !my || !E) { set_error(CC_ERR_INVALID_ARG, "null state pointer"); return CC_ERR_INVALID_ARG; }
    const size_t b = h->N() * 8;
    CC_HIP(hipMemcpyAsync(h->U.rho, rho, b, hipMemcpyHostToDevice, h->stream));
    CC_HIP(hipMemcpyAsync(h->U.mx, mx, b, hipMemcpyHostToDevice, h->stream));
    CC_HIP(hipMemcpyAsync(h->U.my, my, b, hipMemcpyHostToDevice, h->stream));
    CC_HIP(hipMemcpyAsync(h->U.E, E, b, hipMemcpyHostToDevice, h->stream));
    CC_HIP(hipStreamSynchronize(h->stream));
    return CC_OK;
}

cc_status cc_euler2d_download(cc_euler2d* h, double* rho, double* mx, double* my, double* E) {
    CHECK_H();
    if (!rho || !mx || !my || !E) { set_error(CC_ERR_INVALID_ARG, "null state pointer"); return CC_ERR_INVALID_ARG; }
    const size_t b = h->N() * 8;
    CC_HIP(hipMemcpyAsync(rho, h->U.rho, b, hipMemcpyDeviceToHost, h->stream));
    CC_HIP(hipMemcpyAsync(mx, h->U.mx, b, hipMemcpyDeviceToHost, h->stream));
    CC_HIP(hipMemcpyAsync(my, h->U.my, b, hipMemcpyDeviceToHost, h->stream));
    CC_HIP(hipMemcpyAsync(E, h->U.E, b, hipMemcpyDeviceToHost, h->stream));
    CC_HIP(hipStreamSynchronize(h->stream));
    return CC_OK;
}

cc_status cc_euler2d_gravity(cc_euler2d* h, double* gx, double* gy, double* phi, int* sweeps) {
    CHECK_H();
    CC_TRY(h->gravity(h->U));
    const size_t b = h->N() * 8;
    if (!h->p.gravity_on) {  // EH:293: zero field, phi never formed
        CC_HIP(hipMemsetAsync(h->gx, 0, b, h->stream));
        CC_HIP(hipMemsetAsync(h->gy, 0, b, h->stream));
        CC_HIP(hipMemsetAsync(h->phi, 0, b, h->stream));
    }
    if (gx) CC_HIP(hipMemcpyAsync(gx, h->gx, b, hipMemcpyDeviceToHost, h->stream));
    if (gy) CC_HIP(hipMemcpyAsync(gy, h->gy, b, hipMemcpyDeviceToHost, h->stream));
    if (phi) CC_HIP(hipMemcpyAsync(phi, h->phi, b, hipMemcpyDeviceToHost, h->stream));
    int n = 0;
    if (h->p.gravity_on) CC_HIP(hipMemcpyAsync(&n, h->sweeps, sizeof(int), hipMemcpyDeviceToHost, h->stream));
    CC_HIP(hipStreamSynchronize(h->stream));
    if (sweeps) *sweeps = n;
    return CC_OK;
}

cc_status cc_euler2d_compute_dt(cc_euler2d* h, double* dt) {
    CHECK_H();
    if (!dt) { set_error(CC_ERR_INVALID_ARG, "null dt"); return CC_ERR_INVALID_ARG; }
    return h->compute_dt(dt);
}

cc_status cc_euler2d_rhs(cc_euler2d* h, double* drho, double* dmx, double* dmy, double* dE) {
    CHECK_H();
    if (!drho || !dmx || !dmy || !dE) { set_error(CC_ERR_INVALID_ARG, "null output"); return CC_ERR_INVALID_ARG; }
    CC_TRY(h->gravity(h->U));
    CC_TRY(h->stage<0>(h->U, h->U, h->T, 0.0));
    const size_t b = h->N() * 8;
    CC_HIP(hipMemcpyAsync(drho, h->T.rho, b, hipMemcpyDeviceToHost, h->stream));
    CC_HIP(hipMemcpyAsync(dmx, h->T.mx, b, hipMemcpyDeviceToHost, h->stream));
    CC_HIP(hipMemcpyAsync(dmy, h->T.my, b, hipMemcpyDeviceToHost, h->stream));
    CC_HIP(hipMemcpyAsync(dE, h->T.E, b, hipMemcpyDeviceToHost, h->stream));
    CC_HIP(hipStreamSynchronize(h->stream));
    return CC_OK;
}

cc_status cc_euler2d_rk2_step(cc_euler2d* h, double dt) {
    CHECK_H();
    return h->rk2(dt);
}

cc_status cc_euler2d_run(cc_euler2d* h, int nsteps, double* t_inout, double t_end, double* dts) {
    CHECK_H();
    if (!t_inout || nsteps < 0) { set_error(CC_ERR_INVALID_ARG, "null t or negative nsteps"); return CC_ERR_INVALID_ARG; }
    double t = *t_inout;
    for (int s = 0; s < nsteps && t < t_end - 1e-14; ++s) {  // EH:697-708
        double dt = 0.0;
        CC_TRY(h->compute_dt(&dt));
        if (!std::isfinite(dt) || dt <= 0.0) {
            *t_inout = t;
            set_error(CC_ERR_BAD_DT, "Bad dt=" + std::to_string(dt) + " at t=" + std::to_string(t));
            return CC_ERR_BAD_DT;
        }
        if (t + dt > t_end) dt = t_end - t;
        CC_TRY(h->rk2(dt));
        t += dt;
        if (dts) dts[s] = dt;
    }
    CC_HIP(hipStreamSynchronize(h->stream));
    *t_inout = t;
    return CC_OK;
}

cc_status cc_euler2d_sweep_stats(cc_euler2d* h, int* last, long long* total, int* solves) {
    CHECK_H();
    if (last) *last = h->last_sweeps;
    if (total) *total = h->total_sweeps;
    if (solves) *solves = h->solves;
    return CC_OK;
}

cc_status cc_euler2d_frame(cc_euler2d* h, double* out) {
    CHECK_H();
    if (!out) { set_error(CC_ERR_INVALID_ARG, "null output"); return CC_ERR_INVALID_ARG; }
    CC_TRY(h->make_frame());
    CC_HIP(hipMemcpyAsync(out, h->frame.d, sizeof(double) * CC_EULER2D_FRAME_FIELDS * h->N(), hipMemcpyDeviceToHost,
                          h->stream));
    CC_HIP(hipStreamSynchronize(h->stream));
    return CC_OK;
}

cc_status cc_euler2d_totals(cc_euler2d* h, double* out4) {
    CHECK_H();
    if (!out4) { set_error(CC_ERR_INVALID_ARG, "null output"); return CC_ERR_INVALID_ARG; }
    FieldPtrs a{{h->U.rho, h->U.mx, h->U.my, h->U.E}};
    double r[12];
    CC_TRY(h->stats.run(a, nullptr, 4, h->N(), h->stream, r));
    for (int f = 0; f < 4; f++) out4[f] = r[3 * f] * (h->g.dx * h->g.dy);
    return CC_OK;
}

cc_status cc_euler2d_compare(cc_euler2d* a, cc_euler2d* b, double* out) {
    if (!a || !b || !out) { set_error(CC_ERR_INVALID_ARG, "null handle or output"); return CC_ERR_INVALID_ARG; }
    if (a->p.Nx != b->p.Nx || a->p.Ny != b->p.Ny) {
        set_error(CC_ERR_INVALID_ARG, "cc_euler2d_compare: grids differ");
        return CC_ERR_INVALID_ARG;
    }
    CC_TRY(a->make_frame());
    CC_TRY(b->make_frame());
    CC_HIP(hipStreamSynchronize(b->stream));
    return compare_frames(a->stats, a->frame.d, b->frame.d, CC_EULER2D_FRAME_FIELDS, a->N(), a->stream, out);
}

cc_status cc_euler2d_sync(cc_euler2d* h) {
    CHECK_H();
    CC_HIP(hipStreamSynchronize(h->stream));
    return CC_OK;
}

cc_status cc_euler2d_set_profiling(cc_euler2d* h, int on) {
    CHECK_H();
    CC_HIP(hipStreamSynchronize(h->stream));
    h->prof.on = on != 0;
    return CC_OK;
}

cc_status cc_euler2d_kernel_time(cc_euler2d* h, const char* kernel, int* count, double* total_ms) {
    CHECK_H();
    if (!kernel || !count || !total_ms) { set_error(CC_ERR_INVALID_ARG, "null argument"); return CC_ERR_INVALID_ARG; }
    CC_HIP(hipStreamSynchronize(h->stream));
    h->prof.get_total(kernel, count, total_ms);
    return CC_OK;
}

cc_status cc_euler2d_reset_profiling(cc_euler2d* h) {
    CHECK_H();
    CC_HIP(hipStreamSynchronize(h->stream));
    h->prof.reset();
    h->total_sweeps = 0;
    h->solves = 0;
    return CC_OK;
}

cc_status cc_hllc2d_batch(int variant, int axis, int n, const double* WL, const double* WR, double gamma,
                          double floor_rho, int precision, double* out) {
    if (n < 0 || (n > 0 && (!WL || !WR || !out)) || (axis != 0 && axis != 1) || (variant != 0 && variant != 1)) {
        set_error(CC_ERR_INVALID_ARG, "cc_hllc2d_batch: bad arguments");
        return CC_ERR_INVALID_ARG;
    }
    if (n == 0) return CC_OK;
    const size_t b = (size_t)n * 4 * 8;
    double *dL = nullptr, *dR = nullptr, *dO = nullptr;
    if (hipMalloc(&dL, b) != hipSuccess || hipMalloc(&dR, b) != hipSuccess || hipMalloc(&dO, b) != hipSuccess) {
        hipFree(dL); hipFree(dR); hipFree(dO);
        set_error(CC_ERR_OOM, "batch alloc");
        return CC_ERR_OOM;
    }
    cc_status st = CC_OK;
    if (hipMemcpy(dL, WL, b, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(dR, WR, b, hipMemcpyHostToDevice) != hipSuccess) {
        set_error(CC_ERR_HIP, "batch upload");
        st = CC_ERR_HIP;
    }
    if (st == CC_OK) {
        if (precision == CC_MIXED)
            hipLaunchKernelGGL(k_hllc_batch<true>, dim3((n + 255) / 256), dim3(256), 0, 0, variant, axis, n, dL, dR,
                               gamma, floor_rho, dO);
        else
            hipLaunchKernelGGL(k_hllc_batch<false>, dim3((n + 255) / 256), dim3(256), 0, 0, variant, axis, n, dL, dR,
                               gamma, floor_rho, dO);
        if (hipGetLastError() != hipSuccess || hipMemcpy(out, dO, b, hipMemcpyDeviceToHost) != hipSuccess) {
            set_error(CC_ERR_HIP, "batch kernel / download");
            st = CC_ERR_HIP;
        }
    }
    hipFree(dL); hipFree(dR); hipFree(dO);
    return st;
}

}  // extern "C"
