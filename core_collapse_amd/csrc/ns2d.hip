// ns2d.hip -- gfx950 solver step of the 2D Navier-Stokes HLLC + FFT self-gravity code
// (navier_stokes_2d_hllc_poisson.cpp, cited as NS:<line>).  Periodic in both axes, AoS state
// U[q] = {rho, mx, my, E} with q = i + Nx*j as in the reference (NS:88-93).
//
// One right-hand side (compute_rhs, NS:363-518) is:
//   k_ns_rho        rho = max(U.rho, floor) into a plane + per-block partial sums
//   k_ns_mean       fixed-order mean of the partials
//   4 FFT passes    rows fwd (rho - mean) -> columns fwd + Green factor -(4 pi G)/k^2 -> rows inv ->
//                   columns inv -> phi = Re (solve_poisson_periodic_fft, NS:285-335)
//   k_ns_stage      32x8 tile: primitives of tile + 2-cell halo in LDS, central velocity gradients on
//                   tile + 1, first-order HLLC (W[i], W[i+1]) plus viscous stresses on x- and y-faces,
//                   flux divergence + rho g (g = -central grad phi, NS:337-360, read from phi), and the
//                   Heun stage update of NS:758-781 (rho floor only) -- or dU itself (MODE 0).
// The stage-2 instance also emits per-block signal-speed maxima and density minima for the next dt
// (compute_dt, NS:521-546).  Mixed precision runs the HLLC in FP32 on the FP64 primitives.
// This file is compiled without FP contraction so the FP64 path keeps the reference's rounding.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "diag.hpp"
#include "fft.hpp"
#include "hllc.hpp"
#include "profiler.hpp"
#include "stepcommon.hpp"

namespace cc {

struct GeomNS {
    int Nx, Ny, visc, gravity_on;
    double dx, dy, gamma, gm1, floor_rho, floor_P, mu, cfl, cfl_visc;
    double idx1, idy1, idx2c, idy2c;  // 1/dx, 1/dy, 1/(2dx), 1/(2dy)
};

__device__ __forceinline__ double clamp_min(double x, double lo) { return (x < lo) ? lo : x; }  // NS:51

// cons_to_prim NS:119-132 (P clamped, no energy fix)
__device__ __forceinline__ double4 prim_ns(const GeomNS& g, double4 U) {
    double4 W;
    W.x = clamp_min(U.x, g.floor_rho);
    W.y = U.y / W.x;
    W.z = U.z / W.x;
    const double ke = 0.5 * W.x * (W.y * W.y + W.z * W.z);
    const double eint = U.w - ke;
    const double P = g.gm1 * eint;
    W.w = clamp_min(P, g.floor_P);
    return W;
}

__device__ __forceinline__ double sound_ns(const GeomNS& g, const double4& W) {  // NS:147-149
    return sqrt(g.gamma * W.w / clamp_min(W.x, g.floor_rho));
}

// ------------------------------------------------------------------ Poisson helpers
__global__ __launch_bounds__(256) void k_ns_rho(int N, double floor_rho, const double4* __restrict__ U,
                                                double* __restrict__ rho, double* __restrict__ part) {
    __shared__ double red[4];
    double s = 0.0;
    for (int q = blockIdx.x * 256 + threadIdx.x; q < N; q += gridDim.x * 256) {
        const double r = clamp_min(U[q].x, floor_rho);
        rho[q] = r;
        s += r;
    }
    s = wave_sum(s);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0) part[blockIdx.x] = (red[0] + red[1]) + (red[2] + red[3]);
}

__global__ __launch_bounds__(256) void k_sum_partial(int N, const double* __restrict__ x, double* __restrict__ part) {
    __shared__ double red[4];
    double s = 0.0;
    for (int q = blockIdx.x * 256 + threadIdx.x; q < N; q += gridDim.x * 256) s += x[q];
    s = wave_sum(s);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0) part[blockIdx.x] = (red[0] + red[1]) + (red[2] + red[3]);
}

__global__ __launch_bounds__(256) void k_ns_mean(int nb, int N, const double* __restrict__ part,
                                                 double* __restrict__ mean) {
    __shared__ double red[4];
    double s = 0.0;
    for (int k = threadIdx.x; k < nb; k += 256) s += part[k];
    s = wave_sum(s);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0) mean[0] = ((red[0] + red[1]) + (red[2] + red[3])) / (double)N;
}

// ------------------------------------------------------------------ rhs + stage kernel
constexpr int TXN = 32, TYN = 8, RXN = TXN + 4, RYN = TYN + 4;

template <bool MIXED>
__device__ __forceinline__ void face_hllc_ns(const GeomNS& g, const double4& WL, const double4& WR, bool yaxis,
                                             double F[4]) {
    // (rho, u, v, P) -> (rho, un, ut, P); hllc_y swaps u and v (NS:270-282)
    const Hydro4<double> L{WL.x, yaxis ? WL.z : WL.y, yaxis ? WL.y : WL.z, WL.w};
    const Hydro4<double> R{WR.x, yaxis ? WR.z : WR.y, yaxis ? WR.y : WR.z, WR.w};
    if (MIXED) {
        const Hydro4<float> l{(float)L.rho, (float)L.un, (float)L.ut, (float)L.P};
        const Hydro4<float> r{(float)R.rho, (float)R.un, (float)R.ut, (float)R.P};
        float f[4];
        hllc_ns<float>((float)g.gamma, (float)g.floor_rho, l, r, f);
        F[0] = f[0]; F[1] = f[1]; F[2] = f[2]; F[3] = f[3];
    } else {
        hllc_ns<double>(g.gamma, g.floor_rho, L, R, F);
    }
}

// MODE 0: dU = rhs(Uin) -> out.  MODE 1: K = rhs(Uin); out = Uin + dt K (rho floor).  MODE 2:
// out = Uold + dt/2 (K + rhs(Uin)) (rho floor) and the dt partials of the new state.
template <int MODE, bool MIXED>
__global__ __launch_bounds__(256) void k_ns_stage(GeomNS g, const double4* __restrict__ Uin,
                                                  const double4* Uold, double4* K, double4* out,
                                                  const double* __restrict__ phi, double dt,
                                                  double* __restrict__ smaxp, double* __restrict__ rminp) {
    __shared__ double sW[4][RYN][RXN];
    __shared__ double sG[4][RYN][RXN];  // du/dx, du/dy, dv/dx, dv/dy (NS:393-411)
    __shared__ double sFx[4][TYN][TXN + 1];
    __shared__ double sFy[4][TYN + 1][TXN];
    const int tid = threadIdx.x;
    const int i0 = blockIdx.x * TXN, j0 = blockIdx.y * TYN;
    const int Nx = g.Nx, Ny = g.Ny;
    const bool visc = g.visc;

    for (int c = tid; c < RXN * RYN; c += 256) {
        const int lx = c % RXN, ly = c / RXN;
        const int i = imod(i0 - 2 + lx, Nx), j = imod(j0 - 2 + ly, Ny);
        const double4 W = prim_ns(g, Uin[(size_t)j * Nx + i]);
        sW[0][ly][lx] = W.x; sW[1][ly][lx] = W.y; sW[2][ly][lx] = W.z; sW[3][ly][lx] = W.w;
    }
    __syncthreads();
    if (visc) {
        for (int c = tid; c < (RXN - 2) * (RYN - 2); c += 256) {
            const int lx = 1 + c % (RXN - 2), ly = 1 + c / (RXN - 2);
            sG[0][ly][lx] = (sW[1][ly][lx + 1] - sW[1][ly][lx - 1]) * g.idx2c;
            sG[2][ly][lx] = (sW[2][ly][lx + 1] - sW[2][ly][lx - 1]) * g.idx2c;
            sG[1][ly][lx] = (sW[1][ly + 1][lx] - sW[1][ly - 1][lx]) * g.idy2c;
            sG[3][ly][lx] = (sW[2][ly + 1][lx] - sW[2][ly - 1][lx]) * g.idy2c;
        }
        __syncthreads();
    }
    const double mu = g.mu;
    // x-faces i+1/2 for i = i0-1 .. i0+TX-1 (local fx = i - i0 + 1), rows j0 .. j0+TY-1
    for (int f = tid; f < (TXN + 1) * TYN; f += 256) {
        const int fx = f % (TXN + 1), fy = f / (TXN + 1);
        const int ly = fy + 2, lL = fx + 1, lR = fx + 2;
        const double4 WL = make_double4(sW[0][ly][lL], sW[1][ly][lL], sW[2][ly][lL], sW[3][ly][lL]);
        const double4 WR = make_double4(sW[0][ly][lR], sW[1][ly][lR], sW[2][ly][lR], sW[3][ly][lR]);
        double F[4];
        face_hllc_ns<MIXED>(g, WL, WR, false, F);
        if (visc && mu > 0.0) {  // NS:430-455
            const double dudx = 0.5 * (sG[0][ly][lL] + sG[0][ly][lR]);
            const double dudy = 0.5 * (sG[1][ly][lL] + sG[1][ly][lR]);
            const double dvdx = 0.5 * (sG[2][ly][lL] + sG[2][ly][lR]);
            const double dvdy = 0.5 * (sG[3][ly][lL] + sG[3][ly][lR]);
            const double divv = dudx + dvdy;
            const double tau_xx = mu * (2.0 * dudx - (2.0 / 3.0) * divv);
            const double tau_xy = mu * (dudy + dvdx);
            const double uf = 0.5 * (WL.y + WR.y), vf = 0.5 * (WL.z + WR.z);
            F[1] += -tau_xx;
            F[2] += -tau_xy;
            F[3] += -(uf * tau_xx + vf * tau_xy);
        }
        sFx[0][fy][fx] = F[0]; sFx[1][fy][fx] = F[1]; sFx[2][fy][fx] = F[2]; sFx[3][fy][fx] = F[3];
    }
    // y-faces j+1/2 for j = j0-1 .. j0+TY-1, columns i0 .. i0+TX-1
    for (int f = tid; f < TXN * (TYN + 1); f += 256) {
        const int fx = f % TXN, fy = f / TXN;
        const int lx = fx + 2, lD = fy + 1, lU = fy + 2;
        const double4 WL = make_double4(sW[0][lD][lx], sW[1][lD][lx], sW[2][lD][lx], sW[3][lD][lx]);
        const double4 WR = make_double4(sW[0][lU][lx], sW[1][lU][lx], sW[2][lU][lx], sW[3][lU][lx]);
        double F[4];
        face_hllc_ns<MIXED>(g, WL, WR, true, F);
        // back to (f_rho, f_mx, f_my, f_E)
        const double fmx = F[2], fmy = F[1];
        F[1] = fmx;
        F[2] = fmy;
        if (visc && mu > 0.0) {  // NS:457-481
            const double dudx = 0.5 * (sG[0][lD][lx] + sG[0][lU][lx]);
            const double dudy = 0.5 * (sG[1][lD][lx] + sG[1][lU][lx]);
            const double dvdx = 0.5 * (sG[2][lD][lx] + sG[2][lU][lx]);
            const double dvdy = 0.5 * (sG[3][lD][lx] + sG[3][lU][lx]);
            const double divv = dudx + dvdy;
            const double tau_yy = mu * (2.0 * dvdy - (2.0 / 3.0) * divv);
            const double tau_xy = mu * (dudy + dvdx);
            const double uf = 0.5 * (WL.y + WR.y), vf = 0.5 * (WL.z + WR.z);
            F[1] += -tau_xy;
            F[2] += -tau_yy;
            F[3] += -(uf * tau_xy + vf * tau_yy);
        }
        sFy[0][fy][fx] = F[0]; sFy[1][fy][fx] = F[1]; sFy[2][fy][fx] = F[2]; sFy[3][fy][fx] = F[3];
    }
    __syncthreads();
    // flux divergence + gravity (NS:488-515) and the stage update
    const int tx = tid % TXN, ty = tid / TXN;
    const int i = i0 + tx, j = j0 + ty;
    double smax_v = 0.0, rmin_v = 1e300;
    if (i < Nx && j < Ny) {
        const size_t q = (size_t)j * Nx + i;
        double R[4];
#pragma unroll
        for (int c = 0; c < 4; ++c)
            R[c] = -(sFx[c][ty][tx + 1] - sFx[c][ty][tx]) * g.idx1 - (sFy[c][ty + 1][tx] - sFy[c][ty][tx]) * g.idy1;
        if (g.gravity_on) {  // g = -central grad phi (NS:337-360)
            const int ip = i + 1 == Nx ? 0 : i + 1, im = i == 0 ? Nx - 1 : i - 1;
            const int jp = j + 1 == Ny ? 0 : j + 1, jm = j == 0 ? Ny - 1 : j - 1;
            const double gxq = -(phi[(size_t)j * Nx + ip] - phi[(size_t)j * Nx + im]) * g.idx2c;
            const double gyq = -(phi[(size_t)jp * Nx + i] - phi[(size_t)jm * Nx + i]) * g.idy2c;
            const double d = sW[0][ty + 2][tx + 2], u = sW[1][ty + 2][tx + 2], v = sW[2][ty + 2][tx + 2];
            R[1] += d * gxq;
            R[2] += d * gyq;
            R[3] += d * (u * gxq + v * gyq);
        }
        if (MODE == 0) {
            out[q] = make_double4(R[0], R[1], R[2], R[3]);
        } else if (MODE == 1) {  // NS:766-773
            const double4 u0 = Uin[q];
            K[q] = make_double4(R[0], R[1], R[2], R[3]);
            double4 v = make_double4(u0.x + dt * R[0], u0.y + dt * R[1], u0.z + dt * R[2], u0.w + dt * R[3]);
            if (v.x < g.floor_rho) v.x = g.floor_rho;
            out[q] = v;
        } else {  // NS:777-783
            const double4 un = Uold[q], k1 = K[q];
            const double h = 0.5 * dt;
            double4 v = make_double4(un.x + h * (k1.x + R[0]), un.y + h * (k1.y + R[1]), un.z + h * (k1.z + R[2]),
                                     un.w + h * (k1.w + R[3]));
            if (v.x < g.floor_rho) v.x = g.floor_rho;
            out[q] = v;
            const double4 W = prim_ns(g, v);  // compute_dt of the new state (NS:527-533)
            const double a = sound_ns(g, W);
            smax_v = smax(fabs(W.y) + a, fabs(W.z) + a);
            rmin_v = W.x;
        }
    }
    if (MODE == 2) {
        __shared__ double red[2][4];
        smax_v = wave_max(smax_v);
#pragma unroll
        for (int off = 32; off >= 1; off >>= 1) rmin_v = smin(rmin_v, __shfl_xor(rmin_v, off, 64));
        if ((tid & 63) == 0) { red[0][tid >> 6] = smax_v; red[1][tid >> 6] = rmin_v; }
        __syncthreads();
        if (tid == 0) {
            const size_t b = (size_t)blockIdx.y * gridDim.x + blockIdx.x;
            smaxp[b] = smax(smax(red[0][0], red[0][1]), smax(red[0][2], red[0][3]));
            rminp[b] = smin(smin(red[1][0], red[1][1]), smin(red[1][2], red[1][3]));
        }
    }
}

__global__ __launch_bounds__(256) void k_ns_dt_partial(GeomNS g, const double4* __restrict__ U,
                                                       double* __restrict__ smaxp, double* __restrict__ rminp) {
    __shared__ double red[2][4];
    const int q = blockIdx.x * 256 + threadIdx.x;
    double s = 0.0, r = 1e300;
    if (q < g.Nx * g.Ny) {
        const double4 W = prim_ns(g, U[q]);
        const double a = sound_ns(g, W);
        s = smax(fabs(W.y) + a, fabs(W.z) + a);
        r = W.x;
    }
    s = wave_max(s);
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) r = smin(r, __shfl_xor(r, off, 64));
    if ((threadIdx.x & 63) == 0) { red[0][threadIdx.x >> 6] = s; red[1][threadIdx.x >> 6] = r; }
    __syncthreads();
    if (threadIdx.x == 0) {
        smaxp[blockIdx.x] = smax(smax(red[0][0], red[0][1]), smax(red[0][2], red[0][3]));
        rminp[blockIdx.x] = smin(smin(red[1][0], red[1][1]), smin(red[1][2], red[1][3]));
    }
}

// dt of NS:535-545 from the partial maxima / minima
__global__ __launch_bounds__(256) void k_ns_dt(GeomNS g, const double* __restrict__ smaxp,
                                               const double* __restrict__ rminp, int np, double* __restrict__ dt_out) {
    __shared__ double red[2][4];
    double s = 0.0, r = 1e300;
    for (int k = threadIdx.x; k < np; k += 256) { s = smax(s, smaxp[k]); r = smin(r, rminp[k]); }
    s = wave_max(s);
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) r = smin(r, __shfl_xor(r, off, 64));
    if ((threadIdx.x & 63) == 0) { red[0][threadIdx.x >> 6] = s; red[1][threadIdx.x >> 6] = r; }
    __syncthreads();
    if (threadIdx.x == 0) {
        const double sm = smax(smax(red[0][0], red[0][1]), smax(red[0][2], red[0][3]));
        const double rm = smin(smin(red[1][0], red[1][1]), smin(red[1][2], red[1][3]));
        const double h = smin(g.dx, g.dy);
        const double dt_hyp = (sm > 0.0) ? (g.cfl * h / sm) : 1e-6;
        double dt = dt_hyp;
        if (g.visc && g.mu > 0.0) {
            const double nu_max = g.mu / clamp_min(rm, g.floor_rho);
            const double dt_visc = (nu_max > 0.0) ? (g.cfl_visc * h * h / nu_max) : dt_hyp;
            dt = smin(dt, dt_visc);
        }
        dt_out[0] = dt;
    }
}

// The AoS state as four planes {rho, mx, my, E} (for the totals).
__global__ __launch_bounds__(256) void k_ns_unpack(int N, const double4* __restrict__ U, double* __restrict__ out) {
    const int q = blockIdx.x * 256 + threadIdx.x;
    if (q >= N) return;
    const double4 u = U[q];
    out[q] = u.x;
    out[(size_t)N + q] = u.y;
    out[2 * (size_t)N + q] = u.z;
    out[3 * (size_t)N + q] = u.w;
}

// Snapshot frame fields except Phi (the store lambda of NS:698-749): {rho, P, Mach, u, v, momx, momy}.
__global__ __launch_bounds__(256) void k_ns_frame(GeomNS g, const double4* __restrict__ U, double* __restrict__ out) {
    const int q = blockIdx.x * 256 + threadIdx.x;
    const int N = g.Nx * g.Ny;
    if (q >= N) return;
    const double4 W = prim_ns(g, U[q]);
    const double a = sound_ns(g, W);
    const double vmag = sqrt(W.y * W.y + W.z * W.z);
    out[q] = W.x;
    out[(size_t)N + q] = W.w;
    out[2 * (size_t)N + q] = (a > 0.0) ? (vmag / a) : 0.0;
    out[3 * (size_t)N + q] = W.y;
    out[4 * (size_t)N + q] = W.z;
    out[5 * (size_t)N + q] = W.x * W.y;
    out[6 * (size_t)N + q] = W.x * W.z;
}

}  // namespace cc

using namespace cc;

struct cc_ns2d {
    cc_ns2d_params p{};
    GeomNS g{};
    hipStream_t stream = nullptr;
    double4 *U = nullptr, *U1 = nullptr, *K = nullptr, *T = nullptr;
    double *rho = nullptr, *phi = nullptr, *A = nullptr, *part = nullptr, *mean = nullptr;
    double *smaxp = nullptr, *rminp = nullptr, *dtd = nullptr, *gx = nullptr, *gy = nullptr;
    int nb_sum = 0, np_dt = 0;
    Fft2D fft;
    FftGreen green{};
    Profiler prof;
    FrameBuffer frame;
    FieldStats stats;
    std::vector<void*> allocs;

    size_t N() const { return (size_t)p.Nx * p.Ny; }
    Profiler* pr() { return prof.on ? &prof : nullptr; }
    dim3 tiles() const { return dim3((p.Nx + TXN - 1) / TXN, (p.Ny + TYN - 1) / TYN); }

    cc_status alloc(void** ptr, size_t bytes) {
        if (hipMalloc(ptr, bytes) != hipSuccess) {
            set_error(CC_ERR_OOM, "cc_ns2d: device allocation failed");
            return CC_ERR_OOM;
        }
        allocs.push_back(*ptr);
        CC_HIP(hipMemsetAsync(*ptr, 0, bytes, stream));
        return CC_OK;
    }
    // phi from the density plane `rho` (solve_poisson_periodic_fft, NS:285-335)
    cc_status poisson_from_plane() {
        ProfScope ps(pr(), stream, "fft");
        const int N = (int)this->N();
        hipLaunchKernelGGL(k_sum_partial, dim3(nb_sum), dim3(256), 0, stream, N, rho, part);
        CC_CHECK_LAUNCH();
        hipLaunchKernelGGL(k_ns_mean, dim3(1), dim3(256), 0, stream, nb_sum, N, part, mean);
        CC_CHECK_LAUNCH();
        CC_TRY(fft.pass(false, false, FFT_IN_REAL_MINUS_MEAN, FFT_OUT_COMPLEX, A, rho, mean, nullptr, nullptr, stream));
        CC_TRY(fft.pass(true, false, FFT_IN_COMPLEX, FFT_OUT_GREEN, A, nullptr, nullptr, nullptr, &green, stream));
        CC_TRY(fft.pass(false, true, FFT_IN_COMPLEX, FFT_OUT_COMPLEX, A, nullptr, nullptr, nullptr, nullptr, stream));
        CC_TRY(fft.pass(true, true, FFT_IN_COMPLEX, FFT_OUT_REAL, A, nullptr, nullptr, phi, nullptr, stream));
        return CC_OK;
    }
    cc_status poisson_of(const double4* s) {
        if (!p.gravity_on) return CC_OK;
        {
            ProfScope ps(pr(), stream, "fft");
            hipLaunchKernelGGL(k_ns_rho, dim3(nb_sum), dim3(256), 0, stream, (int)N(), p.floor_rho, s, rho, part);
            CC_CHECK_LAUNCH();
        }
        return poisson_from_plane();
    }
    // The snapshot frame of the current state, Phi from the FFT solve of its density (zero without
    // gravity, NS:741-745).  The solve leaves phi as the next step's first solve would.
    cc_status make_frame() {
        const size_t n = N();
        CC_TRY(frame.ensure((size_t)CC_NS2D_FRAME_FIELDS * n));
        hipLaunchKernelGGL(k_ns_frame, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, stream, g, U, frame.d);
        CC_CHECK_LAUNCH();
        double* phi_out = frame.d + 7 * n;
        if (p.gravity_on) {
            CC_TRY(poisson_of(U));
            CC_HIP(hipMemcpyAsync(phi_out, phi, n * sizeof(double), hipMemcpyDeviceToDevice, stream));
        } else {
            CC_HIP(hipMemsetAsync(phi_out, 0, n * sizeof(double), stream));
        }
        return CC_OK;
    }
    template <int MODE>
    cc_status stage(const double4* in, const double4* old, double4* out, double dt) {
        ProfScope ps(pr(), stream, "flux");
        if (p.precision == CC_MIXED)
            hipLaunchKernelGGL((k_ns_stage<MODE, true>), tiles(), dim3(256), 0, stream, g, in, old, K, out, phi, dt,
                               smaxp, rminp);
        else
            hipLaunchKernelGGL((k_ns_stage<MODE, false>), tiles(), dim3(256), 0, stream, g, in, old, K, out, phi, dt,
                               smaxp, rminp);
        CC_CHECK_LAUNCH();
        return CC_OK;
    }
    // Heun update of NS:758-781 with the given dt
    cc_status heun(double dt) {
        CC_TRY(poisson_of(U));
        CC_TRY(stage<1>(U, U, U1, dt));
        CC_TRY(poisson_of(U1));
        CC_TRY(stage<2>(U1, U, U, dt));
        return CC_OK;
    }
    cc_status compute_dt(double* dt, bool fresh) {
        {
            ProfScope ps(pr(), stream, "dt");
            int np = np_dt;
            if (fresh) {
                np = (int)((N() + 255) / 256);
                hipLaunchKernelGGL(k_ns_dt_partial, dim3(np), dim3(256), 0, stream, g, U, smaxp, rminp);
                CC_CHECK_LAUNCH();
            }
            hipLaunchKernelGGL(k_ns_dt, dim3(1), dim3(256), 0, stream, g, smaxp, rminp, np, dtd);
            CC_CHECK_LAUNCH();
        }
        CC_HIP(hipMemcpyAsync(dt, dtd, sizeof(double), hipMemcpyDeviceToHost, stream));
        CC_HIP(hipStreamSynchronize(stream));
        return CC_OK;
    }
    ~cc_ns2d() {
        for (void* a : allocs) hipFree(a);
        fft.destroy();
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

cc_status cc_ns2d_create(const cc_ns2d_params* p, cc_ns2d** out) {
    if (!out || !p) { set_error(CC_ERR_INVALID_ARG, "null params or output handle"); return CC_ERR_INVALID_ARG; }
    *out = nullptr;
    auto pow2 = [](int n) { return n > 0 && (n & (n - 1)) == 0; };
    if (!pow2(p->Nx) || !pow2(p->Ny)) {  // NS:614-617
        set_error(CC_ERR_NON_POW2, "Nx and Ny must be powers of two for the FFT Poisson solver.");
        return CC_ERR_NON_POW2;
    }
    if (p->Nx < 2 || p->Ny < 2 || p->Nx > kFftMaxLen || p->Ny > kFftMaxLen || !(p->x1 > p->x0) ||
        !(p->y1 > p->y0) || !(p->gamma > 1.0)) {
        set_error(CC_ERR_INVALID_ARG, "cc_ns2d: Nx, Ny in [2, 4096], x1 > x0, y1 > y0, gamma > 1 required");
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
    cc_ns2d* h = new cc_ns2d();
    h->p = *p;
    GeomNS& g = h->g;
    g.Nx = p->Nx; g.Ny = p->Ny; g.visc = p->viscosity_on != 0; g.gravity_on = p->gravity_on != 0;
    g.dx = (p->x1 - p->x0) / p->Nx; g.dy = (p->y1 - p->y0) / p->Ny;  // Grid2D NS:95-111
    g.gamma = p->gamma; g.gm1 = p->gamma - 1.0; g.floor_rho = p->floor_rho; g.floor_P = p->floor_P;
    g.mu = p->mu; g.cfl = p->cfl; g.cfl_visc = p->cfl_visc;
    g.idx1 = 1.0 / g.dx; g.idy1 = 1.0 / g.dy; g.idx2c = 1.0 / (2.0 * g.dx); g.idy2c = 1.0 / (2.0 * g.dy);
    h->green.neg4piG = -(4.0 * M_PI * p->G);
    h->green.two_pi = 2.0 * M_PI;
    h->green.Lx = p->x1 - p->x0;
    h->green.Ly = p->y1 - p->y0;
    const size_t N = h->N();
    h->nb_sum = (int)std::min<size_t>((N + 255) / 256, 1024);
    const dim3 t = h->tiles();
    h->np_dt = (int)(t.x * t.y);
    const size_t npart = std::max<size_t>((N + 255) / 256, (size_t)h->np_dt);
    cc_status st = CC_OK;
    if (hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking) != hipSuccess) {
        set_error(CC_ERR_HIP, "hipStreamCreate failed");
        delete h;
        return CC_ERR_HIP;
    }
    if (st == CC_OK) st = h->alloc((void**)&h->U, N * 32);
    if (st == CC_OK) st = h->alloc((void**)&h->U1, N * 32);
    if (st == CC_OK) st = h->alloc((void**)&h->K, N * 32);
    if (st == CC_OK) st = h->alloc((void**)&h->T, N * 32);
    if (st == CC_OK) st = h->alloc((void**)&h->rho, N * 8);
    if (st == CC_OK) st = h->alloc((void**)&h->phi, N * 8);
    if (st == CC_OK) st = h->alloc((void**)&h->A, N * 16);
    if (st == CC_OK) st = h->alloc((void**)&h->part, 1024 * 8);
    if (st == CC_OK) st = h->alloc((void**)&h->mean, 8);
    if (st == CC_OK) st = h->alloc((void**)&h->smaxp, npart * 8);
    if (st == CC_OK) st = h->alloc((void**)&h->rminp, npart * 8);
    if (st == CC_OK) st = h->alloc((void**)&h->dtd, 8);
    if (st == CC_OK) st = h->fft.init(p->Nx, p->Ny);
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

cc_status cc_ns2d_destroy(cc_ns2d* h) {
    if (!h) return CC_OK;
    hipStreamSynchronize(h->stream);
    delete h;
    return CC_OK;
}

cc_status cc_ns2d_upload(cc_ns2d* h, const double* U) {
    CHECK_H();
    if (!U) { set_error(CC_ERR_INVALID_ARG, "null state"); return CC_ERR_INVALID_ARG; }
    CC_HIP(hipMemcpyAsync(h->U, U, h->N() * 32, hipMemcpyHostToDevice, h->stream));
    CC_HIP(hipStreamSynchronize(h->stream));
    return CC_OK;
}

cc_status cc_ns2d_download(cc_ns2d* h, double* U) {
    CHECK_H();
    if (!U) { set_error(CC_ERR_INVALID_ARG, "null state"); return CC_ERR_INVALID_ARG; }
    CC_HIP(hipMemcpyAsync(U, h->U, h->N() * 32, hipMemcpyDeviceToHost, h->stream));
    CC_HIP(hipStreamSynchronize(h->stream));
    return CC_OK;
}

cc_status cc_ns2d_poisson(cc_ns2d* h, const double* rho, double* phi) {
    CHECK_H();
    if (!rho || !phi) { set_error(CC_ERR_INVALID_ARG, "null rho or phi"); return CC_ERR_INVALID_ARG; }
    CC_HIP(hipMemcpyAsync(h->rho, rho, h->N() * 8, hipMemcpyHostToDevice, h->stream));
    CC_TRY(h->poisson_from_plane());
    CC_HIP(hipMemcpyAsync(phi, h->phi, h->N() * 8, hipMemcpyDeviceToHost, h->stream));
    CC_HIP(hipStreamSynchronize(h->stream));
    return CC_OK;
}

cc_status cc_ns2d_rhs(cc_ns2d* h, double* dU) {
    CHECK_H();
    if (!dU) { set_error(CC_ERR_INVALID_ARG, "null dU"); return CC_ERR_INVALID_ARG; }
    CC_TRY(h->poisson_of(h->U));
    CC_TRY(h->stage<0>(h->U, h->U, h->T, 0.0));
    CC_HIP(hipMemcpyAsync(dU, h->T, h->N() * 32, hipMemcpyDeviceToHost, h->stream));
    CC_HIP(hipStreamSynchronize(h->stream));
    return CC_OK;
}

cc_status cc_ns2d_compute_dt(cc_ns2d* h, double* dt) {
    CHECK_H();
    if (!dt) { set_error(CC_ERR_INVALID_ARG, "null dt"); return CC_ERR_INVALID_ARG; }
    return h->compute_dt(dt, true);
}

cc_status cc_ns2d_heun_step(cc_ns2d* h, double dt) {
    CHECK_H();
    return h->heun(dt);
}

cc_status cc_ns2d_run(cc_ns2d* h, int nsteps, double* t_inout, double t_end, double* dts) {
    CHECK_H();
    if (!t_inout || nsteps < 0) { set_error(CC_ERR_INVALID_ARG, "null t or negative nsteps"); return CC_ERR_INVALID_ARG; }
    double t = *t_inout;
    bool fresh = true;  // after a Heun step the stage-2 kernel has left the dt partials of the new state
    for (int s = 0; s < nsteps && t < t_end - 1e-14; ++s) {  // NS:758-785
        double dt = 0.0;
        CC_TRY(h->compute_dt(&dt, fresh));
        if (t + dt > t_end) dt = t_end - t;
        CC_TRY(h->heun(dt));
        fresh = false;
        t += dt;
        if (dts) dts[s] = dt;
    }
    CC_HIP(hipStreamSynchronize(h->stream));
    *t_inout = t;
    return CC_OK;
}

cc_status cc_ns2d_frame(cc_ns2d* h, double* out) {
    CHECK_H();
    if (!out) { set_error(CC_ERR_INVALID_ARG, "null output"); return CC_ERR_INVALID_ARG; }
    CC_TRY(h->make_frame());
    CC_HIP(hipMemcpyAsync(out, h->frame.d, sizeof(double) * CC_NS2D_FRAME_FIELDS * h->N(), hipMemcpyDeviceToHost,
                          h->stream));
    CC_HIP(hipStreamSynchronize(h->stream));
    return CC_OK;
}

cc_status cc_ns2d_totals(cc_ns2d* h, double* out4) {
    CHECK_H();
    if (!out4) { set_error(CC_ERR_INVALID_ARG, "null output"); return CC_ERR_INVALID_ARG; }
    // the AoS state is unpacked into the first four planes of the frame buffer and summed there
    const size_t n = h->N();
    CC_TRY(h->frame.ensure((size_t)CC_NS2D_FRAME_FIELDS * n));
    hipLaunchKernelGGL(k_ns_unpack, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, h->stream, (int)n, h->U, h->frame.d);
    CC_CHECK_LAUNCH();
    FieldPtrs a{{h->frame.d, h->frame.d + n, h->frame.d + 2 * n, h->frame.d + 3 * n}};
    double r[12];
    CC_TRY(h->stats.run(a, nullptr, 4, n, h->stream, r));
    for (int f = 0; f < 4; f++) out4[f] = r[3 * f] * (h->g.dx * h->g.dy);
    return CC_OK;
}

cc_status cc_ns2d_compare(cc_ns2d* a, cc_ns2d* b, double* out) {
    if (!a || !b || !out) { set_error(CC_ERR_INVALID_ARG, "null handle or output"); return CC_ERR_INVALID_ARG; }
    if (a->p.Nx != b->p.Nx || a->p.Ny != b->p.Ny) {
        set_error(CC_ERR_INVALID_ARG, "cc_ns2d_compare: grids differ");
        return CC_ERR_INVALID_ARG;
    }
    CC_TRY(a->make_frame());
    CC_TRY(b->make_frame());
    CC_HIP(hipStreamSynchronize(b->stream));
    return compare_frames(a->stats, a->frame.d, b->frame.d, CC_NS2D_FRAME_FIELDS, a->N(), a->stream, out);
}

cc_status cc_ns2d_sync(cc_ns2d* h) {
    CHECK_H();
    CC_HIP(hipStreamSynchronize(h->stream));
    return CC_OK;
}

cc_status cc_ns2d_set_profiling(cc_ns2d* h, int on) {
    CHECK_H();
    CC_HIP(hipStreamSynchronize(h->stream));
    h->prof.on = on != 0;
    return CC_OK;
}

cc_status cc_ns2d_kernel_time(cc_ns2d* h, const char* kernel, int* count, double* total_ms) {
    CHECK_H();
    if (!kernel || !count || !total_ms) { set_error(CC_ERR_INVALID_ARG, "null argument"); return CC_ERR_INVALID_ARG; }
    CC_HIP(hipStreamSynchronize(h->stream));
    h->prof.get_total(kernel, count, total_ms);
    return CC_OK;
}

cc_status cc_ns2d_reset_profiling(cc_ns2d* h) {
    CHECK_H();
    CC_HIP(hipStreamSynchronize(h->stream));
    h->prof.reset();
    return CC_OK;
}

}  // extern "C"
