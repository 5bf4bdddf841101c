// poisson3d.hpp -- periodic 3D Poisson gravity by FFT with FP64 iterative refinement, on gfx950.
//
// Solves L_h phi = 4 pi G (max(rho, floor) - mean) for the 7-point discrete Laplacian L_h on a
// periodic Nx x Ny x Nz grid (each a power of two in [8, 1024]), then g = -central grad phi.
// One refinement iteration:
//   r = rhs - L_h phi                          FP64, fused into the forward x pass (also max|r|)
//   c = L_h^-1 r                               FFT in R (FP32 or FP64): forward x, forward y,
//                                              z forward * 1/(lambda_x + lambda_y + lambda_z) z inverse
//                                              (one kernel), inverse y, inverse x; r is real, so only
//                                              kx = 0 .. Nx/2 of the x transform is kept and passed on
//                                              (X(Nx - kx) = conj X(kx) rebuilds the rest in the inverse)
//   phi += c                                   FP64, fused into the inverse x pass
// where lambda_a(k) = -(4/h_a^2) sin^2(pi k/N_a) are the discrete Laplacian's eigenvalues, so one
// FP64 iteration is an exact solve and an FP32 one gains about seven digits.  Iterations stop
// when max|r| <= tol max|rhs| or after max_iter corrections: the check runs on the device and the
// remaining kernels of the solve return at once, so a solve never waits for the host and can be
// captured in a graph.  phi is kept between solves (warm start).
//
// Each FFT pass is one kernel over a batch of whole lines: every thread holds 8 points of a line
// in registers and runs Stockham radix-8 stages (one radix-2 or radix-4 stage for log2 N mod 3),
// exchanging points through LDS between stages; input and output are in natural order, inverse
// transforms are conj . forward . conj, and the twiddles exp(-2 pi i m / N) come from a per-axis
// table computed in FP64.
#pragma once
#include "common.hpp"
#include "profiler.hpp"

namespace cc {

// Device-side status of the solver (one per solver; the check kernel updates it).
struct Poisson3DStatus {
    double rhobar, rhsmax, resid;  // mean density, max|rhs|, max|r|/max|rhs| of the last check
    double resid_worst;            // worst final relative residual over all solves
    int iters, done;               // corrections applied in the current solve; solve finished
    int solves, iters_max;         // solves run; most corrections any solve needed
    int unconverged, iters_last;   // solves that stopped at max_iter; corrections of the last solve
    int iters_total, pad;          // corrections over all solves
};

class Poisson3D {
   public:
    cc_status init(int Nx, int Ny, int Nz, double dx, double dy, double dz, double G, double floor_rho, double tol,
                   int max_iter, int precision, hipStream_t s);
    void destroy();
    // phi(rho) and g = -central grad phi; rho, gx, gy, gz are device arrays of Nx*Ny*Nz cells (gx == nullptr:
    // phi only, for callers that difference phi themselves).
    cc_status solve(const double* rho, double* gx, double* gy, double* gz, hipStream_t s, Profiler* prof);
    ~Poisson3D() { destroy(); }

    int Nx = 0, Ny = 0, Nz = 0, max_iter = 0, precision = 0;
    double dx = 1, dy = 1, dz = 1, G = 1, floor_rho = 0, tol = 1e-10;
    double* phi = nullptr;        // N, FP64, mean zero
    void* buf = nullptr;          // (Nx/2+1) Ny Nz complex<R>: the half spectrum along x
    double* part = nullptr;       // block partials
    int npart = 0;
    Poisson3DStatus* st = nullptr;
    void* tw[3] = {nullptr, nullptr, nullptr};   // per axis N_a complex<R> twiddles
    double* lam = nullptr;                       // eigenvalues, Nx + Ny + Nz
};

}  // namespace cc
