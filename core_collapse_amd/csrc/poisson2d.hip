// poisson2d.hip -- SOR Poisson gravity kernels (see poisson2d.hpp for the schedule).
#include <algorithm>
#include <cmath>

#include "poisson2d.hpp"

namespace cc {

static constexpr int kReduceBlocks = 512;
static constexpr int kTileW = 64;  // tile width along i = one wave64 (lane = column)
static constexpr int kTileWJ = 64; // tile height along j
static constexpr int kTileS = 16;  // sweeps per tile = waves per workgroup

// ------------------------------------------------------------------ reductions
__global__ __launch_bounds__(256) void k_partial_sum(const double* __restrict__ x, size_t n,
                                                     double* __restrict__ partial) {
    __shared__ double red[4];
    double acc = 0.0;
    for (size_t q = (size_t)blockIdx.x * 256 + threadIdx.x; q < n; q += (size_t)gridDim.x * 256) acc += x[q];
    acc = wave_sum(acc);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = acc;
    __syncthreads();
    if (threadIdx.x == 0) partial[blockIdx.x] = (red[0] + red[1]) + (red[2] + red[3]);
}

__global__ __launch_bounds__(256) void k_final_mean(const double* __restrict__ partial, int nb, double inv_n,
                                                    double* __restrict__ out) {
    __shared__ double red[4];
    double acc = 0.0;
    for (int b = threadIdx.x; b < nb; b += 256) acc += partial[b];
    acc = wave_sum(acc);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = acc;
    __syncthreads();
    if (threadIdx.x == 0) out[0] = ((red[0] + red[1]) + (red[2] + red[3])) * inv_n;
}

cc_status device_mean(const double* x, size_t n, double* partial, double* out, hipStream_t s) {
    hipLaunchKernelGGL(k_partial_sum, dim3(kReduceBlocks), dim3(256), 0, s, x, n, partial);
    CC_CHECK_LAUNCH();
    hipLaunchKernelGGL(k_final_mean, dim3(1), dim3(256), 0, s, partial, kReduceBlocks, 1.0 / (double)n, out);
    CC_CHECK_LAUNCH();
    return CC_OK;
}

// ------------------------------------------------------------------ tiled lexicographic SOR
// Cell (i,j) belongs to tile (TS,TJ,TI) at sweep s when s in [S*TS, S*TS+S), i+s in
// [W*TI, W*TI+W), j+s in [WJ*TJ, WJ*TJ+WJ).  Within a tile, wave w runs sweep S*TS+w and lane l
// walks column i = W*TI - s + l upward: at local step d it updates row j = WJ*TJ - s + (d - l).
// All updates of one step are on distinct anti-diagonals (one colour of the checkerboard) and
// every value they read was produced at step d-1 or earlier, or by an earlier tile level.
template <int S>
struct TileGeom {
    static constexpr int BI = kTileW + S + 1;           // box width (union of the S shifted squares + halo)
    static constexpr int BJ = kTileWJ + S + 1;
    static constexpr int P = (BI % 2 == 0) ? BI : BI + 1;  // even pitch: lanes of a diagonal hit distinct banks
};

// Does the tile update cell (i,j) at some sweep s in [s0,s1)?  (i in [W*TI-s, W*TI-s+W) <=> s in
// [W*TI-i, W*TI-i+W-1], likewise for j.)
__device__ __forceinline__ bool tile_updates(int i, int j, int s0, int s1, int TI, int TJ) {
    const int lo = max(s0, max(kTileW * TI - i, kTileWJ * TJ - j));
    const int hi = min(s1 - 1, min(kTileW * TI - i + kTileW - 1, kTileWJ * TJ - j + kTileWJ - 1));
    return lo <= hi;
}

template <int S>
__global__ __launch_bounds__(64 * S) void k_sor_tile_fp64(SorArgs a, const int* __restrict__ tiles,
                                                           double* __restrict__ phi, const double* __restrict__ rho,
                                                           const double* __restrict__ rhobar) {
    using G = TileGeom<S>;
    __shared__ double sphi[G::BJ * G::P];
    __shared__ double srhs[G::BJ * G::P];
    const int TS = tiles[3 * blockIdx.x], TJ = tiles[3 * blockIdx.x + 1], TI = tiles[3 * blockIdx.x + 2];
    const int s0 = S * TS, s1 = min(s0 + S, a.nsw);
    const int ib0 = max(0, kTileW * TI - (s1 - 1) - 1), ib1 = min(a.Nx - 1, kTileW * TI - s0 + kTileW);
    const int jb0 = max(0, kTileWJ * TJ - (s1 - 1) - 1), jb1 = min(a.Ny - 1, kTileWJ * TJ - s0 + kTileWJ);
    const int ni = ib1 - ib0 + 1, nj = jb1 - jb0 + 1;
    const double rb = rhobar[0];
    for (int c = threadIdx.x; c < ni * nj; c += blockDim.x) {
        const int li = c % ni, lj = c / ni;
        const size_t g = (size_t)(jb0 + lj) * a.Nx + (ib0 + li);
        sphi[lj * G::P + li] = phi[g];
        srhs[lj * G::P + li] = a.c4piG * (rho[g] - rb);
    }
    __syncthreads();
    const int lane = threadIdx.x & 63;
    const int s = s0 + (threadIdx.x >> 6);
    const int ci = kTileW * TI - s + lane;
    const bool col_ok = (s < s1) && ci >= 0 && ci < a.Nx;
    const int li = ci - ib0;
    const int liE = min(ci + 1, a.Nx - 1) - ib0, liW = max(ci - 1, 0) - ib0;
    for (int d = 0; d < kTileW + kTileWJ - 1; ++d) {
        const int lj = d - lane;
        const int cj = kTileWJ * TJ - s + lj;
        if (col_ok && lj >= 0 && lj < kTileWJ && cj >= 0 && cj < a.Ny) {
            const int r = cj - jb0;
            const int rN = min(cj + 1, a.Ny - 1) - jb0, rS = max(cj - 1, 0) - jb0;
            const double pc = sphi[r * G::P + li];
            const double pnew = ((sphi[r * G::P + liE] + sphi[r * G::P + liW]) * a.idx2 +
                                 (sphi[rN * G::P + li] + sphi[rS * G::P + li]) * a.idy2 - srhs[r * G::P + li]) *
                                a.inv_denom;
            sphi[r * G::P + li] = pc + a.omega * (pnew - pc);
        }
        __syncthreads();
    }
    for (int c = threadIdx.x; c < ni * nj; c += blockDim.x) {
        const int li2 = c % ni, lj2 = c / ni;
        const int i = ib0 + li2, j = jb0 + lj2;
        if (tile_updates(i, j, s0, s1, TI, TJ)) phi[(size_t)j * a.Nx + i] = sphi[lj2 * G::P + li2];
    }
}

// Mixed precision: phi = phi_hi (FP64, fixed for the tile) + delta (FP32 in LDS).  With
// c = (stencil(phi_hi) - rhs)/denom - phi_hi evaluated once in FP64, one SOR update is
// delta += omega * (c + kx (dE + dW) + ky (dN + dS) - delta), exactly the FP64 update rearranged.
template <int S>
__global__ __launch_bounds__(64 * S) void k_sor_tile_mixed(SorArgs a, const int* __restrict__ tiles,
                                                            double* __restrict__ phi, const double* __restrict__ rho,
                                                            const double* __restrict__ rhobar) {
    using G = TileGeom<S>;
    __shared__ float sd[G::BJ * G::P];
    __shared__ float sc[G::BJ * G::P];
    const int TS = tiles[3 * blockIdx.x], TJ = tiles[3 * blockIdx.x + 1], TI = tiles[3 * blockIdx.x + 2];
    const int s0 = S * TS, s1 = min(s0 + S, a.nsw);
    const int ib0 = max(0, kTileW * TI - (s1 - 1) - 1), ib1 = min(a.Nx - 1, kTileW * TI - s0 + kTileW);
    const int jb0 = max(0, kTileWJ * TJ - (s1 - 1) - 1), jb1 = min(a.Ny - 1, kTileWJ * TJ - s0 + kTileWJ);
    const int ni = ib1 - ib0 + 1, nj = jb1 - jb0 + 1;
    const double rb = rhobar[0];
    for (int c = threadIdx.x; c < ni * nj; c += blockDim.x) {
        const int li = c % ni, lj = c / ni;
        const int i = ib0 + li, j = jb0 + lj;
        sd[lj * G::P + li] = 0.f;
        float cv = 0.f;
        if (tile_updates(i, j, s0, s1, TI, TJ)) {
            const size_t g = (size_t)j * a.Nx + i;
            const double pc = phi[g];
            const double pE = phi[(size_t)j * a.Nx + min(i + 1, a.Nx - 1)];
            const double pW = phi[(size_t)j * a.Nx + max(i - 1, 0)];
            const double pN = phi[(size_t)min(j + 1, a.Ny - 1) * a.Nx + i];
            const double pS = phi[(size_t)max(j - 1, 0) * a.Nx + i];
            const double rhs = a.c4piG * (rho[g] - rb);
            cv = (float)(((pE + pW) * a.idx2 + (pN + pS) * a.idy2 - rhs) * a.inv_denom - pc);
        }
        sc[lj * G::P + li] = cv;
    }
    __syncthreads();
    const int lane = threadIdx.x & 63;
    const int s = s0 + (threadIdx.x >> 6);
    const int ci = kTileW * TI - s + lane;
    const bool col_ok = (s < s1) && ci >= 0 && ci < a.Nx;
    const int li = ci - ib0;
    const int liE = min(ci + 1, a.Nx - 1) - ib0, liW = max(ci - 1, 0) - ib0;
    for (int d = 0; d < kTileW + kTileWJ - 1; ++d) {
        const int lj = d - lane;
        const int cj = kTileWJ * TJ - s + lj;
        if (col_ok && lj >= 0 && lj < kTileWJ && cj >= 0 && cj < a.Ny) {
            const int r = cj - jb0;
            const int rN = min(cj + 1, a.Ny - 1) - jb0, rS = max(cj - 1, 0) - jb0;
            const float dc = sd[r * G::P + li];
            const float gs = sc[r * G::P + li] + (sd[r * G::P + liE] + sd[r * G::P + liW]) * a.kxf +
                             (sd[rN * G::P + li] + sd[rS * G::P + li]) * a.kyf;
            sd[r * G::P + li] = dc + a.omegaf * (gs - dc);
        }
        __syncthreads();
    }
    for (int c = threadIdx.x; c < ni * nj; c += blockDim.x) {
        const int li2 = c % ni, lj2 = c / ni;
        const int i = ib0 + li2, j = jb0 + lj2;
        if (tile_updates(i, j, s0, s1, TI, TJ)) {
            const size_t g = (size_t)j * a.Nx + i;
            phi[g] = phi[g] + (double)sd[lj2 * G::P + li2];
        }
    }
}

// ------------------------------------------------------------------ periodic / mixed-BC SOR
// One workgroup; sweeps in order; within a sweep the anti-diagonals i+j = d in increasing d,
// which is the lexicographic order's dependency order including the wrap-around reads
// ((0,j) reads the old (N-1,j); (N-1,j) reads the new (0,j)).
__global__ __launch_bounds__(1024) void k_sor_wavefront_1wg(SorArgs a, double* __restrict__ phi,
                                                            const double* __restrict__ rho,
                                                            const double* __restrict__ rhobar) {
    const double rb = rhobar[0];
    const int Nx = a.Nx, Ny = a.Ny;
    for (int it = 0; it < a.nsw; ++it) {
        for (int d = 0; d < Nx + Ny - 1; ++d) {
            const int ilo = max(0, d - (Ny - 1)), ihi = min(d, Nx - 1);
            for (int i = ilo + (int)threadIdx.x; i <= ihi; i += (int)blockDim.x) {
                const int j = d - i;
                const int iE = bmap(i + 1, Nx, a.px), iW = bmap(i - 1, Nx, a.px);
                const int jN = bmap(j + 1, Ny, a.py), jS = bmap(j - 1, Ny, a.py);
                const size_t k = (size_t)j * Nx + i;
                const double pc = phi[k];
                const double rhs = a.c4piG * (rho[k] - rb);
                const double pnew = ((phi[(size_t)j * Nx + iE] + phi[(size_t)j * Nx + iW]) * a.idx2 +
                                     (phi[(size_t)jN * Nx + i] + phi[(size_t)jS * Nx + i]) * a.idy2 - rhs) *
                                    a.inv_denom;
                phi[k] = pc + a.omega * (pnew - pc);
            }
            __syncthreads();
        }
    }
}

// ------------------------------------------------------------------ gradient, misc
__global__ void k_grad2d(int Nx, int Ny, int px, int py, double dx, double dy, const double* __restrict__ phi,
                         double* __restrict__ gx, double* __restrict__ gy) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x, j = blockIdx.y;
    if (i >= Nx) return;
    const size_t k = (size_t)j * Nx + i;
    const double pE = phi[(size_t)j * Nx + bmap(i + 1, Nx, px)], pW = phi[(size_t)j * Nx + bmap(i - 1, Nx, px)];
    const double pN = phi[(size_t)bmap(j + 1, Ny, py) * Nx + i], pS = phi[(size_t)bmap(j - 1, Ny, py) * Nx + i];
    gx[k] = -(pE - pW) / (2.0 * dx);
    gy[k] = -(pN - pS) / (2.0 * dy);
}

__global__ void k_fill(double* __restrict__ x, size_t n, double v) {
    for (size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x; q < n; q += (size_t)gridDim.x * blockDim.x) x[q] = v;
}

__global__ void k_sub_mean(const double* __restrict__ x, size_t n, const double* __restrict__ mean,
                           double* __restrict__ out) {
    const double m = mean[0];
    for (size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x; q < n; q += (size_t)gridDim.x * blockDim.x)
        out[q] = x[q] - m;
}

// ------------------------------------------------------------------ host side
cc_status Poisson2D::init(int Nx_, int Ny_, bool px_, bool py_, double dx_, double dy_, double G_, int iters_,
                          double omega_, int precision_) {
    destroy();
    Nx = Nx_; Ny = Ny_; px = px_; py = py_; dx = dx_; dy = dy_; G = G_; iters = iters_; omega = omega_;
    precision = precision_;
    tiled_ = !px && !py;
    const size_t N = (size_t)Nx * Ny;
    if (hipMalloc(&phi, N * sizeof(double)) != hipSuccess || hipMalloc(&partial, kReduceBlocks * sizeof(double)) != hipSuccess ||
        hipMalloc(&rhobar, 4 * sizeof(double)) != hipSuccess) {
        set_error(CC_ERR_OOM, "Poisson2D: device allocation failed");
        return CC_ERR_OOM;
    }
    const double idx2 = 1.0 / (dx * dx), idy2 = 1.0 / (dy * dy);
    const double denom = 2.0 * (idx2 + idy2);
    args.Nx = Nx; args.Ny = Ny; args.nsw = iters; args.px = px; args.py = py;
    args.idx2 = idx2; args.idy2 = idy2; args.inv_denom = 1.0 / denom; args.omega = omega;
    args.c4piG = 4.0 * M_PI * G;
    args.kxf = (float)(idx2 / denom); args.kyf = (float)(idy2 / denom); args.omegaf = (float)omega;

    level_off_.assign(1, 0);
    if (tiled_ && iters > 0) {
        const int W = kTileW, WJ = kTileWJ, S = kTileS;
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
        std::vector<int> flat;
        for (auto& v : lv) {
            if (v.empty()) continue;
            flat.insert(flat.end(), v.begin(), v.end());
            level_off_.push_back((int)flat.size() / 3);
        }
        if (hipMalloc(&sched, flat.size() * sizeof(int)) != hipSuccess) {
            set_error(CC_ERR_OOM, "Poisson2D: schedule allocation failed");
            return CC_ERR_OOM;
        }
        CC_HIP(hipMemcpy(sched, flat.data(), flat.size() * sizeof(int), hipMemcpyHostToDevice));
    }
    return CC_OK;
}

void Poisson2D::destroy() {
    if (phi) hipFree(phi);
    if (partial) hipFree(partial);
    if (rhobar) hipFree(rhobar);
    if (sched) hipFree(sched);
    phi = partial = rhobar = nullptr;
    sched = nullptr;
    level_off_.clear();
}

cc_status Poisson2D::solve(const double* rho, double* gx, double* gy, hipStream_t s, Profiler* prof) {
    const size_t N = (size_t)Nx * Ny;
    {
        ProfScope ps(prof, s, "mean");
        CC_TRY(device_mean(rho, N, partial, rhobar, s));
    }
    hipLaunchKernelGGL(k_fill, dim3(std::min<size_t>((N + 255) / 256, 4096)), dim3(256), 0, s, phi, N, 0.0);
    CC_CHECK_LAUNCH();
    if (tiled_) {
        const int nl = levels();
        for (int l = 0; l < nl; ++l) {
            const int nt = level_off_[l + 1] - level_off_[l];
            ProfScope ps(prof, s, "sor");
            if (precision == CC_MIXED)
                hipLaunchKernelGGL(k_sor_tile_mixed<kTileS>, dim3(nt), dim3(64 * kTileS), 0, s, args,
                                   sched + 3 * level_off_[l], phi, rho, rhobar);
            else
                hipLaunchKernelGGL(k_sor_tile_fp64<kTileS>, dim3(nt), dim3(64 * kTileS), 0, s, args,
                                   sched + 3 * level_off_[l], phi, rho, rhobar);
            CC_CHECK_LAUNCH();
        }
    } else if (iters > 0) {
        ProfScope ps(prof, s, "sor");
        hipLaunchKernelGGL(k_sor_wavefront_1wg, dim3(1), dim3(1024), 0, s, args, phi, rho, rhobar);
        CC_CHECK_LAUNCH();
    }
    {
        ProfScope ps(prof, s, "grad");
        hipLaunchKernelGGL(k_grad2d, dim3((Nx + 255) / 256, Ny), dim3(256), 0, s, Nx, Ny, (int)px, (int)py, dx, dy,
                           phi, gx, gy);
        CC_CHECK_LAUNCH();
    }
    return CC_OK;
}

cc_status Poisson2D::phi_mean_removed(double* out_dev, hipStream_t s) {
    const size_t N = (size_t)Nx * Ny;
    CC_TRY(device_mean(phi, N, partial, rhobar + 1, s));
    hipLaunchKernelGGL(k_sub_mean, dim3(std::min<size_t>((N + 255) / 256, 4096)), dim3(256), 0, s, phi, N,
                       rhobar + 1, out_dev);
    CC_CHECK_LAUNCH();
    return CC_OK;
}

}  // namespace cc
