// poisson3d.hip -- periodic 3D FFT Poisson gravity with FP64 iterative refinement (see poisson3d.hpp).
#include <algorithm>
#include <cmath>
#include <type_traits>
#include <vector>

#include "poisson3d.hpp"
#include "stepcommon.hpp"

namespace cc {

namespace {

template <typename R> struct Cx;
template <> struct Cx<float> { using T = float2; };
template <> struct Cx<double> { using T = double2; };

template <typename T>
__device__ __forceinline__ T cmul(T a, T b) {
    return T{a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x};
}

__device__ __forceinline__ int brev(int v, int logn) { return (int)(__builtin_bitreverse32((unsigned)v) >> (32 - logn)); }

struct P3Args {
    int Nx, Ny, Nz, lx, ly, lz;  // sizes and their log2
    double idx2, idy2, idz2, c4piG, floor_rho, tol, inv_n;
    int max_iter;
};

constexpr int kThreads = 256;
constexpr int kStatBlocks = 512;

// In-place radix-2 transforms of 2^lognl lines of 2^logn points in LDS; point p of line l is at
// a[p * ps + l * ls].  LF: consecutive threads take consecutive lines (the ps > 1 layouts), else
// consecutive butterflies of one line.  Twiddle of span 2h, index j: tw[j * n / (2h)] (conjugated
// for the inverse).  DIT expects bit-reversed input and leaves natural order; DIF the reverse.
template <typename T, bool INV, bool LF>
__device__ __forceinline__ void lds_dit(T* a, int logn, int lognl, int ps, int ls, const T* __restrict__ tw) {
    const int nb = 1 << (logn - 1 + lognl);
    const int hmask = (1 << (logn - 1)) - 1;
    for (int s = 0; s < logn; ++s) {
        const int h = 1 << s, tstride = 1 << (logn - 1 - s);
        for (int b = threadIdx.x; b < nb; b += kThreads) {
            const int l = LF ? (b & ((1 << lognl) - 1)) : (b >> (logn - 1));
            const int k = LF ? (b >> lognl) : (b & hmask);
            const int j = k & (h - 1);
            const int i0 = ((k >> s) << (s + 1)) + j, i1 = i0 + h;
            T w = tw[j * tstride];
            if (INV) w.y = -w.y;
            T* p0 = a + i0 * ps + l * ls;
            T* p1 = a + i1 * ps + l * ls;
            const T x0 = *p0, x1 = cmul(*p1, w);
            *p0 = T{x0.x + x1.x, x0.y + x1.y};
            *p1 = T{x0.x - x1.x, x0.y - x1.y};
        }
        __syncthreads();
    }
}

template <typename T, bool INV, bool LF>
__device__ __forceinline__ void lds_dif(T* a, int logn, int lognl, int ps, int ls, const T* __restrict__ tw) {
    const int nb = 1 << (logn - 1 + lognl);
    const int hmask = (1 << (logn - 1)) - 1;
    for (int s = logn - 1; s >= 0; --s) {
        const int h = 1 << s, tstride = 1 << (logn - 1 - s);
        for (int b = threadIdx.x; b < nb; b += kThreads) {
            const int l = LF ? (b & ((1 << lognl) - 1)) : (b >> (logn - 1));
            const int k = LF ? (b >> lognl) : (b & hmask);
            const int j = k & (h - 1);
            const int i0 = ((k >> s) << (s + 1)) + j, i1 = i0 + h;
            T w = tw[j * tstride];
            if (INV) w.y = -w.y;
            T* p0 = a + i0 * ps + l * ls;
            T* p1 = a + i1 * ps + l * ls;
            const T x0 = *p0, x1 = *p1;
            *p0 = T{x0.x + x1.x, x0.y + x1.y};
            *p1 = cmul(T{x0.x - x1.x, x0.y - x1.y}, w);
        }
        __syncthreads();
    }
}

// ---- density statistics: per block sum, max, min of max(rho, floor) (fixed partition and order)
__global__ __launch_bounds__(kThreads) void k_p3_stats(const double* __restrict__ rho, size_t n, double floor_rho,
                                                       double* __restrict__ part) {
    __shared__ double rs[4], rx[4], rn[4];
    double s = 0.0, mx = 0.0, mn = 1e300;
    for (size_t q = (size_t)blockIdx.x * kThreads + threadIdx.x; q < n; q += (size_t)gridDim.x * kThreads) {
        const double r = smax(rho[q], floor_rho);
        s += r;
        mx = smax(mx, r);
        mn = smin(mn, r);
    }
    s = wave_sum(s);
    mx = wave_max(mx);
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) mn = smin(mn, __shfl_xor(mn, off, 64));
    const int w = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) { rs[w] = s; rx[w] = mx; rn[w] = mn; }
    __syncthreads();
    if (threadIdx.x == 0) {
        part[blockIdx.x] = (rs[0] + rs[1]) + (rs[2] + rs[3]);
        part[kStatBlocks + blockIdx.x] = smax(smax(rx[0], rx[1]), smax(rx[2], rx[3]));
        part[2 * kStatBlocks + blockIdx.x] = smin(smin(rn[0], rn[1]), smin(rn[2], rn[3]));
    }
}

__global__ __launch_bounds__(kThreads) void k_p3_stats_final(P3Args a, const double* __restrict__ part,
                                                             Poisson3DStatus* st) {
    __shared__ double rs[4], rx[4], rn[4];
    double s = 0.0, mx = 0.0, mn = 1e300;
    for (int b = threadIdx.x; b < kStatBlocks; b += kThreads) {
        s += part[b];
        mx = smax(mx, part[kStatBlocks + b]);
        mn = smin(mn, part[2 * kStatBlocks + b]);
    }
    s = wave_sum(s);
    mx = wave_max(mx);
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) mn = smin(mn, __shfl_xor(mn, off, 64));
    const int w = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) { rs[w] = s; rx[w] = mx; rn[w] = mn; }
    __syncthreads();
    if (threadIdx.x == 0) {
        const double rhobar = ((rs[0] + rs[1]) + (rs[2] + rs[3])) * a.inv_n;
        const double hi = smax(smax(rx[0], rx[1]), smax(rx[2], rx[3]));
        const double lo = smin(smin(rn[0], rn[1]), smin(rn[2], rn[3]));
        st->rhobar = rhobar;
        st->rhsmax = a.c4piG * smax(hi - rhobar, rhobar - lo);
        st->iters = 0;
        st->done = 0;
    }
}

// A uniform density has rhs = 0: the solution is phi = 0 (also clears a warm start).
__global__ __launch_bounds__(kThreads) void k_p3_flat(double* __restrict__ phi, size_t n,
                                                      const Poisson3DStatus* __restrict__ st) {
    if (st->rhsmax > 0.0) return;
    for (size_t q = (size_t)blockIdx.x * kThreads + threadIdx.x; q < n; q += (size_t)gridDim.x * kThreads) phi[q] = 0.0;
}

// ---- forward x pass with the residual: r = rhs - L_h phi (FP64), max|r| per block, x-FFT of r
template <typename R>
__global__ __launch_bounds__(kThreads) void k_p3_xfwd(P3Args a, int lognl, const double* __restrict__ rho,
                                                      const double* __restrict__ phi, typename Cx<R>::T* __restrict__ buf,
                                                      const typename Cx<R>::T* __restrict__ tw,
                                                      double* __restrict__ part, const Poisson3DStatus* __restrict__ st) {
    using T = typename Cx<R>::T;
    if (st->done) return;
    extern __shared__ unsigned char smem_raw[];
    T* s = reinterpret_cast<T*>(smem_raw);
    const int Nx = a.Nx, Ny = a.Ny, Nz = a.Nz;
    const size_t line0 = (size_t)blockIdx.x << lognl;
    const size_t sxy = (size_t)Nx * Ny;
    const double rhobar = st->rhobar;
    double rmax = 0.0;
    const int ne = Nx << lognl;
    for (int e = threadIdx.x; e < ne; e += kThreads) {
        const int l = e >> a.lx, i = e & (Nx - 1);
        const size_t line = line0 + l;
        const int j = (int)(line & (size_t)(Ny - 1)), k = (int)(line >> a.ly);
        const size_t q = line * Nx + i;
        const double p0 = phi[q];
        const double pe = phi[q - i + ((i + 1) & (Nx - 1))], pw = phi[q - i + ((i - 1) & (Nx - 1))];
        const size_t row = (size_t)k * sxy;
        const double pn = phi[row + (size_t)((j + 1) & (Ny - 1)) * Nx + i];
        const double ps = phi[row + (size_t)((j - 1) & (Ny - 1)) * Nx + i];
        const size_t col = (size_t)j * Nx + i;
        const double pu = phi[(size_t)((k + 1) & (Nz - 1)) * sxy + col];
        const double pd = phi[(size_t)((k - 1) & (Nz - 1)) * sxy + col];
        const double lap = (pe + pw - 2.0 * p0) * a.idx2 + (pn + ps - 2.0 * p0) * a.idy2 + (pu + pd - 2.0 * p0) * a.idz2;
        const double r = a.c4piG * (smax(rho[q], a.floor_rho) - rhobar) - lap;
        rmax = smax(rmax, fabs(r));
        s[(l << a.lx) + brev(i, a.lx)] = T{(R)r, (R)0};
    }
    __syncthreads();
    lds_dit<T, false, false>(s, a.lx, lognl, 1, Nx, tw);
    T* out = buf + line0 * Nx;
    for (int e = threadIdx.x; e < ne; e += kThreads) out[e] = s[e];
    block_max_to(rmax, part);
}

// The refinement check (one block): converged, or out of iterations -> done.
__global__ __launch_bounds__(kThreads) void k_p3_check(P3Args a, const double* __restrict__ part, int npart, int it,
                                                       Poisson3DStatus* st) {
    if (st->done) return;
    const double m = block_reduce_max(part, npart, 0.0);
    if (threadIdx.x == 0) {
        const double rel = st->rhsmax > 0.0 ? m / st->rhsmax : 0.0;
        st->resid = rel;
        const bool conv = m <= a.tol * st->rhsmax;
        if (conv || it >= a.max_iter) {
            st->done = 1;
            st->iters_last = it;
            st->solves += 1;
            st->iters_max = it > st->iters_max ? it : st->iters_max;
            if (!conv) st->unconverged += 1;
            st->resid_worst = smax(st->resid_worst, rel);
        } else {
            st->iters = it + 1;
        }
    }
}

// y pass: 2^logc consecutive x columns of one z plane per block
template <typename R, bool INV>
__global__ __launch_bounds__(kThreads) void k_p3_y(P3Args a, int logc, typename Cx<R>::T* __restrict__ buf,
                                                   const typename Cx<R>::T* __restrict__ tw,
                                                   const Poisson3DStatus* __restrict__ st) {
    using T = typename Cx<R>::T;
    if (st->done) return;
    extern __shared__ unsigned char smem_raw[];
    T* s = reinterpret_cast<T*>(smem_raw);
    const int C = 1 << logc, Nx = a.Nx;
    const int xb = blockIdx.x & ((Nx >> logc) - 1), k = blockIdx.x >> (a.lx - logc);
    T* base = buf + (size_t)k * Nx * a.Ny + (xb << logc);
    const int ne = a.Ny << logc;
    for (int e = threadIdx.x; e < ne; e += kThreads) {
        const int c = e & (C - 1), j = e >> logc;
        s[((INV ? j : brev(j, a.ly)) << logc) + c] = base[(size_t)j * Nx + c];
    }
    __syncthreads();
    if (INV)
        lds_dif<T, true, true>(s, a.ly, logc, C, 1, tw);
    else
        lds_dit<T, false, true>(s, a.ly, logc, C, 1, tw);
    for (int e = threadIdx.x; e < ne; e += kThreads) {
        const int c = e & (C - 1), j = e >> logc;
        base[(size_t)j * Nx + c] = s[((INV ? brev(j, a.ly) : j) << logc) + c];
    }
}

// z pass: forward transform, spectral division by the discrete Laplacian (and 1/N), inverse
// transform, for 2^logc consecutive x columns of one y row per block
template <typename R>
__global__ __launch_bounds__(kThreads) void k_p3_z(P3Args a, int logc, typename Cx<R>::T* __restrict__ buf,
                                                   const typename Cx<R>::T* __restrict__ tw, const double* __restrict__ lam,
                                                   const Poisson3DStatus* __restrict__ st) {
    using T = typename Cx<R>::T;
    if (st->done) return;
    extern __shared__ unsigned char smem_raw[];
    T* s = reinterpret_cast<T*>(smem_raw);
    const int C = 1 << logc, Nx = a.Nx;
    const int xb = blockIdx.x & ((Nx >> logc) - 1), j = blockIdx.x >> (a.lx - logc);
    const size_t zs = (size_t)Nx * a.Ny;
    T* base = buf + (size_t)j * Nx + (xb << logc);
    const int ne = a.Nz << logc;
    for (int e = threadIdx.x; e < ne; e += kThreads) {
        const int c = e & (C - 1), k = e >> logc;
        s[(brev(k, a.lz) << logc) + c] = base[(size_t)k * zs + c];
    }
    __syncthreads();
    lds_dit<T, false, true>(s, a.lz, logc, C, 1, tw);
    const double* lx = lam;
    const double* ly = lam + Nx;
    const double* lz = lam + Nx + a.Ny;
    for (int e = threadIdx.x; e < ne; e += kThreads) {
        const int c = e & (C - 1), k = e >> logc;
        const int kx = (xb << logc) + c;
        const double l = lx[kx] + ly[j] + lz[k];
        const double m = (kx == 0 && j == 0 && k == 0) ? 0.0 : a.inv_n / l;
        const T v = s[e];
        s[e] = T{(R)(v.x * (R)m), (R)(v.y * (R)m)};
    }
    __syncthreads();
    lds_dif<T, true, true>(s, a.lz, logc, C, 1, tw + (size_t)0);
    for (int e = threadIdx.x; e < ne; e += kThreads) {
        const int c = e & (C - 1), k = e >> logc;
        base[(size_t)k * zs + c] = s[(brev(k, a.lz) << logc) + c];
    }
}

// inverse x pass and the FP64 correction phi += Re
template <typename R>
__global__ __launch_bounds__(kThreads) void k_p3_xinv(P3Args a, int lognl, const typename Cx<R>::T* __restrict__ buf,
                                                      const typename Cx<R>::T* __restrict__ tw, double* __restrict__ phi,
                                                      const Poisson3DStatus* __restrict__ st) {
    using T = typename Cx<R>::T;
    if (st->done) return;
    extern __shared__ unsigned char smem_raw[];
    T* s = reinterpret_cast<T*>(smem_raw);
    const int Nx = a.Nx;
    const size_t off = ((size_t)blockIdx.x << lognl) * Nx;
    const int ne = Nx << lognl;
    for (int e = threadIdx.x; e < ne; e += kThreads) s[e] = buf[off + e];
    __syncthreads();
    lds_dif<T, true, false>(s, a.lx, lognl, 1, Nx, tw);
    for (int e = threadIdx.x; e < ne; e += kThreads) {
        const int l = e >> a.lx, i = e & (Nx - 1);
        phi[off + e] += (double)s[(l << a.lx) + brev(i, a.lx)].x;
    }
}

// g = -central grad phi
__global__ __launch_bounds__(kThreads) void k_p3_grad(P3Args a, const double* __restrict__ phi, double* __restrict__ gx,
                                                      double* __restrict__ gy, double* __restrict__ gz, double dx,
                                                      double dy, double dz) {
    const size_t n = (size_t)a.Nx * a.Ny * a.Nz;
    const size_t q = (size_t)blockIdx.x * kThreads + threadIdx.x;
    if (q >= n) return;
    const int Nx = a.Nx, Ny = a.Ny, Nz = a.Nz;
    const int i = (int)(q & (size_t)(Nx - 1));
    const size_t line = q >> a.lx;
    const int j = (int)(line & (size_t)(Ny - 1)), k = (int)(line >> a.ly);
    const size_t sxy = (size_t)Nx * Ny, rowq = q - i, col = (size_t)j * Nx + i;
    gx[q] = -(phi[rowq + ((i + 1) & (Nx - 1))] - phi[rowq + ((i - 1) & (Nx - 1))]) / (2.0 * dx);
    gy[q] = -(phi[(size_t)k * sxy + (size_t)((j + 1) & (Ny - 1)) * Nx + i] -
              phi[(size_t)k * sxy + (size_t)((j - 1) & (Ny - 1)) * Nx + i]) / (2.0 * dy);
    gz[q] = -(phi[(size_t)((k + 1) & (Nz - 1)) * sxy + col] - phi[(size_t)((k - 1) & (Nz - 1)) * sxy + col]) / (2.0 * dz);
}

int ilog2(int n) {
    int l = 0;
    while ((1 << l) < n) ++l;
    return l;
}

struct Geo {
    int lognl_x, logc_y, logc_z, nblk_x, nblk_y, nblk_z;
    size_t lds_x, lds_y, lds_z;
};

Geo geometry(int Nx, int Ny, int Nz, bool mixed) {
    const size_t tsz = mixed ? 8 : 16;
    const int ex = mixed ? 4096 : 2048;   // elements per x-pass block (32 KB)
    const int eyz = mixed ? 8192 : 4096;  // elements per y / z-pass block (64 KB)
    Geo g;
    const int lx = ilog2(Nx), ly = ilog2(Ny), lz = ilog2(Nz);
    g.lognl_x = std::max(0, std::min(ilog2(ex) - lx, ly + lz));
    g.logc_y = std::max(0, std::min(ilog2(eyz) - ly, lx));
    g.logc_z = std::max(0, std::min(ilog2(eyz) - lz, lx));
    g.nblk_x = (Ny * Nz) >> g.lognl_x;
    g.nblk_y = (Nx >> g.logc_y) * Nz;
    g.nblk_z = (Nx >> g.logc_z) * Ny;
    g.lds_x = ((size_t)Nx << g.lognl_x) * tsz;
    g.lds_y = ((size_t)Ny << g.logc_y) * tsz;
    g.lds_z = ((size_t)Nz << g.logc_z) * tsz;
    return g;
}

P3Args make_args(const Poisson3D& P) {
    P3Args a;
    a.Nx = P.Nx; a.Ny = P.Ny; a.Nz = P.Nz;
    a.lx = ilog2(P.Nx); a.ly = ilog2(P.Ny); a.lz = ilog2(P.Nz);
    a.idx2 = 1.0 / (P.dx * P.dx); a.idy2 = 1.0 / (P.dy * P.dy); a.idz2 = 1.0 / (P.dz * P.dz);
    a.c4piG = 4.0 * M_PI * P.G;
    a.floor_rho = P.floor_rho;
    a.tol = P.tol;
    a.inv_n = 1.0 / ((double)P.Nx * P.Ny * P.Nz);
    a.max_iter = P.max_iter;
    return a;
}

template <typename R>
cc_status solve_t(Poisson3D& P, const double* rho, double* gx, double* gy, double* gz, hipStream_t s, Profiler* prof) {
    using T = typename Cx<R>::T;
    const P3Args a = make_args(P);
    const Geo g = geometry(P.Nx, P.Ny, P.Nz, std::is_same<R, float>::value);
    const size_t n = (size_t)P.Nx * P.Ny * P.Nz;
    T* buf = static_cast<T*>(P.buf);
    const T* twx = static_cast<const T*>(P.tw[0]);
    const T* twy = static_cast<const T*>(P.tw[1]);
    const T* twz = static_cast<const T*>(P.tw[2]);
    {
        ProfScope ps(prof, s, "mean");
        hipLaunchKernelGGL(k_p3_stats, dim3(kStatBlocks), dim3(kThreads), 0, s, rho, n, P.floor_rho, P.part);
        CC_CHECK_LAUNCH();
        hipLaunchKernelGGL(k_p3_stats_final, dim3(1), dim3(kThreads), 0, s, a, P.part, P.st);
        CC_CHECK_LAUNCH();
        hipLaunchKernelGGL(k_p3_flat, dim3((unsigned)std::min<size_t>((n + kThreads - 1) / kThreads, 4096)),
                           dim3(kThreads), 0, s, P.phi, n, P.st);
        CC_CHECK_LAUNCH();
    }
    for (int it = 0; it <= P.max_iter; ++it) {
        {
            ProfScope ps(prof, s, "fft_x");
            hipLaunchKernelGGL((k_p3_xfwd<R>), dim3(g.nblk_x), dim3(kThreads), g.lds_x, s, a, g.lognl_x, rho, P.phi, buf,
                               twx, P.part, P.st);
            CC_CHECK_LAUNCH();
        }
        {
            ProfScope ps(prof, s, "check");
            hipLaunchKernelGGL(k_p3_check, dim3(1), dim3(kThreads), 0, s, a, P.part, g.nblk_x, it, P.st);
            CC_CHECK_LAUNCH();
        }
        if (it == P.max_iter) break;
        {
            ProfScope ps(prof, s, "fft_y");
            hipLaunchKernelGGL((k_p3_y<R, false>), dim3(g.nblk_y), dim3(kThreads), g.lds_y, s, a, g.logc_y, buf, twy,
                               P.st);
            CC_CHECK_LAUNCH();
        }
        {
            ProfScope ps(prof, s, "fft_z");
            hipLaunchKernelGGL((k_p3_z<R>), dim3(g.nblk_z), dim3(kThreads), g.lds_z, s, a, g.logc_z, buf, twz, P.lam,
                               P.st);
            CC_CHECK_LAUNCH();
        }
        {
            ProfScope ps(prof, s, "fft_y");
            hipLaunchKernelGGL((k_p3_y<R, true>), dim3(g.nblk_y), dim3(kThreads), g.lds_y, s, a, g.logc_y, buf, twy,
                               P.st);
            CC_CHECK_LAUNCH();
        }
        {
            ProfScope ps(prof, s, "fft_x");
            hipLaunchKernelGGL((k_p3_xinv<R>), dim3(g.nblk_x), dim3(kThreads), g.lds_x, s, a, g.lognl_x, buf, twx,
                               P.phi, P.st);
            CC_CHECK_LAUNCH();
        }
    }
    {
        ProfScope ps(prof, s, "grad");
        hipLaunchKernelGGL(k_p3_grad, dim3((unsigned)((n + kThreads - 1) / kThreads)), dim3(kThreads), 0, s, a, P.phi,
                           gx, gy, gz, P.dx, P.dy, P.dz);
        CC_CHECK_LAUNCH();
    }
    return CC_OK;
}

bool pow2_in(int n, int lo, int hi) { return n >= lo && n <= hi && (n & (n - 1)) == 0; }

}  // namespace

cc_status Poisson3D::init(int Nx_, int Ny_, int Nz_, double dx_, double dy_, double dz_, double G_, double floor_rho_,
                          double tol_, int max_iter_, int precision_, hipStream_t s) {
    destroy();
    if (!pow2_in(Nx_, 8, 1024) || !pow2_in(Ny_, 8, 1024) || !pow2_in(Nz_, 8, 1024)) {
        set_error(CC_ERR_NON_POW2, "Poisson3D: Nx, Ny, Nz must be powers of two in [8, 1024]");
        return CC_ERR_NON_POW2;
    }
    if (max_iter_ < 0 || max_iter_ > 64 || !(tol_ >= 0.0)) {
        set_error(CC_ERR_INVALID_ARG, "Poisson3D: max_iter must be in [0, 64] and tol >= 0");
        return CC_ERR_INVALID_ARG;
    }
    Nx = Nx_; Ny = Ny_; Nz = Nz_; dx = dx_; dy = dy_; dz = dz_; G = G_; floor_rho = floor_rho_; tol = tol_;
    max_iter = max_iter_; precision = precision_;
    const bool mixed = precision == CC_MIXED;
    const size_t n = (size_t)Nx * Ny * Nz, tsz = mixed ? 8 : 16;
    const Geo g = geometry(Nx, Ny, Nz, mixed);
    npart = std::max(g.nblk_x, 3 * kStatBlocks);
    bool ok = hipMalloc(&phi, n * sizeof(double)) == hipSuccess && hipMalloc(&buf, n * tsz) == hipSuccess &&
              hipMalloc(&part, (size_t)npart * sizeof(double)) == hipSuccess &&
              hipMalloc(&st, sizeof(Poisson3DStatus)) == hipSuccess &&
              hipMalloc(&lam, (size_t)(Nx + Ny + Nz) * sizeof(double)) == hipSuccess;
    const int dims[3] = {Nx, Ny, Nz};
    for (int ax = 0; ax < 3 && ok; ++ax) ok = hipMalloc(&tw[ax], (size_t)(dims[ax] / 2) * tsz) == hipSuccess;
    if (!ok) {
        destroy();
        set_error(CC_ERR_OOM, "Poisson3D: device allocation failed");
        return CC_ERR_OOM;
    }
    CC_HIP(hipMemsetAsync(phi, 0, n * sizeof(double), s));
    CC_HIP(hipMemsetAsync(st, 0, sizeof(Poisson3DStatus), s));
    std::vector<double> lh;
    const double hs[3] = {dx, dy, dz};
    for (int ax = 0; ax < 3; ++ax)
        for (int m = 0; m < dims[ax]; ++m) {
            const double sn = std::sin(M_PI * (double)m / (double)dims[ax]);
            lh.push_back(-4.0 / (hs[ax] * hs[ax]) * sn * sn);
        }
    CC_HIP(hipMemcpy(lam, lh.data(), lh.size() * sizeof(double), hipMemcpyHostToDevice));
    for (int ax = 0; ax < 3; ++ax) {
        const int na = dims[ax];
        std::vector<double> t(na);  // (re, im) pairs of exp(-2 pi i m / N)
        for (int m = 0; m < na / 2; ++m) {
            const double ang = -2.0 * M_PI * (double)m / (double)na;
            t[2 * m] = std::cos(ang);
            t[2 * m + 1] = std::sin(ang);
        }
        if (mixed) {
            std::vector<float> f(t.begin(), t.end());
            CC_HIP(hipMemcpy(tw[ax], f.data(), f.size() * sizeof(float), hipMemcpyHostToDevice));
        } else {
            CC_HIP(hipMemcpy(tw[ax], t.data(), t.size() * sizeof(double), hipMemcpyHostToDevice));
        }
    }
    return CC_OK;
}

void Poisson3D::destroy() {
    if (phi) hipFree(phi);
    if (buf) hipFree(buf);
    if (part) hipFree(part);
    if (st) hipFree(st);
    if (lam) hipFree(lam);
    for (auto& t : tw) {
        if (t) hipFree(t);
        t = nullptr;
    }
    phi = nullptr; buf = nullptr; part = nullptr; st = nullptr; lam = nullptr;
}

cc_status Poisson3D::solve(const double* rho, double* gx, double* gy, double* gz, hipStream_t s, Profiler* prof) {
    if (precision == CC_MIXED) return solve_t<float>(*this, rho, gx, gy, gz, s, prof);
    return solve_t<double>(*this, rho, gx, gy, gz, s, prof);
}

}  // namespace cc
