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

template <typename T> __device__ __forceinline__ T cadd(T a, T b) { return T{a.x + b.x, a.y + b.y}; }
template <typename T> __device__ __forceinline__ T csub(T a, T b) { return T{a.x - b.x, a.y - b.y}; }
template <typename T> __device__ __forceinline__ T cmul(T a, T b) { return T{a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x}; }
template <typename T> __device__ __forceinline__ T cmi(T a) { return T{a.y, -a.x}; }  // a * (-i)
template <typename T> __device__ __forceinline__ T cconj(T a) { return T{a.x, -a.y}; }

struct P3Args {
    int Nx, Ny, Nz, lx, ly, lz;  // sizes and their log2
    int nh;                      // Nx/2 + 1: the x frequencies kept (the spectrum of real data is Hermitian)
    double idx2, idy2, idz2, c4piG, floor_rho, tol, inv_n;
    int max_iter;
};

constexpr int kThreads = 256;
constexpr int kStatBlocks = 2048;

// Forward DFTs of 2, 4 and 8 points in registers (natural order in and out).
template <typename T> __device__ __forceinline__ void dft2(T* y) {
    const T a = y[0];
    y[0] = cadd(a, y[1]);
    y[1] = csub(a, y[1]);
}
template <typename T> __device__ __forceinline__ void dft4(T* y) {
    const T a0 = cadd(y[0], y[2]), a1 = csub(y[0], y[2]), a2 = cadd(y[1], y[3]), a3 = cmi(csub(y[1], y[3]));
    y[0] = cadd(a0, a2); y[2] = csub(a0, a2);
    y[1] = cadd(a1, a3); y[3] = csub(a1, a3);
}
template <typename T> __device__ __forceinline__ void dft8(T* y) {
    using R = decltype(y[0].x);
    T e[4] = {y[0], y[2], y[4], y[6]}, o[4] = {y[1], y[3], y[5], y[7]};
    dft4(e);
    dft4(o);
    const R h = (R)0.70710678118654752440;  // X_k = E_k + w^k O_k, X_{k+4} = E_k - w^k O_k, w = exp(-i pi/4)
    const T o1 = T{(o[1].x + o[1].y) * h, (o[1].y - o[1].x) * h};
    const T o2 = cmi(o[2]);
    const T o3 = T{(o[3].y - o[3].x) * h, (-o[3].x - o[3].y) * h};
    y[0] = cadd(e[0], o[0]); y[4] = csub(e[0], o[0]);
    y[1] = cadd(e[1], o1); y[5] = csub(e[1], o1);
    y[2] = cadd(e[2], o2); y[6] = csub(e[2], o2);
    y[3] = cadd(e[3], o3); y[7] = csub(e[3], o3);
}

// One Stockham stage of radix 2^LR on a line of n = 2^logn points held 8 per thread: thread j owns
// elements j + m*T (T = n/8 = 2^lt threads per line, m = 0..7) and runs the butterflies
// b = j + u*T (u < 8/R) on inputs b + m*n/R (= its registers u + m*8/R) with twiddles
// exp(-2 pi i m k / (Ns R)), k = b mod Ns, from the n-entry table tw.  Outputs go to
// (b / Ns) Ns R + k + m Ns: through the LDS line sl (element p at sl[p * ps]), or -- in the last
// stage, where that is the thread's own register u + m*8/R -- straight back into x.
template <typename T, int LR>
__device__ __forceinline__ void stock_stage(T (&x)[8], int j, int lt, int lns, int logn, const T* __restrict__ tw,
                                            T* sl, int ps, bool last) {
    constexpr int R = 1 << LR, U = 8 / R;
    const int Ns = 1 << lns;
#pragma unroll
    for (int u = 0; u < U; ++u) {
        const int b = j + (u << lt);
        const int k = b & (Ns - 1);
        T y[R];
#pragma unroll
        for (int m = 0; m < R; ++m) y[m] = x[u + m * U];
        if (lns > 0) {
#pragma unroll
            for (int m = 1; m < R; ++m) y[m] = cmul(y[m], tw[(m * k) << (logn - lns - LR)]);
        }
        if (R == 8) dft8(y);
        else if (R == 4) dft4(y);
        else dft2(y);
        if (last) {
#pragma unroll
            for (int m = 0; m < R; ++m) x[u + m * U] = y[m];
        } else {
            const int base = ((b >> lns) << (lns + LR)) + k;
#pragma unroll
            for (int m = 0; m < R; ++m) sl[(base + (m << lns)) * ps] = y[m];
        }
    }
}

// Forward FFT of the line whose elements j + m*T the thread holds in x (natural order in and out):
// radix-8 stages, then one radix-2 or radix-4 stage for log2 n mod 3.  Every thread of the block
// must call it (it synchronises the block between stages).
template <typename T>
__device__ __forceinline__ void fft_line(T (&x)[8], int j, int lt, int logn, const T* __restrict__ tw, T* sl, int ps) {
    const int n8 = logn / 3, rem = logn % 3;
    int lns = 0;
    for (int st = 0; st < n8; ++st) {
        const bool last = st == n8 - 1 && rem == 0;
        stock_stage<T, 3>(x, j, lt, lns, logn, tw, sl, ps, last);
        lns += 3;
        if (!last) {
            __syncthreads();
#pragma unroll
            for (int m = 0; m < 8; ++m) x[m] = sl[(j + (m << lt)) * ps];
            __syncthreads();
        }
    }
    if (rem == 1) stock_stage<T, 1>(x, j, lt, lns, logn, tw, sl, ps, true);
    else if (rem == 2) stock_stage<T, 2>(x, j, lt, lns, logn, tw, sl, ps, true);
}

// Maximum over the block (blockDim a multiple of 64, at most 1024) into part[blockIdx.x].
__device__ __forceinline__ void block_max_any(double v, double* part) {
    __shared__ double red[16];
    v = wave_max(v);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
        double m = red[0];
        for (int w = 1; w < (int)(blockDim.x >> 6); ++w) m = smax(m, red[w]);
        part[blockIdx.x] = m;
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

// ---- forward x pass with the residual: r = rhs - L_h phi (FP64), max|r| per block, x-FFT of r.
// Block: 2^lnl lines of Nx points, Nx/8 threads per line.
template <typename R>
__global__ __launch_bounds__(1024) void k_p3_xfwd(P3Args a, int lnl, const double* __restrict__ rho,
                                                  const double* __restrict__ phi, typename Cx<R>::T* __restrict__ buf,
                                                  const typename Cx<R>::T* __restrict__ tw, double* __restrict__ part,
                                                  const Poisson3DStatus* __restrict__ st) {
    using T = typename Cx<R>::T;
    if (st->done) return;
    extern __shared__ unsigned char smem_raw[];
    T* s = reinterpret_cast<T*>(smem_raw);
    const int Nx = a.Nx, Ny = a.Ny, Nz = a.Nz, lt = a.lx - 3;
    const int j = threadIdx.x & ((1 << lt) - 1), l = threadIdx.x >> lt;
    const size_t line = ((size_t)blockIdx.x << lnl) + l;
    const int jy = (int)(line & (size_t)(Ny - 1)), kz = (int)(line >> a.ly);
    const size_t sxy = (size_t)Nx * Ny, row = line * Nx;
    const size_t rn = (size_t)kz * sxy + (size_t)((jy + 1) & (Ny - 1)) * Nx;
    const size_t rs = (size_t)kz * sxy + (size_t)((jy - 1) & (Ny - 1)) * Nx;
    const size_t ru = (size_t)((kz + 1) & (Nz - 1)) * sxy + (size_t)jy * Nx;
    const size_t rd = (size_t)((kz - 1) & (Nz - 1)) * sxy + (size_t)jy * Nx;
    const double rhobar = st->rhobar;
    double rmax = 0.0;
    T x[8];
#pragma unroll
    for (int m = 0; m < 8; ++m) {
        const int i = j + (m << lt);
        const size_t q = row + i;
        const double p0 = phi[q];
        const double lap = (phi[row + ((i + 1) & (Nx - 1))] + phi[row + ((i - 1) & (Nx - 1))] - 2.0 * p0) * a.idx2 +
                           (phi[rn + i] + phi[rs + i] - 2.0 * p0) * a.idy2 + (phi[ru + i] + phi[rd + i] - 2.0 * p0) * a.idz2;
        const double r = a.c4piG * (smax(rho[q], a.floor_rho) - rhobar) - lap;
        rmax = smax(rmax, fabs(r));
        x[m] = T{(R)r, (R)0};
    }
    fft_line(x, j, lt, a.lx, tw, s + ((size_t)l << a.lx), 1);
    // the line is real, so X(Nx - kx) = conj X(kx): keep kx = 0 .. Nx/2 (elements m < 4, and m = 4 of j = 0)
    const size_t rowh = line * (size_t)a.nh;
#pragma unroll
    for (int m = 0; m < 5; ++m)
        if (m < 4 || j == 0) buf[rowh + j + (m << lt)] = x[m];
    block_max_any(rmax, part);
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
            st->iters_total += it;
            st->solves += 1;
            st->iters_max = it > st->iters_max ? it : st->iters_max;
            if (!conv) st->unconverged += 1;
            st->resid_worst = smax(st->resid_worst, rel);
        } else {
            st->iters = it + 1;
        }
    }
}

// y pass: 2^lc consecutive x columns of one z plane per block, Ny/8 threads per column.
template <typename R, bool INV>
__global__ __launch_bounds__(1024) void k_p3_y(P3Args a, int lc, typename Cx<R>::T* __restrict__ buf,
                                               const typename Cx<R>::T* __restrict__ tw,
                                               const Poisson3DStatus* __restrict__ st) {
    using T = typename Cx<R>::T;
    if (st->done) return;
    extern __shared__ unsigned char smem_raw[];
    T* s = reinterpret_cast<T*>(smem_raw);
    const int nh = a.nh, lt = a.ly - 3, C = 1 << lc, nbc = (nh + C - 1) >> lc;
    const int c = threadIdx.x & (C - 1), j = threadIdx.x >> lc;
    const int xb = blockIdx.x % nbc, k = blockIdx.x / nbc;
    // columns past nh (the last block of a plane) transform a copy of the last column and store nothing:
    // every thread takes part in the block's barriers
    const int col = (xb << lc) + c;
    const bool valid = col < nh;
    T* base = buf + (size_t)k * nh * a.Ny + (valid ? col : nh - 1);
    T x[8];
#pragma unroll
    for (int m = 0; m < 8; ++m) {
        const T v = base[(size_t)(j + (m << lt)) * nh];
        x[m] = INV ? cconj(v) : v;
    }
    fft_line(x, j, lt, a.ly, tw, s + c, C);
    if (valid) {
#pragma unroll
        for (int m = 0; m < 8; ++m) base[(size_t)(j + (m << lt)) * nh] = INV ? cconj(x[m]) : x[m];
    }
}

// z pass: forward transform, spectral division by the discrete Laplacian (and 1/N), inverse
// transform (conj . forward . conj), for 2^lc consecutive x columns of one y row per block.
template <typename R>
__global__ __launch_bounds__(1024) void k_p3_z(P3Args a, int lc, typename Cx<R>::T* __restrict__ buf,
                                               const typename Cx<R>::T* __restrict__ tw, const double* __restrict__ lam,
                                               const Poisson3DStatus* __restrict__ st) {
    using T = typename Cx<R>::T;
    if (st->done) return;
    extern __shared__ unsigned char smem_raw[];
    T* s = reinterpret_cast<T*>(smem_raw);
    const int Nx = a.Nx, nh = a.nh, lt = a.lz - 3, C = 1 << lc, nbc = (nh + C - 1) >> lc;
    const int c = threadIdx.x & (C - 1), j = threadIdx.x >> lc;
    const int xb = blockIdx.x % nbc, jy = blockIdx.x / nbc;
    const int col = (xb << lc) + c;
    const bool valid = col < nh;  // as in the y pass
    const size_t zs = (size_t)nh * a.Ny;
    T* base = buf + (size_t)jy * nh + (valid ? col : nh - 1);
    T x[8];
#pragma unroll
    for (int m = 0; m < 8; ++m) x[m] = base[(size_t)(j + (m << lt)) * zs];
    fft_line(x, j, lt, a.lz, tw, s + c, C);
    const int kx = valid ? col : nh - 1;
    const double lxy = lam[kx] + lam[Nx + jy];
#pragma unroll
    for (int m = 0; m < 8; ++m) {
        const int kz = j + (m << lt);
        const double l = lxy + lam[Nx + a.Ny + kz];
        const R f = (kx == 0 && jy == 0 && kz == 0) ? (R)0 : (R)(a.inv_n / l);
        x[m] = T{x[m].x * f, -(x[m].y * f)};  // scaled and conjugated for the inverse
    }
    fft_line(x, j, lt, a.lz, tw, s + c, C);
    if (valid) {
#pragma unroll
        for (int m = 0; m < 8; ++m) base[(size_t)(j + (m << lt)) * zs] = cconj(x[m]);
    }
}

// inverse x pass and the FP64 correction phi += Re
template <typename R>
__global__ __launch_bounds__(1024) void k_p3_xinv(P3Args a, int lnl, const typename Cx<R>::T* __restrict__ buf,
                                                  const typename Cx<R>::T* __restrict__ tw, double* __restrict__ phi,
                                                  const Poisson3DStatus* __restrict__ st) {
    using T = typename Cx<R>::T;
    if (st->done) return;
    extern __shared__ unsigned char smem_raw[];
    T* s = reinterpret_cast<T*>(smem_raw);
    const int lt = a.lx - 3, Nx = a.Nx;
    const int j = threadIdx.x & ((1 << lt) - 1), l = threadIdx.x >> lt;
    const size_t line = ((size_t)blockIdx.x << lnl) + l, row = line * Nx, rowh = line * (size_t)a.nh;
    T x[8];
#pragma unroll
    for (int m = 0; m < 8; ++m) {  // kx > Nx/2 from the Hermitian symmetry X(kx) = conj X(Nx - kx)
        const int p = j + (m << lt);
        x[m] = p <= (Nx >> 1) ? cconj(buf[rowh + p]) : buf[rowh + (Nx - p)];
    }
    fft_line(x, j, lt, a.lx, tw, s + ((size_t)l << a.lx), 1);
#pragma unroll
    for (int m = 0; m < 8; ++m) phi[row + j + (m << lt)] += (double)x[m].x;  // Re(conj(.)) = Re(.)
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
    int lnl_x, lc_y, lc_z, nblk_x, nblk_y, nblk_z, thr_x, thr_y, thr_z;
    size_t lds_x, lds_y, lds_z;
};

// x pass: 2^lnl lines of Nx/8 threads (<= 256 threads); y / z passes: 2^lc columns of N/8 threads,
// as many columns as keep the block <= 1024 threads, its LDS <= 64 KB and the columns <= 16.
Geo geometry(int Nx, int Ny, int Nz, bool mixed) {
    const int tsz = mixed ? 8 : 16;
    Geo g;
    const int lx = ilog2(Nx), ly = ilog2(Ny), lz = ilog2(Nz);
    g.lnl_x = std::min(8 - (lx - 3), ly + lz);
    auto cols = [&](int ln) {
        int lc = std::min(lx, 4);
        while (lc > 0 && ((1 << lc) * (1 << (ln - 3)) > 1024 || ((size_t)1 << lc) * ((size_t)1 << ln) * tsz > 65536)) --lc;
        return lc;
    };
    g.lc_y = cols(ly);
    g.lc_z = cols(lz);
    g.thr_x = (1 << g.lnl_x) << (lx - 3);
    g.thr_y = (1 << g.lc_y) << (ly - 3);
    g.thr_z = (1 << g.lc_z) << (lz - 3);
    g.nblk_x = (Ny * Nz) >> g.lnl_x;
    const int nh = Nx / 2 + 1;
    g.nblk_y = ((nh + (1 << g.lc_y) - 1) >> g.lc_y) * Nz;
    g.nblk_z = ((nh + (1 << g.lc_z) - 1) >> g.lc_z) * Ny;
    g.lds_x = ((size_t)Nx << g.lnl_x) * tsz;
    g.lds_y = ((size_t)Ny << g.lc_y) * tsz;
    g.lds_z = ((size_t)Nz << g.lc_z) * tsz;
    return g;
}

P3Args make_args(const Poisson3D& P) {
    P3Args a;
    a.Nx = P.Nx; a.Ny = P.Ny; a.Nz = P.Nz;
    a.lx = ilog2(P.Nx); a.ly = ilog2(P.Ny); a.lz = ilog2(P.Nz);
    a.nh = P.Nx / 2 + 1;
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
            hipLaunchKernelGGL((k_p3_xfwd<R>), dim3(g.nblk_x), dim3(g.thr_x), g.lds_x, s, a, g.lnl_x, rho, P.phi, buf,
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
            hipLaunchKernelGGL((k_p3_y<R, false>), dim3(g.nblk_y), dim3(g.thr_y), g.lds_y, s, a, g.lc_y, buf, twy,
                               P.st);
            CC_CHECK_LAUNCH();
        }
        {
            ProfScope ps(prof, s, "fft_z");
            hipLaunchKernelGGL((k_p3_z<R>), dim3(g.nblk_z), dim3(g.thr_z), g.lds_z, s, a, g.lc_z, buf, twz, P.lam,
                               P.st);
            CC_CHECK_LAUNCH();
        }
        {
            ProfScope ps(prof, s, "fft_y");
            hipLaunchKernelGGL((k_p3_y<R, true>), dim3(g.nblk_y), dim3(g.thr_y), g.lds_y, s, a, g.lc_y, buf, twy,
                               P.st);
            CC_CHECK_LAUNCH();
        }
        {
            ProfScope ps(prof, s, "fft_x");
            hipLaunchKernelGGL((k_p3_xinv<R>), dim3(g.nblk_x), dim3(g.thr_x), g.lds_x, s, a, g.lnl_x, buf, twx,
                               P.phi, P.st);
            CC_CHECK_LAUNCH();
        }
    }
    if (gx) {  // without g arrays the caller takes g from phi itself
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
    const size_t nspec = (size_t)(Nx / 2 + 1) * Ny * Nz;  // the half spectrum along x
    bool ok = hipMalloc(&phi, n * sizeof(double)) == hipSuccess && hipMalloc(&buf, nspec * tsz) == hipSuccess &&
              hipMalloc(&part, (size_t)npart * sizeof(double)) == hipSuccess &&
              hipMalloc(&st, sizeof(Poisson3DStatus)) == hipSuccess &&
              hipMalloc(&lam, (size_t)(Nx + Ny + Nz) * sizeof(double)) == hipSuccess;
    const int dims[3] = {Nx, Ny, Nz};
    for (int ax = 0; ax < 3 && ok; ++ax) ok = hipMalloc(&tw[ax], (size_t)dims[ax] * tsz) == hipSuccess;
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
        std::vector<double> t(2 * na);  // (re, im) pairs of exp(-2 pi i m / N), m < N
        for (int m = 0; m < na; ++m) {
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
