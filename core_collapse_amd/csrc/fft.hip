// fft.hip -- batched radix-2 FFT lines in LDS; see fft.hpp.
#include <cmath>
#include <vector>

#include "fft.hpp"

namespace cc {

struct FftLaunch {
    int n, logn, lpb, nlines, Nx, Ny, cols, inverse, in, out;
    const double2* tw;
};

// One workgroup (256 threads) transforms lpb lines.  Element e of line l sits at q = line*Nx + e
// (rows) or q = e*Nx + line (columns; consecutive lines are adjacent in memory, so loads coalesce).
__global__ __launch_bounds__(256) void k_fft_lines(FftLaunch a, double2* __restrict__ data,
                                                   const double* __restrict__ in_real,
                                                   const double* __restrict__ mean, double* __restrict__ out_real,
                                                   FftGreen gr) {
    extern __shared__ double2 buf[];
    const int n = a.n, lpb = a.lpb, tid = threadIdx.x;
    const int line0 = blockIdx.x * lpb;
    const double mu = (a.in == FFT_IN_REAL_MINUS_MEAN) ? mean[0] : 0.0;
    // load in bit-reversed order (the swap loop of NS:153-159 is the bit-reversal permutation)
    for (int idx = tid; idx < lpb * n; idx += 256) {
        const int l = a.cols ? idx % lpb : idx / n, e = a.cols ? idx / lpb : idx % n;
        const int line = line0 + l;
        if (line >= a.nlines) continue;
        const size_t q = a.cols ? (size_t)e * a.Nx + line : (size_t)line * a.Nx + e;
        double2 v;
        if (a.in == FFT_IN_REAL_MINUS_MEAN) v = make_double2(in_real[q] - mu, 0.0);
        else v = data[q];
        const int r = n > 1 ? (int)(__brev((unsigned)e) >> (32 - a.logn)) : 0;
        buf[l * n + r] = v;
    }
    __syncthreads();
    // stages len = 2, 4, ..., n (NS:160-173)
    for (int len = 2; len <= n; len <<= 1) {
        const int half = len >> 1;
        for (int t = tid; t < lpb * (n >> 1); t += 256) {
            const int l = t / (n >> 1), b = t - l * (n >> 1);
            const int grp = b / half, j = b - grp * half;
            const int ia = l * n + grp * len + j, ib = ia + half;
            const double2 w = a.tw[half - 1 + j];
            const double2 u = buf[ia], x = buf[ib];
            const double vr = x.x * w.x - x.y * w.y;
            const double vi = x.x * w.y + x.y * w.x;
            buf[ia] = make_double2(u.x + vr, u.y + vi);
            buf[ib] = make_double2(u.x - vr, u.y - vi);
        }
        __syncthreads();
    }
    for (int idx = tid; idx < lpb * n; idx += 256) {
        const int l = a.cols ? idx % lpb : idx / n, e = a.cols ? idx / lpb : idx % n;
        const int line = line0 + l;
        if (line >= a.nlines) continue;
        const size_t q = a.cols ? (size_t)e * a.Nx + line : (size_t)line * a.Nx + e;
        double2 v = buf[l * n + e];
        if (a.inverse) {  // a[i] /= n (NS:175-177)
            v.x /= (double)n;
            v.y /= (double)n;
        }
        if (a.out == FFT_OUT_GREEN) {  // NS:300-320
            const int i = a.cols ? line : e, j = a.cols ? e : line;
            const int nj = (j <= a.Ny / 2) ? j : j - a.Ny;
            const double ky = gr.two_pi * (double)nj / gr.Ly;
            const int ni = (i <= a.Nx / 2) ? i : i - a.Nx;
            const double kx = gr.two_pi * (double)ni / gr.Lx;
            const double k2 = kx * kx + ky * ky;
            if (k2 == 0.0) {
                v = make_double2(0.0, 0.0);
            } else {
                const double sc = gr.neg4piG / k2;
                v.x *= sc;
                v.y *= sc;
            }
        }
        if (a.out == FFT_OUT_REAL) out_real[q] = v.x;
        else data[q] = v;
    }
}

// Twiddles of every stage, w_j = wlen^j by repeated multiplication from w_0 = 1 (NS:162-171),
// stage len stored at offset len/2 - 1 (n - 1 entries in all).
void Fft2D::build_twiddles(int n, bool inverse, std::vector<double>& out) {
    for (int len = 2; len <= n; len <<= 1) {
        const double ang = 2.0 * M_PI / len * (inverse ? +1.0 : -1.0);
        const double lr = std::cos(ang), li = std::sin(ang);
        double wr = 1.0, wi = 0.0;
        for (int j = 0; j < len / 2; j++) {
            out.push_back(wr);
            out.push_back(wi);
            const double nr = wr * lr - wi * li;
            const double ni = wr * li + wi * lr;
            wr = nr;
            wi = ni;
        }
    }
}

static int ilog2(int n) {
    int l = 0;
    while ((1 << l) < n) ++l;
    return l;
}

cc_status Fft2D::init(int Nx_, int Ny_) {
    destroy();
    Nx = Nx_;
    Ny = Ny_;
    auto pow2 = [](int n) { return n > 0 && (n & (n - 1)) == 0; };
    if (!pow2(Nx) || !pow2(Ny)) {
        set_error(CC_ERR_NON_POW2, "Nx and Ny must be powers of two for the FFT Poisson solver.");
        return CC_ERR_NON_POW2;
    }
    if (Nx > kFftMaxLen || Ny > kFftMaxLen) {
        set_error(CC_ERR_INVALID_ARG, "FFT line length above 4096 is not supported");
        return CC_ERR_INVALID_ARG;
    }
    std::vector<double> t;
    ax.n = Nx; ax.logn = ilog2(Nx); ax.lpb = std::max(1, std::min(16, 2048 / Nx));
    ay.n = Ny; ay.logn = ilog2(Ny); ay.lpb = std::max(1, std::min(16, 2048 / Ny));
    ax.tw_fwd = t.size() / 2; build_twiddles(Nx, false, t);
    ax.tw_inv = t.size() / 2; build_twiddles(Nx, true, t);
    ay.tw_fwd = t.size() / 2; build_twiddles(Ny, false, t);
    ay.tw_inv = t.size() / 2; build_twiddles(Ny, true, t);
    t.push_back(0.0); t.push_back(0.0);  // keeps the table non-empty for 1-point axes
    if (hipMalloc(&tw, t.size() * sizeof(double)) != hipSuccess) {
        set_error(CC_ERR_OOM, "FFT twiddle allocation failed");
        return CC_ERR_OOM;
    }
    CC_HIP(hipMemcpy(tw, t.data(), t.size() * sizeof(double), hipMemcpyHostToDevice));
    return CC_OK;
}

void Fft2D::destroy() {
    if (tw) hipFree(tw);
    tw = nullptr;
}

cc_status Fft2D::pass(bool cols, bool inverse, FftIn in, FftOut out, double* data, const double* in_real,
                      const double* mean, double* out_real, const FftGreen* green, hipStream_t s) {
    const FftAxis& A = cols ? ay : ax;
    FftLaunch L;
    L.n = A.n; L.logn = A.logn; L.lpb = A.lpb; L.nlines = cols ? Nx : Ny; L.Nx = Nx; L.Ny = Ny;
    L.cols = cols; L.inverse = inverse; L.in = in; L.out = out;
    L.tw = reinterpret_cast<const double2*>(tw) + (inverse ? A.tw_inv : A.tw_fwd);
    FftGreen gr = green ? *green : FftGreen{0, 0, 1, 1};
    const int blocks = (L.nlines + L.lpb - 1) / L.lpb;
    const size_t lds = (size_t)L.lpb * L.n * sizeof(double2);
    hipLaunchKernelGGL(k_fft_lines, dim3(blocks), dim3(256), lds, s, L, reinterpret_cast<double2*>(data), in_real,
                       mean, out_real, gr);
    CC_CHECK_LAUNCH();
    return CC_OK;
}

}  // namespace cc

extern "C" cc_status cc_fft2d(int Nx, int Ny, int inverse, double* data) {
    using namespace cc;
    if (!data) { set_error(CC_ERR_INVALID_ARG, "null data"); return CC_ERR_INVALID_ARG; }
    Fft2D f;
    CC_TRY(f.init(Nx, Ny));
    const size_t b = (size_t)Nx * Ny * 16;
    double* d = nullptr;
    if (hipMalloc(&d, b) != hipSuccess) { set_error(CC_ERR_OOM, "fft alloc"); return CC_ERR_OOM; }
    cc_status st = CC_OK;
    if (hipMemcpy(d, data, b, hipMemcpyHostToDevice) != hipSuccess) { set_error(CC_ERR_HIP, "fft upload"); st = CC_ERR_HIP; }
    if (st == CC_OK) st = f.pass(false, inverse != 0, FFT_IN_COMPLEX, FFT_OUT_COMPLEX, d, nullptr, nullptr, nullptr,
                                 nullptr, 0);
    if (st == CC_OK) st = f.pass(true, inverse != 0, FFT_IN_COMPLEX, FFT_OUT_COMPLEX, d, nullptr, nullptr, nullptr,
                                 nullptr, 0);
    if (st == CC_OK && hipMemcpy(data, d, b, hipMemcpyDeviceToHost) != hipSuccess) {
        set_error(CC_ERR_HIP, "fft download");
        st = CC_ERR_HIP;
    }
    hipFree(d);
    return st;
}
