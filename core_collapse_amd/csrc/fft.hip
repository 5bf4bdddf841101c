// fft.hip -- batched radix-2 FFT lines in LDS; see fft.hpp.
#include <cmath>
#include <mutex>
#include <string>
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

cc_status Fft2D::plan(int Nx_, int Ny_) {
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
    ax.n = Nx; ax.logn = ilog2(Nx); ax.lpb = std::max(1, std::min(16, 2048 / Nx));
    ay.n = Ny; ay.logn = ilog2(Ny); ay.lpb = std::max(1, std::min(16, 2048 / Ny));
    ax.tw_fwd = 0;
    ax.tw_inv = ax.tw_fwd + (size_t)(Nx - 1);
    ay.tw_fwd = ax.tw_inv + (size_t)(Nx - 1);
    ay.tw_inv = ay.tw_fwd + (size_t)(Ny - 1);
    return CC_OK;
}

cc_status Fft2D::init(int Nx_, int Ny_) {
    destroy();
    CC_TRY(plan(Nx_, Ny_));
    std::vector<double> t;
    build_twiddles(Nx, false, t);
    build_twiddles(Nx, true, t);
    build_twiddles(Ny, false, t);
    build_twiddles(Ny, true, t);
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

// ------------------------------------------------------------------ allocation-free entry point
// The twiddle tables of cc_fft2d_noalloc live in `scratch`, a file-scope device array: every GPU the
// library runs on holds one copy until the process ends.  Host code never passes the array's name to
// a kernel (on the host it is only a stand-in symbol); it uses the per-device address that
// hipGetSymbolAddress returns on the entry point's first call on that device.
namespace cc {

constexpr int kFftScratchEntries = 4 * kFftMaxLen;  // complex FP64 entries: 4 tables of < 4096 (256 KiB)
static __device__ __align__(256) double2 scratch[kFftScratchEntries];

constexpr int kFftMaxStages = 12;  // log2(4096)
// Per table (x fwd, x inv, y fwd, y inv): length n, offset, and w_len = (cos, sin)(+-2 pi / len) of
// every stage from the host's libm (NS:164).
struct FftTwiddlePlan {
    int n[4], off[4];
    double wr[4][kFftMaxStages], wi[4][kFftMaxStages];
};

// One thread per (table, stage): w_j = w_len^j by the same sequential complex products as
// Fft2D::build_twiddles (NS:162-171), compiled without contraction, so the tables are bit-identical.
__global__ void k_fft_twiddles(FftTwiddlePlan pl, double2* __restrict__ tw) {
    const int tab = (int)threadIdx.x / kFftMaxStages, st = (int)threadIdx.x % kFftMaxStages;
    if (tab >= 4) return;
    const int len = 2 << st;
    if (len > pl.n[tab]) return;
    const double lr = pl.wr[tab][st], li = pl.wi[tab][st];
    double wr = 1.0, wi = 0.0;
    double2* out = tw + pl.off[tab] + (len / 2 - 1);
    for (int j = 0; j < len / 2; j++) {
        out[j] = make_double2(wr, wi);
        const double nr = wr * lr - wi * li;
        const double ni = wr * li + wi * lr;
        wr = nr;
        wi = ni;
    }
}

constexpr int kMaxDevices = 64;
static std::mutex g_scratch_mu;
static double2* g_scratch_ptr[kMaxDevices] = {};

// The device address of `scratch` on the current device (looked up once per device).
static cc_status fft_scratch(double2** out) {
    int dev = 0;
    CC_HIP(hipGetDevice(&dev));
    if (dev < 0 || dev >= kMaxDevices) {
        set_error(CC_ERR_INVALID_ARG, "cc_fft2d_noalloc: device index out of range");
        return CC_ERR_INVALID_ARG;
    }
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void* ptr = nullptr;
        const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !ptr) {
            set_error(CC_ERR_HIP, std::string("cc_fft2d_noalloc: hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
            return CC_ERR_HIP;
        }
        g_scratch_ptr[dev] = static_cast<double2*>(ptr);
    }
    *out = g_scratch_ptr[dev];
    return CC_OK;
}

}  // namespace cc

extern "C" cc_status cc_fft2d_noalloc(int Nx, int Ny, int inverse, double* data, void* stream) {
    using namespace cc;
    if (!data) { set_error(CC_ERR_INVALID_ARG, "null data"); return CC_ERR_INVALID_ARG; }
    Fft2D f;
    CC_TRY(f.plan(Nx, Ny));  // powers of two <= 4096: the tables fit kFftScratchEntries
    if (f.twiddle_entries() > (size_t)kFftScratchEntries) {
        set_error(CC_ERR_INVALID_ARG, "cc_fft2d_noalloc: twiddle tables exceed the scratch array");
        return CC_ERR_INVALID_ARG;
    }
    double2* tw = nullptr;
    CC_TRY(fft_scratch(&tw));
    FftTwiddlePlan pl{};
    const int ns[4] = {Nx, Nx, Ny, Ny};
    const size_t offs[4] = {f.ax.tw_fwd, f.ax.tw_inv, f.ay.tw_fwd, f.ay.tw_inv};
    for (int t = 0; t < 4; ++t) {
        pl.n[t] = ns[t];
        pl.off[t] = (int)offs[t];
        for (int st = 0; st < kFftMaxStages; ++st) {
            const int len = 2 << st;
            const double ang = 2.0 * M_PI / len * ((t & 1) ? +1.0 : -1.0);
            pl.wr[t][st] = std::cos(ang);
            pl.wi[t][st] = std::sin(ang);
        }
    }
    hipStream_t s = static_cast<hipStream_t>(stream);
    hipLaunchKernelGGL(k_fft_twiddles, dim3(1), dim3(64), 0, s, pl, tw);
    CC_CHECK_LAUNCH();
    f.tw = reinterpret_cast<double*>(tw);
    cc_status st = f.pass(false, inverse != 0, FFT_IN_COMPLEX, FFT_OUT_COMPLEX, data, nullptr, nullptr, nullptr, nullptr, s);
    if (st == CC_OK)
        st = f.pass(true, inverse != 0, FFT_IN_COMPLEX, FFT_OUT_COMPLEX, data, nullptr, nullptr, nullptr, nullptr, s);
    f.tw = nullptr;  // the plan does not own the scratch array
    return st;
}
