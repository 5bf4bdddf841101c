// diag.hip -- device reductions for the in-situ diagnostics (totals and frame comparisons).
#include <vector>

#include "diag.hpp"

namespace cc {

namespace {

constexpr int kStatBlocks = 1024;  // fixed partial count: the reduction order depends on n only

__device__ __forceinline__ void block_reduce3(double& s0, double& s1, double& m2) {
    __shared__ double red[3][4];
    s0 = wave_sum(s0);
    s1 = wave_sum(s1);
    m2 = wave_max(m2);
    const int w = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) {
        red[0][w] = s0;
        red[1][w] = s1;
        red[2][w] = m2;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        s0 = (red[0][0] + red[0][1]) + (red[0][2] + red[0][3]);
        s1 = (red[1][0] + red[1][1]) + (red[1][2] + red[1][3]);
        m2 = smax(smax(red[2][0], red[2][1]), smax(red[2][2], red[2][3]));
    }
}

// grid (nblk, nf), 256 threads: block b of field f reduces the cells b*256 + k*nblk*256 + t.
__global__ __launch_bounds__(256) void k_field_partial(FieldPtrs a, FieldPtrs b, int with_b, size_t n,
                                                       double* __restrict__ part) {
    const int f = blockIdx.y;
    const double* __restrict__ x = a.p[f];
    const double* __restrict__ y = b.p[f];
    double s0 = 0.0, s1 = 0.0, m2 = 0.0;
    const size_t stride = (size_t)gridDim.x * 256;
    for (size_t q = (size_t)blockIdx.x * 256 + threadIdx.x; q < n; q += stride) {
        const double v = x[q];
        if (with_b) {
            const double w = y[q];
            const double d = fabs(v - w);
            s0 += d;
            s1 += fabs(w);
            m2 = smax(m2, d);
        } else {
            s0 += v;
            s1 += fabs(v);
            m2 = smax(m2, fabs(v));
        }
    }
    block_reduce3(s0, s1, m2);
    if (threadIdx.x == 0) {
        double* o = part + ((size_t)f * gridDim.x + blockIdx.x) * 3;
        o[0] = s0;
        o[1] = s1;
        o[2] = m2;
    }
}

// one block per field: reduce its nblk partials
__global__ __launch_bounds__(256) void k_field_final(int nblk, const double* __restrict__ part, double* __restrict__ res) {
    const int f = blockIdx.x;
    double s0 = 0.0, s1 = 0.0, m2 = 0.0;
    for (int k = threadIdx.x; k < nblk; k += 256) {
        const double* o = part + ((size_t)f * nblk + k) * 3;
        s0 += o[0];
        s1 += o[1];
        m2 = smax(m2, o[2]);
    }
    block_reduce3(s0, s1, m2);
    if (threadIdx.x == 0) {
        res[3 * f + 0] = s0;
        res[3 * f + 1] = s1;
        res[3 * f + 2] = m2;
    }
}

}  // namespace

cc_status FieldStats::run(const FieldPtrs& a, const FieldPtrs* b, int nf, size_t n, hipStream_t s, double* host_out) {
    if (nf <= 0 || nf > kDiagMaxFields || n == 0) {
        set_error(CC_ERR_INVALID_ARG, "diagnostics: bad field count or size");
        return CC_ERR_INVALID_ARG;
    }
    if (nf > cap_fields_) {
        destroy();
        if (hipMalloc((void**)&part_, sizeof(double) * 3 * kStatBlocks * kDiagMaxFields) != hipSuccess ||
            hipMalloc((void**)&res_, sizeof(double) * 3 * kDiagMaxFields) != hipSuccess) {
            destroy();
            set_error(CC_ERR_OOM, "diagnostics: device allocation failed");
            return CC_ERR_OOM;
        }
        cap_fields_ = kDiagMaxFields;
    }
    const size_t want = (n + 255) / 256;
    const int nblk = (int)(want < (size_t)kStatBlocks ? want : (size_t)kStatBlocks);
    FieldPtrs bb = b ? *b : a;
    hipLaunchKernelGGL(k_field_partial, dim3(nblk, nf), dim3(256), 0, s, a, bb, b ? 1 : 0, n, part_);
    CC_CHECK_LAUNCH();
    hipLaunchKernelGGL(k_field_final, dim3(nf), dim3(256), 0, s, nblk, part_, res_);
    CC_CHECK_LAUNCH();
    CC_HIP(hipMemcpyAsync(host_out, res_, sizeof(double) * 3 * nf, hipMemcpyDeviceToHost, s));
    CC_HIP(hipStreamSynchronize(s));
    return CC_OK;
}

void FieldStats::destroy() {
    if (part_) hipFree(part_);
    if (res_) hipFree(res_);
    part_ = res_ = nullptr;
    cap_fields_ = 0;
}

cc_status FrameBuffer::ensure(size_t count) {
    if (count <= cap) return CC_OK;
    destroy();
    if (hipMalloc((void**)&d, count * sizeof(double)) != hipSuccess) {
        d = nullptr;
        set_error(CC_ERR_OOM, "frame buffer: device allocation failed");
        return CC_ERR_OOM;
    }
    cap = count;
    return CC_OK;
}

void FrameBuffer::destroy() {
    if (d) hipFree(d);
    d = nullptr;
    cap = 0;
}

cc_status compare_frames(FieldStats& st, const double* fa, const double* fb, int nf, size_t n, hipStream_t s,
                         double* out) {
    FieldPtrs a{}, b{};
    for (int f = 0; f < nf; f++) {
        a.p[f] = fa + (size_t)f * n;
        b.p[f] = fb + (size_t)f * n;
    }
    std::vector<double> r(3 * (size_t)nf);
    CC_TRY(st.run(a, &b, nf, n, s, r.data()));
    for (int f = 0; f < nf; f++) {
        out[3 * f + 0] = r[3 * f + 0] / (double)n;
        out[3 * f + 1] = (r[3 * f + 1] > 0.0) ? r[3 * f + 0] / r[3 * f + 1] : 0.0;
        out[3 * f + 2] = r[3 * f + 2];
    }
    return CC_OK;
}

}  // namespace cc
