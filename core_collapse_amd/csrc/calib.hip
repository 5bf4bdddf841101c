// calib.hip -- a plain device copy of known size, used to scale rocprofv3 FETCH_SIZE / WRITE_SIZE
// counters to bytes (cc_copy_calibration in ccamd.h).
#include "common.hpp"

namespace cc {
__global__ __launch_bounds__(256) void k_plain_copy(const double4* __restrict__ src, double4* __restrict__ dst,
                                                    size_t n) {
    for (size_t q = (size_t)blockIdx.x * 256 + threadIdx.x; q < n; q += (size_t)gridDim.x * 256) dst[q] = src[q];
}
}  // namespace cc

extern "C" cc_status cc_copy_calibration(size_t nbytes, int reps, double* avg_ms) {
    if (nbytes < 32 || reps < 1 || !avg_ms) {
        cc::set_error(CC_ERR_INVALID_ARG, "copy calibration: bad arguments");
        return CC_ERR_INVALID_ARG;
    }
    const size_t n = nbytes / sizeof(double4);
    double4 *a = nullptr, *b = nullptr;
    if (hipMalloc(&a, n * sizeof(double4)) != hipSuccess || hipMalloc(&b, n * sizeof(double4)) != hipSuccess) {
        if (a) hipFree(a);
        cc::set_error(CC_ERR_OOM, "copy calibration: allocation failed");
        return CC_ERR_OOM;
    }
    hipMemset(a, 0, n * sizeof(double4));
    hipEvent_t e0, e1;
    hipEventCreate(&e0);
    hipEventCreate(&e1);
    hipLaunchKernelGGL(cc::k_plain_copy, dim3(4096), dim3(256), 0, 0, a, b, n);  // warm-up
    hipEventRecord(e0, 0);
    for (int r = 0; r < reps; ++r) hipLaunchKernelGGL(cc::k_plain_copy, dim3(4096), dim3(256), 0, 0, a, b, n);
    hipEventRecord(e1, 0);
    hipError_t err = hipEventSynchronize(e1);
    float ms = 0.f;
    hipEventElapsedTime(&ms, e0, e1);
    *avg_ms = ms / reps;
    hipEventDestroy(e0);
    hipEventDestroy(e1);
    hipFree(a);
    hipFree(b);
    if (err != hipSuccess) {
        cc::set_error(CC_ERR_HIP, hipGetErrorString(err));
        return CC_ERR_HIP;
    }
    return CC_OK;
}
