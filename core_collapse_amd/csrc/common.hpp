// common.hpp -- shared host/device helpers for the gfx950 solver-step library.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <string>

#include "../../include/ccamd.h"

namespace cc {

// ---------------------------------------------------------------- errors
struct Error {
    cc_status code;
    std::string msg;
};
void set_error(cc_status code, const std::string& msg);  // thread-local last error (ccamd.cpp)

#define CC_HIP(expr)                                                                             \
    do {                                                                                         \
        hipError_t _e = (expr);                                                                  \
        if (_e != hipSuccess) {                                                                  \
            ::cc::set_error(CC_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(_e));      \
            return CC_ERR_HIP;                                                                   \
        }                                                                                        \
    } while (0)

#define CC_CHECK_LAUNCH()                                                                        \
    do {                                                                                         \
        hipError_t _e = hipGetLastError();                                                       \
        if (_e != hipSuccess) {                                                                  \
            ::cc::set_error(CC_ERR_HIP, std::string("kernel launch: ") + hipGetErrorString(_e)); \
            return CC_ERR_HIP;                                                                   \
        }                                                                                        \
    } while (0)

#define CC_TRY(expr)                      \
    do {                                  \
        cc_status _s = (expr);            \
        if (_s != CC_OK) return _s;       \
    } while (0)

// ---------------------------------------------------------------- device math
// std::max / std::min semantics of the reference (max(a,b) = a<b ? b : a).
template <typename T>
__host__ __device__ __forceinline__ T smax(T a, T b) { return (a < b) ? b : a; }
template <typename T>
__host__ __device__ __forceinline__ T smin(T a, T b) { return (b < a) ? b : a; }
// Periodic wrap.  Callers pass in-range indices shifted by a few cells, so the common cases are
// resolved without an integer division (which expands to a long float-reciprocal sequence on the GPU).
__host__ __device__ __forceinline__ int imod(int i, int n) {
    if (i >= 0) {
        if (i < n) return i;
        if (i < 2 * n) return i - n;
    } else if (i >= -n) {
        return i + n;
    }
    const int r = i % n;
    return (r < 0) ? r + n : r;
}
__host__ __device__ __forceinline__ int iclamp(int v, int lo, int hi) { return v < lo ? lo : (v > hi ? hi : v); }
// Boundary map of a cell (or face) index: periodic wrap modulo n, or clamp to [0, n-1].
__host__ __device__ __forceinline__ int bmap(int i, int n, bool periodic) {
    return periodic ? imod(i, n) : iclamp(i, 0, n - 1);
}

// Smallest positive value used where the reference writes 1e-300 (M2:155, M2:444); it underflows in FP32.
// a / b in the arithmetic type R.  FP64 divides exactly (the reference's arithmetic).  FP32 in device code
// multiplies by the hardware reciprocal (v_rcp_f32, 1 ulp), about a tenth of the instructions of a correctly
// rounded FP32 quotient; the FP32 solvers only produce FP32-rounded fluxes whose ill-conditioned cases go to
// an FP64 solver, so the extra ulp is immaterial.
template <typename R>
__host__ __device__ __forceinline__ R qdiv(R a, R b) { return a / b; }
template <>
__host__ __device__ __forceinline__ float qdiv<float>(float a, float b) {
#if defined(__HIP_DEVICE_COMPILE__)
    return a * __builtin_amdgcn_rcpf(b);
#else
    return a / b;
#endif
}

template <typename R> struct Tiny;
template <> struct Tiny<double> { static __host__ __device__ __forceinline__ double v() { return 1e-300; } };
template <> struct Tiny<float> { static __host__ __device__ __forceinline__ float v() { return 1e-30f; } };

// Ordered-bits atomic max for non-negative doubles (positive IEEE doubles order like their bit patterns).
__device__ __forceinline__ void atomic_max_pos(unsigned long long* addr, double v) {
    atomicMax(addr, (unsigned long long)__double_as_longlong(v));
}

// Wave64 max reduction through DPP/permute (returns the max in every lane).
__device__ __forceinline__ double wave_max(double v) {
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) v = smax(v, __shfl_xor(v, off, 64));
    return v;
}
__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

}  // namespace cc
