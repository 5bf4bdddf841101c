// stepcommon.hpp -- device scalars and reductions shared by the step drivers.
#pragma once
#include "common.hpp"

namespace cc {

// Device-resident step scalars: dt is produced and consumed on the device, so a step never
// waits for the host.
struct Scalars {
    unsigned long long spare;
    int step;
    int pad;
    double dt, t, t_end;
    double divb;
};

// Block (256 threads) maximum of a non-negative value, written to part[block] by thread 0.
__device__ __forceinline__ void block_max_to(double v, double* part) {
    __shared__ double red[4];
    v = wave_max(v);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
        const size_t b = ((size_t)blockIdx.z * gridDim.y + blockIdx.y) * gridDim.x + blockIdx.x;
        part[b] = smax(smax(red[0], red[1]), smax(red[2], red[3]));
    }
}

// Maximum of npart partial maxima together with the seed (256 threads, one block); valid in thread 0.
__device__ __forceinline__ double block_reduce_max(const double* part, int npart, double seed) {
    __shared__ double red[4];
    double m = seed;
    for (int k = threadIdx.x; k < npart; k += 256) m = smax(m, part[k]);
    m = wave_max(m);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = m;
    __syncthreads();
    return smax(smax(red[0], red[1]), smax(red[2], red[3]));
}

// dt = min(cfl h / maxs, 0.25 h^2 / eta) (M2:590-597, M3:640-647), clamped to t_end (M2:1114,
// M3:1287); the result and t + dt are left on the device and appended to the dt log.
__device__ __forceinline__ void finalize_dt(Scalars* sc, double maxs, double h, double cfl, double eta, double* dts) {
    const double dt_h = cfl * h / maxs;
    double dt_r = 1e100;
    if (eta > 0.0) dt_r = 0.25 * h * h / eta;
    double dt = smin(dt_h, dt_r);
    if (sc->t + dt > sc->t_end) dt = sc->t_end - sc->t;
    sc->dt = dt;
    sc->t = sc->t + dt;
    if (dts) dts[sc->step] = dt;
    sc->step += 1;
}

__global__ void k_set_scalars(Scalars* sc, double dt, double t, double t_end, int reset_step);

// ------------------------------------------------------------------ compensated face-field update
// With compensation on, each face carries the rounding error of its stored FP64 value as an FP32 low
// word (hi + lo is the field to about 1e-24 relative), so the discrete divergence of the stored field
// no longer accumulates one rounding per face per step: hi stays the correctly rounded face value and
// max |div B| stays at the size of one rounding (<= 2 ulp(B) / h per axis) over any number of steps.
// The increments are the reference's (M2:807-830, M3:941-1015); only the additions are exact.
// TwoSum (Knuth): s + e == a + b exactly, s = fl(a + b).  Compiled code must not reassociate these.
__host__ __device__ __forceinline__ void two_sum(double a, double b, double& s, double& e) {
    s = a + b;
    const double bb = s - a;
    e = (a - (s - bb)) + (b - bb);
}
// One CT stage: (hi, lo) + delta.
__host__ __device__ __forceinline__ void ct_add(double hi, float lo, double delta, double& hi1, float& lo1) {
    double e;
    two_sum(hi, (double)lo + delta, hi1, e);
    lo1 = (float)e;
}
// The Heun average 1/2 (B^n + B^2) of the face field (rk2_step M2:850-866, M3:1042-1065) from the two
// compensated fields B^n = (hn, ln) and B^2 = (h2, l2), exact up to the rounding of the low words.
__host__ __device__ __forceinline__ void ct_avg(double hn, float ln, double h2, float l2, double& hi, float& lo) {
    double s, e, h, l;
    two_sum(hn, h2, s, e);
    const double t = e + ((double)ln + (double)l2);
    two_sum(s, t, h, l);
    hi = 0.5 * h;
    lo = (float)(0.5 * l);
}

}  // namespace cc
