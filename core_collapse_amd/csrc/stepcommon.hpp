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

}  // namespace cc
