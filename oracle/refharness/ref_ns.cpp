// ref_ns.cpp -- exposes the reference 2D Navier-Stokes program's own functions
// (navier_stokes_2d_hllc_poisson.cpp, compiled from its source in place) through
// the oracle C API of oracle/oracle.h.  Test infrastructure: built only where the
// reference source is present, output only into oracle/_ref/.  The reference's
// main() is renamed so the translation unit links as a library.  The Heun update
// lives inside the reference's main (NS:758-781); or_ns_heun_step below drives the
// reference's own compute_rhs through those same update lines.
#define main cc_reference_program_main
#include REF_SOURCE
#undef main
#include "../oracle.h"
#include <cstring>

static Params mk(const or_ns_params* q) {
    Params p;
    p.gamma = q->gamma; p.cfl = q->cfl; p.x0 = q->x0; p.x1 = q->x1; p.y0 = q->y0; p.y1 = q->y1;
    p.Nx = q->Nx; p.Ny = q->Ny; p.floor_rho = q->floor_rho; p.floor_P = q->floor_P;
    p.viscosity_on = q->viscosity_on != 0; p.mu = q->mu; p.cfl_visc = q->cfl_visc;
    p.gravity_on = q->gravity_on != 0; p.G = q->G;
    return p;
}
static Grid2D grid(const Params& p) { return Grid2D(p.Nx, p.Ny, p.x0, p.x1, p.y0, p.y1); }
static std::vector<U4> load(const Params& p, const double* U) {
    std::vector<U4> v((size_t)p.Nx * p.Ny);
    std::memcpy(v.data(), U, v.size() * sizeof(U4));
    return v;
}

extern "C" {
void or_ns_hllc_x(const or_ns_params* q, const double* a, const double* b, double* o) {
    Params p = mk(q);
    U4 F = hllc_x(W4{a[0], a[1], a[2], a[3]}, W4{b[0], b[1], b[2], b[3]}, p);
    o[0] = F.d; o[1] = F.mx; o[2] = F.my; o[3] = F.E;
}
void or_ns_hllc_y(const or_ns_params* q, const double* a, const double* b, double* o) {
    Params p = mk(q);
    U4 F = hllc_y(W4{a[0], a[1], a[2], a[3]}, W4{b[0], b[1], b[2], b[3]}, p);
    o[0] = F.d; o[1] = F.mx; o[2] = F.my; o[3] = F.E;
}
void or_ns_fft2d(int Nx, int Ny, int inverse, double* data) {
    std::vector<std::complex<double>> A((size_t)Nx * Ny);
    std::memcpy((void*)A.data(), data, A.size() * 16);
    fft2d(A, Nx, Ny, inverse != 0);
    std::memcpy(data, (void*)A.data(), A.size() * 16);
}
int or_ns_poisson_fft(const or_ns_params* q, const double* rho, double* phi) {
    Params p = mk(q);
    if (!is_pow2(p.Nx) || !is_pow2(p.Ny)) return 1;
    std::vector<double> r(rho, rho + (size_t)p.Nx * p.Ny), out;
    solve_poisson_periodic_fft(r, out, grid(p), p);
    std::memcpy(phi, out.data(), out.size() * 8);
    return 0;
}
void or_ns_gravity_accel(const or_ns_params* q, const double* phi, double* gx, double* gy) {
    Params p = mk(q);
    std::vector<double> f(phi, phi + (size_t)p.Nx * p.Ny), a, b;
    compute_gravity_accel(f, a, b, grid(p));
    std::memcpy(gx, a.data(), a.size() * 8); std::memcpy(gy, b.data(), b.size() * 8);
}
void or_ns_compute_rhs(const or_ns_params* q, const double* U, double* dU) {
    Params p = mk(q);
    std::vector<U4> d;
    compute_rhs(load(p, U), d, grid(p), p);
    std::memcpy(dU, d.data(), d.size() * sizeof(U4));
}
double or_ns_compute_dt(const or_ns_params* q, const double* U) {
    Params p = mk(q);
    return compute_dt(load(p, U), grid(p), p);
}
void or_ns_heun_step(const or_ns_params* q, double* Uh, double dt) {
    Params p = mk(q);
    Grid2D g = grid(p);
    std::vector<U4> U = load(p, Uh), k1, k2, U1;
    const int N = p.Nx * p.Ny;
    compute_rhs(U, k1, g, p);
    U1.resize(N);
    for (int c = 0; c < N; c++) {
        U1[c].d = U[c].d + dt * k1[c].d; U1[c].mx = U[c].mx + dt * k1[c].mx;
        U1[c].my = U[c].my + dt * k1[c].my; U1[c].E = U[c].E + dt * k1[c].E;
        if (U1[c].d < p.floor_rho) U1[c].d = p.floor_rho;
    }
    compute_rhs(U1, k2, g, p);
    for (int c = 0; c < N; c++) {
        U[c].d = U[c].d + 0.5 * dt * (k1[c].d + k2[c].d); U[c].mx = U[c].mx + 0.5 * dt * (k1[c].mx + k2[c].mx);
        U[c].my = U[c].my + 0.5 * dt * (k1[c].my + k2[c].my); U[c].E = U[c].E + 0.5 * dt * (k1[c].E + k2[c].E);
        if (U[c].d < p.floor_rho) U[c].d = p.floor_rho;
    }
    std::memcpy(Uh, U.data(), U.size() * sizeof(U4));
}
double or_ns_run(const or_ns_params* q, double* U, int nsteps, double t, double t_end, double* dts) {
    Params p = mk(q);
    for (int s = 0; s < nsteps && t < t_end - 1e-14; ++s) {
        std::vector<U4> Uv = load(p, U);
        double dt = compute_dt(Uv, grid(p), p);
        if (t + dt > t_end) dt = t_end - t;
        or_ns_heun_step(q, U, dt);
        t += dt;
        if (dts) dts[s] = dt;
    }
    return t;
}
// libstdc++'s distributions over std::mt19937 as the programs' ICs draw them
// (EH:585-587 uniform_real_distribution; NS:638-639 normal_distribution).
void or_std_normal(unsigned seed, int n, double* out) {
    std::mt19937 rng(seed);
    std::normal_distribution<double> g(0.0, 1.0);
    for (int k = 0; k < n; ++k) out[k] = g(rng);
}
void or_std_uniform(unsigned seed, int n, double a, double b, double* out) {
    std::mt19937 rng(seed);
    std::uniform_real_distribution<double> u(a, b);
    for (int k = 0; k < n; ++k) out[k] = u(rng);
}
}
