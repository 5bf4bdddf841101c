// ref_eh.cpp -- exposes the reference 2D Euler program's own step functions
// (euler2d_gravity_hllc.cpp, compiled from its source in place) through the
// oracle C API of oracle/oracle.h.  Test infrastructure: built only where the
// reference source is present, output only into oracle/_ref/.  The reference's
// main() is renamed so the translation unit links as a library.
#define main cc_reference_program_main
#include REF_SOURCE
#undef main
#include "../oracle.h"
#include <cstring>

static Params mk(const or_eh_params* q) {
    Params p;
    p.gamma = q->gamma; p.cfl = q->cfl; p.x0 = q->x0; p.x1 = q->x1; p.y0 = q->y0; p.y1 = q->y1;
    p.Nx = q->Nx; p.Ny = q->Ny; p.floor_rho = q->floor_rho; p.floor_P = q->floor_P;
    p.periodic_x = q->periodic_x != 0; p.periodic_y = q->periodic_y != 0;
    p.gravity_on = q->gravity_on != 0; p.G = q->G;
    return p;  // max_iter / tol / omega are constants inside the reference solver (EH:329-331)
}
static State load(const Params& p, const or_eh_state* S) {
    size_t N = (size_t)p.Nx * p.Ny;
    State U;
    U.rho.assign(S->rho, S->rho + N); U.mx.assign(S->mx, S->mx + N);
    U.my.assign(S->my, S->my + N); U.E.assign(S->E, S->E + N);
    return U;
}
static void store(const State& U, or_eh_state* S) {
    size_t n = U.rho.size() * 8;
    std::memcpy(S->rho, U.rho.data(), n); std::memcpy(S->mx, U.mx.data(), n);
    std::memcpy(S->my, U.my.data(), n); std::memcpy(S->E, U.E.data(), n);
}

extern "C" {
void or_eh_hllc_x(const or_eh_params* q, const double* a, const double* b, double* o) {
    Params p = mk(q);
    Flux F = hllc_x(p, Prim{a[0], a[1], a[2], a[3]}, Prim{b[0], b[1], b[2], b[3]});
    o[0] = F.f1; o[1] = F.f2; o[2] = F.f3; o[3] = F.f4;
}
void or_eh_hllc_y(const or_eh_params* q, const double* a, const double* b, double* o) {
    Params p = mk(q);
    Flux F = hllc_y(p, Prim{a[0], a[1], a[2], a[3]}, Prim{b[0], b[1], b[2], b[3]});
    o[0] = F.f1; o[1] = F.f2; o[2] = F.f3; o[3] = F.f4;
}
int or_eh_poisson(const or_eh_params* q, const or_eh_state* S, double* gx, double* gy, double* phi_out) {
    Params p = mk(q);
    GravField g = compute_poisson_gravity_periodic(p, load(p, S));
    std::memcpy(gx, g.gx.data(), g.gx.size() * 8); std::memcpy(gy, g.gy.data(), g.gy.size() * 8);
    (void)phi_out;  // the reference keeps phi local
    return -1;      // and does not report its sweep count
}
double or_eh_compute_dt(const or_eh_params* q, const or_eh_state* S) {
    Params p = mk(q);
    return compute_dt(p, load(p, S));
}
void or_eh_rhs(const or_eh_params* q, const or_eh_state* S, or_eh_state* dS) {
    Params p = mk(q);
    store(rhs(p, load(p, S)), dS);
}
void or_eh_rk2_step(const or_eh_params* q, or_eh_state* S, double dt) {
    Params p = mk(q);
    State U = load(p, S);
    rk2_step(p, U, dt);
    store(U, S);
}
double or_eh_run(const or_eh_params* q, or_eh_state* S, int nsteps, double t, double t_end, double* dts) {
    Params p = mk(q);
    State U = load(p, S);
    for (int s = 0; s < nsteps && t < t_end - 1e-14; ++s) {  // EH:697-708
        double dt = compute_dt(p, U);
        if (!std::isfinite(dt) || dt <= 0.0) return -1.0;
        if (t + dt > t_end) dt = t_end - t;
        rk2_step(p, U, dt);
        t += dt;
        if (dts) dts[s] = dt;
    }
    store(U, S);
    return t;
}
// The program's default initial condition (init_random_blob, EH:574-642), for IC parity.
void or_eh_init_random_blob(const or_eh_params* q, or_eh_state* S) {
    Params p = mk(q);
    store(init_random_blob(p), S);
}
}
