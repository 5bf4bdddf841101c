// ref_m2.cpp -- exposes the reference 2D MHD program's own step functions
// (mhd2d_resistive_hlld_ct_gravity.cpp, compiled from its source in place)
// through the oracle C API of oracle/oracle.h.  Test infrastructure: built
// only where the reference source is present, output only into oracle/_ref/.
// The reference's main() is renamed so the translation unit links as a
// library; nothing of the reference is copied into this repository.
#define main cc_reference_program_main
#include REF_SOURCE
#undef main
#include "../oracle.h"
#include <cstring>

static Params mk(const or_m2_params* q) {
    Params p;
    p.gamma = q->gamma; p.cfl = q->cfl; p.x0 = q->x0; p.x1 = q->x1; p.y0 = q->y0; p.y1 = q->y1;
    p.Nx = q->Nx; p.Ny = q->Ny; p.floor_rho = q->floor_rho; p.floor_P = q->floor_P;
    p.periodic_x = q->periodic_x != 0; p.periodic_y = q->periodic_y != 0; p.eta = q->eta;
    p.gravity_on = q->gravity_on != 0; p.G = q->G; p.sor_iters = q->sor_iters; p.sor_omega = q->sor_omega;
    return p;
}
static void load(const Params& p, const or_m2_state* S, Ucc& U, Bfaces& B) {
    size_t N = (size_t)p.Nx * p.Ny, NFx = (size_t)(p.Nx + 1) * p.Ny, NFy = (size_t)p.Nx * (p.Ny + 1);
    U.rho.assign(S->rho, S->rho + N); U.mx.assign(S->mx, S->mx + N); U.my.assign(S->my, S->my + N);
    U.mz.assign(S->mz, S->mz + N); U.E.assign(S->E, S->E + N); U.Bz.assign(S->Bz, S->Bz + N);
    B.Bx_f.assign(S->Bx_f, S->Bx_f + NFx); B.By_f.assign(S->By_f, S->By_f + NFy);
}
static void store(const Ucc& U, const Bfaces& B, or_m2_state* S) {
    std::memcpy(S->rho, U.rho.data(), U.rho.size() * 8); std::memcpy(S->mx, U.mx.data(), U.mx.size() * 8);
    std::memcpy(S->my, U.my.data(), U.my.size() * 8); std::memcpy(S->mz, U.mz.data(), U.mz.size() * 8);
    std::memcpy(S->E, U.E.data(), U.E.size() * 8); std::memcpy(S->Bz, U.Bz.data(), U.Bz.size() * 8);
    std::memcpy(S->Bx_f, B.Bx_f.data(), B.Bx_f.size() * 8); std::memcpy(S->By_f, B.By_f.data(), B.By_f.size() * 8);
}
static Prim P8(const double* a) { return Prim{a[0], a[1], a[2], a[3], a[4], a[5], a[6], a[7]}; }

extern "C" {
void or_m2_hlld_x(const or_m2_params* q, const double* WL, const double* WR, double Bn, double* o) {
    Params p = mk(q);
    FluxX F = hlld_x(p, P8(WL), P8(WR), Bn);
    o[0] = F.frho; o[1] = F.fmx; o[2] = F.fmy; o[3] = F.fmz; o[4] = F.fE; o[5] = F.fBy; o[6] = F.fBz; o[7] = F.Ez;
}
void or_m2_hlld_y(const or_m2_params* q, const double* WL, const double* WR, double Bn, double* o) {
    Params p = mk(q);
    FluxY G = hlld_y(p, P8(WL), P8(WR), Bn);
    o[0] = G.frho; o[1] = G.fmx; o[2] = G.fmy; o[3] = G.fmz; o[4] = G.fE; o[5] = G.fBx; o[6] = G.fBz; o[7] = G.Ez;
}
void or_m2_cons_to_prim(const or_m2_params* q, const double* c, double* W8) {
    Params p = mk(q);
    Prim W = cons_to_prim(p, c[0], c[1], c[2], c[3], c[4], c[5], c[6], c[7]);
    W8[0] = W.rho; W8[1] = W.u; W8[2] = W.v; W8[3] = W.w; W8[4] = W.P; W8[5] = W.Bx; W8[6] = W.By; W8[7] = W.Bz;
}
double or_m2_mc_limiter(double dm, double dp) { return mc_limiter(dm, dp); }
double or_m2_fast_speed_x(const or_m2_params* q, const double* W8) { Params p = mk(q); return fast_speed_x(p, P8(W8)); }
double or_m2_fast_speed_y(const or_m2_params* q, const double* W8) { Params p = mk(q); return fast_speed_y(p, P8(W8)); }
void or_m2_poisson(const or_m2_params* q, const double* rho, double dx, double dy, double* gx, double* gy) {
    Params p = mk(q);
    std::vector<double> r(rho, rho + (size_t)p.Nx * p.Ny);
    GravField g = compute_poisson_gravity_periodic(p, r, dx, dy);
    std::memcpy(gx, g.gx.data(), g.gx.size() * 8); std::memcpy(gy, g.gy.data(), g.gy.size() * 8);
}
double or_m2_compute_dt(const or_m2_params* q, const or_m2_state* S, double dx, double dy) {
    Params p = mk(q); Ucc U; Bfaces B; load(p, S, U, B);
    return compute_dt(p, U, B, dx, dy);
}
void or_m2_advance_stage(const or_m2_params* q, const or_m2_state* S0, or_m2_state* S1, const double* gx,
                         const double* gy, double dx, double dy, double dt) {
    Params p = mk(q); Ucc U0, U1; Bfaces B0, B1; load(p, S0, U0, B0);
    GravField g; size_t N = (size_t)p.Nx * p.Ny; g.gx.assign(gx, gx + N); g.gy.assign(gy, gy + N);
    advance_stage(p, U0, B0, U1, B1, g, dx, dy, dt);
    store(U1, B1, S1);
}
void or_m2_rk2_step(const or_m2_params* q, or_m2_state* S, const double* gx, const double* gy, double dx, double dy,
                    double dt) {
    Params p = mk(q); Ucc U; Bfaces B; load(p, S, U, B);
    GravField g; size_t N = (size_t)p.Nx * p.Ny; g.gx.assign(gx, gx + N); g.gy.assign(gy, gy + N);
    rk2_step(p, U, B, g, dx, dy, dt);
    store(U, B, S);
}
void or_m2_divb(const or_m2_params* q, const double* Bx_f, const double* By_f, double dx, double dy, double* out) {
    Params p = mk(q); Bfaces B;
    B.Bx_f.assign(Bx_f, Bx_f + (size_t)(p.Nx + 1) * p.Ny); B.By_f.assign(By_f, By_f + (size_t)p.Nx * (p.Ny + 1));
    for (int j = 0; j < p.Ny; j++) for (int i = 0; i < p.Nx; i++) out[idx(i, j, p.Nx)] = divB_cc(p, B, i, j, dx, dy);
}
// The main loop body of M2:1111-1117 for a fixed number of steps, without output.
double or_m2_run(const or_m2_params* q, or_m2_state* S, double dx, double dy, int nsteps, double t0, double t_end,
                 double* dts) {
    Params p = mk(q); Ucc U; Bfaces B; load(p, S, U, B);
    double t = t0;
    for (int s = 0; s < nsteps && t < t_end; s++) {
        GravField g = compute_poisson_gravity_periodic(p, U.rho, dx, dy);
        double dt = compute_dt(p, U, B, dx, dy);
        if (t + dt > t_end) dt = t_end - t;
        rk2_step(p, U, B, g, dx, dy, dt);
        t += dt;
        if (dts) dts[s] = dt;
    }
    store(U, B, S);
    return t;
}
}
