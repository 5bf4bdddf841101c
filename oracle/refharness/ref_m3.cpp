// ref_m3.cpp -- exposes the reference 3D MHD program's own step functions
// (mhd3d_resistive_hlld_ct.cpp, compiled from its source in place) through the
// oracle C API of oracle/oracle.h.  Test infrastructure only; output only into
// oracle/_ref/.  The reference has only its hybrid EOS, so ideal_eos is ignored
// here (the hybrid EOS with K1 = Kbounce = 0 is the ideal gas with gamma_th).
#define main cc_reference_program_main
#include REF_SOURCE
#undef main
#include "../oracle.h"
#include <cstring>

static Params mk(const or_m3_params* q) {
    Params p;
    p.gamma1 = q->gamma1; p.gamma2 = q->gamma2; p.K1 = q->K1; p.rho_nuc = q->rho_nuc; p.gamma_th = q->gamma_th;
    p.rho_bounce = q->rho_bounce; p.gamma_bounce = q->gamma_bounce; p.Kbounce = q->Kbounce;
    p.gamma_init = q->gamma_init; p.cfl = q->cfl;
    p.x0 = q->x0; p.x1 = q->x1; p.y0 = q->y0; p.y1 = q->y1; p.z0 = q->z0; p.z1 = q->z1;
    p.Nx = q->Nx; p.Ny = q->Ny; p.Nz = q->Nz; p.floor_rho = q->floor_rho; p.floor_P = q->floor_P;
    p.periodic_x = q->periodic_x != 0; p.periodic_y = q->periodic_y != 0; p.periodic_z = q->periodic_z != 0;
    p.eta = q->eta;
    return p;
}
static void load(const Params& p, const or_m3_state* S, Ucc& U, Bfaces& B) {
    size_t N = (size_t)p.Nx * p.Ny * p.Nz;
    U.rho.assign(S->rho, S->rho + N); U.mx.assign(S->mx, S->mx + N); U.my.assign(S->my, S->my + N);
    U.mz.assign(S->mz, S->mz + N); U.E.assign(S->E, S->E + N);
    B.Bx_f.assign(S->Bx_f, S->Bx_f + (size_t)(p.Nx + 1) * p.Ny * p.Nz);
    B.By_f.assign(S->By_f, S->By_f + (size_t)p.Nx * (p.Ny + 1) * p.Nz);
    B.Bz_f.assign(S->Bz_f, S->Bz_f + (size_t)p.Nx * p.Ny * (p.Nz + 1));
}
static void store(const Ucc& U, const Bfaces& B, or_m3_state* S) {
    std::memcpy(S->rho, U.rho.data(), U.rho.size() * 8); std::memcpy(S->mx, U.mx.data(), U.mx.size() * 8);
    std::memcpy(S->my, U.my.data(), U.my.size() * 8); std::memcpy(S->mz, U.mz.data(), U.mz.size() * 8);
    std::memcpy(S->E, U.E.data(), U.E.size() * 8);
    std::memcpy(S->Bx_f, B.Bx_f.data(), B.Bx_f.size() * 8); std::memcpy(S->By_f, B.By_f.data(), B.By_f.size() * 8);
    std::memcpy(S->Bz_f, B.Bz_f.data(), B.Bz_f.size() * 8);
}
static Prim P8(const double* a) { return Prim{a[0], a[1], a[2], a[3], a[4], a[5], a[6], a[7]}; }

extern "C" {
void or_m3_hlld_x(const or_m3_params* q, const double* WL, const double* WR, double Bn, double* o) {
    Params p = mk(q);
    FluxX F = hlld_x(p, P8(WL), P8(WR), Bn);
    o[0] = F.frho; o[1] = F.fmx; o[2] = F.fmy; o[3] = F.fmz; o[4] = F.fE; o[5] = F.fBy; o[6] = F.fBz;
}
void or_m3_cons_to_prim(const or_m3_params* q, const double* c, double* W8) {
    Params p = mk(q);
    Prim W = cons_to_prim(p, c[0], c[1], c[2], c[3], c[4], c[5], c[6], c[7]);
    W8[0] = W.rho; W8[1] = W.u; W8[2] = W.v; W8[3] = W.w; W8[4] = W.P; W8[5] = W.Bx; W8[6] = W.By; W8[7] = W.Bz;
}
void or_m3_eos_cold(const or_m3_params* q, double rho, double* o) {
    Params p = mk(q);
    ColdEOS c = eos_cold(p, rho);
    o[0] = c.Pcold; o[1] = c.ecold; o[2] = c.gamma_c;
}
double or_m3_sound_speed2(const or_m3_params* q, const double* W) { Params p = mk(q); return sound_speed2(p, P8(W)); }
double or_m3_total_energy(const or_m3_params* q, const double* W) {
    Params p = mk(q);
    return total_energy_from_prim(p, P8(W));
}
double or_m3_compute_dt(const or_m3_params* q, const or_m3_state* S, double dx, double dy, double dz) {
    Params p = mk(q); Ucc U; Bfaces B; load(p, S, U, B);
    return compute_dt(p, U, B, dx, dy, dz);
}
void or_m3_advance_stage(const or_m3_params* q, const or_m3_state* S0, or_m3_state* S1, double dx, double dy,
                         double dz, double dt) {
    Params p = mk(q); Ucc U0, U1; Bfaces B0, B1; load(p, S0, U0, B0);
    advance_stage(p, U0, B0, U1, B1, dx, dy, dz, dt);
    store(U1, B1, S1);
}
void or_m3_rk2_step(const or_m3_params* q, or_m3_state* S, double dx, double dy, double dz, double dt) {
    Params p = mk(q); Ucc U; Bfaces B; load(p, S, U, B);
    rk2_step(p, U, B, dx, dy, dz, dt);
    store(U, B, S);
}
void or_m3_divb(const or_m3_params* q, const or_m3_state* S, double dx, double dy, double dz, double* out) {
    Params p = mk(q); Ucc U; Bfaces B; load(p, S, U, B);
    for (int k = 0; k < p.Nz; k++) for (int j = 0; j < p.Ny; j++) for (int i = 0; i < p.Nx; i++)
        out[idx(i, j, k, p.Nx, p.Ny)] = divB_cc(p, B, i, j, k, dx, dy, dz);
}
// The loop of M3:1285-1291 for a fixed number of steps, without output.
double or_m3_run(const or_m3_params* q, or_m3_state* S, double dx, double dy, double dz, int nsteps, double t0,
                 double t_end, double* dts) {
    Params p = mk(q); Ucc U; Bfaces B; load(p, S, U, B);
    double t = t0;
    for (int s = 0; s < nsteps && t < t_end; s++) {
        double dt = compute_dt(p, U, B, dx, dy, dz);
        if (t + dt > t_end) dt = t_end - t;
        rk2_step(p, U, B, dx, dy, dz, dt);
        t += dt;
        if (dts) dts[s] = dt;
    }
    store(U, B, S);
    return t;
}
}
