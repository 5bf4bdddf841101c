/*
 * oracle_eh.c -- CPU restatement of the 2D Euler HLLC + SOR-gravity step of
 * euler2d_gravity_hllc.cpp (EH), in FP64 with the reference's operation order.
 *
 * Test infrastructure only: loaded by tests/, __graft_entry__.smoke() and bench.py's CPU leg
 * as the checker; the product library never uses it.
 *
 * Pinned by tests/golden/eh_*.npz (written from the reference's own functions by
 * tests/golden/gen_golden.py) and by closed forms in tests/test_oracle_euler2d.py.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "oracle.h"

static double dmax(double a, double b) { return (a < b) ? b : a; } /* std::max */
static double dmin(double a, double b) { return (b < a) ? b : a; } /* std::min */
static int wrapi(int i, int n) { int r = i % n; if (r < 0) r += n; return r; } /* EH:225 */
static int clampi(int v, int lo, int hi) { return v < lo ? lo : (v > hi ? hi : v); }
static size_t idx(int i, int j, int Nx) { return (size_t)j * (size_t)Nx + (size_t)i; } /* EH:27 */

typedef struct { double rho, u, v, P; } Prim;
typedef struct { double f1, f2, f3, f4; } Flux;

/* EH:62-72 */
static void enforce_floors(const or_eh_params* p, double* rho, double* mx, double* my, double* E) {
    *rho = dmax(*rho, p->floor_rho);
    double u = *mx / *rho, v = *my / *rho;
    double kin = 0.5 * *rho * (u * u + v * v);
    double P = (p->gamma - 1.0) * (*E - kin);
    if (P < p->floor_P) {
        double eint = p->floor_P / (p->gamma - 1.0);
        *E = kin + eint;
    }
}
/* EH:74-83 */
static void cons_to_prim_full(const or_eh_params* p, double rho, double mx, double my, double E, double* u, double* v,
                              double* P, double* a) {
    rho = dmax(rho, p->floor_rho);
    *u = mx / rho;
    *v = my / rho;
    double kin = 0.5 * rho * (*u * *u + *v * *v);
    *P = (p->gamma - 1.0) * (E - kin);
    if (*P < p->floor_P) *P = p->floor_P;
    *a = sqrt(p->gamma * *P / rho);
}
/* EH:85-92 */
static double minmod(double a, double b) {
    if (a * b <= 0.0) return 0.0;
    return (fabs(a) < fabs(b)) ? a : b;
}
static double mc_limiter(double dm, double dp) { return minmod(0.5 * (dm + dp), minmod(2.0 * dm, 2.0 * dp)); }
/* EH:96-100 */
static Prim cons_to_prim(const or_eh_params* p, double rho, double mx, double my, double E) {
    double u, v, P, a;
    cons_to_prim_full(p, rho, mx, my, E, &u, &v, &P, &a);
    Prim W = {dmax(rho, p->floor_rho), u, v, P};
    return W;
}
/* EH:102-109 */
static void prim_to_cons(const or_eh_params* p, Prim W, double* c) {
    c[0] = dmax(W.rho, p->floor_rho);
    c[1] = c[0] * W.u;
    c[2] = c[0] * W.v;
    double kin = 0.5 * c[0] * (W.u * W.u + W.v * W.v);
    double eint = W.P / (p->gamma - 1.0);
    c[3] = kin + eint;
}
/* EH:113-123 */
static Flux euler_flux(const or_eh_params* p, Prim W, int axis) {
    double E = W.P / (p->gamma - 1.0) + 0.5 * W.rho * (W.u * W.u + W.v * W.v);
    Flux F;
    if (axis == 0) {
        F.f1 = W.rho * W.u; F.f2 = W.rho * W.u * W.u + W.P; F.f3 = W.rho * W.u * W.v; F.f4 = (E + W.P) * W.u;
    } else {
        F.f1 = W.rho * W.v; F.f2 = W.rho * W.u * W.v; F.f3 = W.rho * W.v * W.v + W.P; F.f4 = (E + W.P) * W.v;
    }
    return F;
}
/* hllc_x EH:125-170 (axis 0) and hllc_y EH:172-219 (axis 1); the star energy carries +P*SM (EH:154, 202). */
static Flux hllc(const or_eh_params* p, Prim WL, Prim WR, int axis) {
    double aL = sqrt(p->gamma * WL.P / dmax(WL.rho, p->floor_rho));
    double aR = sqrt(p->gamma * WR.P / dmax(WR.rho, p->floor_rho));
    double unL = axis == 0 ? WL.u : WL.v, unR = axis == 0 ? WR.u : WR.v;
    double SL = dmin(unL - aL, unR - aR);
    double SR = dmax(unL + aL, unR + aR);
    double cL[4], cR[4];
    prim_to_cons(p, WL, cL);
    prim_to_cons(p, WR, cR);
    Flux FL = euler_flux(p, WL, axis), FR = euler_flux(p, WR, axis);
    double num = (WR.P - WL.P) + cL[0] * unL * (SL - unL) - cR[0] * unR * (SR - unR);
    double den = cL[0] * (SL - unL) - cR[0] * (SR - unR);
    double SM = (fabs(den) > 1e-14) ? (num / den) : 0.0;
    if (0.0 <= SL) return FL;
    if (0.0 >= SR) return FR;
    int left = 0.0 < SM;
    Prim W = left ? WL : WR;
    const double* c = left ? cL : cR;
    double S = left ? SL : SR;
    Flux F0 = left ? FL : FR;
    double un = axis == 0 ? W.u : W.v;
    double rS = c[0] * (S - un) / (S - SM);
    double mnS = rS * SM;
    double mtS = rS * (axis == 0 ? W.v : W.u);
    double ES = ((S - un) * c[3] - W.P * un + W.P * SM) / (S - SM);
    double mxS = axis == 0 ? mnS : mtS, myS = axis == 0 ? mtS : mnS;
    Flux F = {F0.f1 + S * (rS - c[0]), F0.f2 + S * (mxS - c[1]), F0.f3 + S * (myS - c[2]), F0.f4 + S * (ES - c[3])};
    return F;
}

void or_eh_hllc_x(const or_eh_params* p, const double* a, const double* b, double* out) {
    Prim L = {a[0], a[1], a[2], a[3]}, R = {b[0], b[1], b[2], b[3]};
    Flux F = hllc(p, L, R, 0);
    out[0] = F.f1; out[1] = F.f2; out[2] = F.f3; out[3] = F.f4;
}
void or_eh_hllc_y(const or_eh_params* p, const double* a, const double* b, double* out) {
    Prim L = {a[0], a[1], a[2], a[3]}, R = {b[0], b[1], b[2], b[3]};
    Flux F = hllc(p, L, R, 1);
    out[0] = F.f1; out[1] = F.f2; out[2] = F.f3; out[3] = F.f4;
}

/* compute_poisson_gravity_periodic EH:288-386: periodic SOR (always wrapped, whatever the hydro
 * boundary flags) until the in-sweep residual max|lap(phi) - rhs| < tol, mean removed after every
 * sweep.  Returns the number of sweeps run; phi_out (nullable) receives phi. */
int or_eh_poisson(const or_eh_params* p, const or_eh_state* U, double* gx, double* gy, double* phi_out) {
    const int Nx = p->Nx, Ny = p->Ny;
    const size_t N = (size_t)Nx * Ny;
    memset(gx, 0, N * 8);
    memset(gy, 0, N * 8);
    if (!p->gravity_on) return 0;
    const double dx = (p->x1 - p->x0) / Nx, dy = (p->y1 - p->y0) / Ny;
    const double idx2 = 1.0 / (dx * dx), idy2 = 1.0 / (dy * dy);
    double* rho = (double*)malloc(N * 8);
    double* rhs = (double*)malloc(N * 8);
    double* phi = (double*)calloc(N, 8);
    double rhobar = 0.0;
    for (int j = 0; j < Ny; ++j)
        for (int i = 0; i < Nx; ++i) {
            size_t k = idx(i, j, Nx);
            double r = dmax(U->rho[k], p->floor_rho);
            rho[k] = r;
            rhobar += r;
        }
    rhobar /= (double)(Nx * Ny);
    const double FOURPI = 4.0 * M_PI;
    for (size_t k = 0; k < N; ++k) rhs[k] = FOURPI * p->G * (rho[k] - rhobar);
    const double denom = 2.0 * (idx2 + idy2);
    int it = 0;
    for (; it < p->max_iter;) {
        double max_res = 0.0;
        for (int j = 0; j < Ny; ++j)
            for (int i = 0; i < Nx; ++i) {
                size_t k = idx(i, j, Nx);
                double phiE = phi[idx(wrapi(i + 1, Nx), j, Nx)];
                double phiW = phi[idx(wrapi(i - 1, Nx), j, Nx)];
                double phiN = phi[idx(i, wrapi(j + 1, Ny), Nx)];
                double phiS = phi[idx(i, wrapi(j - 1, Ny), Nx)];
                double phi_new = ((phiE + phiW) * idx2 + (phiN + phiS) * idy2 - rhs[k]) / denom;
                double delta = phi_new - phi[k];
                phi[k] += p->omega * delta;
                double lap = (phiE - 2.0 * phi[k] + phiW) * idx2 + (phiN - 2.0 * phi[k] + phiS) * idy2;
                double res = lap - rhs[k];
                max_res = dmax(max_res, fabs(res));
            }
        double phibar = 0.0;
        for (size_t k = 0; k < N; ++k) phibar += phi[k];
        phibar /= (double)N;
        for (size_t k = 0; k < N; ++k) phi[k] -= phibar;
        ++it;
        if (max_res < p->tol) break;
    }
    for (int j = 0; j < Ny; ++j)
        for (int i = 0; i < Nx; ++i) {
            size_t k = idx(i, j, Nx);
            double phiE = phi[idx(wrapi(i + 1, Nx), j, Nx)], phiW = phi[idx(wrapi(i - 1, Nx), j, Nx)];
            double phiN = phi[idx(i, wrapi(j + 1, Ny), Nx)], phiS = phi[idx(i, wrapi(j - 1, Ny), Nx)];
            gx[k] = -(phiE - phiW) / (2.0 * dx);
            gy[k] = -(phiN - phiS) / (2.0 * dy);
        }
    if (phi_out) memcpy(phi_out, phi, N * 8);
    free(rho); free(rhs); free(phi);
    return it;
}

/* EH:390-404 */
double or_eh_compute_dt(const or_eh_params* p, const or_eh_state* U) {
    const int Nx = p->Nx, Ny = p->Ny;
    const double dx = (p->x1 - p->x0) / Nx, dy = (p->y1 - p->y0) / Ny;
    double max_speed = 1e-30;
    for (int j = 0; j < Ny; ++j)
        for (int i = 0; i < Nx; ++i) {
            size_t k = idx(i, j, Nx);
            double u, v, P, a;
            cons_to_prim_full(p, U->rho[k], U->mx[k], U->my[k], U->E[k], &u, &v, &P, &a);
            max_speed = dmax(max_speed, fabs(u) + a);
            max_speed = dmax(max_speed, fabs(v) + a);
        }
    return p->cfl * dmin(dx, dy) / max_speed;
}

static int bc(int i, int n, int periodic) { return periodic ? wrapi(i, n) : clampi(i, 0, n - 1); }
static double comp(Prim W, int c) { return c == 0 ? W.rho : c == 1 ? W.u : c == 2 ? W.v : W.P; }

/* MUSCL face states of EH:437-451 (x) / EH:476-490 (y): four cells along the axis. */
static void face_states(const or_eh_params* p, const Prim* W, int i, int j, int axis, Prim* WLf, Prim* WRf) {
    const int Nx = p->Nx, Ny = p->Ny;
    double L[4], R[4];
    if (axis == 0) {
        int iL = bc(i - 1, Nx, p->periodic_x), iR = bc(i, Nx, p->periodic_x);
        const Prim wm = W[idx(bc(iL - 1, Nx, p->periodic_x), j, Nx)], w0 = W[idx(bc(iL, Nx, p->periodic_x), j, Nx)];
        const Prim wp = W[idx(bc(iR, Nx, p->periodic_x), j, Nx)], wpp = W[idx(bc(iR + 1, Nx, p->periodic_x), j, Nx)];
        for (int c = 0; c < 4; ++c) {
            double a = comp(wm, c), b = comp(w0, c), d = comp(wp, c), e = comp(wpp, c);
            double sL = mc_limiter(b - a, d - b), sR = mc_limiter(d - b, e - d);
            L[c] = b + 0.5 * sL;
            R[c] = d - 0.5 * sR;
        }
    } else {
        int jL = bc(j - 1, Ny, p->periodic_y), jR = bc(j, Ny, p->periodic_y);
        const Prim wm = W[idx(i, bc(jL - 1, Ny, p->periodic_y), Nx)], w0 = W[idx(i, bc(jL, Ny, p->periodic_y), Nx)];
        const Prim wp = W[idx(i, bc(jR, Ny, p->periodic_y), Nx)], wpp = W[idx(i, bc(jR + 1, Ny, p->periodic_y), Nx)];
        for (int c = 0; c < 4; ++c) {
            double a = comp(wm, c), b = comp(w0, c), d = comp(wp, c), e = comp(wpp, c);
            double sL = mc_limiter(b - a, d - b), sR = mc_limiter(d - b, e - d);
            L[c] = b + 0.5 * sL;
            R[c] = d - 0.5 * sR;
        }
    }
    WLf->rho = dmax(L[0], p->floor_rho); WLf->u = L[1]; WLf->v = L[2]; WLf->P = dmax(L[3], p->floor_P);
    WRf->rho = dmax(R[0], p->floor_rho); WRf->u = R[1]; WRf->v = R[2]; WRf->P = dmax(R[3], p->floor_P);
}

/* rhs EH:406-527 (gravity solved inside, EH:411) */
static void rhs_with(const or_eh_params* p, const or_eh_state* U, or_eh_state* dU, const double* gx, const double* gy) {
    const int Nx = p->Nx, Ny = p->Ny;
    const size_t N = (size_t)Nx * Ny;
    const double dx = (p->x1 - p->x0) / Nx, dy = (p->y1 - p->y0) / Ny;
    Prim* W = (Prim*)malloc(N * sizeof(Prim));
    for (int j = 0; j < Ny; ++j)
        for (int i = 0; i < Nx; ++i) {
            size_t k = idx(i, j, Nx);
            double rho = U->rho[k], mx = U->mx[k], my = U->my[k], E = U->E[k];
            enforce_floors(p, &rho, &mx, &my, &E);
            W[k] = cons_to_prim(p, rho, mx, my, E);
        }
    Flux* Fx = (Flux*)malloc((size_t)(Nx + 1) * Ny * sizeof(Flux));
    Flux* Fy = (Flux*)malloc((size_t)Nx * (Ny + 1) * sizeof(Flux));
    for (int j = 0; j < Ny; ++j)
        for (int i = 0; i <= Nx; ++i) {
            Prim L, R;
            face_states(p, W, i, j, 0, &L, &R);
            Fx[(size_t)j * (Nx + 1) + i] = hllc(p, L, R, 0);
        }
    for (int j = 0; j <= Ny; ++j)
        for (int i = 0; i < Nx; ++i) {
            Prim L, R;
            face_states(p, W, i, j, 1, &L, &R);
            Fy[(size_t)j * Nx + i] = hllc(p, L, R, 1);
        }
    for (int j = 0; j < Ny; ++j)
        for (int i = 0; i < Nx; ++i) {
            size_t k = idx(i, j, Nx);
            Flux xR = Fx[(size_t)j * (Nx + 1) + i + 1], xL = Fx[(size_t)j * (Nx + 1) + i];
            Flux yU = Fy[(size_t)(j + 1) * Nx + i], yD = Fy[(size_t)j * Nx + i];
            dU->rho[k] = -(xR.f1 - xL.f1) / dx - (yU.f1 - yD.f1) / dy;
            dU->mx[k] = -(xR.f2 - xL.f2) / dx - (yU.f2 - yD.f2) / dy;
            dU->my[k] = -(xR.f3 - xL.f3) / dx - (yU.f3 - yD.f3) / dy;
            dU->E[k] = -(xR.f4 - xL.f4) / dx - (yU.f4 - yD.f4) / dy;
            if (p->gravity_on) {
                double rho = dmax(U->rho[k], p->floor_rho);
                double u = U->mx[k] / rho, v = U->my[k] / rho;
                dU->mx[k] += rho * gx[k];
                dU->my[k] += rho * gy[k];
                dU->E[k] += rho * (u * gx[k] + v * gy[k]);
            }
        }
    free(W); free(Fx); free(Fy);
}

void or_eh_rhs_g(const or_eh_params* p, const or_eh_state* U, or_eh_state* dU, const double* gx, const double* gy) {
    rhs_with(p, U, dU, gx, gy);
}

void or_eh_rhs(const or_eh_params* p, const or_eh_state* U, or_eh_state* dU) {
    const size_t N = (size_t)p->Nx * p->Ny;
    double* gx = (double*)malloc(N * 8);
    double* gy = (double*)malloc(N * 8);
    or_eh_poisson(p, U, gx, gy, NULL);
    rhs_with(p, U, dU, gx, gy);
    free(gx); free(gy);
}

typedef struct { double *rho, *mx, *my, *E; } Buf;
static void buf_alloc(or_eh_state* s, size_t N) {
    s->rho = (double*)malloc(N * 8); s->mx = (double*)malloc(N * 8);
    s->my = (double*)malloc(N * 8); s->E = (double*)malloc(N * 8);
}
static void buf_free(or_eh_state* s) { free(s->rho); free(s->mx); free(s->my); free(s->E); }

/* EH:529-550 */
void or_eh_rk2_step(const or_eh_params* p, or_eh_state* U, double dt) {
    const size_t N = (size_t)p->Nx * p->Ny;
    or_eh_state k1, U1, k2;
    buf_alloc(&k1, N); buf_alloc(&U1, N); buf_alloc(&k2, N);
    or_eh_rhs(p, U, &k1);
    memcpy(U1.rho, U->rho, N * 8); memcpy(U1.mx, U->mx, N * 8);
    memcpy(U1.my, U->my, N * 8); memcpy(U1.E, U->E, N * 8);
    for (size_t q = 0; q < N; ++q) {
        U1.rho[q] += dt * k1.rho[q];
        U1.mx[q] += dt * k1.mx[q];
        U1.my[q] += dt * k1.my[q];
        U1.E[q] += dt * k1.E[q];
        enforce_floors(p, &U1.rho[q], &U1.mx[q], &U1.my[q], &U1.E[q]);
    }
    or_eh_rhs(p, &U1, &k2);
    for (size_t q = 0; q < N; ++q) {
        U->rho[q] = 0.5 * U->rho[q] + 0.5 * (U1.rho[q] + dt * k2.rho[q]);
        U->mx[q] = 0.5 * U->mx[q] + 0.5 * (U1.mx[q] + dt * k2.mx[q]);
        U->my[q] = 0.5 * U->my[q] + 0.5 * (U1.my[q] + dt * k2.my[q]);
        U->E[q] = 0.5 * U->E[q] + 0.5 * (U1.E[q] + dt * k2.E[q]);
        enforce_floors(p, &U->rho[q], &U->mx[q], &U->my[q], &U->E[q]);
    }
    buf_free(&k1); buf_free(&U1); buf_free(&k2);
}

/* Driver loop EH:697-708 without the snapshots: dt, "Bad dt" check, clamp to t_end, rk2_step.
 * Returns the final t, or -1 when dt is not finite and positive (the program returns 1). */
double or_eh_run(const or_eh_params* p, or_eh_state* U, int nsteps, double t, double t_end, double* dts_out) {
    for (int s = 0; s < nsteps && t < t_end - 1e-14; ++s) {
        double dt = or_eh_compute_dt(p, U);
        if (!isfinite(dt) || dt <= 0.0) return -1.0;
        if (t + dt > t_end) dt = t_end - t;
        or_eh_rk2_step(p, U, dt);
        t += dt;
        if (dts_out) dts_out[s] = dt;
    }
    return t;
}

/* The snapshot frame (the store lambda of EH:661-691): floors on a copy, cons_to_prim_full, planes
 * {rho, P, Mach}. */
void or_eh_frame(const or_eh_params* p, const or_eh_state* U, double* out) {
    const size_t N = (size_t)p->Nx * (size_t)p->Ny;
    for (int j = 0; j < p->Ny; j++)
        for (int i = 0; i < p->Nx; i++) {
            const size_t q = idx(i, j, p->Nx);
            double rho = U->rho[q], mx = U->mx[q], my = U->my[q], E = U->E[q];
            enforce_floors(p, &rho, &mx, &my, &E);
            double u, v, P, a;
            cons_to_prim_full(p, rho, mx, my, E, &u, &v, &P, &a);
            const double vmag = sqrt(u * u + v * v);
            out[q] = rho;
            out[N + q] = P;
            out[2 * N + q] = (a > 0.0) ? (vmag / a) : 0.0;
        }
}
