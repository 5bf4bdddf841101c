/*
 * oracle_e3.c -- CPU restatement of the 3D Euler HLLC + FFT self-gravity step (the periodic
 * uniform-sphere collapse benchmark), in FP64.
 *
 * Test infrastructure only: loaded by tests/, __graft_entry__.smoke() and bench.py's CPU leg
 * as the checker; the product library never uses it.
 *
 * No reference program has this step; it is the 2D Euler step of euler2d_gravity_hllc.cpp
 * (EH:406-550) extended to three axes, with the SOR solve replaced by a periodic spectral solve:
 *   - prims from floored cons (EH:62-100, kinetic energy 0.5 rho ((u^2 + v^2) + w^2));
 *   - per axis, 4-point MUSCL-MC face states (EH:437-455) and the HLLC of EH:125-219 with two
 *     transverse velocities (star energy ((S-u)E - P u + P S_M)/(S - S_M), S_M guard 1e-14); the
 *     cross momentum flux of components a < b is rho*a*b;
 *   - dU = -(dF_x/dx) - (dF_y/dy) - (dF_z/dz) + (0, rho g, rho ((u gx + v gy) + w gz));
 *   - Heun with floors after each stage (EH:529-550); dt = cfl min(dx,dy,dz)/max(|u|+a,|v|+a,|w|+a);
 *   - gravity: rhs = 4 pi G (max(rho,floor) - mean); phi solves the 7-point discrete Laplacian
 *     L_h phi = rhs exactly (spectral division by the discrete eigenvalues
 *     -(4/h^2) sin^2(pi k/N) per axis, the k = 0 mode set to zero); g = -central grad phi.
 * On a state that is uniform along z with w = 0 and gz = 0 every operation reduces to the EH
 * step's in the same order, so the two agree bit for bit (tests/test_oracle_euler3d.py).
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "oracle.h"

static double dmax(double a, double b) { return (a < b) ? b : a; }
static double dmin(double a, double b) { return (b < a) ? b : a; }
static int wrapi(int i, int n) { int r = i % n; if (r < 0) r += n; return r; }

typedef struct { double rho, u, v, w, P; } Prim5;

static size_t cidx(const or_e3_params* p, int i, int j, int k) {
    return ((size_t)wrapi(k, p->Nz) * p->Ny + wrapi(j, p->Ny)) * p->Nx + wrapi(i, p->Nx);
}

static void floors5(const or_e3_params* p, double* rho, double* mx, double* my, double* mz, double* E) {
    *rho = dmax(*rho, p->floor_rho);
    double u = *mx / *rho, v = *my / *rho, w = *mz / *rho;
    double kin = 0.5 * *rho * ((u * u + v * v) + w * w);
    double P = (p->gamma - 1.0) * (*E - kin);
    if (P < p->floor_P) *E = kin + p->floor_P / (p->gamma - 1.0);
}

static void prim_full(const or_e3_params* p, double rho, double mx, double my, double mz, double E, double* u,
                      double* v, double* w, double* P, double* a) {
    rho = dmax(rho, p->floor_rho);
    *u = mx / rho; *v = my / rho; *w = mz / rho;
    double kin = 0.5 * rho * ((*u * *u + *v * *v) + *w * *w);
    *P = (p->gamma - 1.0) * (E - kin);
    if (*P < p->floor_P) *P = p->floor_P;
    *a = sqrt(p->gamma * *P / rho);
}

static double minmod(double a, double b) {
    if (a * b <= 0.0) return 0.0;
    return (fabs(a) < fabs(b)) ? a : b;
}
static double mc_limiter(double dm, double dp) { return minmod(0.5 * (dm + dp), minmod(2.0 * dm, 2.0 * dp)); }

/* velocity component c (0,1,2) of a primitive */
static double vel(const Prim5* W, int c) { return c == 0 ? W->u : (c == 1 ? W->v : W->w); }

static void cons5(const or_e3_params* p, const Prim5* W, double c[5]) {
    c[0] = dmax(W->rho, p->floor_rho);
    c[1] = c[0] * W->u;
    c[2] = c[0] * W->v;
    c[3] = c[0] * W->w;
    double kin = 0.5 * c[0] * ((W->u * W->u + W->v * W->v) + W->w * W->w);
    c[4] = kin + W->P / (p->gamma - 1.0);
}

/* physical flux along axis ax: {rho un, momenta (normal rho un un + P, cross rho a b for a < b), (E+P) un} */
static void flux5(const or_e3_params* p, const Prim5* W, int ax, double F[5]) {
    double E = W->P / (p->gamma - 1.0) + 0.5 * W->rho * ((W->u * W->u + W->v * W->v) + W->w * W->w);
    double un = vel(W, ax);
    F[0] = W->rho * un;
    for (int c = 0; c < 3; ++c) {
        if (c == ax) F[1 + c] = W->rho * un * un + W->P;
        else if (c < ax) F[1 + c] = W->rho * vel(W, c) * un;
        else F[1 + c] = W->rho * un * vel(W, c);
    }
    F[4] = (E + W->P) * un;
}

static void hllc5(const or_e3_params* p, const Prim5* WL, const Prim5* WR, int ax, double out[5]) {
    double aL = sqrt(p->gamma * WL->P / dmax(WL->rho, p->floor_rho));
    double aR = sqrt(p->gamma * WR->P / dmax(WR->rho, p->floor_rho));
    double unL = vel(WL, ax), unR = vel(WR, ax);
    double SL = dmin(unL - aL, unR - aR);
    double SR = dmax(unL + aL, unR + aR);
    double cL[5], cR[5], FL[5], FR[5];
    cons5(p, WL, cL);
    cons5(p, WR, cR);
    flux5(p, WL, ax, FL);
    flux5(p, WR, ax, FR);
    double num = (WR->P - WL->P) + cL[0] * unL * (SL - unL) - cR[0] * unR * (SR - unR);
    double den = cL[0] * (SL - unL) - cR[0] * (SR - unR);
    double SM = (fabs(den) > 1e-14) ? (num / den) : 0.0;
    if (0.0 <= SL) { memcpy(out, FL, sizeof FL); return; }
    if (0.0 >= SR) { memcpy(out, FR, sizeof FR); return; }
    int left = 0.0 < SM;
    const Prim5* W = left ? WL : WR;
    const double* c = left ? cL : cR;
    const double* F0 = left ? FL : FR;
    double S = left ? SL : SR;
    double un = vel(W, ax);
    double rS = c[0] * (S - un) / (S - SM);
    double ES = ((S - un) * c[4] - W->P * un + W->P * SM) / (S - SM);
    out[0] = F0[0] + S * (rS - c[0]);
    for (int m = 0; m < 3; ++m) {
        double mS = (m == ax) ? rS * SM : rS * vel(W, m);
        out[1 + m] = F0[1 + m] + S * (mS - c[1 + m]);
    }
    out[4] = F0[4] + S * (ES - c[4]);
}

void or_e3_hllc(const or_e3_params* p, int axis, const double* WL5, const double* WR5, double* out5) {
    Prim5 L = {WL5[0], WL5[1], WL5[2], WL5[3], WL5[4]}, R = {WR5[0], WR5[1], WR5[2], WR5[3], WR5[4]};
    hllc5(p, &L, &R, axis, out5);
}

/* MUSCL face states of the face between cells (c-1) and c along ax, from cells c-2 .. c+1 */
static void face_states(const or_e3_params* p, const Prim5* W, int i, int j, int k, int ax, Prim5* L, Prim5* R) {
    const int di = ax == 0, dj = ax == 1, dk = ax == 2;
    const Prim5* wm = &W[cidx(p, i - 2 * di, j - 2 * dj, k - 2 * dk)];
    const Prim5* w0 = &W[cidx(p, i - di, j - dj, k - dk)];
    const Prim5* wp = &W[cidx(p, i, j, k)];
    const Prim5* wpp = &W[cidx(p, i + di, j + dj, k + dk)];
    const double* am = &wm->rho; const double* a0 = &w0->rho;
    const double* ap = &wp->rho; const double* app = &wpp->rho;
    double l[5], r[5];
    for (int c = 0; c < 5; ++c) {
        double sL = mc_limiter(a0[c] - am[c], ap[c] - a0[c]);
        double sR = mc_limiter(ap[c] - a0[c], app[c] - ap[c]);
        l[c] = a0[c] + 0.5 * sL;
        r[c] = ap[c] - 0.5 * sR;
    }
    L->rho = dmax(l[0], p->floor_rho); L->u = l[1]; L->v = l[2]; L->w = l[3]; L->P = dmax(l[4], p->floor_P);
    R->rho = dmax(r[0], p->floor_rho); R->u = r[1]; R->v = r[2]; R->w = r[3]; R->P = dmax(r[4], p->floor_P);
}

void or_e3_rhs_g(const or_e3_params* p, const or_e3_state* U, or_e3_state* dU, const double* gx, const double* gy,
                 const double* gz) {
    const int Nx = p->Nx, Ny = p->Ny, Nz = p->Nz;
    const size_t N = (size_t)Nx * Ny * Nz;
    const double dx = (p->x1 - p->x0) / Nx, dy = (p->y1 - p->y0) / Ny, dz = (p->z1 - p->z0) / Nz;
    Prim5* W = (Prim5*)malloc(N * sizeof(Prim5));
    for (size_t q = 0; q < N; ++q) {
        double rho = U->rho[q], mx = U->mx[q], my = U->my[q], mz = U->mz[q], E = U->E[q];
        floors5(p, &rho, &mx, &my, &mz, &E);
        double u, v, w, P, a;
        prim_full(p, rho, mx, my, mz, E, &u, &v, &w, &P, &a);
        W[q].rho = dmax(rho, p->floor_rho); W[q].u = u; W[q].v = v; W[q].w = w; W[q].P = P;
    }
    /* face fluxes: F[ax][q][5], face q = lower face of cell q along ax (periodic: N faces per axis) */
    double* F[3];
    for (int ax = 0; ax < 3; ++ax) {
        F[ax] = (double*)malloc(N * 5 * sizeof(double));
        for (int k = 0; k < Nz; ++k)
            for (int j = 0; j < Ny; ++j)
                for (int i = 0; i < Nx; ++i) {
                    Prim5 L, R;
                    face_states(p, W, i, j, k, ax, &L, &R);
                    hllc5(p, &L, &R, ax, &F[ax][cidx(p, i, j, k) * 5]);
                }
    }
    double* out[5] = {dU->rho, dU->mx, dU->my, dU->mz, dU->E};
    for (int k = 0; k < Nz; ++k)
        for (int j = 0; j < Ny; ++j)
            for (int i = 0; i < Nx; ++i) {
                const size_t q = cidx(p, i, j, k);
                const double* xL = &F[0][q * 5]; const double* xR = &F[0][cidx(p, i + 1, j, k) * 5];
                const double* yD = &F[1][q * 5]; const double* yU = &F[1][cidx(p, i, j + 1, k) * 5];
                const double* zD = &F[2][q * 5]; const double* zU = &F[2][cidx(p, i, j, k + 1) * 5];
                double d[5];
                for (int c = 0; c < 5; ++c)
                    d[c] = -(xR[c] - xL[c]) / dx - (yU[c] - yD[c]) / dy - (zU[c] - zD[c]) / dz;
                if (p->gravity_on) {
                    double rho = dmax(U->rho[q], p->floor_rho);
                    double u = U->mx[q] / rho, v = U->my[q] / rho, w = U->mz[q] / rho;
                    d[1] += rho * gx[q];
                    d[2] += rho * gy[q];
                    d[3] += rho * gz[q];
                    d[4] += rho * ((u * gx[q] + v * gy[q]) + w * gz[q]);
                }
                for (int c = 0; c < 5; ++c) out[c][q] = d[c];
            }
    for (int ax = 0; ax < 3; ++ax) free(F[ax]);
    free(W);
}

/* ---- spectral Poisson ---- */
typedef struct { double re, im; } cplx;

/* in-place radix-2 FFT of n points at stride st (sign -1 forward, +1 inverse, unnormalised);
 * twiddles exp(sign 2 pi i m / n) from cos / sin of the exact angle */
static void fft1(cplx* a, int n, size_t st, int sign) {
    for (int i = 1, j = 0; i < n; ++i) {
        int bit = n >> 1;
        for (; j & bit; bit >>= 1) j ^= bit;
        j ^= bit;
        if (i < j) { cplx t = a[i * st]; a[i * st] = a[j * st]; a[j * st] = t; }
    }
    for (int len = 2; len <= n; len <<= 1) {
        for (int i = 0; i < n; i += len)
            for (int m = 0; m < len / 2; ++m) {
                const double ang = sign * 2.0 * M_PI * (double)m / (double)len;
                const double wr = cos(ang), wi = sin(ang);
                cplx* x0 = &a[(size_t)(i + m) * st];
                cplx* x1 = &a[(size_t)(i + m + len / 2) * st];
                const double br = x1->re * wr - x1->im * wi, bi = x1->re * wi + x1->im * wr;
                x1->re = x0->re - br; x1->im = x0->im - bi;
                x0->re += br; x0->im += bi;
            }
    }
}

static void fft3(const or_e3_params* p, cplx* a, int sign) {
    const int Nx = p->Nx, Ny = p->Ny, Nz = p->Nz;
    for (int k = 0; k < Nz; ++k)
        for (int j = 0; j < Ny; ++j) fft1(&a[((size_t)k * Ny + j) * Nx], Nx, 1, sign);
    for (int k = 0; k < Nz; ++k)
        for (int i = 0; i < Nx; ++i) fft1(&a[(size_t)k * Ny * Nx + i], Ny, (size_t)Nx, sign);
    for (int j = 0; j < Ny; ++j)
        for (int i = 0; i < Nx; ++i) fft1(&a[(size_t)j * Nx + i], Nz, (size_t)Nx * Ny, sign);
}

static double eig(int m, int n, double h) {
    const double s = sin(M_PI * (double)m / (double)n);
    return -4.0 / (h * h) * s * s;
}

void or_e3_poisson(const or_e3_params* p, const double* rho_in, double* gx, double* gy, double* gz, double* phi_out) {
    const int Nx = p->Nx, Ny = p->Ny, Nz = p->Nz;
    const size_t N = (size_t)Nx * Ny * Nz;
    if (!p->gravity_on) {
        memset(gx, 0, N * 8); memset(gy, 0, N * 8); memset(gz, 0, N * 8);
        if (phi_out) memset(phi_out, 0, N * 8);
        return;
    }
    const double dx = (p->x1 - p->x0) / Nx, dy = (p->y1 - p->y0) / Ny, dz = (p->z1 - p->z0) / Nz;
    double rhobar = 0.0;
    for (size_t q = 0; q < N; ++q) rhobar += dmax(rho_in[q], p->floor_rho);
    rhobar /= (double)N;
    cplx* a = (cplx*)malloc(N * sizeof(cplx));
    for (size_t q = 0; q < N; ++q) {
        a[q].re = 4.0 * M_PI * p->G * (dmax(rho_in[q], p->floor_rho) - rhobar);
        a[q].im = 0.0;
    }
    fft3(p, a, -1);
    for (int k = 0; k < Nz; ++k)
        for (int j = 0; j < Ny; ++j)
            for (int i = 0; i < Nx; ++i) {
                const size_t q = ((size_t)k * Ny + j) * Nx + i;
                const double lam = eig(i, Nx, dx) + eig(j, Ny, dy) + eig(k, Nz, dz);
                if (i == 0 && j == 0 && k == 0) { a[q].re = 0.0; a[q].im = 0.0; continue; }
                a[q].re /= lam; a[q].im /= lam;
            }
    fft3(p, a, +1);
    double* phi = (double*)malloc(N * 8);
    for (size_t q = 0; q < N; ++q) phi[q] = a[q].re / (double)N;
    free(a);
    for (int k = 0; k < Nz; ++k)
        for (int j = 0; j < Ny; ++j)
            for (int i = 0; i < Nx; ++i) {
                const size_t q = cidx(p, i, j, k);
                gx[q] = -(phi[cidx(p, i + 1, j, k)] - phi[cidx(p, i - 1, j, k)]) / (2.0 * dx);
                gy[q] = -(phi[cidx(p, i, j + 1, k)] - phi[cidx(p, i, j - 1, k)]) / (2.0 * dy);
                gz[q] = -(phi[cidx(p, i, j, k + 1)] - phi[cidx(p, i, j, k - 1)]) / (2.0 * dz);
            }
    if (phi_out) memcpy(phi_out, phi, N * 8);
    free(phi);
}

void or_e3_rhs(const or_e3_params* p, const or_e3_state* U, or_e3_state* dU) {
    const size_t N = (size_t)p->Nx * p->Ny * p->Nz;
    double* g = (double*)malloc(3 * N * 8);
    or_e3_poisson(p, U->rho, g, g + N, g + 2 * N, NULL);
    or_e3_rhs_g(p, U, dU, g, g + N, g + 2 * N);
    free(g);
}

double or_e3_compute_dt(const or_e3_params* p, const or_e3_state* U) {
    const size_t N = (size_t)p->Nx * p->Ny * p->Nz;
    const double dx = (p->x1 - p->x0) / p->Nx, dy = (p->y1 - p->y0) / p->Ny, dz = (p->z1 - p->z0) / p->Nz;
    double ms = 1e-30;
    for (size_t q = 0; q < N; ++q) {
        double u, v, w, P, a;
        prim_full(p, U->rho[q], U->mx[q], U->my[q], U->mz[q], U->E[q], &u, &v, &w, &P, &a);
        ms = dmax(ms, fabs(u) + a);
        ms = dmax(ms, fabs(v) + a);
        ms = dmax(ms, fabs(w) + a);
    }
    return p->cfl * dmin(dmin(dx, dy), dz) / ms;
}

static void alloc5(or_e3_state* s, size_t N) {
    s->rho = (double*)malloc(N * 8); s->mx = (double*)malloc(N * 8); s->my = (double*)malloc(N * 8);
    s->mz = (double*)malloc(N * 8); s->E = (double*)malloc(N * 8);
}
static void free5(or_e3_state* s) { free(s->rho); free(s->mx); free(s->my); free(s->mz); free(s->E); }

void or_e3_rk2_step(const or_e3_params* p, or_e3_state* U, double dt) {
    const size_t N = (size_t)p->Nx * p->Ny * p->Nz;
    or_e3_state k1, U1, k2;
    alloc5(&k1, N); alloc5(&U1, N); alloc5(&k2, N);
    or_e3_rhs(p, U, &k1);
    for (size_t q = 0; q < N; ++q) {
        U1.rho[q] = U->rho[q] + dt * k1.rho[q];
        U1.mx[q] = U->mx[q] + dt * k1.mx[q];
        U1.my[q] = U->my[q] + dt * k1.my[q];
        U1.mz[q] = U->mz[q] + dt * k1.mz[q];
        U1.E[q] = U->E[q] + dt * k1.E[q];
        floors5(p, &U1.rho[q], &U1.mx[q], &U1.my[q], &U1.mz[q], &U1.E[q]);
    }
    or_e3_rhs(p, &U1, &k2);
    for (size_t q = 0; q < N; ++q) {
        U->rho[q] = 0.5 * U->rho[q] + 0.5 * (U1.rho[q] + dt * k2.rho[q]);
        U->mx[q] = 0.5 * U->mx[q] + 0.5 * (U1.mx[q] + dt * k2.mx[q]);
        U->my[q] = 0.5 * U->my[q] + 0.5 * (U1.my[q] + dt * k2.my[q]);
        U->mz[q] = 0.5 * U->mz[q] + 0.5 * (U1.mz[q] + dt * k2.mz[q]);
        U->E[q] = 0.5 * U->E[q] + 0.5 * (U1.E[q] + dt * k2.E[q]);
        floors5(p, &U->rho[q], &U->mx[q], &U->my[q], &U->mz[q], &U->E[q]);
    }
    free5(&k1); free5(&U1); free5(&k2);
}

double or_e3_run(const or_e3_params* p, or_e3_state* U, int nsteps, double t, double t_end, double* dts_out) {
    for (int s = 0; s < nsteps && t < t_end; ++s) {
        double dt = or_e3_compute_dt(p, U);
        if (!isfinite(dt) || dt <= 0.0) return -1.0;
        if (t + dt > t_end) dt = t_end - t;
        or_e3_rk2_step(p, U, dt);
        t += dt;
        if (dts_out) dts_out[s] = dt;
    }
    return t;
}
