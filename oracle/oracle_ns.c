/*
 * oracle_ns.c -- CPU restatement of the 2D Navier-Stokes HLLC + FFT-gravity step of
 * navier_stokes_2d_hllc_poisson.cpp (NS), in FP64 with the reference's operation order.
 *
 * Test infrastructure only: loaded by tests/, __graft_entry__.smoke() and bench.py's CPU leg
 * as the checker; the product library never uses it.
 *
 * Complex arithmetic follows what the reference's std::complex<double> evaluates without
 * -ffast-math or FMA: (a+ib)(c+id) = (ac - bd) + i(ad + bc); complex / real and complex * real
 * scale both parts.  Pinned by tests/golden/ns_*.npz and tests/test_oracle_ns2d.py.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "oracle.h"

static double dmax(double a, double b) { return (a < b) ? b : a; }
static double dmin(double a, double b) { return (b < a) ? b : a; }
static double clamp_min(double x, double lo) { return (x < lo) ? lo : x; } /* NS:51 */
static int imod(int a, int n) { int r = a % n; return (r < 0) ? (r + n) : r; } /* NS:113-116 */

typedef struct { double d, mx, my, E; } U4;
typedef struct { double d, u, v, P; } W4;

/* NS:119-132 */
static W4 cons_to_prim(const or_ns_params* p, const double* U) {
    W4 W;
    W.d = clamp_min(U[0], p->floor_rho);
    W.u = U[1] / W.d;
    W.v = U[2] / W.d;
    double ke = 0.5 * W.d * (W.u * W.u + W.v * W.v);
    double eint = U[3] - ke;
    double P = (p->gamma - 1.0) * eint;
    W.P = clamp_min(P, p->floor_P);
    return W;
}
/* NS:147-149 */
static double sound_speed(const or_ns_params* p, W4 W) { return sqrt(p->gamma * W.P / clamp_min(W.d, p->floor_rho)); }

/* ---------------- FFT NS:152-195 ---------------- */
static void fft1d(double* re, double* im, int n, int inverse) {
    for (int i = 1, j = 0; i < n; i++) {
        int bit = n >> 1;
        for (; j & bit; bit >>= 1) j ^= bit;
        j ^= bit;
        if (i < j) {
            double t = re[i]; re[i] = re[j]; re[j] = t;
            t = im[i]; im[i] = im[j]; im[j] = t;
        }
    }
    for (int len = 2; len <= n; len <<= 1) {
        const double ang = 2.0 * M_PI / len * (inverse ? +1.0 : -1.0);
        const double wr_len = cos(ang), wi_len = sin(ang);
        for (int i = 0; i < n; i += len) {
            double wr = 1.0, wi = 0.0;
            for (int j = 0; j < len / 2; j++) {
                int a = i + j, b = i + j + len / 2;
                double ur = re[a], ui = im[a];
                double vr = re[b] * wr - im[b] * wi;
                double vi = re[b] * wi + im[b] * wr;
                re[a] = ur + vr; im[a] = ui + vi;
                re[b] = ur - vr; im[b] = ui - vi;
                double nr = wr * wr_len - wi * wi_len;
                double ni = wr * wi_len + wi * wr_len;
                wr = nr; wi = ni;
            }
        }
    }
    if (inverse)
        for (int i = 0; i < n; i++) { re[i] /= (double)n; im[i] /= (double)n; }
}

/* fft2d NS:182-195 on interleaved complex data[2*(i + Nx*j) + {0,1}]: rows, then columns. */
void or_ns_fft2d(int Nx, int Ny, int inverse, double* data) {
    int L = Nx > Ny ? Nx : Ny;
    double* re = (double*)malloc(L * 8);
    double* im = (double*)malloc(L * 8);
    for (int j = 0; j < Ny; j++) {
        for (int i = 0; i < Nx; i++) { re[i] = data[2 * (i + Nx * j)]; im[i] = data[2 * (i + Nx * j) + 1]; }
        fft1d(re, im, Nx, inverse);
        for (int i = 0; i < Nx; i++) { data[2 * (i + Nx * j)] = re[i]; data[2 * (i + Nx * j) + 1] = im[i]; }
    }
    for (int i = 0; i < Nx; i++) {
        for (int j = 0; j < Ny; j++) { re[j] = data[2 * (i + Nx * j)]; im[j] = data[2 * (i + Nx * j) + 1]; }
        fft1d(re, im, Ny, inverse);
        for (int j = 0; j < Ny; j++) { data[2 * (i + Nx * j)] = re[j]; data[2 * (i + Nx * j) + 1] = im[j]; }
    }
    free(re); free(im);
}

/* ---------------- HLLC NS:198-282 (textbook p*, SM guard |den| > 0) ---------------- */
static U4 flux_x_from_prim(const or_ns_params* p, W4 W) {
    double E = W.P / (p->gamma - 1.0) + 0.5 * W.d * (W.u * W.u + W.v * W.v);
    U4 F = {W.d * W.u, W.d * W.u * W.u + W.P, W.d * W.u * W.v, (E + W.P) * W.u};
    return F;
}
static U4 hllc_x(const or_ns_params* p, W4 WL, W4 WR) {
    const double aL = sound_speed(p, WL), aR = sound_speed(p, WR);
    const double uL = WL.u, vL = WL.v, dL = WL.d, pL = WL.P;
    const double uR = WR.u, vR = WR.v, dR = WR.d, pR = WR.P;
    const double EL = pL / (p->gamma - 1.0) + 0.5 * dL * (uL * uL + vL * vL);
    const double ER = pR / (p->gamma - 1.0) + 0.5 * dR * (uR * uR + vR * vR);
    const double SL = dmin(uL - aL, uR - aR);
    const double SR = dmax(uL + aL, uR + aR);
    const double num = (pR - pL) + dL * uL * (SL - uL) - dR * uR * (SR - uR);
    const double den = dL * (SL - uL) - dR * (SR - uR);
    const double SM = (fabs(den) > 0.0) ? (num / den) : 0.0;
    const U4 FL = flux_x_from_prim(p, WL), FR = flux_x_from_prim(p, WR);
    if (0.0 <= SL) return FL;
    if (0.0 >= SR) return FR;
    const int left = 0.0 < SM;
    const W4 W = left ? WL : WR;
    const double S = left ? SL : SR, E = left ? EL : ER;
    const double fac = W.d * (S - W.u) / (S - SM);
    const double pstar = W.P + W.d * (S - W.u) * (SM - W.u);
    const double Es = ((S - W.u) * E - W.P * W.u + pstar * SM) / (S - SM);
    const U4 F0 = left ? FL : FR;
    U4 F = {F0.d + S * (fac - W.d), F0.mx + S * (fac * SM - W.d * W.u), F0.my + S * (fac * W.v - W.d * W.v),
            F0.E + S * (Es - E)};
    return F;
}
static U4 hllc_y(const or_ns_params* p, W4 WL, W4 WR) {
    W4 AL = WL, AR = WR;
    AL.u = WL.v; AL.v = WL.u;
    AR.u = WR.v; AR.v = WR.u;
    U4 Fx = hllc_x(p, AL, AR);
    U4 G = {Fx.d, Fx.my, Fx.mx, Fx.E};
    return G;
}
void or_ns_hllc_x(const or_ns_params* p, const double* a, const double* b, double* o) {
    W4 L = {a[0], a[1], a[2], a[3]}, R = {b[0], b[1], b[2], b[3]};
    U4 F = hllc_x(p, L, R);
    o[0] = F.d; o[1] = F.mx; o[2] = F.my; o[3] = F.E;
}
void or_ns_hllc_y(const or_ns_params* p, const double* a, const double* b, double* o) {
    W4 L = {a[0], a[1], a[2], a[3]}, R = {b[0], b[1], b[2], b[3]};
    U4 F = hllc_y(p, L, R);
    o[0] = F.d; o[1] = F.mx; o[2] = F.my; o[3] = F.E;
}

static int is_pow2(int n) { return (n > 0) && ((n & (n - 1)) == 0); }

/* solve_poisson_periodic_fft NS:285-335: continuous k = 2 pi n / L, k = 0 mode zeroed. */
int or_ns_poisson_fft(const or_ns_params* p, const double* rho, double* phi) {
    const int Nx = p->Nx, Ny = p->Ny, N = Nx * Ny;
    if (!is_pow2(Nx) || !is_pow2(Ny)) return 1;
    double mean = 0.0;
    for (int q = 0; q < N; q++) mean += rho[q];
    mean /= (double)N;
    double* A = (double*)malloc((size_t)N * 16);
    for (int q = 0; q < N; q++) { A[2 * q] = rho[q] - mean; A[2 * q + 1] = 0.0; }
    or_ns_fft2d(Nx, Ny, 0, A);
    const double Lx = p->x1 - p->x0, Ly = p->y1 - p->y0;
    const double two_pi = 2.0 * M_PI;
    for (int j = 0; j < Ny; j++) {
        int nj = (j <= Ny / 2) ? j : j - Ny;
        const double ky = two_pi * (double)nj / Ly;
        for (int i = 0; i < Nx; i++) {
            int ni = (i <= Nx / 2) ? i : i - Nx;
            const double kx = two_pi * (double)ni / Lx;
            const double k2 = kx * kx + ky * ky;
            const int q = i + Nx * j;
            if (k2 == 0.0) {
                A[2 * q] = 0.0; A[2 * q + 1] = 0.0;
            } else {
                const double s = (-(4.0 * M_PI * p->G) / k2);
                A[2 * q] *= s; A[2 * q + 1] *= s;
            }
        }
    }
    or_ns_fft2d(Nx, Ny, 1, A);
    for (int q = 0; q < N; q++) phi[q] = A[2 * q];
    free(A);
    return 0;
}

/* compute_gravity_accel NS:337-360 */
void or_ns_gravity_accel(const or_ns_params* p, const double* phi, double* gx, double* gy) {
    const int Nx = p->Nx, Ny = p->Ny;
    const double dx = (p->x1 - p->x0) / Nx, dy = (p->y1 - p->y0) / Ny;
    const double idx2 = 1.0 / (2.0 * dx), idy2 = 1.0 / (2.0 * dy);
    for (int j = 0; j < Ny; j++) {
        int jp = imod(j + 1, Ny), jm = imod(j - 1, Ny);
        for (int i = 0; i < Nx; i++) {
            int ip = imod(i + 1, Nx), im = imod(i - 1, Nx);
            gx[i + Nx * j] = -(phi[ip + Nx * j] - phi[im + Nx * j]) * idx2;
            gy[i + Nx * j] = -(phi[i + Nx * jp] - phi[i + Nx * jm]) * idy2;
        }
    }
}

/* compute_rhs NS:363-518 on AoS U[4*q + c], q = i + Nx*j. */
void or_ns_compute_rhs(const or_ns_params* p, const double* U, double* dU) {
    const int Nx = p->Nx, Ny = p->Ny, N = Nx * Ny;
    const double dx = (p->x1 - p->x0) / Nx, dy = (p->y1 - p->y0) / Ny;
    W4* W = (W4*)malloc((size_t)N * sizeof(W4));
    double* rho = (double*)malloc((size_t)N * 8);
    for (int q = 0; q < N; q++) { W[q] = cons_to_prim(p, U + 4 * q); rho[q] = W[q].d; }
    double *phi = NULL, *gx = NULL, *gy = NULL;
    if (p->gravity_on) {
        phi = (double*)malloc((size_t)N * 8); gx = (double*)malloc((size_t)N * 8); gy = (double*)malloc((size_t)N * 8);
        or_ns_poisson_fft(p, rho, phi);
        or_ns_gravity_accel(p, phi, gx, gy);
    }
    U4* Fx = (U4*)malloc((size_t)N * sizeof(U4));
    U4* Fy = (U4*)malloc((size_t)N * sizeof(U4));
    for (int j = 0; j < Ny; j++)
        for (int i = 0; i < Nx; i++) Fx[i + Nx * j] = hllc_x(p, W[i + Nx * j], W[imod(i + 1, Nx) + Nx * j]);
    for (int j = 0; j < Ny; j++) {
        int jp = imod(j + 1, Ny);
        for (int i = 0; i < Nx; i++) Fy[i + Nx * j] = hllc_y(p, W[i + Nx * j], W[i + Nx * jp]);
    }
    if (p->viscosity_on && p->mu > 0.0) {
        double* du_dx = (double*)malloc((size_t)N * 8); double* du_dy = (double*)malloc((size_t)N * 8);
        double* dv_dx = (double*)malloc((size_t)N * 8); double* dv_dy = (double*)malloc((size_t)N * 8);
        const double idx2 = 1.0 / (2.0 * dx), idy2 = 1.0 / (2.0 * dy);
        for (int j = 0; j < Ny; j++) {
            int jp = imod(j + 1, Ny), jm = imod(j - 1, Ny);
            for (int i = 0; i < Nx; i++) {
                int ip = imod(i + 1, Nx), im = imod(i - 1, Nx);
                int q = i + Nx * j, qip = ip + Nx * j, qim = im + Nx * j, qjp = i + Nx * jp, qjm = i + Nx * jm;
                du_dx[q] = (W[qip].u - W[qim].u) * idx2;
                dv_dx[q] = (W[qip].v - W[qim].v) * idx2;
                du_dy[q] = (W[qjp].u - W[qjm].u) * idy2;
                dv_dy[q] = (W[qjp].v - W[qjm].v) * idy2;
            }
        }
        for (int pass = 0; pass < 2; pass++)
            for (int j = 0; j < Ny; j++) {
                int jp = imod(j + 1, Ny);
                for (int i = 0; i < Nx; i++) {
                    int qL = i + Nx * j, qR = pass == 0 ? imod(i + 1, Nx) + Nx * j : i + Nx * jp;
                    const double dudx = 0.5 * (du_dx[qL] + du_dx[qR]);
                    const double dudy = 0.5 * (du_dy[qL] + du_dy[qR]);
                    const double dvdx = 0.5 * (dv_dx[qL] + dv_dx[qR]);
                    const double dvdy = 0.5 * (dv_dy[qL] + dv_dy[qR]);
                    const double divv = dudx + dvdy;
                    const double mu = p->mu;
                    const double uf = 0.5 * (W[qL].u + W[qR].u);
                    const double vf = 0.5 * (W[qL].v + W[qR].v);
                    const double tau_xy = mu * (dudy + dvdx);
                    if (pass == 0) { /* x-faces NS:430-455 */
                        const double tau_xx = mu * (2.0 * dudx - (2.0 / 3.0) * divv);
                        U4* F = &Fx[i + Nx * j];
                        F->mx += -tau_xx;
                        F->my += -tau_xy;
                        F->E += -(uf * tau_xx + vf * tau_xy);
                    } else { /* y-faces NS:457-481 */
                        const double tau_yy = mu * (2.0 * dvdy - (2.0 / 3.0) * divv);
                        U4* G = &Fy[i + Nx * j];
                        G->mx += -tau_xy;
                        G->my += -tau_yy;
                        G->E += -(uf * tau_xy + vf * tau_yy);
                    }
                }
            }
        free(du_dx); free(du_dy); free(dv_dx); free(dv_dy);
    }
    const double idx1 = 1.0 / dx, idy1 = 1.0 / dy;
    for (int j = 0; j < Ny; j++) {
        int jm = imod(j - 1, Ny);
        for (int i = 0; i < Nx; i++) {
            int im = imod(i - 1, Nx), q = i + Nx * j;
            const U4 Fp = Fx[q], Fm = Fx[im + Nx * j], Gp = Fy[q], Gm = Fy[i + Nx * jm];
            double R[4];
            R[0] = -(Fp.d - Fm.d) * idx1 - (Gp.d - Gm.d) * idy1;
            R[1] = -(Fp.mx - Fm.mx) * idx1 - (Gp.mx - Gm.mx) * idy1;
            R[2] = -(Fp.my - Fm.my) * idx1 - (Gp.my - Gm.my) * idy1;
            R[3] = -(Fp.E - Fm.E) * idx1 - (Gp.E - Gm.E) * idy1;
            if (p->gravity_on) {
                const double d = W[q].d, u = W[q].u, v = W[q].v;
                R[1] += d * gx[q];
                R[2] += d * gy[q];
                R[3] += d * (u * gx[q] + v * gy[q]);
            }
            memcpy(dU + 4 * q, R, 32);
        }
    }
    free(W); free(rho); free(Fx); free(Fy);
    if (phi) { free(phi); free(gx); free(gy); }
}

/* compute_dt NS:521-546 */
double or_ns_compute_dt(const or_ns_params* p, const double* U) {
    const int Nx = p->Nx, Ny = p->Ny, N = Nx * Ny;
    const double dx = (p->x1 - p->x0) / Nx, dy = (p->y1 - p->y0) / Ny;
    double smax = 0.0, rho_min = 1e300;
    for (int q = 0; q < N; q++) {
        W4 W = cons_to_prim(p, U + 4 * q);
        const double a = sound_speed(p, W);
        smax = dmax(smax, fabs(W.u) + a);
        smax = dmax(smax, fabs(W.v) + a);
        rho_min = dmin(rho_min, W.d);
    }
    const double h = dmin(dx, dy);
    double dt_hyp = (smax > 0.0) ? (p->cfl * h / smax) : 1e-6;
    double dt = dt_hyp;
    if (p->viscosity_on && p->mu > 0.0) {
        const double nu_max = p->mu / clamp_min(rho_min, p->floor_rho);
        const double dt_visc = (nu_max > 0.0) ? (p->cfl_visc * h * h / nu_max) : dt_hyp;
        dt = dmin(dt, dt_visc);
    }
    return dt;
}

/* Heun step of the driver loop NS:758-781 (rho floor only, no energy floor). */
void or_ns_heun_step(const or_ns_params* p, double* U, double dt) {
    const int N = p->Nx * p->Ny;
    double* k1 = (double*)malloc((size_t)N * 32);
    double* k2 = (double*)malloc((size_t)N * 32);
    double* U1 = (double*)malloc((size_t)N * 32);
    or_ns_compute_rhs(p, U, k1);
    for (int q = 0; q < N; q++) {
        for (int c = 0; c < 4; c++) U1[4 * q + c] = U[4 * q + c] + dt * k1[4 * q + c];
        if (U1[4 * q] < p->floor_rho) U1[4 * q] = p->floor_rho;
    }
    or_ns_compute_rhs(p, U1, k2);
    for (int q = 0; q < N; q++) {
        for (int c = 0; c < 4; c++) U[4 * q + c] = U[4 * q + c] + 0.5 * dt * (k1[4 * q + c] + k2[4 * q + c]);
        if (U[4 * q] < p->floor_rho) U[4 * q] = p->floor_rho;
    }
    free(k1); free(k2); free(U1);
}

/* Driver loop NS:758-781 without the snapshots. */
double or_ns_run(const or_ns_params* p, double* U, int nsteps, double t, double t_end, double* dts_out) {
    for (int s = 0; s < nsteps && t < t_end - 1e-14; ++s) {
        double dt = or_ns_compute_dt(p, U);
        if (t + dt > t_end) dt = t_end - t;
        or_ns_heun_step(p, U, dt);
        t += dt;
        if (dts_out) dts_out[s] = dt;
    }
    return t;
}

/* The snapshot frame (the store lambda of NS:698-749): planes {rho, P, Mach, u, v, momx, momy, Phi},
 * Phi by the FFT solve of the frame's density (zero without gravity).  Returns the solver's status. */
int or_ns_frame(const or_ns_params* p, const double* U, double* out) {
    const size_t N = (size_t)p->Nx * (size_t)p->Ny;
    for (size_t q = 0; q < N; q++) {
        const W4 W = cons_to_prim(p, U + 4 * q);
        const double a = sound_speed(p, W);
        const double vmag = sqrt(W.u * W.u + W.v * W.v);
        out[q] = W.d;
        out[N + q] = W.P;
        out[2 * N + q] = (a > 0.0) ? (vmag / a) : 0.0;
        out[3 * N + q] = W.u;
        out[4 * N + q] = W.v;
        out[5 * N + q] = W.d * W.u;
        out[6 * N + q] = W.d * W.v;
    }
    if (!p->gravity_on) {
        for (size_t q = 0; q < N; q++) out[7 * N + q] = 0.0;
        return 0;
    }
    return or_ns_poisson_fft(p, out, out + 7 * N);
}
