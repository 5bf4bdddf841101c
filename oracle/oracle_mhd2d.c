/*
 * oracle_mhd2d.c -- CPU FP64 restatement of the 2D resistive MHD solver step.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the parity oracle: only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it, and
 * only as the checker.  The product path (core_collapse_amd) never links,
 * loads or falls back to it.
 *
 * Every function restates, formula by formula and in the same operation
 * order, the reference program mhd2d_resistive_hlld_ct_gravity.cpp (cited
 * as M2:<line>).  Built with -O2 -ffp-contract=off and no -march (as the
 * reference Makefile builds, which gives no FMA on x86-64), so results are
 * expected to be bit-identical to the compiled reference; tests check this
 * against golden vectors made from the reference (tests/golden/) and, where
 * the reference builds, against oracle/_ref directly.
 *
 * Pinning (see tests/test_oracle_*.py): every public function here is
 * checked against reference-generated golden vectors; in addition
 * closed-form / invariant checks (uniform-state flux, supersonic upwinding,
 * CT divergence, brute-force matrix-splitting SOR) are applied.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include "oracle.h"

/* M2:41-42 */
static inline size_t idx2(int i, int j, int Nx) { return (size_t)j * (size_t)Nx + (size_t)i; }
static inline int imod(int i, int n) { int r = i % n; return (r < 0) ? r + n : r; }
static inline int iclamp(int v, int lo, int hi) { return v < lo ? lo : (v > hi ? hi : v); }
/* std::max / std::min semantics: max(a,b) = (a<b)?b:a ; min(a,b) = (b<a)?b:a */
static inline double dmax(double a, double b) { return (a < b) ? b : a; }
static inline double dmin(double a, double b) { return (b < a) ? b : a; }

typedef struct { double rho, u, v, w, P, Bx, By, Bz; } Prim;
typedef struct { double frho, fmx, fmy, fmz, fE, fBy, fBz, Ez; } FluxX;

/* M2:106-112 */
static inline double minmod(double a, double b) {
    if (a * b <= 0.0) return 0.0;
    return (fabs(a) < fabs(b)) ? a : b;
}
static inline double mc_limiter(double dm, double dp) {
    return minmod(0.5 * (dm + dp), minmod(2.0 * dm, 2.0 * dp));
}

/* M2:114-124 */
static inline void enforce_floors(const or_m2_params* p, double* rho, double* mx, double* my, double* mz,
                                  double* E, double Bx, double By, double Bz) {
    *rho = dmax(*rho, p->floor_rho);
    double u = *mx / *rho, v = *my / *rho, w = *mz / *rho;
    double kin = 0.5 * *rho * (u * u + v * v + w * w);
    double mag = 0.5 * (Bx * Bx + By * By + Bz * Bz);
    double P = (p->gamma - 1.0) * (*E - kin - mag);
    if (P < p->floor_P) {
        double eint = p->floor_P / (p->gamma - 1.0);
        *E = kin + mag + eint;
    }
}

/* M2:126-135 */
static inline Prim cons_to_prim(const or_m2_params* p, double rho, double mx, double my, double mz, double E,
                                double Bx, double By, double Bz) {
    rho = dmax(rho, p->floor_rho);
    double u = mx / rho, v = my / rho, w = mz / rho;
    double kin = 0.5 * rho * (u * u + v * v + w * w);
    double mag = 0.5 * (Bx * Bx + By * By + Bz * Bz);
    double P = (p->gamma - 1.0) * (E - kin - mag);
    if (P < p->floor_P) P = p->floor_P;
    Prim W = {rho, u, v, w, P, Bx, By, Bz};
    return W;
}

/* M2:148-174 */
static inline double sound_speed2(const or_m2_params* p, const Prim* W) {
    return p->gamma * W->P / dmax(W->rho, p->floor_rho);
}
static inline double fast_speed_n(const or_m2_params* p, const Prim* W, double Bn) {
    double a2 = sound_speed2(p, W);
    double b2 = W->Bx * W->Bx + W->By * W->By + W->Bz * W->Bz;
    double rho = dmax(W->rho, 1e-300);
    double vA2 = b2 / rho;
    double vAn2 = (Bn * Bn) / rho;
    double disc = (a2 + vA2) * (a2 + vA2) - 4.0 * a2 * vAn2;
    disc = dmax(0.0, disc);
    double cf2 = 0.5 * ((a2 + vA2) + sqrt(disc));
    return sqrt(dmax(0.0, cf2));
}

/* ------------------------------------------------------------------ */
/* Poisson gravity, lexicographic SOR (M2:178-237)                     */
/* ------------------------------------------------------------------ */
void or_m2_poisson(const or_m2_params* p, const double* rho_cc, double dx, double dy, double* gx, double* gy) {
    int Nx = p->Nx, Ny = p->Ny;
    size_t N = (size_t)Nx * (size_t)Ny;
    for (size_t q = 0; q < N; q++) { gx[q] = 0.0; gy[q] = 0.0; }
    if (!p->gravity_on) return;

    double rhobar = 0.0;
    for (size_t q = 0; q < N; q++) rhobar += rho_cc[q];
    rhobar /= (double)N;

    double* rhs = (double*)malloc(N * sizeof(double));
    double* phi = (double*)calloc(N, sizeof(double));
    for (size_t q = 0; q < N; q++) rhs[q] = 4.0 * M_PI * p->G * (rho_cc[q] - rhobar);

#define WIDX(ii, jj) idx2(p->periodic_x ? imod((ii), Nx) : iclamp((ii), 0, Nx - 1), \
                          p->periodic_y ? imod((jj), Ny) : iclamp((jj), 0, Ny - 1), Nx)
    double idx2v = 1.0 / (dx * dx), idy2v = 1.0 / (dy * dy);
    double denom = 2.0 * (idx2v + idy2v);
    double omega = p->sor_omega;

    for (int it = 0; it < p->sor_iters; ++it) {
        for (int j = 0; j < Ny; j++) {
            for (int i = 0; i < Nx; i++) {
                size_t k = WIDX(i, j);
                double phiE = phi[WIDX(i + 1, j)];
                double phiW = phi[WIDX(i - 1, j)];
                double phiN = phi[WIDX(i, j + 1)];
                double phiS = phi[WIDX(i, j - 1)];
                double phi_new = ((phiE + phiW) * idx2v + (phiN + phiS) * idy2v - rhs[k]) / denom;
                phi[k] += omega * (phi_new - phi[k]);
            }
        }
        double phibar = 0.0;
        for (size_t q = 0; q < N; q++) phibar += phi[q];
        phibar /= (double)N;
        for (size_t q = 0; q < N; q++) phi[q] -= phibar;
    }
    for (int j = 0; j < Ny; j++) {
        for (int i = 0; i < Nx; i++) {
            size_t k = WIDX(i, j);
            double phiE = phi[WIDX(i + 1, j)];
            double phiW = phi[WIDX(i - 1, j)];
            double phiN = phi[WIDX(i, j + 1)];
            double phiS = phi[WIDX(i, j - 1)];
            gx[k] = -(phiE - phiW) / (2.0 * dx);
            gy[k] = -(phiN - phiS) / (2.0 * dy);
        }
    }
#undef WIDX
    free(rhs);
    free(phi);
}

/* Same iteration, but returns phi (mean removed every sweep) for SOR-iterate tests. */
void or_m2_sor_phi(const or_m2_params* p, const double* rho_cc, double dx, double dy, double* phi_out) {
    int Nx = p->Nx, Ny = p->Ny;
    size_t N = (size_t)Nx * (size_t)Ny;
    double rhobar = 0.0;
    for (size_t q = 0; q < N; q++) rhobar += rho_cc[q];
    rhobar /= (double)N;
    double* rhs = (double*)malloc(N * sizeof(double));
    double* phi = phi_out;
    for (size_t q = 0; q < N; q++) { rhs[q] = 4.0 * M_PI * p->G * (rho_cc[q] - rhobar); phi[q] = 0.0; }
#define WIDX(ii, jj) idx2(p->periodic_x ? imod((ii), Nx) : iclamp((ii), 0, Nx - 1), \
                          p->periodic_y ? imod((jj), Ny) : iclamp((jj), 0, Ny - 1), Nx)
    double idx2v = 1.0 / (dx * dx), idy2v = 1.0 / (dy * dy);
    double denom = 2.0 * (idx2v + idy2v);
    double omega = p->sor_omega;
    for (int it = 0; it < p->sor_iters; ++it) {
        for (int j = 0; j < Ny; j++)
            for (int i = 0; i < Nx; i++) {
                size_t k = WIDX(i, j);
                double phiE = phi[WIDX(i + 1, j)], phiW = phi[WIDX(i - 1, j)];
                double phiN = phi[WIDX(i, j + 1)], phiS = phi[WIDX(i, j - 1)];
                double phi_new = ((phiE + phiW) * idx2v + (phiN + phiS) * idy2v - rhs[k]) / denom;
                phi[k] += omega * (phi_new - phi[k]);
            }
        double phibar = 0.0;
        for (size_t q = 0; q < N; q++) phibar += phi[q];
        phibar /= (double)N;
        for (size_t q = 0; q < N; q++) phi[q] -= phibar;
    }
#undef WIDX
    free(rhs);
}

/* ------------------------------------------------------------------ */
/* Face / cell helpers (M2:241-286)                                    */
/* ------------------------------------------------------------------ */
static inline double get_Bx_face(const or_m2_params* p, const double* Bx_f, int i_face, int j) {
    int Nx = p->Nx, Ny = p->Ny;
    if (p->periodic_x) i_face = imod(i_face, Nx + 1); else i_face = iclamp(i_face, 0, Nx);
    if (p->periodic_y) j = imod(j, Ny); else j = iclamp(j, 0, Ny - 1);
    return Bx_f[(size_t)j * (size_t)(Nx + 1) + (size_t)i_face];
}
static inline double get_By_face(const or_m2_params* p, const double* By_f, int i, int j_face) {
    int Nx = p->Nx, Ny = p->Ny;
    if (p->periodic_x) i = imod(i, Nx); else i = iclamp(i, 0, Nx - 1);
    if (p->periodic_y) j_face = imod(j_face, Ny + 1); else j_face = iclamp(j_face, 0, Ny);
    return By_f[(size_t)j_face * (size_t)Nx + (size_t)i];
}
static inline double bx_cc(const or_m2_params* p, const double* Bx_f, int i, int j) {
    int Nx = p->Nx, Ny = p->Ny;
    int fi0 = p->periodic_x ? imod(i, Nx + 1) : iclamp(i, 0, Nx);
    int fi1 = p->periodic_x ? imod(i + 1, Nx + 1) : iclamp(i + 1, 0, Nx);
    int fj = p->periodic_y ? imod(j, Ny) : iclamp(j, 0, Ny - 1);
    double bL = Bx_f[(size_t)fj * (size_t)(Nx + 1) + (size_t)fi0];
    double bR = Bx_f[(size_t)fj * (size_t)(Nx + 1) + (size_t)fi1];
    return 0.5 * (bL + bR);
}
static inline double by_cc(const or_m2_params* p, const double* By_f, int i, int j) {
    int Nx = p->Nx, Ny = p->Ny;
    int fi = p->periodic_x ? imod(i, Nx) : iclamp(i, 0, Nx - 1);
    int fj0 = p->periodic_y ? imod(j, Ny + 1) : iclamp(j, 0, Ny);
    int fj1 = p->periodic_y ? imod(j + 1, Ny + 1) : iclamp(j + 1, 0, Ny);
    double bD = By_f[(size_t)fj0 * (size_t)Nx + (size_t)fi];
    double bU = By_f[(size_t)fj1 * (size_t)Nx + (size_t)fi];
    return 0.5 * (bD + bU);
}
static inline Prim get_prim_cc(const or_m2_params* p, const or_m2_state* U, int i, int j) {
    int Nx = p->Nx, Ny = p->Ny;
    if (p->periodic_x) i = imod(i, Nx); else i = iclamp(i, 0, Nx - 1);
    if (p->periodic_y) j = imod(j, Ny); else j = iclamp(j, 0, Ny - 1);
    size_t q = idx2(i, j, Nx);
    double Bx = bx_cc(p, U->Bx_f, i, j);
    double By = by_cc(p, U->By_f, i, j);
    return cons_to_prim(p, U->rho[q], U->mx[q], U->my[q], U->mz[q], U->E[q], Bx, By, U->Bz[q]);
}

/* ------------------------------------------------------------------ */
/* PLM-MC reconstruction (M2:290-336)                                  */
/* ------------------------------------------------------------------ */
static void finish_recon(const or_m2_params* p, const Prim* Wc, const double s[8], Prim* WL, Prim* WR) {
    /* s[] holds slopes for rho,u,v,w,P,Bx,By,Bz; the normal component's slope is 0 and
       the reference keeps Wc's value unchanged for it (M2:305-308, 329-332). */
    WL->rho = Wc->rho - 0.5 * s[0]; WR->rho = Wc->rho + 0.5 * s[0];
    WL->u = Wc->u - 0.5 * s[1];     WR->u = Wc->u + 0.5 * s[1];
    WL->v = Wc->v - 0.5 * s[2];     WR->v = Wc->v + 0.5 * s[2];
    WL->w = Wc->w - 0.5 * s[3];     WR->w = Wc->w + 0.5 * s[3];
    WL->P = Wc->P - 0.5 * s[4];     WR->P = Wc->P + 0.5 * s[4];
    WL->rho = dmax(WL->rho, p->floor_rho); WR->rho = dmax(WR->rho, p->floor_rho);
    WL->P = dmax(WL->P, p->floor_P);       WR->P = dmax(WR->P, p->floor_P);
}
static void recon_x(const or_m2_params* p, const or_m2_state* U, int i, int j, Prim* WL, Prim* WR) {
    Prim Wm = get_prim_cc(p, U, i - 1, j), Wc = get_prim_cc(p, U, i, j), Wp = get_prim_cc(p, U, i + 1, j);
    double s[8];
    s[0] = mc_limiter(Wc.rho - Wm.rho, Wp.rho - Wc.rho);
    s[1] = mc_limiter(Wc.u - Wm.u, Wp.u - Wc.u);
    s[2] = mc_limiter(Wc.v - Wm.v, Wp.v - Wc.v);
    s[3] = mc_limiter(Wc.w - Wm.w, Wp.w - Wc.w);
    s[4] = mc_limiter(Wc.P - Wm.P, Wp.P - Wc.P);
    double sBy = mc_limiter(Wc.By - Wm.By, Wp.By - Wc.By);
    double sBz = mc_limiter(Wc.Bz - Wm.Bz, Wp.Bz - Wc.Bz);
    finish_recon(p, &Wc, s, WL, WR);
    WL->Bx = Wc.Bx; WR->Bx = Wc.Bx;
    WL->By = Wc.By - 0.5 * sBy; WR->By = Wc.By + 0.5 * sBy;
    WL->Bz = Wc.Bz - 0.5 * sBz; WR->Bz = Wc.Bz + 0.5 * sBz;
}
static void recon_y(const or_m2_params* p, const or_m2_state* U, int i, int j, Prim* WL, Prim* WR) {
    Prim Wm = get_prim_cc(p, U, i, j - 1), Wc = get_prim_cc(p, U, i, j), Wp = get_prim_cc(p, U, i, j + 1);
    double s[8];
    s[0] = mc_limiter(Wc.rho - Wm.rho, Wp.rho - Wc.rho);
    s[1] = mc_limiter(Wc.u - Wm.u, Wp.u - Wc.u);
    s[2] = mc_limiter(Wc.v - Wm.v, Wp.v - Wc.v);
    s[3] = mc_limiter(Wc.w - Wm.w, Wp.w - Wc.w);
    s[4] = mc_limiter(Wc.P - Wm.P, Wp.P - Wc.P);
    double sBx = mc_limiter(Wc.Bx - Wm.Bx, Wp.Bx - Wc.Bx);
    double sBz = mc_limiter(Wc.Bz - Wm.Bz, Wp.Bz - Wc.Bz);
    finish_recon(p, &Wc, s, WL, WR);
    WL->Bx = Wc.Bx - 0.5 * sBx; WR->Bx = Wc.Bx + 0.5 * sBx;
    WL->By = Wc.By; WR->By = Wc.By;
    WL->Bz = Wc.Bz - 0.5 * sBz; WR->Bz = Wc.Bz + 0.5 * sBz;
}

/* ------------------------------------------------------------------ */
/* HLLD, x-normal (M2:340-530)                                         */
/* ------------------------------------------------------------------ */
static inline double total_pressure(const Prim* W) {
    double B2 = W->Bx * W->Bx + W->By * W->By + W->Bz * W->Bz;
    return W->P + 0.5 * B2;
}
static FluxX flux_ideal_x(const or_m2_params* p, const Prim* W) {
    double rho = W->rho, u = W->u, v = W->v, w = W->w, P = W->P;
    double Bx = W->Bx, By = W->By, Bz = W->Bz;
    double v2 = u * u + v * v + w * w;
    double B2 = Bx * Bx + By * By + Bz * Bz;
    double pt = P + 0.5 * B2;
    double E = P / (p->gamma - 1.0) + 0.5 * rho * v2 + 0.5 * B2;
    double vdotB = u * Bx + v * By + w * Bz;
    FluxX F;
    F.frho = rho * u;
    F.fmx = rho * u * u + pt - Bx * Bx;
    F.fmy = rho * u * v - Bx * By;
    F.fmz = rho * u * w - Bx * Bz;
    F.fE = (E + pt) * u - Bx * vdotB;
    F.fBy = u * By - v * Bx;
    F.fBz = u * Bz - w * Bx;
    F.Ez = -(F.fBy);
    return F;
}
static void state7(const or_m2_params* p, const Prim* W, double out[7]) {
    /* M2:393-399 (cons7) and M2:480-484 (state7) are the same expression */
    double B2 = W->Bx * W->Bx + W->By * W->By + W->Bz * W->Bz;
    double E = W->P / (p->gamma - 1.0) + 0.5 * W->rho * (W->u * W->u + W->v * W->v + W->w * W->w) + 0.5 * B2;
    out[0] = W->rho; out[1] = W->rho * W->u; out[2] = W->rho * W->v; out[3] = W->rho * W->w;
    out[4] = E; out[5] = W->By; out[6] = W->Bz;
}

/* Branch taken by the last hlld_x call (test bookkeeping for stratified golden vectors):
 * 0 F_L (S_L >= 0), 1 F_R (S_R <= 0), 2 HLL (|Bn| < 1e-12), 3 U*_L, 4 U**_L, 5 U**_R, 6 U*_R. */
static int g_last_branch = -1;
int or_m2_last_branch(void) { return g_last_branch; }

static FluxX hlld_x(const or_m2_params* p, Prim WL, Prim WR, double Bn) {
    WL.Bx = Bn; WR.Bx = Bn;
    double rhoL = WL.rho, rhoR = WR.rho, uL = WL.u, uR = WR.u, vL = WL.v, vR = WR.v, wL = WL.w, wR = WR.w;
    double ByL = WL.By, ByR = WR.By, BzL = WL.Bz, BzR = WR.Bz;

    FluxX FL = flux_ideal_x(p, &WL);
    FluxX FR = flux_ideal_x(p, &WR);
    double cfL = fast_speed_n(p, &WL, WL.Bx);
    double cfR = fast_speed_n(p, &WR, WR.Bx);
    double SL = dmin(uL - cfL, uR - cfR);
    double SR = dmax(uL + cfL, uR + cfR);
    if (0.0 <= SL) { g_last_branch = 0; return FL; }
    if (0.0 >= SR) { g_last_branch = 1; return FR; }

    /* M2:393-401: cons7 uses rho*W.u etc. identical to state7 */
    double UL[7], UR[7];
    {
        double rho = WL.rho, mx = rho * WL.u, my = rho * WL.v, mz = rho * WL.w;
        double B2 = WL.Bx * WL.Bx + WL.By * WL.By + WL.Bz * WL.Bz;
        double E = WL.P / (p->gamma - 1.0) + 0.5 * rho * (WL.u * WL.u + WL.v * WL.v + WL.w * WL.w) + 0.5 * B2;
        UL[0] = rho; UL[1] = mx; UL[2] = my; UL[3] = mz; UL[4] = E; UL[5] = WL.By; UL[6] = WL.Bz;
    }
    {
        double rho = WR.rho, mx = rho * WR.u, my = rho * WR.v, mz = rho * WR.w;
        double B2 = WR.Bx * WR.Bx + WR.By * WR.By + WR.Bz * WR.Bz;
        double E = WR.P / (p->gamma - 1.0) + 0.5 * rho * (WR.u * WR.u + WR.v * WR.v + WR.w * WR.w) + 0.5 * B2;
        UR[0] = rho; UR[1] = mx; UR[2] = my; UR[3] = mz; UR[4] = E; UR[5] = WR.By; UR[6] = WR.Bz;
    }
    double ptL = total_pressure(&WL), ptR = total_pressure(&WR);

    double denom = rhoL * (SL - uL) - rhoR * (SR - uR);
    double SM = 0.0;
    if (fabs(denom) > 1e-14) SM = (ptR - ptL + rhoL * uL * (SL - uL) - rhoR * uR * (SR - uR)) / denom;

    double rhoL_s = rhoL * (SL - uL) / (SL - SM);
    double rhoR_s = rhoR * (SR - uR) / (SR - SM);
    double pt_s = ptL + rhoL * (SL - uL) * (SM - uL);

    double sL = SL - uL, sR = SR - uR;
    double dL = rhoL * sL * (SL - SM) - Bn * Bn;
    double dR = rhoR * sR * (SR - SM) - Bn * Bn;

    double ByL_s = ByL, BzL_s = BzL, ByR_s = ByR, BzR_s = BzR;
    if (fabs(dL) > 1e-14) {
        ByL_s = ByL * (rhoL * sL * (SL - uL) - Bn * Bn) / dL;
        BzL_s = BzL * (rhoL * sL * (SL - uL) - Bn * Bn) / dL;
    }
    if (fabs(dR) > 1e-14) {
        ByR_s = ByR * (rhoR * sR * (SR - uR) - Bn * Bn) / dR;
        BzR_s = BzR * (rhoR * sR * (SR - uR) - Bn * Bn) / dR;
    }
    double vL_s = vL, wL_s = wL, vR_s = vR, wR_s = wR;
    if (fabs(dL) > 1e-14) {
        vL_s = vL - Bn * (ByL_s - ByL) / (rhoL * (SL - uL) * (SL - SM));
        wL_s = wL - Bn * (BzL_s - BzL) / (rhoL * (SL - uL) * (SL - SM));
    }
    if (fabs(dR) > 1e-14) {
        vR_s = vR - Bn * (ByR_s - ByR) / (rhoR * (SR - uR) * (SR - SM));
        wR_s = wR - Bn * (BzR_s - BzR) / (rhoR * (SR - uR) * (SR - SM));
    }
    double sqrt_rhoL = sqrt(dmax(rhoL_s, 1e-300));
    double sqrt_rhoR = sqrt(dmax(rhoR_s, 1e-300));
    double SLL = SM - fabs(Bn) / sqrt_rhoL;
    double SRR = SM + fabs(Bn) / sqrt_rhoR;

    const double gd = p->hlld_guard;
    if (fabs(Bn) < 1e-12 || (gd > 0.0 && (fabs(dL) < gd * fmax(Bn * Bn, fabs(rhoL * sL * (SL - SM))) ||
                                          fabs(dR) < gd * fmax(Bn * Bn, fabs(rhoR * sR * (SR - SM)))))) {
        g_last_branch = 2;
        double inv = 1.0 / (SR - SL);
        double FUL[7] = {FL.frho, FL.fmx, FL.fmy, FL.fmz, FL.fE, FL.fBy, FL.fBz};
        double FUR[7] = {FR.frho, FR.fmx, FR.fmy, FR.fmz, FR.fE, FR.fBy, FR.fBz};
        double f[7];
        for (int k = 0; k < 7; k++) f[k] = (SR * FUL[k] - SL * FUR[k] + SL * SR * (UR[k] - UL[k])) * inv;
        FluxX FH = {f[0], f[1], f[2], f[3], f[4], f[5], f[6], 0.0};
        FH.Ez = -(FH.fBy);
        return FH;
    }

    double signBn = (Bn >= 0.0) ? 1.0 : -1.0;
    double denom2 = sqrt_rhoL + sqrt_rhoR;
    double v_ss = (sqrt_rhoL * vL_s + sqrt_rhoR * vR_s + (ByR_s - ByL_s) * signBn) / denom2;
    double w_ss = (sqrt_rhoL * wL_s + sqrt_rhoR * wR_s + (BzR_s - BzL_s) * signBn) / denom2;
    double By_ss = (sqrt_rhoL * ByR_s + sqrt_rhoR * ByL_s + sqrt_rhoL * sqrt_rhoR * (vR_s - vL_s) * signBn) / denom2;
    double Bz_ss = (sqrt_rhoL * BzR_s + sqrt_rhoR * BzL_s + sqrt_rhoL * sqrt_rhoR * (wR_s - wL_s) * signBn) / denom2;

    /* U_star (M2:486-505) evaluated only for the branch that is returned (pure function). */
    const Prim* Wref;
    double S, rho_s, v_s, w_s, By_s, Bz_s;
    const FluxX* Fref;
    if (0.0 < SLL)      { g_last_branch = 3; Wref = &WL; S = SL; rho_s = rhoL_s; v_s = vL_s; w_s = wL_s; By_s = ByL_s; Bz_s = BzL_s; Fref = &FL; }
    else if (0.0 < SM)  { g_last_branch = 4; Wref = &WL; S = SL; rho_s = rhoL_s; v_s = v_ss; w_s = w_ss; By_s = By_ss; Bz_s = Bz_ss; Fref = &FL; }
    else if (0.0 < SRR) { g_last_branch = 5; Wref = &WR; S = SR; rho_s = rhoR_s; v_s = v_ss; w_s = w_ss; By_s = By_ss; Bz_s = Bz_ss; Fref = &FR; }
    else                { g_last_branch = 6; Wref = &WR; S = SR; rho_s = rhoR_s; v_s = vR_s; w_s = wR_s; By_s = ByR_s; Bz_s = BzR_s; Fref = &FR; }

    double Ust[7];
    {
        double Bx = Bn, By = By_s, Bz = Bz_s;
        double B2W = Wref->Bx * Wref->Bx + Wref->By * Wref->By + Wref->Bz * Wref->Bz;
        double EW = Wref->P / (p->gamma - 1.0) +
                    0.5 * Wref->rho * (Wref->u * Wref->u + Wref->v * Wref->v + Wref->w * Wref->w) + 0.5 * B2W;
        double vdotB_s = SM * Bx + v_s * By + w_s * Bz;
        double vdotB_w = Wref->u * Wref->Bx + Wref->v * Wref->By + Wref->w * Wref->Bz;
        double E = ((S - Wref->u) * EW - (total_pressure(Wref) * Wref->u - Bn * vdotB_w) + (pt_s * SM - Bn * vdotB_s)) /
                   (S - SM);
        Ust[0] = rho_s; Ust[1] = rho_s * SM; Ust[2] = rho_s * v_s; Ust[3] = rho_s * w_s;
        Ust[4] = E; Ust[5] = By; Ust[6] = Bz;
    }
    double U0[7];
    state7(p, Wref, U0);
    FluxX F;
    F.frho = Fref->frho + S * (Ust[0] - U0[0]);
    F.fmx = Fref->fmx + S * (Ust[1] - U0[1]);
    F.fmy = Fref->fmy + S * (Ust[2] - U0[2]);
    F.fmz = Fref->fmz + S * (Ust[3] - U0[3]);
    F.fE = Fref->fE + S * (Ust[4] - U0[4]);
    F.fBy = Fref->fBy + S * (Ust[5] - U0[5]);
    F.fBz = Fref->fBz + S * (Ust[6] - U0[6]);
    F.Ez = -(F.fBy);
    return F;
}

/* M2:540-561: y-normal by permutation (u,v,w)->(v,u,w), (Bx,By,Bz)->(By,Bx,Bz) */
static FluxX hlld_y(const or_m2_params* p, Prim WL, Prim WR, double Bn, double* out_fBx) {
    WL.By = Bn; WR.By = Bn;
    Prim aL = {WL.rho, WL.v, WL.u, WL.w, WL.P, WL.By, WL.Bx, WL.Bz};
    Prim aR = {WR.rho, WR.v, WR.u, WR.w, WR.P, WR.By, WR.Bx, WR.Bz};
    FluxX Fx = hlld_x(p, aL, aR, Bn);
    FluxX G;
    G.frho = Fx.frho; G.fmx = Fx.fmy; G.fmy = Fx.fmx; G.fmz = Fx.fmz; G.fE = Fx.fE;
    *out_fBx = Fx.fBy;
    G.fBy = 0.0;
    G.fBz = Fx.fBz;
    G.Ez = Fx.fBy;
    return G;
}

/* Public flux wrappers: W = {rho,u,v,w,P,Bx,By,Bz}; out = 7 fluxes + Ez.
 * x: out = {frho,fmx,fmy,fmz,fE,fBy,fBz,Ez};  y: out = {frho,fmx,fmy,fmz,fE,fBx,fBz,Ez}. */
void or_m2_hlld_x(const or_m2_params* p, const double* WLa, const double* WRa, double Bn, double* out) {
    Prim WL = {WLa[0], WLa[1], WLa[2], WLa[3], WLa[4], WLa[5], WLa[6], WLa[7]};
    Prim WR = {WRa[0], WRa[1], WRa[2], WRa[3], WRa[4], WRa[5], WRa[6], WRa[7]};
    FluxX F = hlld_x(p, WL, WR, Bn);
    out[0] = F.frho; out[1] = F.fmx; out[2] = F.fmy; out[3] = F.fmz; out[4] = F.fE;
    out[5] = F.fBy; out[6] = F.fBz; out[7] = F.Ez;
}
void or_m2_hlld_y(const or_m2_params* p, const double* WLa, const double* WRa, double Bn, double* out) {
    Prim WL = {WLa[0], WLa[1], WLa[2], WLa[3], WLa[4], WLa[5], WLa[6], WLa[7]};
    Prim WR = {WRa[0], WRa[1], WRa[2], WRa[3], WRa[4], WRa[5], WRa[6], WRa[7]};
    double fBx;
    FluxX G = hlld_y(p, WL, WR, Bn, &fBx);
    out[0] = G.frho; out[1] = G.fmx; out[2] = G.fmy; out[3] = G.fmz; out[4] = G.fE;
    out[5] = fBx; out[6] = G.fBz; out[7] = G.Ez;
}
void or_m2_cons_to_prim(const or_m2_params* p, const double* c8, double* W8) {
    Prim W = cons_to_prim(p, c8[0], c8[1], c8[2], c8[3], c8[4], c8[5], c8[6], c8[7]);
    W8[0] = W.rho; W8[1] = W.u; W8[2] = W.v; W8[3] = W.w; W8[4] = W.P; W8[5] = W.Bx; W8[6] = W.By; W8[7] = W.Bz;
}
double or_m2_mc_limiter(double dm, double dp) { return mc_limiter(dm, dp); }
double or_m2_fast_speed_x(const or_m2_params* p, const double* W8) {
    Prim W = {W8[0], W8[1], W8[2], W8[3], W8[4], W8[5], W8[6], W8[7]};
    return fast_speed_n(p, &W, W.Bx);
}
double or_m2_fast_speed_y(const or_m2_params* p, const double* W8) {
    Prim W = {W8[0], W8[1], W8[2], W8[3], W8[4], W8[5], W8[6], W8[7]};
    return fast_speed_n(p, &W, W.By);
}

/* ------------------------------------------------------------------ */
/* Jz at edge (M2:565-575), dt (M2:579-598)                            */
/* ------------------------------------------------------------------ */
static inline double Jz_edge(const or_m2_params* p, const double* Bx_f, const double* By_f, int ie, int je,
                             double dx, double dy) {
    double ByR = get_By_face(p, By_f, ie, je);
    double ByL = get_By_face(p, By_f, ie - 1, je);
    double dBydx = (ByR - ByL) / dx;
    double BxU = get_Bx_face(p, Bx_f, ie, je);
    double BxD = get_Bx_face(p, Bx_f, ie, je - 1);
    double dBxdy = (BxU - BxD) / dy;
    return dBydx - dBxdy;
}

double or_m2_compute_dt(const or_m2_params* p, const or_m2_state* U, double dx, double dy) {
    int Nx = p->Nx, Ny = p->Ny;
    double maxs = 1e-14;
    for (int j = 0; j < Ny; j++)
        for (int i = 0; i < Nx; i++) {
            Prim W = get_prim_cc(p, U, i, j);
            maxs = dmax(maxs, fabs(W.u) + fast_speed_n(p, &W, W.Bx));
            maxs = dmax(maxs, fabs(W.v) + fast_speed_n(p, &W, W.By));
        }
    double dt_h = p->cfl * dmin(dx, dy) / maxs;
    double dt_r = 1e100;
    if (p->eta > 0.0) {
        double h = dmin(dx, dy);
        dt_r = 0.25 * h * h / p->eta;
    }
    return dmin(dt_h, dt_r);
}

/* ------------------------------------------------------------------ */
/* Sources (M2:602-686)                                                */
/* ------------------------------------------------------------------ */
static void add_gravity_sources(const or_m2_params* p, or_m2_state* U, const double* gx, const double* gy,
                                double dt) {
    if (!p->gravity_on) return;
    int Nx = p->Nx, Ny = p->Ny;
    for (int j = 0; j < Ny; j++)
        for (int i = 0; i < Nx; i++) {
            size_t q = idx2(i, j, Nx);
            double rho = U->rho[q];
            double gxq = gx[q], gyq = gy[q];
            double mx = U->mx[q], my = U->my[q];
            double u = mx / dmax(rho, 1e-300);
            double v = my / dmax(rho, 1e-300);
            U->mx[q] += dt * rho * gxq;
            U->my[q] += dt * rho * gyq;
            U->E[q] += dt * rho * (u * gxq + v * gyq);
        }
    for (int j = 0; j < Ny; j++)
        for (int i = 0; i < Nx; i++) {
            size_t q = idx2(i, j, Nx);
            double Bx = bx_cc(p, U->Bx_f, i, j), By = by_cc(p, U->By_f, i, j);
            enforce_floors(p, &U->rho[q], &U->mx[q], &U->my[q], &U->mz[q], &U->E[q], Bx, By, U->Bz[q]);
        }
}

static void add_joule_heating_and_Bz_resistive(const or_m2_params* p, or_m2_state* U, double dx, double dy,
                                               double dt) {
    if (p->eta <= 0.0) return;
    int Nx = p->Nx, Ny = p->Ny;
#define ATBZ(ii, jj) U->Bz[idx2(p->periodic_x ? imod((ii), Nx) : iclamp((ii), 0, Nx - 1), \
                                p->periodic_y ? imod((jj), Ny) : iclamp((jj), 0, Ny - 1), Nx)]
    for (int j = 0; j < Ny; j++)
        for (int i = 0; i < Nx; i++) {
            size_t q = idx2(i, j, Nx);
            double dBzdy = (ATBZ(i, j + 1) - ATBZ(i, j - 1)) / (2 * dy);
            double dBzdx = (ATBZ(i + 1, j) - ATBZ(i - 1, j)) / (2 * dx);
            double ByR = get_By_face(p, U->By_f, i + 1, j);
            double ByL = get_By_face(p, U->By_f, i, j);
            double dBydx = (ByR - ByL) / dx;
            double BxU = get_Bx_face(p, U->Bx_f, i, j + 1);
            double BxD = get_Bx_face(p, U->Bx_f, i, j);
            double dBxdy = (BxU - BxD) / dy;
            double Jx = dBzdy, Jy = -dBzdx, Jz = dBydx - dBxdy;
            double J2 = Jx * Jx + Jy * Jy + Jz * Jz;
            U->E[q] += dt * p->eta * J2;
        }
    size_t N = (size_t)Nx * (size_t)Ny;
    double* nBz = (double*)calloc(N, sizeof(double));
    double idx2v = 1.0 / (dx * dx), idy2v = 1.0 / (dy * dy);
    for (int j = 0; j < Ny; j++)
        for (int i = 0; i < Nx; i++) {
            double lap = (ATBZ(i + 1, j) - 2 * ATBZ(i, j) + ATBZ(i - 1, j)) * idx2v +
                         (ATBZ(i, j + 1) - 2 * ATBZ(i, j) + ATBZ(i, j - 1)) * idy2v;
            nBz[idx2(i, j, Nx)] = ATBZ(i, j) + dt * p->eta * lap;
        }
#undef ATBZ
    memcpy(U->Bz, nBz, N * sizeof(double));
    free(nBz);
    for (int j = 0; j < Ny; j++)
        for (int i = 0; i < Nx; i++) {
            size_t q = idx2(i, j, Nx);
            double Bx = bx_cc(p, U->Bx_f, i, j), By = by_cc(p, U->By_f, i, j);
            enforce_floors(p, &U->rho[q], &U->mx[q], &U->my[q], &U->mz[q], &U->E[q], Bx, By, U->Bz[q]);
        }
}

/* ------------------------------------------------------------------ */
/* One forward-Euler stage (M2:690-834)                                */
/* ------------------------------------------------------------------ */
void or_m2_advance_stage(const or_m2_params* p, const or_m2_state* U0, or_m2_state* U1, const double* gx,
                         const double* gy, double dx, double dy, double dt) {
    int Nx = p->Nx, Ny = p->Ny;
    size_t N = (size_t)Nx * (size_t)Ny;
    size_t NFx = (size_t)(Nx + 1) * (size_t)Ny, NFy = (size_t)Nx * (size_t)(Ny + 1);
    memcpy(U1->rho, U0->rho, N * sizeof(double));
    memcpy(U1->mx, U0->mx, N * sizeof(double));
    memcpy(U1->my, U0->my, N * sizeof(double));
    memcpy(U1->mz, U0->mz, N * sizeof(double));
    memcpy(U1->E, U0->E, N * sizeof(double));
    memcpy(U1->Bz, U0->Bz, N * sizeof(double));
    memcpy(U1->Bx_f, U0->Bx_f, NFx * sizeof(double));
    memcpy(U1->By_f, U0->By_f, NFy * sizeof(double));

    double* Fx = (double*)malloc(NFx * 7 * sizeof(double));
    double* Gy = (double*)malloc(NFy * 7 * sizeof(double));
    double* Ez_x = (double*)calloc(NFx, sizeof(double));
    double* Ez_y = (double*)calloc(NFy, sizeof(double));

    for (int j = 0; j < Ny; j++)
        for (int i = 0; i <= Nx; i++) {
            int iL = i - 1, iR = i;
            if (p->periodic_x) { iL = imod(iL, Nx); iR = imod(iR, Nx); }
            else { iL = iclamp(iL, 0, Nx - 1); iR = iclamp(iR, 0, Nx - 1); }
            Prim WLL, WLR, WRL, WRR;
            recon_x(p, U0, iL, j, &WLL, &WLR);
            recon_x(p, U0, iR, j, &WRL, &WRR);
            double Bn = get_Bx_face(p, U0->Bx_f, i, j);
            FluxX F = hlld_x(p, WLR, WRL, Bn);
            double* f = Fx + ((size_t)j * (size_t)(Nx + 1) + (size_t)i) * 7;
            f[0] = F.frho; f[1] = F.fmx; f[2] = F.fmy; f[3] = F.fmz; f[4] = F.fE; f[5] = F.fBy; f[6] = F.fBz;
            Ez_x[(size_t)j * (size_t)(Nx + 1) + (size_t)i] = F.Ez;
        }
    for (int j = 0; j <= Ny; j++)
        for (int i = 0; i < Nx; i++) {
            int jD = j - 1, jU = j;
            if (p->periodic_y) { jD = imod(jD, Ny); jU = imod(jU, Ny); }
            else { jD = iclamp(jD, 0, Ny - 1); jU = iclamp(jU, 0, Ny - 1); }
            Prim WDD, WDU, WUD, WUU;
            recon_y(p, U0, i, jD, &WDD, &WDU);
            recon_y(p, U0, i, jU, &WUD, &WUU);
            double Bn = get_By_face(p, U0->By_f, i, j);
            double fBx;
            FluxX G = hlld_y(p, WDU, WUD, Bn, &fBx);
            double* g = Gy + ((size_t)j * (size_t)Nx + (size_t)i) * 7;
            g[0] = G.frho; g[1] = G.fmx; g[2] = G.fmy; g[3] = G.fmz; g[4] = G.fE; g[5] = fBx; g[6] = G.fBz;
            Ez_y[(size_t)j * (size_t)Nx + (size_t)i] = G.Ez;
        }

    double invdx = 1.0 / dx, invdy = 1.0 / dy;
    for (int j = 0; j < Ny; j++)
        for (int i = 0; i < Nx; i++) {
            size_t q = idx2(i, j, Nx);
            const double* FxR = Fx + ((size_t)j * (size_t)(Nx + 1) + (size_t)(i + 1)) * 7;
            const double* FxL = Fx + ((size_t)j * (size_t)(Nx + 1) + (size_t)i) * 7;
            const double* GyU = Gy + ((size_t)(j + 1) * (size_t)Nx + (size_t)i) * 7;
            const double* GyD = Gy + ((size_t)j * (size_t)Nx + (size_t)i) * 7;
            double drho = -(FxR[0] - FxL[0]) * invdx - (GyU[0] - GyD[0]) * invdy;
            double dmx = -(FxR[1] - FxL[1]) * invdx - (GyU[1] - GyD[1]) * invdy;
            double dmy = -(FxR[2] - FxL[2]) * invdx - (GyU[2] - GyD[2]) * invdy;
            double dmz = -(FxR[3] - FxL[3]) * invdx - (GyU[3] - GyD[3]) * invdy;
            double dE = -(FxR[4] - FxL[4]) * invdx - (GyU[4] - GyD[4]) * invdy;
            double dBz = -(FxR[6] - FxL[6]) * invdx - (GyU[6] - GyD[6]) * invdy;
            U1->rho[q] += dt * drho;
            U1->mx[q] += dt * dmx;
            U1->my[q] += dt * dmy;
            U1->mz[q] += dt * dmz;
            U1->E[q] += dt * dE;
            U1->Bz[q] += dt * dBz;
            double Bx = bx_cc(p, U0->Bx_f, i, j), By = by_cc(p, U0->By_f, i, j);
            enforce_floors(p, &U1->rho[q], &U1->mx[q], &U1->my[q], &U1->mz[q], &U1->E[q], Bx, By, U1->Bz[q]);
        }

    double* Ez_edge = (double*)calloc((size_t)(Nx + 1) * (size_t)(Ny + 1), sizeof(double));
#define EX(ii, jj) Ez_x[(size_t)(p->periodic_y ? imod((jj), Ny) : iclamp((jj), 0, Ny - 1)) * (size_t)(Nx + 1) + \
                        (size_t)(p->periodic_x ? imod((ii), Nx + 1) : iclamp((ii), 0, Nx))]
#define EY(ii, jj) Ez_y[(size_t)(p->periodic_y ? imod((jj), Ny + 1) : iclamp((jj), 0, Ny)) * (size_t)Nx + \
                        (size_t)(p->periodic_x ? imod((ii), Nx) : iclamp((ii), 0, Nx - 1))]
    for (int je = 0; je <= Ny; ++je)
        for (int ie = 0; ie <= Nx; ++ie) {
            double ez_ideal = 0.25 * (EX(ie, je - 1) + EX(ie, je) + EY(ie - 1, je) + EY(ie, je));
            double ez_res = 0.0;
            if (p->eta > 0.0) ez_res = p->eta * Jz_edge(p, U0->Bx_f, U0->By_f, ie, je, dx, dy);
            Ez_edge[(size_t)je * (size_t)(Nx + 1) + (size_t)ie] = ez_ideal + ez_res;
        }
#undef EX
#undef EY
    for (int j = 0; j < Ny; j++)
        for (int i = 0; i <= Nx; i++) {
            int je0 = j, je1 = j + 1;
            if (p->periodic_y) { je0 = imod(je0, Ny + 1); je1 = imod(je1, Ny + 1); }
            else { je0 = iclamp(je0, 0, Ny); je1 = iclamp(je1, 0, Ny); }
            size_t f = (size_t)j * (size_t)(Nx + 1) + (size_t)i;
            double ezU = Ez_edge[(size_t)je1 * (size_t)(Nx + 1) + (size_t)i];
            double ezD = Ez_edge[(size_t)je0 * (size_t)(Nx + 1) + (size_t)i];
            const double d = -(dt / dy) * (ezU - ezD);
            if (p->ct_comp) or_ct_add(U0->Bx_f[f], U0->Bx_lo[f], d, &U1->Bx_f[f], &U1->Bx_lo[f]);
            else U1->Bx_f[f] += d;
        }
    for (int j = 0; j <= Ny; j++)
        for (int i = 0; i < Nx; i++) {
            int ie0 = i, ie1 = i + 1;
            if (p->periodic_x) { ie0 = imod(ie0, Nx + 1); ie1 = imod(ie1, Nx + 1); }
            else { ie0 = iclamp(ie0, 0, Nx); ie1 = iclamp(ie1, 0, Nx); }
            size_t f = (size_t)j * (size_t)Nx + (size_t)i;
            double ezR = Ez_edge[(size_t)j * (size_t)(Nx + 1) + (size_t)ie1];
            double ezL = Ez_edge[(size_t)j * (size_t)(Nx + 1) + (size_t)ie0];
            const double d = +(dt / dx) * (ezR - ezL);
            if (p->ct_comp) or_ct_add(U0->By_f[f], U0->By_lo[f], d, &U1->By_f[f], &U1->By_lo[f]);
            else U1->By_f[f] += d;
        }
    free(Ez_edge);
    free(Fx); free(Gy); free(Ez_x); free(Ez_y);

    add_gravity_sources(p, U1, gx, gy, dt);
    add_joule_heating_and_Bz_resistive(p, U1, dx, dy, dt);
}

static void state_alloc(or_m2_state* S, int Nx, int Ny) {
    size_t N = (size_t)Nx * (size_t)Ny;
    S->Bx_lo = (float*)calloc((size_t)(Nx + 1) * (size_t)Ny, sizeof(float));
    S->By_lo = (float*)calloc((size_t)Nx * (size_t)(Ny + 1), sizeof(float));
    S->rho = (double*)malloc(N * sizeof(double)); S->mx = (double*)malloc(N * sizeof(double));
    S->my = (double*)malloc(N * sizeof(double));  S->mz = (double*)malloc(N * sizeof(double));
    S->E = (double*)malloc(N * sizeof(double));   S->Bz = (double*)malloc(N * sizeof(double));
    S->Bx_f = (double*)malloc((size_t)(Nx + 1) * (size_t)Ny * sizeof(double));
    S->By_f = (double*)malloc((size_t)Nx * (size_t)(Ny + 1) * sizeof(double));
}
static void state_free(or_m2_state* S) {
    free(S->rho); free(S->mx); free(S->my); free(S->mz); free(S->E); free(S->Bz); free(S->Bx_f); free(S->By_f);
    free(S->Bx_lo); free(S->By_lo);
}

/* M2:838-869 */
void or_m2_rk2_step(const or_m2_params* p, or_m2_state* U, const double* gx, const double* gy, double dx,
                    double dy, double dt) {
    int Nx = p->Nx, Ny = p->Ny;
    size_t N = (size_t)Nx * (size_t)Ny;
    or_m2_state U1, U2;
    state_alloc(&U1, Nx, Ny);
    state_alloc(&U2, Nx, Ny);
    or_m2_advance_stage(p, U, &U1, gx, gy, dx, dy, dt);
    double* g1x = (double*)malloc(N * sizeof(double));
    double* g1y = (double*)malloc(N * sizeof(double));
    or_m2_poisson(p, U1.rho, dx, dy, g1x, g1y);
    or_m2_advance_stage(p, &U1, &U2, g1x, g1y, dx, dy, dt);
    for (size_t q = 0; q < N; q++) {
        U->rho[q] = 0.5 * (U->rho[q] + U2.rho[q]);
        U->mx[q] = 0.5 * (U->mx[q] + U2.mx[q]);
        U->my[q] = 0.5 * (U->my[q] + U2.my[q]);
        U->mz[q] = 0.5 * (U->mz[q] + U2.mz[q]);
        U->E[q] = 0.5 * (U->E[q] + U2.E[q]);
        U->Bz[q] = 0.5 * (U->Bz[q] + U2.Bz[q]);
    }
    size_t NFx = (size_t)(Nx + 1) * (size_t)Ny, NFy = (size_t)Nx * (size_t)(Ny + 1);
    if (p->ct_comp) {
        for (size_t k = 0; k < NFx; k++) or_ct_avg(U->Bx_f[k], U->Bx_lo[k], U2.Bx_f[k], U2.Bx_lo[k], &U->Bx_f[k], &U->Bx_lo[k]);
        for (size_t k = 0; k < NFy; k++) or_ct_avg(U->By_f[k], U->By_lo[k], U2.By_f[k], U2.By_lo[k], &U->By_f[k], &U->By_lo[k]);
    } else {
        for (size_t k = 0; k < NFx; k++) U->Bx_f[k] = 0.5 * (U->Bx_f[k] + U2.Bx_f[k]);
        for (size_t k = 0; k < NFy; k++) U->By_f[k] = 0.5 * (U->By_f[k] + U2.By_f[k]);
    }
    for (int j = 0; j < Ny; j++)
        for (int i = 0; i < Nx; i++) {
            size_t q = idx2(i, j, Nx);
            double Bx = bx_cc(p, U->Bx_f, i, j), By = by_cc(p, U->By_f, i, j);
            enforce_floors(p, &U->rho[q], &U->mx[q], &U->my[q], &U->mz[q], &U->E[q], Bx, By, U->Bz[q]);
        }
    free(g1x); free(g1y);
    state_free(&U1);
    state_free(&U2);
}

/* M2:873-879 */
void or_m2_divb(const or_m2_params* p, const double* Bx_f, const double* By_f, double dx, double dy, double* out) {
    for (int j = 0; j < p->Ny; j++)
        for (int i = 0; i < p->Nx; i++) {
            double BxR = get_Bx_face(p, Bx_f, i + 1, j), BxL = get_Bx_face(p, Bx_f, i, j);
            double ByU = get_By_face(p, By_f, i, j + 1), ByD = get_By_face(p, By_f, i, j);
            out[idx2(i, j, p->Nx)] = (BxR - BxL) / dx + (ByU - ByD) / dy;
        }
}

/* The driver loop of M2:1111-1117 for a fixed number of steps (no output):
 *   g = Poisson(U.rho); dt = compute_dt; clamp to t_end; rk2_step; t += dt.
 * dts_out (may be NULL) receives each step's dt.  Returns the final t. */
double or_m2_run(const or_m2_params* p, or_m2_state* U, double dx, double dy, int nsteps, double t0, double t_end,
                 double* dts_out) {
    size_t N = (size_t)p->Nx * (size_t)p->Ny;
    double* gx = (double*)malloc(N * sizeof(double));
    double* gy = (double*)malloc(N * sizeof(double));
    double t = t0;
    for (int s = 0; s < nsteps && t < t_end; s++) {
        or_m2_poisson(p, U->rho, dx, dy, gx, gy);
        double dt = or_m2_compute_dt(p, U, dx, dy);
        if (t + dt > t_end) dt = t_end - t;
        or_m2_rk2_step(p, U, gx, gy, dx, dy, dt);
        t += dt;
        if (dts_out) dts_out[s] = dt;
    }
    free(gx); free(gy);
    return t;
}

/* The snapshot frame (the store lambda of M2:1065-1106): 10 planes {rho, P, Mach, u, v, vabs, Bx, By,
 * Bz, divB} of Nx*Ny cells. */
void or_m2_frame(const or_m2_params* p, const or_m2_state* U, double dx, double dy, double* out) {
    const size_t N = (size_t)p->Nx * (size_t)p->Ny;
    for (int j = 0; j < p->Ny; j++)
        for (int i = 0; i < p->Nx; i++) {
            const size_t q = idx2(i, j, p->Nx);
            const double Bx_cc = bx_cc(p, U->Bx_f, i, j), By_cc = by_cc(p, U->By_f, i, j);
            const Prim W = cons_to_prim(p, U->rho[q], U->mx[q], U->my[q], U->mz[q], U->E[q], Bx_cc, By_cc, U->Bz[q]);
            const double cs = sqrt(dmax(0.0, sound_speed2(p, &W)));
            const double vabs = sqrt(W.u * W.u + W.v * W.v + W.w * W.w);
            const double Mach = vabs / dmax(cs, 1e-14);
            const double BxR = get_Bx_face(p, U->Bx_f, i + 1, j), BxL = get_Bx_face(p, U->Bx_f, i, j);
            const double ByU = get_By_face(p, U->By_f, i, j + 1), ByD = get_By_face(p, U->By_f, i, j);
            const double f[10] = {W.rho, W.P, Mach, W.u, W.v, vabs, Bx_cc, By_cc, W.Bz,
                                  (BxR - BxL) / dx + (ByU - ByD) / dy};
            for (int k = 0; k < 10; k++) out[(size_t)k * N + q] = f[k];
        }
}
