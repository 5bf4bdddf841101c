/*
 * oracle_mhd3d.c -- CPU FP64 restatement of the 3D resistive MHD solver step.
 *
 * TEST INFRASTRUCTURE ONLY (see oracle_mhd2d.c for the rules): the parity oracle for the 3D step of
 * mhd3d_resistive_hlld_ct.cpp (cited as M3:<line>), restated formula by formula in the same
 * operation order, built without FMA so it can match the compiled reference bit for bit.
 *
 * EOS: ideal_eos = 0 is the reference's hybrid cold + thermal + bounce EOS (M3:114-212).
 * ideal_eos = 1 is an ideal gas with gamma_ideal for the 3D benchmark problem, which the reference
 * does not have; its EOS functions are the reference's 2D ideal-gas functions
 * (mhd2d_resistive_hlld_ct_gravity.cpp:114-150) in 3D.  The hybrid EOS with K1 = Kbounce = 0 is
 * the same ideal gas (gamma = gamma_th), which is how tests pin that mode against the reference.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include "oracle.h"

static inline size_t idx3(int i, int j, int k, int Nx, int Ny) {
    return ((size_t)k * (size_t)Ny + (size_t)j) * (size_t)Nx + (size_t)i;
}
static inline int imod(int i, int n) { int r = i % n; return (r < 0) ? r + n : r; }
static inline int iclamp(int v, int lo, int hi) { return v < lo ? lo : (v > hi ? hi : v); }
static inline double dmax(double a, double b) { return (a < b) ? b : a; }
static inline double dmin(double a, double b) { return (b < a) ? b : a; }

typedef struct { double rho, u, v, w, P, Bx, By, Bz; } Prim;
typedef struct { double Pcold, ecold, gamma_c; } ColdEOS;

static inline double minmod(double a, double b) {
    if (a * b <= 0.0) return 0.0;
    return (fabs(a) < fabs(b)) ? a : b;
}
static inline double mc_limiter(double dm, double dp) { return minmod(0.5 * (dm + dp), minmod(2.0 * dm, 2.0 * dp)); }

/* ---------------- hybrid EOS (M3:116-212) ---------------- */
static ColdEOS eos_cold(const or_m3_params* p, double rho) {
    rho = dmax(rho, p->floor_rho);
    double K2 = p->K1 * pow(p->rho_nuc, p->gamma1 - p->gamma2);
    if (rho < p->rho_nuc) {
        double Pc = p->K1 * pow(rho, p->gamma1);
        double ec = p->K1 * pow(rho, p->gamma1 - 1.0) / (p->gamma1 - 1.0);
        ColdEOS c = {Pc, ec, p->gamma1};
        return c;
    }
    double Pc = K2 * pow(rho, p->gamma2);
    double ec1_n = p->K1 * pow(p->rho_nuc, p->gamma1 - 1.0) / (p->gamma1 - 1.0);
    double ec2_n = K2 * pow(p->rho_nuc, p->gamma2 - 1.0) / (p->gamma2 - 1.0);
    double ec = ec1_n + (K2 * pow(rho, p->gamma2 - 1.0) / (p->gamma2 - 1.0) - ec2_n);
    double gamma_c = p->gamma2;
    if (rho >= p->rho_bounce) {
        double pb = p->Kbounce * (pow(rho, p->gamma_bounce) - pow(p->rho_bounce, p->gamma_bounce));
        double eb = p->Kbounce * (pow(rho, p->gamma_bounce - 1.0) - pow(p->rho_bounce, p->gamma_bounce - 1.0)) /
                    (p->gamma_bounce - 1.0);
        Pc += pb;
        ec += eb;
        gamma_c = (p->gamma2 * (Pc - pb) + p->gamma_bounce * pb) / dmax(Pc, p->floor_P);
    }
    ColdEOS c = {Pc, ec, gamma_c};
    return c;
}
static double pressure_from_rho_espec(const or_m3_params* p, double rho, double e_spec) {
    rho = dmax(rho, p->floor_rho);
    ColdEOS cold = eos_cold(p, rho);
    double eth = dmax(0.0, e_spec - cold.ecold);
    double Pth = (p->gamma_th - 1.0) * rho * eth;
    return dmax(cold.Pcold + Pth, p->floor_P);
}
static double total_energy_from_prim(const or_m3_params* p, const Prim* W) {
    double rho = dmax(W->rho, p->floor_rho);
    double kin = 0.5 * rho * (W->u * W->u + W->v * W->v + W->w * W->w);
    double mag = 0.5 * (W->Bx * W->Bx + W->By * W->By + W->Bz * W->Bz);
    if (p->ideal_eos) return kin + mag + W->P / (p->gamma_ideal - 1.0);
    ColdEOS cold = eos_cold(p, rho);
    double Pth = dmax(0.0, W->P - cold.Pcold);
    double eth = (p->gamma_th > 1.0) ? (Pth / ((p->gamma_th - 1.0) * rho)) : 0.0;
    double eint = rho * (cold.ecold + eth);
    return kin + mag + eint;
}
static void enforce_floors(const or_m3_params* p, double* rho, double* mx, double* my, double* mz, double* E,
                           double Bx, double By, double Bz) {
    *rho = dmax(*rho, p->floor_rho);
    double u = *mx / *rho, v = *my / *rho, w = *mz / *rho;
    double kin = 0.5 * *rho * (u * u + v * v + w * w);
    double mag = 0.5 * (Bx * Bx + By * By + Bz * Bz);
    if (p->ideal_eos) { /* M2:114-124 */
        double P = (p->gamma_ideal - 1.0) * (*E - kin - mag);
        if (P < p->floor_P) {
            double eint = p->floor_P / (p->gamma_ideal - 1.0);
            *E = kin + mag + eint;
        }
        return;
    }
    double eint_tot = *E - kin - mag;
    if (eint_tot < 0.0) eint_tot = 0.0;
    double e_spec = eint_tot / *rho;
    double P = pressure_from_rho_espec(p, *rho, e_spec);
    if (P < p->floor_P) P = p->floor_P;
    ColdEOS cold = eos_cold(p, *rho);
    double Pth = dmax(0.0, P - cold.Pcold);
    double eth = (p->gamma_th > 1.0) ? (Pth / ((p->gamma_th - 1.0) * *rho)) : 0.0;
    double eint_new = *rho * (cold.ecold + eth);
    *E = kin + mag + eint_new;
}
static Prim cons_to_prim(const or_m3_params* p, double rho, double mx, double my, double mz, double E, double Bx,
                         double By, double Bz) {
    rho = dmax(rho, p->floor_rho);
    double u = mx / rho, v = my / rho, w = mz / rho;
    double kin = 0.5 * rho * (u * u + v * v + w * w);
    double mag = 0.5 * (Bx * Bx + By * By + Bz * Bz);
    double P;
    if (p->ideal_eos) { /* M2:126-135 */
        P = (p->gamma_ideal - 1.0) * (E - kin - mag);
        if (P < p->floor_P) P = p->floor_P;
    } else {
        double eint_tot = E - kin - mag;
        if (eint_tot < 0.0) eint_tot = 0.0;
        double e_spec = eint_tot / rho;
        P = pressure_from_rho_espec(p, rho, e_spec);
        if (P < p->floor_P) P = p->floor_P;
    }
    Prim W = {rho, u, v, w, P, Bx, By, Bz};
    return W;
}
static double sound_speed2(const or_m3_params* p, const Prim* W) {
    if (p->ideal_eos) return p->gamma_ideal * W->P / dmax(W->rho, p->floor_rho);
    double rho = dmax(W->rho, p->floor_rho);
    ColdEOS cold = eos_cold(p, rho);
    double Pth = dmax(0.0, W->P - cold.Pcold);
    double num = cold.gamma_c * cold.Pcold + p->gamma_th * Pth;
    return dmax(0.0, num / rho);
}
static double fast_speed_n(const or_m3_params* p, const Prim* W, double Bn) { /* M3:214-248 */
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

/* ---------------- face / cell helpers (M3:250-308) ---------------- */
#define MAPC(v, n, per) ((per) ? imod((v), (n)) : iclamp((v), 0, (n) - 1))
#define MAPF(v, n, per) ((per) ? imod((v), (n) + 1) : iclamp((v), 0, (n)))
static inline double get_Bx_face(const or_m3_params* p, const double* B, int i, int j, int k) {
    i = MAPF(i, p->Nx, p->periodic_x); j = MAPC(j, p->Ny, p->periodic_y); k = MAPC(k, p->Nz, p->periodic_z);
    return B[((size_t)k * p->Ny + j) * (size_t)(p->Nx + 1) + i];
}
static inline double get_By_face(const or_m3_params* p, const double* B, int i, int j, int k) {
    i = MAPC(i, p->Nx, p->periodic_x); j = MAPF(j, p->Ny, p->periodic_y); k = MAPC(k, p->Nz, p->periodic_z);
    return B[((size_t)k * (p->Ny + 1) + j) * (size_t)p->Nx + i];
}
static inline double get_Bz_face(const or_m3_params* p, const double* B, int i, int j, int k) {
    i = MAPC(i, p->Nx, p->periodic_x); j = MAPC(j, p->Ny, p->periodic_y); k = MAPF(k, p->Nz, p->periodic_z);
    return B[((size_t)k * p->Ny + j) * (size_t)p->Nx + i];
}
static inline double bx_cc(const or_m3_params* p, const or_m3_state* S, int i, int j, int k) {
    return 0.5 * (get_Bx_face(p, S->Bx_f, i, j, k) + get_Bx_face(p, S->Bx_f, i + 1, j, k));
}
static inline double by_cc(const or_m3_params* p, const or_m3_state* S, int i, int j, int k) {
    return 0.5 * (get_By_face(p, S->By_f, i, j, k) + get_By_face(p, S->By_f, i, j + 1, k));
}
static inline double bz_cc(const or_m3_params* p, const or_m3_state* S, int i, int j, int k) {
    return 0.5 * (get_Bz_face(p, S->Bz_f, i, j, k) + get_Bz_face(p, S->Bz_f, i, j, k + 1));
}
static Prim get_prim_cc(const or_m3_params* p, const or_m3_state* U, int i, int j, int k) {
    i = MAPC(i, p->Nx, p->periodic_x); j = MAPC(j, p->Ny, p->periodic_y); k = MAPC(k, p->Nz, p->periodic_z);
    size_t q = idx3(i, j, k, p->Nx, p->Ny);
    double Bx = bx_cc(p, U, i, j, k), By = by_cc(p, U, i, j, k), Bz = bz_cc(p, U, i, j, k);
    return cons_to_prim(p, U->rho[q], U->mx[q], U->my[q], U->mz[q], U->E[q], Bx, By, Bz);
}

/* ---------------- PLM (M3:312-382) ---------------- */
static void recon(const or_m3_params* p, const or_m3_state* U, int i, int j, int k, int ax, Prim* WL, Prim* WR) {
    int di = (ax == 0), dj = (ax == 1), dk = (ax == 2);
    Prim Wm = get_prim_cc(p, U, i - di, j - dj, k - dk);
    Prim Wc = get_prim_cc(p, U, i, j, k);
    Prim Wp = get_prim_cc(p, U, i + di, j + dj, k + dk);
    double sr = mc_limiter(Wc.rho - Wm.rho, Wp.rho - Wc.rho);
    double su = mc_limiter(Wc.u - Wm.u, Wp.u - Wc.u);
    double sv = mc_limiter(Wc.v - Wm.v, Wp.v - Wc.v);
    double sw = mc_limiter(Wc.w - Wm.w, Wp.w - Wc.w);
    double sP = mc_limiter(Wc.P - Wm.P, Wp.P - Wc.P);
    WL->rho = Wc.rho - 0.5 * sr; WR->rho = Wc.rho + 0.5 * sr;
    WL->u = Wc.u - 0.5 * su; WR->u = Wc.u + 0.5 * su;
    WL->v = Wc.v - 0.5 * sv; WR->v = Wc.v + 0.5 * sv;
    WL->w = Wc.w - 0.5 * sw; WR->w = Wc.w + 0.5 * sw;
    WL->P = Wc.P - 0.5 * sP; WR->P = Wc.P + 0.5 * sP;
    if (ax == 0) {
        double sBy = mc_limiter(Wc.By - Wm.By, Wp.By - Wc.By), sBz = mc_limiter(Wc.Bz - Wm.Bz, Wp.Bz - Wc.Bz);
        WL->Bx = Wc.Bx; WR->Bx = Wc.Bx;
        WL->By = Wc.By - 0.5 * sBy; WR->By = Wc.By + 0.5 * sBy;
        WL->Bz = Wc.Bz - 0.5 * sBz; WR->Bz = Wc.Bz + 0.5 * sBz;
    } else if (ax == 1) {
        double sBx = mc_limiter(Wc.Bx - Wm.Bx, Wp.Bx - Wc.Bx), sBz = mc_limiter(Wc.Bz - Wm.Bz, Wp.Bz - Wc.Bz);
        WL->Bx = Wc.Bx - 0.5 * sBx; WR->Bx = Wc.Bx + 0.5 * sBx;
        WL->By = Wc.By; WR->By = Wc.By;
        WL->Bz = Wc.Bz - 0.5 * sBz; WR->Bz = Wc.Bz + 0.5 * sBz;
    } else {
        double sBx = mc_limiter(Wc.Bx - Wm.Bx, Wp.Bx - Wc.Bx), sBy = mc_limiter(Wc.By - Wm.By, Wp.By - Wc.By);
        WL->Bx = Wc.Bx - 0.5 * sBx; WR->Bx = Wc.Bx + 0.5 * sBx;
        WL->By = Wc.By - 0.5 * sBy; WR->By = Wc.By + 0.5 * sBy;
        WL->Bz = Wc.Bz; WR->Bz = Wc.Bz;
    }
    WL->rho = dmax(WL->rho, p->floor_rho); WR->rho = dmax(WR->rho, p->floor_rho);
    WL->P = dmax(WL->P, p->floor_P); WR->P = dmax(WR->P, p->floor_P);
}

/* ---------------- HLLD (M3:386-568), EOS-aware energies ---------------- */
static inline double total_pressure(const Prim* W) {
    double B2 = W->Bx * W->Bx + W->By * W->By + W->Bz * W->Bz;
    return W->P + 0.5 * B2;
}
static void flux_ideal_x(const or_m3_params* p, const Prim* W, double F[7]) {
    double rho = W->rho, u = W->u, v = W->v, w = W->w, P = W->P, Bx = W->Bx, By = W->By, Bz = W->Bz;
    double B2 = Bx * Bx + By * By + Bz * Bz;
    double pt = P + 0.5 * B2;
    double E = total_energy_from_prim(p, W);
    double vdotB = u * Bx + v * By + w * Bz;
    F[0] = rho * u;
    F[1] = rho * u * u + pt - Bx * Bx;
    F[2] = rho * u * v - Bx * By;
    F[3] = rho * u * w - Bx * Bz;
    F[4] = (E + pt) * u - Bx * vdotB;
    F[5] = u * By - v * Bx;
    F[6] = u * Bz - w * Bx;
}
static void hlld_x(const or_m3_params* p, Prim WL, Prim WR, double Bn, double out[7]) {
    WL.Bx = Bn; WR.Bx = Bn;
    double rhoL = WL.rho, rhoR = WR.rho, uL = WL.u, uR = WR.u, vL = WL.v, vR = WR.v, wL = WL.w, wR = WR.w;
    double ByL = WL.By, ByR = WR.By, BzL = WL.Bz, BzR = WR.Bz;
    double FL[7], FR[7];
    flux_ideal_x(p, &WL, FL);
    flux_ideal_x(p, &WR, FR);
    double cfL = fast_speed_n(p, &WL, WL.Bx), cfR = fast_speed_n(p, &WR, WR.Bx);
    double SL = dmin(uL - cfL, uR - cfR), SR = dmax(uL + cfL, uR + cfR);
    if (0.0 <= SL) { memcpy(out, FL, sizeof FL); return; }
    if (0.0 >= SR) { memcpy(out, FR, sizeof FR); return; }
    double UL[7], UR[7];
    UL[0] = WL.rho; UL[1] = WL.rho * WL.u; UL[2] = WL.rho * WL.v; UL[3] = WL.rho * WL.w;
    UL[4] = total_energy_from_prim(p, &WL); UL[5] = WL.By; UL[6] = WL.Bz;
    UR[0] = WR.rho; UR[1] = WR.rho * WR.u; UR[2] = WR.rho * WR.v; UR[3] = WR.rho * WR.w;
    UR[4] = total_energy_from_prim(p, &WR); UR[5] = WR.By; UR[6] = WR.Bz;
    double ptL = total_pressure(&WL), ptR = total_pressure(&WR);
    double denom = rhoL * (SL - uL) - rhoR * (SR - uR);
    double SM = 0.0;
    if (fabs(denom) > 1e-14) SM = (ptR - ptL + rhoL * uL * (SL - uL) - rhoR * uR * (SR - uR)) / denom;
    double rhoL_s = rhoL * (SL - uL) / (SL - SM), rhoR_s = rhoR * (SR - uR) / (SR - SM);
    double pt_s = ptL + rhoL * (SL - uL) * (SM - uL);
    double sL = SL - uL, sR = SR - uR;
    double dL = rhoL * sL * (SL - SM) - Bn * Bn, dR = rhoR * sR * (SR - SM) - Bn * Bn;
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
    double sqrt_rhoL = sqrt(dmax(rhoL_s, 1e-300)), sqrt_rhoR = sqrt(dmax(rhoR_s, 1e-300));
    double SLL = SM - fabs(Bn) / sqrt_rhoL, SRR = SM + fabs(Bn) / sqrt_rhoR;
    const double gd = p->hlld_guard;
    if (fabs(Bn) < 1e-12 || (gd > 0.0 && (fabs(dL) < gd * fmax(Bn * Bn, fabs(rhoL * sL * (SL - SM))) ||
                                          fabs(dR) < gd * fmax(Bn * Bn, fabs(rhoR * sR * (SR - SM)))))) {
        double inv = 1.0 / (SR - SL);
        for (int c = 0; c < 7; c++) out[c] = (SR * FL[c] - SL * FR[c] + SL * SR * (UR[c] - UL[c])) * inv;
        return;
    }
    double signBn = (Bn >= 0.0) ? 1.0 : -1.0;
    double denom2 = sqrt_rhoL + sqrt_rhoR;
    double v_ss = (sqrt_rhoL * vL_s + sqrt_rhoR * vR_s + (ByR_s - ByL_s) * signBn) / denom2;
    double w_ss = (sqrt_rhoL * wL_s + sqrt_rhoR * wR_s + (BzR_s - BzL_s) * signBn) / denom2;
    double By_ss = (sqrt_rhoL * ByR_s + sqrt_rhoR * ByL_s + sqrt_rhoL * sqrt_rhoR * (vR_s - vL_s) * signBn) / denom2;
    double Bz_ss = (sqrt_rhoL * BzR_s + sqrt_rhoR * BzL_s + sqrt_rhoL * sqrt_rhoR * (wR_s - wL_s) * signBn) / denom2;
    const Prim* Wref;
    const double* Fref;
    double S, rho_s, v_s, w_s, By_s, Bz_s;
    if (0.0 < SLL)      { Wref = &WL; Fref = FL; S = SL; rho_s = rhoL_s; v_s = vL_s; w_s = wL_s; By_s = ByL_s; Bz_s = BzL_s; }
    else if (0.0 < SM)  { Wref = &WL; Fref = FL; S = SL; rho_s = rhoL_s; v_s = v_ss; w_s = w_ss; By_s = By_ss; Bz_s = Bz_ss; }
    else if (0.0 < SRR) { Wref = &WR; Fref = FR; S = SR; rho_s = rhoR_s; v_s = v_ss; w_s = w_ss; By_s = By_ss; Bz_s = Bz_ss; }
    else                { Wref = &WR; Fref = FR; S = SR; rho_s = rhoR_s; v_s = vR_s; w_s = wR_s; By_s = ByR_s; Bz_s = BzR_s; }
    /* U_star M3:526-544 */
    double EW = total_energy_from_prim(p, Wref);
    double vdotB_s = SM * Bn + v_s * By_s + w_s * Bz_s;
    double vdotB_w = Wref->u * Wref->Bx + Wref->v * Wref->By + Wref->w * Wref->Bz;
    double E = ((S - Wref->u) * EW - (total_pressure(Wref) * Wref->u - Bn * vdotB_w) + (pt_s * SM - Bn * vdotB_s)) /
               (S - SM);
    double Ust[7] = {rho_s, rho_s * SM, rho_s * v_s, rho_s * w_s, E, By_s, Bz_s};
    /* state7 M3:521-524 */
    double U0[7] = {Wref->rho, Wref->rho * Wref->u, Wref->rho * Wref->v, Wref->rho * Wref->w,
                    total_energy_from_prim(p, Wref), Wref->By, Wref->Bz};
    for (int c = 0; c < 7; c++) out[c] = Fref[c] + S * (Ust[c] - U0[c]);
}
/* y / z by permutation (M3:577-622); out in the x-layout of the permuted call */
static void hlld_y(const or_m3_params* p, Prim WL, Prim WR, double Bn, double G[7]) {
    WL.By = Bn; WR.By = Bn;
    Prim aL = {WL.rho, WL.v, WL.u, WL.w, WL.P, WL.By, WL.Bx, WL.Bz};
    Prim aR = {WR.rho, WR.v, WR.u, WR.w, WR.P, WR.By, WR.Bx, WR.Bz};
    double F[7];
    hlld_x(p, aL, aR, Bn, F);
    G[0] = F[0]; G[1] = F[2]; G[2] = F[1]; G[3] = F[3]; G[4] = F[4]; G[5] = F[5] /* fBx */; G[6] = F[6] /* fBz */;
}
static void hlld_z(const or_m3_params* p, Prim WL, Prim WR, double Bn, double H[7]) {
    WL.Bz = Bn; WR.Bz = Bn;
    Prim aL = {WL.rho, WL.w, WL.u, WL.v, WL.P, WL.Bz, WL.Bx, WL.By};
    Prim aR = {WR.rho, WR.w, WR.u, WR.v, WR.P, WR.Bz, WR.Bx, WR.By};
    double F[7];
    hlld_x(p, aL, aR, Bn, F);
    H[0] = F[0]; H[1] = F[2]; H[2] = F[3]; H[3] = F[1]; H[4] = F[4]; H[5] = F[5] /* fBx */; H[6] = F[6] /* fBy */;
}

void or_m3_hlld_x(const or_m3_params* p, const double* a, const double* b, double Bn, double* out7) {
    Prim WL = {a[0], a[1], a[2], a[3], a[4], a[5], a[6], a[7]}, WR = {b[0], b[1], b[2], b[3], b[4], b[5], b[6], b[7]};
    hlld_x(p, WL, WR, Bn, out7);
}
void or_m3_cons_to_prim(const or_m3_params* p, const double* c, double* W8) {
    Prim W = cons_to_prim(p, c[0], c[1], c[2], c[3], c[4], c[5], c[6], c[7]);
    W8[0] = W.rho; W8[1] = W.u; W8[2] = W.v; W8[3] = W.w; W8[4] = W.P; W8[5] = W.Bx; W8[6] = W.By; W8[7] = W.Bz;
}
void or_m3_eos_cold(const or_m3_params* p, double rho, double* out3) {
    ColdEOS c = eos_cold(p, rho);
    out3[0] = c.Pcold; out3[1] = c.ecold; out3[2] = c.gamma_c;
}
double or_m3_sound_speed2(const or_m3_params* p, const double* a) {
    Prim W = {a[0], a[1], a[2], a[3], a[4], a[5], a[6], a[7]};
    return sound_speed2(p, &W);
}
double or_m3_total_energy(const or_m3_params* p, const double* a) {
    Prim W = {a[0], a[1], a[2], a[3], a[4], a[5], a[6], a[7]};
    return total_energy_from_prim(p, &W);
}

/* ---------------- dt (M3:626-648) ---------------- */
double or_m3_compute_dt(const or_m3_params* p, const or_m3_state* U, double dx, double dy, double dz) {
    double maxs = 1e-14;
    for (int k = 0; k < p->Nz; k++)
        for (int j = 0; j < p->Ny; j++)
            for (int i = 0; i < p->Nx; i++) {
                Prim W = get_prim_cc(p, U, i, j, k);
                maxs = dmax(maxs, fabs(W.u) + fast_speed_n(p, &W, W.Bx));
                maxs = dmax(maxs, fabs(W.v) + fast_speed_n(p, &W, W.By));
                maxs = dmax(maxs, fabs(W.w) + fast_speed_n(p, &W, W.Bz));
            }
    double h = dmin(dmin(dx, dy), dz);
    double dt_h = p->cfl * h / maxs;
    double dt_r = 1e100;
    if (p->eta > 0.0) dt_r = 0.25 * h * h / p->eta;
    return dmin(dt_h, dt_r);
}

/* ---------------- Joule heating (M3:652-692) ---------------- */
static void add_joule_heating(const or_m3_params* p, or_m3_state* U, double dx, double dy, double dz, double dt) {
    if (p->eta <= 0.0) return;
    int Nx = p->Nx, Ny = p->Ny, Nz = p->Nz;
    for (int k = 0; k < Nz; k++)
        for (int j = 0; j < Ny; j++)
            for (int i = 0; i < Nx; i++) {
                size_t q = idx3(i, j, k, Nx, Ny);
                double Bxp[3] = {bx_cc(p, U, i + 1, j, k), by_cc(p, U, i + 1, j, k), bz_cc(p, U, i + 1, j, k)};
                double Bxm[3] = {bx_cc(p, U, i - 1, j, k), by_cc(p, U, i - 1, j, k), bz_cc(p, U, i - 1, j, k)};
                double Byp[3] = {bx_cc(p, U, i, j + 1, k), by_cc(p, U, i, j + 1, k), bz_cc(p, U, i, j + 1, k)};
                double Bym[3] = {bx_cc(p, U, i, j - 1, k), by_cc(p, U, i, j - 1, k), bz_cc(p, U, i, j - 1, k)};
                double Bzp[3] = {bx_cc(p, U, i, j, k + 1), by_cc(p, U, i, j, k + 1), bz_cc(p, U, i, j, k + 1)};
                double Bzm[3] = {bx_cc(p, U, i, j, k - 1), by_cc(p, U, i, j, k - 1), bz_cc(p, U, i, j, k - 1)};
                double dBzdy = (Byp[2] - Bym[2]) / (2 * dy);
                double dBydz = (Bzp[1] - Bzm[1]) / (2 * dz);
                double dBxdz = (Bzp[0] - Bzm[0]) / (2 * dz);
                double dBzdx = (Bxp[2] - Bxm[2]) / (2 * dx);
                double dBydx = (Bxp[1] - Bxm[1]) / (2 * dx);
                double dBxdy = (Byp[0] - Bym[0]) / (2 * dy);
                double Jx = dBzdy - dBydz, Jy = dBxdz - dBzdx, Jz = dBydx - dBxdy;
                double J2 = Jx * Jx + Jy * Jy + Jz * Jz;
                U->E[q] += dt * p->eta * J2;
            }
}

static void state_copy(const or_m3_params* p, const or_m3_state* A, or_m3_state* B) {
    size_t N = (size_t)p->Nx * p->Ny * p->Nz;
    memcpy(B->rho, A->rho, N * 8); memcpy(B->mx, A->mx, N * 8); memcpy(B->my, A->my, N * 8);
    memcpy(B->mz, A->mz, N * 8); memcpy(B->E, A->E, N * 8);
    memcpy(B->Bx_f, A->Bx_f, (size_t)(p->Nx + 1) * p->Ny * p->Nz * 8);
    memcpy(B->By_f, A->By_f, (size_t)p->Nx * (p->Ny + 1) * p->Nz * 8);
    memcpy(B->Bz_f, A->Bz_f, (size_t)p->Nx * p->Ny * (p->Nz + 1) * 8);
}

/* ---------------- one stage (M3:694-1030) ---------------- */
void or_m3_advance_stage(const or_m3_params* p, const or_m3_state* U0, or_m3_state* U1, double dx, double dy,
                         double dz, double dt) {
    int Nx = p->Nx, Ny = p->Ny, Nz = p->Nz;
    int px = p->periodic_x, py = p->periodic_y, pz = p->periodic_z;
    state_copy(p, U0, U1);
    size_t NFx = (size_t)(Nx + 1) * Ny * Nz, NFy = (size_t)Nx * (Ny + 1) * Nz, NFz = (size_t)Nx * Ny * (Nz + 1);
    double* Fx = (double*)malloc(NFx * 5 * 8);
    double* Gy = (double*)malloc(NFy * 5 * 8);
    double* Hz = (double*)malloc(NFz * 5 * 8);
    double* Ey_x = (double*)calloc(NFx, 8); double* Ez_x = (double*)calloc(NFx, 8);
    double* Ex_y = (double*)calloc(NFy, 8); double* Ez_y = (double*)calloc(NFy, 8);
    double* Ex_z = (double*)calloc(NFz, 8); double* Ey_z = (double*)calloc(NFz, 8);
#define FX(i, j, k) (((size_t)(k) * Ny + (j)) * (size_t)(Nx + 1) + (i))
#define FY(i, j, k) (((size_t)(k) * (Ny + 1) + (j)) * (size_t)Nx + (i))
#define FZ(i, j, k) (((size_t)(k) * Ny + (j)) * (size_t)Nx + (i))
    for (int k = 0; k < Nz; k++)
        for (int j = 0; j < Ny; j++)
            for (int i = 0; i <= Nx; i++) {
                int iL = px ? imod(i - 1, Nx) : iclamp(i - 1, 0, Nx - 1), iR = px ? imod(i, Nx) : iclamp(i, 0, Nx - 1);
                Prim a, WL, WR, b;
                recon(p, U0, iL, j, k, 0, &a, &WL);
                recon(p, U0, iR, j, k, 0, &WR, &b);
                double F[7];
                hlld_x(p, WL, WR, get_Bx_face(p, U0->Bx_f, i, j, k), F);
                size_t f = FX(i, j, k);
                for (int c = 0; c < 5; c++) Fx[f * 5 + c] = F[c];
                Ey_x[f] = F[6];
                Ez_x[f] = -F[5];
            }
    for (int k = 0; k < Nz; k++)
        for (int j = 0; j <= Ny; j++)
            for (int i = 0; i < Nx; i++) {
                int jD = py ? imod(j - 1, Ny) : iclamp(j - 1, 0, Ny - 1), jU = py ? imod(j, Ny) : iclamp(j, 0, Ny - 1);
                Prim a, WL, WR, b;
                recon(p, U0, i, jD, k, 1, &a, &WL);
                recon(p, U0, i, jU, k, 1, &WR, &b);
                double G[7];
                hlld_y(p, WL, WR, get_By_face(p, U0->By_f, i, j, k), G);
                size_t f = FY(i, j, k);
                for (int c = 0; c < 5; c++) Gy[f * 5 + c] = G[c];
                Ez_y[f] = G[5];
                Ex_y[f] = -G[6];
            }
    for (int k = 0; k <= Nz; k++)
        for (int j = 0; j < Ny; j++)
            for (int i = 0; i < Nx; i++) {
                int kB = pz ? imod(k - 1, Nz) : iclamp(k - 1, 0, Nz - 1), kF = pz ? imod(k, Nz) : iclamp(k, 0, Nz - 1);
                Prim a, WL, WR, b;
                recon(p, U0, i, j, kB, 2, &a, &WL);
                recon(p, U0, i, j, kF, 2, &WR, &b);
                double H[7];
                hlld_z(p, WL, WR, get_Bz_face(p, U0->Bz_f, i, j, k), H);
                size_t f = FZ(i, j, k);
                for (int c = 0; c < 5; c++) Hz[f * 5 + c] = H[c];
                Ex_z[f] = H[6];
                Ey_z[f] = -H[5];
            }
    double invdx = 1.0 / dx, invdy = 1.0 / dy, invdz = 1.0 / dz;
    for (int k = 0; k < Nz; k++)
        for (int j = 0; j < Ny; j++)
            for (int i = 0; i < Nx; i++) {
                size_t q = idx3(i, j, k, Nx, Ny);
                const double *FxR = Fx + FX(i + 1, j, k) * 5, *FxL = Fx + FX(i, j, k) * 5;
                const double *GyU = Gy + FY(i, j + 1, k) * 5, *GyD = Gy + FY(i, j, k) * 5;
                const double *HzF = Hz + FZ(i, j, k + 1) * 5, *HzB = Hz + FZ(i, j, k) * 5;
                double d[5];
                for (int c = 0; c < 5; c++)
                    d[c] = -(FxR[c] - FxL[c]) * invdx - (GyU[c] - GyD[c]) * invdy - (HzF[c] - HzB[c]) * invdz;
                U1->rho[q] += dt * d[0];
                U1->mx[q] += dt * d[1];
                U1->my[q] += dt * d[2];
                U1->mz[q] += dt * d[3];
                U1->E[q] += dt * d[4];
                double Bx = bx_cc(p, U0, i, j, k), By = by_cc(p, U0, i, j, k), Bz = bz_cc(p, U0, i, j, k);
                enforce_floors(p, &U1->rho[q], &U1->mx[q], &U1->my[q], &U1->mz[q], &U1->E[q], Bx, By, Bz);
            }
    /* edge EMFs (M3:836-939) */
    size_t NEx = (size_t)Nx * (Ny + 1) * (Nz + 1), NEy = (size_t)(Nx + 1) * Ny * (Nz + 1),
           NEz = (size_t)(Nx + 1) * (Ny + 1) * Nz;
    double* Ex_e = (double*)calloc(NEx, 8); double* Ey_e = (double*)calloc(NEy, 8); double* Ez_e = (double*)calloc(NEz, 8);
#define EXY(i, j, k) Ex_y[FY(MAPC(i, Nx, px), MAPF(j, Ny, py), MAPC(k, Nz, pz))]
#define EZX(i, j, k) Ez_x[FX(MAPF(i, Nx, px), MAPC(j, Ny, py), MAPC(k, Nz, pz))]
#define EYX(i, j, k) Ey_x[FX(MAPF(i, Nx, px), MAPC(j, Ny, py), MAPC(k, Nz, pz))]
#define EXZ(i, j, k) Ex_z[FZ(MAPC(i, Nx, px), MAPC(j, Ny, py), MAPF(k, Nz, pz))]
#define EYZ(i, j, k) Ey_z[FZ(MAPC(i, Nx, px), MAPC(j, Ny, py), MAPF(k, Nz, pz))]
#define EZY(i, j, k) Ez_y[FY(MAPC(i, Nx, px), MAPF(j, Ny, py), MAPC(k, Nz, pz))]
    for (int kf = 0; kf <= Nz; kf++)
        for (int jf = 0; jf <= Ny; jf++)
            for (int i = 0; i < Nx; i++) {
                double ex_ideal = 0.25 * (EXY(i, jf, kf - 1) + EXY(i, jf, kf) + EXZ(i, jf - 1, kf) + EXZ(i, jf, kf));
                double ex_res = 0.0;
                if (p->eta > 0.0) {
                    double Bzp = bz_cc(p, U0, i, jf, kf), Bzm = bz_cc(p, U0, i, jf - 1, kf);
                    double Byp = by_cc(p, U0, i, jf, kf), Bym = by_cc(p, U0, i, jf, kf - 1);
                    ex_res = p->eta * ((Bzp - Bzm) / dy - (Byp - Bym) / dz);
                }
                Ex_e[((size_t)kf * (Ny + 1) + jf) * (size_t)Nx + i] = ex_ideal + ex_res;
            }
    for (int kf = 0; kf <= Nz; kf++)
        for (int j = 0; j < Ny; j++)
            for (int i = 0; i <= Nx; i++) {
                double ey_ideal = 0.25 * (EYX(i, j, kf - 1) + EYX(i, j, kf) + EYZ(i - 1, j, kf) + EYZ(i, j, kf));
                double ey_res = 0.0;
                if (p->eta > 0.0) {
                    double Bxp = bx_cc(p, U0, i, j, kf), Bxm = bx_cc(p, U0, i, j, kf - 1);
                    double Bzp = bz_cc(p, U0, i, j, kf), Bzm = bz_cc(p, U0, i - 1, j, kf);
                    ey_res = p->eta * ((Bxp - Bxm) / dz - (Bzp - Bzm) / dx);
                }
                Ey_e[((size_t)kf * Ny + j) * (size_t)(Nx + 1) + i] = ey_ideal + ey_res;
            }
    for (int k = 0; k < Nz; k++)
        for (int jf = 0; jf <= Ny; jf++)
            for (int i = 0; i <= Nx; i++) {
                double ez_ideal = 0.25 * (EZX(i, jf - 1, k) + EZX(i, jf, k) + EZY(i - 1, jf, k) + EZY(i, jf, k));
                double ez_res = 0.0;
                if (p->eta > 0.0) {
                    double Byp = by_cc(p, U0, i, jf, k), Bym = by_cc(p, U0, i - 1, jf, k);
                    double Bxp = bx_cc(p, U0, i, jf, k), Bxm = bx_cc(p, U0, i, jf - 1, k);
                    ez_res = p->eta * ((Byp - Bym) / dx - (Bxp - Bxm) / dy);
                }
                Ez_e[((size_t)k * (Ny + 1) + jf) * (size_t)(Nx + 1) + i] = ez_ideal + ez_res;
            }
    /* CT (M3:941-1015) */
    for (int k = 0; k < Nz; k++)
        for (int j = 0; j < Ny; j++)
            for (int i = 0; i <= Nx; i++) {
                int jf0 = j, jf1 = j + 1, kf0 = k, kf1 = k + 1;
                if (py) { jf0 = imod(jf0, Ny + 1); jf1 = imod(jf1, Ny + 1); } else { jf0 = iclamp(jf0, 0, Ny); jf1 = iclamp(jf1, 0, Ny); }
                if (pz) { kf0 = imod(kf0, Nz + 1); kf1 = imod(kf1, Nz + 1); } else { kf0 = iclamp(kf0, 0, Nz); kf1 = iclamp(kf1, 0, Nz); }
                double ezU = Ez_e[((size_t)k * (Ny + 1) + jf1) * (size_t)(Nx + 1) + i];
                double ezD = Ez_e[((size_t)k * (Ny + 1) + jf0) * (size_t)(Nx + 1) + i];
                double eyF = Ey_e[((size_t)kf1 * Ny + j) * (size_t)(Nx + 1) + i];
                double eyB = Ey_e[((size_t)kf0 * Ny + j) * (size_t)(Nx + 1) + i];
                const double d = -(dt / dy) * (ezU - ezD) + (dt / dz) * (eyF - eyB);
                const size_t f = FX(i, j, k);
                if (p->ct_comp) or_ct_add(U0->Bx_f[f], U0->Bx_lo[f], d, &U1->Bx_f[f], &U1->Bx_lo[f]);
                else U1->Bx_f[f] += d;
            }
    for (int k = 0; k < Nz; k++)
        for (int j = 0; j <= Ny; j++)
            for (int i = 0; i < Nx; i++) {
                int if0 = i, if1 = i + 1, kf0 = k, kf1 = k + 1;
                if (px) { if0 = imod(if0, Nx + 1); if1 = imod(if1, Nx + 1); } else { if0 = iclamp(if0, 0, Nx); if1 = iclamp(if1, 0, Nx); }
                if (pz) { kf0 = imod(kf0, Nz + 1); kf1 = imod(kf1, Nz + 1); } else { kf0 = iclamp(kf0, 0, Nz); kf1 = iclamp(kf1, 0, Nz); }
                double ezR = Ez_e[((size_t)k * (Ny + 1) + j) * (size_t)(Nx + 1) + if1];
                double ezL = Ez_e[((size_t)k * (Ny + 1) + j) * (size_t)(Nx + 1) + if0];
                double exF = Ex_e[((size_t)kf1 * (Ny + 1) + j) * (size_t)Nx + i];
                double exB = Ex_e[((size_t)kf0 * (Ny + 1) + j) * (size_t)Nx + i];
                const double d = +(dt / dx) * (ezR - ezL) - (dt / dz) * (exF - exB);
                const size_t f = FY(i, j, k);
                if (p->ct_comp) or_ct_add(U0->By_f[f], U0->By_lo[f], d, &U1->By_f[f], &U1->By_lo[f]);
                else U1->By_f[f] += d;
            }
    for (int k = 0; k <= Nz; k++)
        for (int j = 0; j < Ny; j++)
            for (int i = 0; i < Nx; i++) {
                int if0 = i, if1 = i + 1, jf0 = j, jf1 = j + 1;
                if (px) { if0 = imod(if0, Nx + 1); if1 = imod(if1, Nx + 1); } else { if0 = iclamp(if0, 0, Nx); if1 = iclamp(if1, 0, Nx); }
                if (py) { jf0 = imod(jf0, Ny + 1); jf1 = imod(jf1, Ny + 1); } else { jf0 = iclamp(jf0, 0, Ny); jf1 = iclamp(jf1, 0, Ny); }
                double eyR = Ey_e[((size_t)k * Ny + j) * (size_t)(Nx + 1) + if1];
                double eyL = Ey_e[((size_t)k * Ny + j) * (size_t)(Nx + 1) + if0];
                double exU = Ex_e[((size_t)k * (Ny + 1) + jf1) * (size_t)Nx + i];
                double exD = Ex_e[((size_t)k * (Ny + 1) + jf0) * (size_t)Nx + i];
                const double d = -(dt / dx) * (eyR - eyL) + (dt / dy) * (exU - exD);
                const size_t f = FZ(i, j, k);
                if (p->ct_comp) or_ct_add(U0->Bz_f[f], U0->Bz_lo[f], d, &U1->Bz_f[f], &U1->Bz_lo[f]);
                else U1->Bz_f[f] += d;
            }
    add_joule_heating(p, U1, dx, dy, dz, dt);
    for (int k = 0; k < Nz; k++)
        for (int j = 0; j < Ny; j++)
            for (int i = 0; i < Nx; i++) {
                size_t q = idx3(i, j, k, Nx, Ny);
                double Bx = bx_cc(p, U1, i, j, k), By = by_cc(p, U1, i, j, k), Bz = bz_cc(p, U1, i, j, k);
                enforce_floors(p, &U1->rho[q], &U1->mx[q], &U1->my[q], &U1->mz[q], &U1->E[q], Bx, By, Bz);
            }
    free(Fx); free(Gy); free(Hz); free(Ey_x); free(Ez_x); free(Ex_y); free(Ez_y); free(Ex_z); free(Ey_z);
    free(Ex_e); free(Ey_e); free(Ez_e);
#undef FX
#undef FY
#undef FZ
}

static void state_alloc(const or_m3_params* p, or_m3_state* S) {
    size_t N = (size_t)p->Nx * p->Ny * p->Nz;
    S->rho = malloc(N * 8); S->mx = malloc(N * 8); S->my = malloc(N * 8); S->mz = malloc(N * 8); S->E = malloc(N * 8);
    S->Bx_f = malloc((size_t)(p->Nx + 1) * p->Ny * p->Nz * 8);
    S->By_f = malloc((size_t)p->Nx * (p->Ny + 1) * p->Nz * 8);
    S->Bz_f = malloc((size_t)p->Nx * p->Ny * (p->Nz + 1) * 8);
    S->Bx_lo = calloc((size_t)(p->Nx + 1) * p->Ny * p->Nz, sizeof(float));
    S->By_lo = calloc((size_t)p->Nx * (p->Ny + 1) * p->Nz, sizeof(float));
    S->Bz_lo = calloc((size_t)p->Nx * p->Ny * (p->Nz + 1), sizeof(float));
}
static void state_free(or_m3_state* S) {
    free(S->rho); free(S->mx); free(S->my); free(S->mz); free(S->E); free(S->Bx_f); free(S->By_f); free(S->Bz_f);
    free(S->Bx_lo); free(S->By_lo); free(S->Bz_lo);
}

/* M3:1034-1066 */
void or_m3_rk2_step(const or_m3_params* p, or_m3_state* U, double dx, double dy, double dz, double dt) {
    or_m3_state U1, U2;
    state_alloc(p, &U1);
    state_alloc(p, &U2);
    or_m3_advance_stage(p, U, &U1, dx, dy, dz, dt);
    or_m3_advance_stage(p, &U1, &U2, dx, dy, dz, dt);
    size_t N = (size_t)p->Nx * p->Ny * p->Nz;
    for (size_t q = 0; q < N; q++) {
        U->rho[q] = 0.5 * (U->rho[q] + U2.rho[q]);
        U->mx[q] = 0.5 * (U->mx[q] + U2.mx[q]);
        U->my[q] = 0.5 * (U->my[q] + U2.my[q]);
        U->mz[q] = 0.5 * (U->mz[q] + U2.mz[q]);
        U->E[q] = 0.5 * (U->E[q] + U2.E[q]);
    }
    size_t NFx = (size_t)(p->Nx + 1) * p->Ny * p->Nz, NFy = (size_t)p->Nx * (p->Ny + 1) * p->Nz,
           NFz = (size_t)p->Nx * p->Ny * (p->Nz + 1);
    if (p->ct_comp) {
        for (size_t f = 0; f < NFx; f++) or_ct_avg(U->Bx_f[f], U->Bx_lo[f], U2.Bx_f[f], U2.Bx_lo[f], &U->Bx_f[f], &U->Bx_lo[f]);
        for (size_t f = 0; f < NFy; f++) or_ct_avg(U->By_f[f], U->By_lo[f], U2.By_f[f], U2.By_lo[f], &U->By_f[f], &U->By_lo[f]);
        for (size_t f = 0; f < NFz; f++) or_ct_avg(U->Bz_f[f], U->Bz_lo[f], U2.Bz_f[f], U2.Bz_lo[f], &U->Bz_f[f], &U->Bz_lo[f]);
    } else {
        for (size_t f = 0; f < NFx; f++) U->Bx_f[f] = 0.5 * (U->Bx_f[f] + U2.Bx_f[f]);
        for (size_t f = 0; f < NFy; f++) U->By_f[f] = 0.5 * (U->By_f[f] + U2.By_f[f]);
        for (size_t f = 0; f < NFz; f++) U->Bz_f[f] = 0.5 * (U->Bz_f[f] + U2.Bz_f[f]);
    }
    for (int k = 0; k < p->Nz; k++)
        for (int j = 0; j < p->Ny; j++)
            for (int i = 0; i < p->Nx; i++) {
                size_t q = idx3(i, j, k, p->Nx, p->Ny);
                double Bx = bx_cc(p, U, i, j, k), By = by_cc(p, U, i, j, k), Bz = bz_cc(p, U, i, j, k);
                enforce_floors(p, &U->rho[q], &U->mx[q], &U->my[q], &U->mz[q], &U->E[q], Bx, By, Bz);
            }
    state_free(&U1);
    state_free(&U2);
}

/* M3:1070-1078 */
void or_m3_divb(const or_m3_params* p, const or_m3_state* U, double dx, double dy, double dz, double* out) {
    for (int k = 0; k < p->Nz; k++)
        for (int j = 0; j < p->Ny; j++)
            for (int i = 0; i < p->Nx; i++) {
                double BxR = get_Bx_face(p, U->Bx_f, i + 1, j, k), BxL = get_Bx_face(p, U->Bx_f, i, j, k);
                double ByU = get_By_face(p, U->By_f, i, j + 1, k), ByD = get_By_face(p, U->By_f, i, j, k);
                double BzF = get_Bz_face(p, U->Bz_f, i, j, k + 1), BzB = get_Bz_face(p, U->Bz_f, i, j, k);
                out[idx3(i, j, k, p->Nx, p->Ny)] = (BxR - BxL) / dx + (ByU - ByD) / dy + (BzF - BzB) / dz;
            }
}

/* The loop of M3:1285-1291 for a fixed number of steps (no output). */
double or_m3_run(const or_m3_params* p, or_m3_state* U, double dx, double dy, double dz, int nsteps, double t0,
                 double t_end, double* dts_out) {
    double t = t0;
    for (int s = 0; s < nsteps && t < t_end; s++) {
        double dt = or_m3_compute_dt(p, U, dx, dy, dz);
        if (t + dt > t_end) dt = t_end - t;
        or_m3_rk2_step(p, U, dx, dy, dz, dt);
        t += dt;
        if (dts_out) dts_out[s] = dt;
    }
    return t;
}

/* The snapshot frame (the store block of M3:1293-1325): 11 volumes {rho, P, Mach, u, v, w, vabs, Bx,
 * By, Bz, divB} of Nx*Ny*Nz cells. */
void or_m3_frame(const or_m3_params* p, const or_m3_state* U, double dx, double dy, double dz, double* out) {
    const size_t N = (size_t)p->Nx * (size_t)p->Ny * (size_t)p->Nz;
    for (int k = 0; k < p->Nz; k++)
        for (int j = 0; j < p->Ny; j++)
            for (int i = 0; i < p->Nx; i++) {
                const size_t q = idx3(i, j, k, p->Nx, p->Ny);
                const double Bx = bx_cc(p, U, i, j, k), By = by_cc(p, U, i, j, k), Bz = bz_cc(p, U, i, j, k);
                const Prim W = cons_to_prim(p, U->rho[q], U->mx[q], U->my[q], U->mz[q], U->E[q], Bx, By, Bz);
                const double cs = sqrt(sound_speed2(p, &W));
                const double vabs = sqrt(W.u * W.u + W.v * W.v + W.w * W.w);
                const double Mach = vabs / dmax(cs, 1e-12);
                const double BxR = get_Bx_face(p, U->Bx_f, i + 1, j, k), BxL = get_Bx_face(p, U->Bx_f, i, j, k);
                const double ByU = get_By_face(p, U->By_f, i, j + 1, k), ByD = get_By_face(p, U->By_f, i, j, k);
                const double BzF = get_Bz_face(p, U->Bz_f, i, j, k + 1), BzB = get_Bz_face(p, U->Bz_f, i, j, k);
                const double f[11] = {W.rho, W.P, Mach, W.u, W.v, W.w, vabs, W.Bx, W.By, W.Bz,
                                      (BxR - BxL) / dx + (ByU - ByD) / dy + (BzF - BzB) / dz};
                for (int c = 0; c < 11; c++) out[(size_t)c * N + q] = f[c];
            }
}
