// hllc.hpp -- the two HLLC Riemann solvers of the hydro codes, host + device, FP64 or FP32.
//
//   hllc_eh<R>: euler2d_gravity_hllc.cpp hllc_x / hllc_y (EH:125-219): cons from floored prims,
//               SM guard |den| > 1e-14, star energy ((S-u)E - P u + P S_M)/(S - S_M) (EH:154, 202).
//   hllc_ns<R>: navier_stokes_2d_hllc_poisson.cpp hllc_x (NS:212-268): textbook p* = P + rho(S-u)(S_M-u)
//               in the star energy, SM guard |den| > 0; hllc_y swaps (u, v) and the momentum
//               fluxes (NS:270-282).
// Both are written for the x direction on (rho, un, ut, P); the y direction passes (rho, v, u, P)
// and swaps the two momentum outputs.  Operation order follows the reference so the FP64 result is
// bit-identical when compiled without FP contraction.
#pragma once
#include "common.hpp"

namespace cc {

template <typename R>
struct Hydro4 {
    R rho, un, ut, P;  // density, normal and tangential velocity, pressure
};

// Output {f_rho, f_mn, f_mt, f_E} (normal / tangential momentum fluxes).  YAXIS keeps the
// reference's product order rho*u*v = (rho*u)*v for the cross flux of the y sweep (EH:119).
template <typename R, bool YAXIS>
__host__ __device__ __forceinline__ void hllc_eh(R gamma, R floor_rho, const Hydro4<R>& L, const Hydro4<R>& Rt,
                                                 R out[4]) {
    const R gm1 = gamma - R(1);
    const R aL = sqrt(gamma * L.P / smax(L.rho, floor_rho));
    const R aR = sqrt(gamma * Rt.P / smax(Rt.rho, floor_rho));
    const R SL = smin(L.un - aL, Rt.un - aR);
    const R SR = smax(L.un + aL, Rt.un + aR);
    // prim_to_cons (EH:102-109) of both states
    const R rL = smax(L.rho, floor_rho), rR = smax(Rt.rho, floor_rho);
    const R mnL = rL * L.un, mtL = rL * L.ut, mnR = rR * Rt.un, mtR = rR * Rt.ut;
    const R EL = R(0.5) * rL * (L.un * L.un + L.ut * L.ut) + L.P / gm1;
    const R ER = R(0.5) * rR * (Rt.un * Rt.un + Rt.ut * Rt.ut) + Rt.P / gm1;
    const R num = (Rt.P - L.P) + rL * L.un * (SL - L.un) - rR * Rt.un * (SR - Rt.un);
    const R den = rL * (SL - L.un) - rR * (SR - Rt.un);
    const R SM = (fabs(den) > R(1e-14)) ? (num / den) : R(0);
    // physical fluxes (EH:113-123); the energy there is P/(g-1) + kin, the same sum as EL / ER
    const bool up_left = R(0) <= SL, up_right = R(0) >= SR, star_left = R(0) < SM;
    const bool use_left = up_left || (!up_right && star_left);
    const R rho = use_left ? L.rho : Rt.rho, un = use_left ? L.un : Rt.un, ut = use_left ? L.ut : Rt.ut;
    const R P = use_left ? L.P : Rt.P;
    const R c0 = use_left ? rL : rR, cn = use_left ? mnL : mnR, ct = use_left ? mtL : mtR, cE = use_left ? EL : ER;
    const R Ef = P / gm1 + R(0.5) * rho * (un * un + ut * ut);
    const R F0 = rho * un, F1 = rho * un * un + P, F2 = YAXIS ? rho * ut * un : rho * un * ut, F3 = (Ef + P) * un;
    if (up_left || up_right) {
        out[0] = F0; out[1] = F1; out[2] = F2; out[3] = F3;
        return;
    }
    const R S = use_left ? SL : SR;
    const R rS = qdiv<R>(c0 * (S - un), S - SM);
    const R ES = qdiv<R>((S - un) * cE - P * un + P * SM, S - SM);
    out[0] = F0 + S * (rS - c0);
    out[1] = F1 + S * (rS * SM - cn);
    out[2] = F2 + S * (rS * ut - ct);
    out[3] = F3 + S * (ES - cE);
}

template <typename R>
__host__ __device__ __forceinline__ void hllc_ns(R gamma, R floor_rho, const Hydro4<R>& L, const Hydro4<R>& Rt,
                                                 R out[4]) {
    const R gm1 = gamma - R(1);
    const R aL = sqrt(gamma * L.P / smax(L.rho, floor_rho));  // sound_speed NS:147-149
    const R aR = sqrt(gamma * Rt.P / smax(Rt.rho, floor_rho));
    const R EL = L.P / gm1 + R(0.5) * L.rho * (L.un * L.un + L.ut * L.ut);
    const R ER = Rt.P / gm1 + R(0.5) * Rt.rho * (Rt.un * Rt.un + Rt.ut * Rt.ut);
    const R SL = smin(L.un - aL, Rt.un - aR);
    const R SR = smax(L.un + aL, Rt.un + aR);
    const R num = (Rt.P - L.P) + L.rho * L.un * (SL - L.un) - Rt.rho * Rt.un * (SR - Rt.un);
    const R den = L.rho * (SL - L.un) - Rt.rho * (SR - Rt.un);
    const R SM = (fabs(den) > R(0)) ? (num / den) : R(0);
    const bool up_left = R(0) <= SL, up_right = R(0) >= SR, star_left = R(0) < SM;
    const bool use_left = up_left || (!up_right && star_left);
    const R d = use_left ? L.rho : Rt.rho, u = use_left ? L.un : Rt.un, v = use_left ? L.ut : Rt.ut;
    const R P = use_left ? L.P : Rt.P, E = use_left ? EL : ER;
    // flux_x_from_prim NS:198-210 (its energy is the same sum as EL / ER)
    const R F0 = d * u, F1 = d * u * u + P, F2 = d * u * v, F3 = (E + P) * u;
    if (up_left || up_right) {
        out[0] = F0; out[1] = F1; out[2] = F2; out[3] = F3;
        return;
    }
    const R S = use_left ? SL : SR;
    const R fac = d * (S - u) / (S - SM);
    const R pstar = P + d * (S - u) * (SM - u);
    const R Es = ((S - u) * E - P * u + pstar * SM) / (S - SM);
    out[0] = F0 + S * (fac - d);
    out[1] = F1 + S * (fac * SM - d * u);
    out[2] = F2 + S * (fac * v - d * v);
    out[3] = F3 + S * (Es - E);
}

// The EH HLLC with two transverse velocities, for the 3D Euler step along axis AX (0, 1, 2) on
// primitives (rho, u, v, w, P); out = {f_rho, f_mx, f_my, f_mz, f_E}.  Kinetic energy
// 0.5 rho ((u^2 + v^2) + w^2) and the cross momentum flux rho*a*b (a < b in the order u, v, w), so
// that with w = 0 on x / y faces every value equals hllc_eh's bit for bit.
template <typename R, int AX>
__host__ __device__ __forceinline__ void hllc_e3(R gamma, R floor_rho, const R L[5], const R Rt[5], R out[5]) {
    const R gm1 = gamma - R(1);
    const R unL = L[1 + AX], unR = Rt[1 + AX];
    const R aL = sqrt(qdiv<R>(gamma * L[4], smax(L[0], floor_rho)));
    const R aR = sqrt(qdiv<R>(gamma * Rt[4], smax(Rt[0], floor_rho)));
    const R SL = smin(unL - aL, unR - aR);
    const R SR = smax(unL + aL, unR + aR);
    const R rL = smax(L[0], floor_rho), rR = smax(Rt[0], floor_rho);
    const R EL = R(0.5) * rL * ((L[1] * L[1] + L[2] * L[2]) + L[3] * L[3]) + qdiv<R>(L[4], gm1);
    const R ER = R(0.5) * rR * ((Rt[1] * Rt[1] + Rt[2] * Rt[2]) + Rt[3] * Rt[3]) + qdiv<R>(Rt[4], gm1);
    const R num = (Rt[4] - L[4]) + rL * unL * (SL - unL) - rR * unR * (SR - unR);
    const R den = rL * (SL - unL) - rR * (SR - unR);
    const R SM = (fabs(den) > R(1e-14)) ? qdiv<R>(num, den) : R(0);
    const bool up_left = R(0) <= SL, up_right = R(0) >= SR, star_left = R(0) < SM;
    const bool use_left = up_left || (!up_right && star_left);
    const R* W = use_left ? L : Rt;
    const R rho = W[0], un = W[1 + AX], P = W[4];
    const R Ef = qdiv<R>(P, gm1) + R(0.5) * rho * ((W[1] * W[1] + W[2] * W[2]) + W[3] * W[3]);
    R F[5];
    F[0] = rho * un;
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        if (c == AX) F[1 + c] = rho * un * un + P;
        else if (c < AX) F[1 + c] = rho * W[1 + c] * un;
        else F[1 + c] = rho * un * W[1 + c];
    }
    F[4] = (Ef + P) * un;
    if (up_left || up_right) {
#pragma unroll
        for (int c = 0; c < 5; ++c) out[c] = F[c];
        return;
    }
    const R S = use_left ? SL : SR;
    const R c0 = use_left ? rL : rR, cE = use_left ? EL : ER;
    const R rS = qdiv<R>(c0 * (S - un), S - SM);
    const R ES = qdiv<R>((S - un) * cE - P * un + P * SM, S - SM);
    out[0] = F[0] + S * (rS - c0);
#pragma unroll
    for (int m = 0; m < 3; ++m) {
        const R mS = (m == AX) ? rS * SM : rS * W[1 + m];
        out[1 + m] = F[1 + m] + S * (mS - c0 * W[1 + m]);
    }
    out[4] = F[4] + S * (ES - cE);
}

}  // namespace cc
