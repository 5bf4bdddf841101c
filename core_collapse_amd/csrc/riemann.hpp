// riemann.hpp -- device-side equations of state, PLM-MC limiter and the HLLD / HLLC Riemann solvers.
//
// The solvers are templated on the arithmetic type R (double for the FP64 step, float for the
// mixed-precision step) and on an EOS policy that supplies the total energy and sound speed, so
// the 2D ideal-gas code (mhd2d_resistive_hlld_ct_gravity.cpp), the 3D hybrid-EOS code
// (mhd3d_resistive_hlld_ct.cpp) and the 3D ideal-gas variant share one HLLD.
#pragma once
#include "common.hpp"

namespace cc {

template <typename R>
struct Prim {
    R rho, u, v, w, P, Bx, By, Bz;
};

template <typename R>
struct Flux7 {  // x-normal layout: rho, mx, my, mz, E, By, Bz
    R f[7];
};

// MC limiter, M2:106-112 (same in M3:106-112, EH:85-92).
template <typename R>
__host__ __device__ __forceinline__ R minmod(R a, R b) {
    if (a * b <= R(0)) return R(0);
    return (fabs(a) < fabs(b)) ? a : b;
}
template <typename R>
__host__ __device__ __forceinline__ R mc_limiter(R dm, R dp) {
    return minmod<R>(R(0.5) * (dm + dp), minmod<R>(R(2) * dm, R(2) * dp));
}

// ---------------------------------------------------------------- EOS policies
// An EOS policy provides aux(W) (per-state data shared by the energy and sound-speed calls, e.g.
// the cold-curve evaluation of the hybrid EOS), total_energy(W, aux) and sound_speed2(W, aux).
struct NoAux {};

// Ideal gas with the expression order of M2 (E = P/(g-1) + 0.5 rho v^2 + 0.5 B^2, M2:357 / M2:397 / M2:482).
template <typename R>
struct IdealEOS2D {
    using Aux = NoAux;
    R gm1;        // gamma - 1
    R gamma;
    R floor_rho;
    __host__ __device__ __forceinline__ Aux aux(const Prim<R>&) const { return Aux{}; }
    __host__ __device__ __forceinline__ R total_energy(const Prim<R>& W, const Aux&) const {
        R B2 = W.Bx * W.Bx + W.By * W.By + W.Bz * W.Bz;
        return qdiv<R>(W.P, gm1) + R(0.5) * W.rho * (W.u * W.u + W.v * W.v + W.w * W.w) + R(0.5) * B2;
    }
    __host__ __device__ __forceinline__ R sound_speed2(const Prim<R>& W, const Aux&) const {  // M2:148-150
        return qdiv<R>(gamma * W.P, smax(W.rho, floor_rho));
    }
};

// Ideal gas in the 3D code's form (total energy as kin + mag + eint, M3:154-163 with a gamma-law
// internal energy; sound speed gamma P / rho as M2:148-150).
template <typename R>
struct IdealEOS3D {
    using Aux = NoAux;
    R gm1, gamma, floor_rho;
    __host__ __device__ __forceinline__ Aux aux(const Prim<R>&) const { return Aux{}; }
    __host__ __device__ __forceinline__ R total_energy(const Prim<R>& W, const Aux&) const {
        const R rho = smax(W.rho, floor_rho);
        const R kin = R(0.5) * rho * (W.u * W.u + W.v * W.v + W.w * W.w);
        const R mag = R(0.5) * (W.Bx * W.Bx + W.By * W.By + W.Bz * W.Bz);
        return kin + mag + qdiv<R>(W.P, gm1);
    }
    __host__ __device__ __forceinline__ R sound_speed2(const Prim<R>& W, const Aux&) const {
        return qdiv<R>(gamma * W.P, smax(W.rho, floor_rho));
    }
};

template <typename R> __host__ __device__ __forceinline__ R rpow(R x, R y);
template <> __host__ __device__ __forceinline__ double rpow<double>(double x, double y) { return pow(x, y); }
template <> __host__ __device__ __forceinline__ float rpow<float>(float x, float y) { return powf(x, y); }

// The hybrid cold + thermal + bounce EOS of M3:114-212.  K2, the cold energies at rho_nuc and the
// bounce powers are constants of the parameters and are evaluated once on the host.
template <typename R>
struct HybridEOS {
    struct Aux { R Pc, ec, gc; };
    R gamma1, gamma2, K1, K2, rho_nuc, gamma_th, rho_bounce, gamma_bounce, Kbounce;
    R ec1_n, ec2_n, rb_g, rb_g1;  // K1 rho_nuc^(g1-1)/(g1-1), K2 rho_nuc^(g2-1)/(g2-1), rho_b^gb, rho_b^(gb-1)
    R floor_rho, floor_P;
    __host__ __device__ __forceinline__ Aux cold(R rho) const {  // eos_cold, M3:118-144
        rho = smax(rho, floor_rho);
        if (rho < rho_nuc) {
            return Aux{K1 * rpow<R>(rho, gamma1), K1 * rpow<R>(rho, gamma1 - R(1)) / (gamma1 - R(1)), gamma1};
        }
        R Pc = K2 * rpow<R>(rho, gamma2);
        R ec = ec1_n + (K2 * rpow<R>(rho, gamma2 - R(1)) / (gamma2 - R(1)) - ec2_n);
        R gc = gamma2;
        if (rho >= rho_bounce) {
            const R pb = Kbounce * (rpow<R>(rho, gamma_bounce) - rb_g);
            const R eb = Kbounce * (rpow<R>(rho, gamma_bounce - R(1)) - rb_g1) / (gamma_bounce - R(1));
            Pc += pb;
            ec += eb;
            gc = (gamma2 * (Pc - pb) + gamma_bounce * pb) / smax(Pc, floor_P);
        }
        return Aux{Pc, ec, gc};
    }
    __host__ __device__ __forceinline__ Aux aux(const Prim<R>& W) const { return cold(smax(W.rho, floor_rho)); }
    __host__ __device__ __forceinline__ R total_energy(const Prim<R>& W, const Aux& c) const {  // M3:154-163
        const R rho = smax(W.rho, floor_rho);
        const R kin = R(0.5) * rho * (W.u * W.u + W.v * W.v + W.w * W.w);
        const R mag = R(0.5) * (W.Bx * W.Bx + W.By * W.By + W.Bz * W.Bz);
        const R Pth = smax(R(0), W.P - c.Pc);
        const R eth = (gamma_th > R(1)) ? (Pth / ((gamma_th - R(1)) * rho)) : R(0);
        return kin + mag + rho * (c.ec + eth);
    }
    __host__ __device__ __forceinline__ R sound_speed2(const Prim<R>& W, const Aux& c) const {  // M3:206-212
        const R rho = smax(W.rho, floor_rho);
        const R Pth = smax(R(0), W.P - c.Pc);
        return smax(R(0), (c.gc * c.Pc + gamma_th * Pth) / rho);
    }
};

// Fast magnetosonic speed along the axis whose field component is Bn (M2:152-174, M3:214-248).
template <typename R, typename EOS>
__host__ __device__ __forceinline__ R fast_speed(const EOS& eos, const Prim<R>& W, R Bn, const typename EOS::Aux& ax) {
    R a2 = eos.sound_speed2(W, ax);
    R b2 = W.Bx * W.Bx + W.By * W.By + W.Bz * W.Bz;
    R rho = smax(W.rho, Tiny<R>::v());
    R vA2 = qdiv<R>(b2, rho);
    R vAn2 = qdiv<R>(Bn * Bn, rho);
    R disc = (a2 + vA2) * (a2 + vA2) - R(4) * a2 * vAn2;
    disc = smax(R(0), disc);
    R cf2 = R(0.5) * ((a2 + vA2) + sqrt(disc));
    return sqrt(smax(R(0), cf2));
}
template <typename R, typename EOS>
__host__ __device__ __forceinline__ R fast_speed(const EOS& eos, const Prim<R>& W, R Bn) {
    return fast_speed<R>(eos, W, Bn, eos.aux(W));
}

template <typename R>
__host__ __device__ __forceinline__ void flux_ideal_x(const Prim<R>& W, R E, R F[7]) {  // M2:351-370 / M3:396-413
    R B2 = W.Bx * W.Bx + W.By * W.By + W.Bz * W.Bz;
    R pt = W.P + R(0.5) * B2;
    R vdotB = W.u * W.Bx + W.v * W.By + W.w * W.Bz;
    F[0] = W.rho * W.u;
    F[1] = W.rho * W.u * W.u + pt - W.Bx * W.Bx;
    F[2] = W.rho * W.u * W.v - W.Bx * W.By;
    F[3] = W.rho * W.u * W.w - W.Bx * W.Bz;
    F[4] = (E + pt) * W.u - W.Bx * vdotB;
    F[5] = W.u * W.By - W.v * W.Bx;
    F[6] = W.u * W.Bz - W.w * W.Bx;
}

// HLLD, x-normal, with the reference's formulas (M2:372-530 / M3:415-568), including its
// non-textbook v*, w* denominators, the single-star energy formula reused for the double-star
// states, the double-star flux through S_L / S_R, and the HLL fallback when |Bn| < 1e-12.
// bn_small carries that fallback test evaluated on the FP64 face field.
// If `degenerate` is non-null, it is set for a face whose low-precision evaluation cannot be trusted;
// the caller then recomputes the face in FP64 (the outputs of a degenerate face are zeros: the solver
// stops there).  Without the guard that is a face within the relative margin `tau` of the Alfven
// degeneracy dL = 0 / dR = 0 (M2:419-442), where the star states are ill-conditioned.  With the guard
// (below), faces with |d| under guard * m take HLL, which does not divide by d, and the others have star
// states whose condition number in d is at most 1 / guard: the one decision that rounding can flip is
// the guard test itself, so only a face with |d| within the relative band `tau` of guard * m (either
// side) is recomputed.
// `guard` > 0 (no reference counterpart; 0 is the reference) sends a face whose d_L or d_R is below
// guard * max(Bn^2, |rho s (S - S_M)|) to the reference's own HLL flux (the |Bn| < 1e-12 branch):
// the reference divides the star states by d with only |d| > 1e-14 as a guard, so a face whose outer
// wave speed nearly meets the star-region Alfven speed gets star states of O(100) and more.
template <typename R, typename EOS>
__host__ __device__ __forceinline__ void hlld_x(const EOS& eos, Prim<R> WL, Prim<R> WR, R Bn, bool bn_small, R out[7],
                                                bool* degenerate = nullptr, R tau = R(0), R guard = R(0)) {
    WL.Bx = Bn;
    WR.Bx = Bn;
    const R rhoL = WL.rho, rhoR = WR.rho, uL = WL.u, uR = WR.u;
    const typename EOS::Aux axL = eos.aux(WL), axR = eos.aux(WR);
    const R EL = eos.total_energy(WL, axL), ER = eos.total_energy(WR, axR);
    R FL[7], FR[7];
    flux_ideal_x<R>(WL, EL, FL);
    flux_ideal_x<R>(WR, ER, FR);
    R cfL = fast_speed<R>(eos, WL, WL.Bx, axL);
    R cfR = fast_speed<R>(eos, WR, WR.Bx, axR);
    R SL = smin(uL - cfL, uR - cfR);
    R SR = smax(uL + cfL, uR + cfR);
    if (R(0) <= SL) {
#pragma unroll
        for (int k = 0; k < 7; k++) out[k] = FL[k];
        return;
    }
    if (R(0) >= SR) {
#pragma unroll
        for (int k = 0; k < 7; k++) out[k] = FR[k];
        return;
    }
    auto hll = [&]() {  // M2:449-467 (HLL)
        R UL[7] = {WL.rho, WL.rho * WL.u, WL.rho * WL.v, WL.rho * WL.w, EL, WL.By, WL.Bz};
        R UR[7] = {WR.rho, WR.rho * WR.u, WR.rho * WR.v, WR.rho * WR.w, ER, WR.By, WR.Bz};
        R inv = qdiv<R>(R(1), SR - SL);
#pragma unroll
        for (int k = 0; k < 7; k++) out[k] = (SR * FL[k] - SL * FR[k] + SL * SR * (UR[k] - UL[k])) * inv;
    };
    if (bn_small) {
        hll();
        return;
    }
    const R BL2 = WL.Bx * WL.Bx + WL.By * WL.By + WL.Bz * WL.Bz;
    const R BR2 = WR.Bx * WR.Bx + WR.By * WR.By + WR.Bz * WR.Bz;
    const R ptL = WL.P + R(0.5) * BL2, ptR = WR.P + R(0.5) * BR2;

    R denom = rhoL * (SL - uL) - rhoR * (SR - uR);
    R SM = R(0);
    if (fabs(denom) > R(1e-14)) SM = qdiv<R>(ptR - ptL + rhoL * uL * (SL - uL) - rhoR * uR * (SR - uR), denom);

    const R rhoL_s = qdiv<R>(rhoL * (SL - uL), SL - SM);
    const R rhoR_s = qdiv<R>(rhoR * (SR - uR), SR - SM);
    const R pt_s = ptL + rhoL * (SL - uL) * (SM - uL);
    const R sL = SL - uL, sR = SR - uR;
    const R Bn2 = Bn * Bn;
    const R dL = rhoL * sL * (SL - SM) - Bn2;
    const R dR = rhoR * sR * (SR - SM) - Bn2;
    if (degenerate) {
        if (guard > R(0)) {
            const R mL = guard * smax(Bn2, R(fabs(rhoL * sL * (SL - SM)))), mR = guard * smax(Bn2, R(fabs(rhoR * sR * (SR - SM))));
            *degenerate = fabs(fabs(dL) - mL) <= tau * mL || fabs(fabs(dR) - mR) <= tau * mR ||
                          fabs(denom) < tau * guard * (fabs(rhoL * (SL - uL)) + fabs(rhoR * (SR - uR)));
        } else {
            const R tL = fabs(rhoL * sL * (SL - SM)) + Bn2, tR = fabs(rhoR * sR * (SR - SM)) + Bn2;
            *degenerate = fabs(dL) < tau * tL || fabs(dR) < tau * tR ||
                          fabs(denom) < tau * (fabs(rhoL * (SL - uL)) + fabs(rhoR * (SR - uR)));
        }
        if (*degenerate) {  // the caller recomputes this face in FP64: the rest would be discarded
#pragma unroll
            for (int k = 0; k < 7; k++) out[k] = R(0);
            return;
        }
    }
    if (guard > R(0) && (fabs(dL) < guard * smax(Bn2, R(fabs(rhoL * sL * (SL - SM)))) ||
                         fabs(dR) < guard * smax(Bn2, R(fabs(rhoR * sR * (SR - SM)))))) {
        hll();
        return;
    }

    R ByL_s = WL.By, BzL_s = WL.Bz, ByR_s = WR.By, BzR_s = WR.Bz;
    R vL_s = WL.v, wL_s = WL.w, vR_s = WR.v, wR_s = WR.w;
    if (fabs(dL) > R(1e-14)) {
        R num = rhoL * sL * (SL - uL) - Bn2;
        ByL_s = qdiv<R>(WL.By * num, dL);
        BzL_s = qdiv<R>(WL.Bz * num, dL);
        R den = rhoL * (SL - uL) * (SL - SM);
        vL_s = WL.v - qdiv<R>(Bn * (ByL_s - WL.By), den);
        wL_s = WL.w - qdiv<R>(Bn * (BzL_s - WL.Bz), den);
    }
    if (fabs(dR) > R(1e-14)) {
        R num = rhoR * sR * (SR - uR) - Bn2;
        ByR_s = qdiv<R>(WR.By * num, dR);
        BzR_s = qdiv<R>(WR.Bz * num, dR);
        R den = rhoR * (SR - uR) * (SR - SM);
        vR_s = WR.v - qdiv<R>(Bn * (ByR_s - WR.By), den);
        wR_s = WR.w - qdiv<R>(Bn * (BzR_s - WR.Bz), den);
    }
    const R sqrt_rhoL = sqrt(smax(rhoL_s, Tiny<R>::v()));
    const R sqrt_rhoR = sqrt(smax(rhoR_s, Tiny<R>::v()));
    const R SLL = SM - qdiv<R>(fabs(Bn), sqrt_rhoL);
    const R SRR = SM + qdiv<R>(fabs(Bn), sqrt_rhoR);

    // Select the branch first (M2:526-529) and evaluate only its star state; the others are unused.
    const bool left = (R(0) < SM);
    const bool outer = left ? (R(0) < SLL) : !(R(0) < SRR);
    // component-wise selection keeps everything in registers
    const R Wr_rho = left ? WL.rho : WR.rho, Wr_u = left ? WL.u : WR.u, Wr_v = left ? WL.v : WR.v;
    const R Wr_w = left ? WL.w : WR.w, Wr_P = left ? WL.P : WR.P;
    const R Wr_By = left ? WL.By : WR.By, Wr_Bz = left ? WL.Bz : WR.Bz;
    const R S = left ? SL : SR;
    const R Er = left ? EL : ER;
    const R rho_s = left ? rhoL_s : rhoR_s;
    R v_s, w_s, By_s, Bz_s;
    if (outer) {
        v_s = left ? vL_s : vR_s;
        w_s = left ? wL_s : wR_s;
        By_s = left ? ByL_s : ByR_s;
        Bz_s = left ? BzL_s : BzR_s;
    } else {  // M2:469-478
        const R signBn = (Bn >= R(0)) ? R(1) : R(-1);
        const R denom2 = sqrt_rhoL + sqrt_rhoR;
        v_s = qdiv<R>(sqrt_rhoL * vL_s + sqrt_rhoR * vR_s + (ByR_s - ByL_s) * signBn, denom2);
        w_s = qdiv<R>(sqrt_rhoL * wL_s + sqrt_rhoR * wR_s + (BzR_s - BzL_s) * signBn, denom2);
        By_s = qdiv<R>(sqrt_rhoL * ByR_s + sqrt_rhoR * ByL_s + sqrt_rhoL * sqrt_rhoR * (vR_s - vL_s) * signBn, denom2);
        Bz_s = qdiv<R>(sqrt_rhoL * BzR_s + sqrt_rhoR * BzL_s + sqrt_rhoL * sqrt_rhoR * (wR_s - wL_s) * signBn, denom2);
    }
    // U_star (M2:486-505) and flux_from_state (M2:507-519); Wr.Bx = Bn on both sides
    const R Br2 = Bn * Bn + Wr_By * Wr_By + Wr_Bz * Wr_Bz;
    const R ptr = Wr_P + R(0.5) * Br2;
    const R vdotB_s = SM * Bn + v_s * By_s + w_s * Bz_s;
    const R vdotB_w = Wr_u * Bn + Wr_v * Wr_By + Wr_w * Wr_Bz;
    const R Es = qdiv<R>((S - Wr_u) * Er - (ptr * Wr_u - Bn * vdotB_w) + (pt_s * SM - Bn * vdotB_s), S - SM);
    const R Ust[7] = {rho_s, rho_s * SM, rho_s * v_s, rho_s * w_s, Es, By_s, Bz_s};
    const R U0[7] = {Wr_rho, Wr_rho * Wr_u, Wr_rho * Wr_v, Wr_rho * Wr_w, Er, Wr_By, Wr_Bz};
#pragma unroll
    for (int k = 0; k < 7; k++) out[k] = (left ? FL[k] : FR[k]) + S * (Ust[k] - U0[k]);
}

// PLM-MC reconstruction of the centre cell c from its neighbours m, p along the sweep axis
// (M2:290-336).  'ax' selects the normal field component (0: Bx, 1: By, 2: Bz), which is not
// reconstructed.  Returns the left (minus) and right (plus) face states of the centre cell.
template <typename R>
__host__ __device__ __forceinline__ void plm_mc(const Prim<R>& Wm, const Prim<R>& Wc, const Prim<R>& Wp, int ax, R floor_rho,
                                       R floor_P, Prim<R>& WL, Prim<R>& WR) {
    R sr = mc_limiter<R>(Wc.rho - Wm.rho, Wp.rho - Wc.rho);
    R su = mc_limiter<R>(Wc.u - Wm.u, Wp.u - Wc.u);
    R sv = mc_limiter<R>(Wc.v - Wm.v, Wp.v - Wc.v);
    R sw = mc_limiter<R>(Wc.w - Wm.w, Wp.w - Wc.w);
    R sP = mc_limiter<R>(Wc.P - Wm.P, Wp.P - Wc.P);
    R sBx = (ax == 0) ? R(0) : mc_limiter<R>(Wc.Bx - Wm.Bx, Wp.Bx - Wc.Bx);
    R sBy = (ax == 1) ? R(0) : mc_limiter<R>(Wc.By - Wm.By, Wp.By - Wc.By);
    R sBz = (ax == 2) ? R(0) : mc_limiter<R>(Wc.Bz - Wm.Bz, Wp.Bz - Wc.Bz);
    WL.rho = Wc.rho - R(0.5) * sr; WR.rho = Wc.rho + R(0.5) * sr;
    WL.u = Wc.u - R(0.5) * su;     WR.u = Wc.u + R(0.5) * su;
    WL.v = Wc.v - R(0.5) * sv;     WR.v = Wc.v + R(0.5) * sv;
    WL.w = Wc.w - R(0.5) * sw;     WR.w = Wc.w + R(0.5) * sw;
    WL.P = Wc.P - R(0.5) * sP;     WR.P = Wc.P + R(0.5) * sP;
    // the normal component keeps the centre value (its slope is zero above; M2:306, 330)
    WL.Bx = Wc.Bx - R(0.5) * sBx;  WR.Bx = Wc.Bx + R(0.5) * sBx;
    WL.By = Wc.By - R(0.5) * sBy;  WR.By = Wc.By + R(0.5) * sBy;
    WL.Bz = Wc.Bz - R(0.5) * sBz;  WR.Bz = Wc.Bz + R(0.5) * sBz;
    WL.rho = smax(WL.rho, floor_rho); WR.rho = smax(WR.rho, floor_rho);
    WL.P = smax(WL.P, floor_P);       WR.P = smax(WR.P, floor_P);
}

}  // namespace cc

namespace cc {
// Mixed-precision HLLD: FP32 evaluation from FP64 face states, with the face recomputed in FP64 when
// it lies within the relative margin tau of the Alfven degeneracy (tau >= 1 recomputes every face).
constexpr float kDegenerateTau = 2e-2f;
// With the degeneracy guard on: the relative band around the guard threshold inside which the FP32
// decision may differ from the FP64 one.  The FP32 error of d is a few tens of ulps of
// max(Bn^2, |rho s (S - S_M)|), i.e. below 1e-5 of it, against a band half-width of 1e-2 x guard of it.
constexpr float kGuardBand = 1e-2f;
// The fallback margin of hlld_x's degenerate test for a guard value.
__host__ __device__ inline float default_degen_tau(double guard) { return guard > 0.0 ? kGuardBand : kDegenerateTau; }

template <typename EOS32, typename EOS64>
__host__ __device__ __forceinline__ void hlld_x_mixed(const EOS32& e32, const EOS64& e64, const Prim<double>& WL,
                                                      const Prim<double>& WR, double Bn, double out[7],
                                                      float tau = -1.f, double guard = 0.0) {
    if (tau < 0.f) tau = default_degen_tau(guard);
    const bool bn_small = fabs(Bn) < 1e-12;
    const Prim<float> L{(float)WL.rho, (float)WL.u, (float)WL.v, (float)WL.w, (float)WL.P, (float)WL.Bx, (float)WL.By, (float)WL.Bz};
    const Prim<float> R{(float)WR.rho, (float)WR.u, (float)WR.v, (float)WR.w, (float)WR.P, (float)WR.Bx, (float)WR.By, (float)WR.Bz};
    float F[7];
    bool deg = false;
    hlld_x<float>(e32, L, R, (float)Bn, bn_small, F, &deg, tau, (float)guard);
    if (deg) {
        hlld_x<double>(e64, WL, WR, Bn, bn_small, out, nullptr, 0.0, guard);
    } else {
#pragma unroll
        for (int k = 0; k < 7; k++) out[k] = (double)F[k];
    }
}
}  // namespace cc
