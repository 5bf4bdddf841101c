"""The oracle's two options with no reference counterpart, pinned on the CPU.

hlld_guard: an HLLD face whose star-state denominator d = rho (S - u)(S - S_M) - Bn^2 is below
hlld_guard * max(Bn^2, |rho (S - u)(S - S_M)|) on either side takes the reference's own HLL flux
(its |Bn| < 1e-12 branch, M2:449 / M3:491).  Pinned by: the reference's arithmetic at guard 0 (golden
vectors elsewhere), a guarded face equal either to the unguarded one or to an HLL flux restated here
from the textbook formulas, a face taken from a real run where the reference's 1e-14 guard lets
d = -2.9e-9 through, and the 3D turbulence problem, which the reference arithmetic blows up.

ct_comp: each face carries the FP32 rounding error of its stored value through CT (M2:807-830,
M3:941-1015) and the Heun average (M2:850-866, M3:1042-1065).  Pinned by: the divergence of the
stored field staying at one rounding while the plain update's grows, the divergence of hi + lo
(extended precision) staying at the increments' rounding, and agreement with the plain update to
rounding over a few steps.
"""
import ctypes as C

import numpy as np

from core_collapse_amd import ic
from core_collapse_amd.mhd2d import Params
from oracle import bind as OB
from tests.helpers import m2_arrays, m2_oparams, m3_arrays, m3_oparams, oracle, random_mhd2d_state

G53 = 5.0 / 3.0


def _hll_textbook(WL, WR, Bn, gamma=G53):
    """HLL flux of ideal MHD (x-normal), from the physical fluxes and Davis wave speeds
    S_L = min(u_L - c_fL, u_R - c_fR), S_R = max(u_L + c_fL, u_R + c_fR); 7 components
    (rho, mx, my, mz, E, By, Bz).  W = (rho, u, v, w, P, Bx, By, Bz) with Bx replaced by Bn."""
    def phys(W):
        rho, u, v, w, P, _, By, Bz = W
        Bx = Bn
        B2 = Bx * Bx + By * By + Bz * Bz
        pt = P + 0.5 * B2
        E = P / (gamma - 1) + 0.5 * rho * (u * u + v * v + w * w) + 0.5 * B2
        vB = u * Bx + v * By + w * Bz
        U = np.array([rho, rho * u, rho * v, rho * w, E, By, Bz])
        F = np.array([rho * u, rho * u * u + pt - Bx * Bx, rho * u * v - Bx * By, rho * u * w - Bx * Bz,
                      (E + pt) * u - Bx * vB, u * By - v * Bx, u * Bz - w * Bx])
        a2 = gamma * P / rho
        vA2 = B2 / rho
        cf = np.sqrt(0.5 * ((a2 + vA2) + np.sqrt(max((a2 + vA2) ** 2 - 4 * a2 * Bx * Bx / rho, 0.0))))
        return U, F, u, cf
    UL, FL, uL, cL = phys(WL)
    UR, FR, uR, cR = phys(WR)
    SL, SR = min(uL - cL, uR - cR), max(uL + cL, uR + cR)
    return (SR * FL - SL * FR + SL * SR * (UR - UL)) / (SR - SL)


def _m2_hlld(p, WL, WR, Bn):
    out = np.zeros(8)
    oracle().or_m2_hlld_x(C.byref(p), OB.ptr(np.ascontiguousarray(WL, dtype=np.float64)),
                          OB.ptr(np.ascontiguousarray(WR, dtype=np.float64)), C.c_double(Bn), OB.ptr(out))
    return out


# An x-face of the 2D default configuration (192^2, t = 0.2310, one-ulp change of the initial density):
# d_R = -2.9e-9 passes the reference's |d| > 1e-14 test, its star states reach By* = 110, v* = -791 and the
# energy flux -4.1e3, which drives a neighbouring cell's density negative and the time step to 5e-9.
STALL_WL = np.array([0.19203785403939233, 0.022900874132435169, -1.9588677404300183, 0.0, 0.0088144554388325056,
                     0.13995694751134413, -0.0030659160355990785, 0.0])
STALL_WR = np.array([0.19164299882760077, 0.02194225825901483, -1.9593106895924011, 0.0, 0.0086446715007738682,
                     0.13995694751134413, -0.0027893855729160455, 0.0])
STALL_BN = 0.13995694751134413


def test_reference_hlld_amplifies_the_stall_face_and_the_guard_takes_hll():
    ref = _m2_hlld(OB.m2_params(), STALL_WL, STALL_WR, STALL_BN)
    assert abs(ref[4]) > 1e3, ref          # the reference's energy flux at this face
    assert oracle().or_m2_last_branch() == 4  # U**_L
    got = _m2_hlld(OB.m2_params(hlld_guard=1e-2), STALL_WL, STALL_WR, STALL_BN)
    assert oracle().or_m2_last_branch() == 2  # the HLL branch
    want = _hll_textbook(STALL_WL, STALL_WR, STALL_BN)
    assert np.allclose(got[:7], want, rtol=1e-13, atol=1e-15), (got[:7], want)
    assert got[7] == -got[5]                  # Ez = -F(By)


def test_guarded_hlld_is_the_reference_face_or_hll():
    """Random and near-degenerate faces: with the guard, every face is either bit-identical to the
    reference's or the textbook HLL flux; near-degenerate faces (field along the normal, c_s < v_A)
    are the ones that change."""
    rng = np.random.default_rng(4)
    n = 600
    changed = 0
    for k in range(n):
        WL = np.array([rng.uniform(0.2, 3), *rng.normal(0, 0.5, 3), rng.uniform(0.05, 2), 0, *rng.normal(0, 1, 2)])
        WR = np.array([rng.uniform(0.2, 3), *rng.normal(0, 0.5, 3), rng.uniform(0.05, 2), 0, *rng.normal(0, 1, 2)])
        Bn = rng.normal(0, 1.0)
        if k % 3 == 0:  # near the degeneracy: B along the normal, Alfven speed above the sound speed
            WL[6:8] = 1e-3 * rng.normal(0, 1, 2)
            WR[:] = WL
            WR[1] += 1e-3 * rng.normal()
            WL[4] = WR[4] = 0.3 * WL[0]
            Bn = 1.5 * np.sqrt(WL[0])
        ref = _m2_hlld(OB.m2_params(), WL, WR, Bn)
        got = _m2_hlld(OB.m2_params(hlld_guard=1e-2), WL, WR, Bn)
        if np.array_equal(ref, got):
            continue
        changed += 1
        want = _hll_textbook(WL, WR, Bn)
        assert np.allclose(got[:7], want, rtol=1e-12, atol=1e-14 * np.abs(want).max()), (k, got[:7], want)
    assert 100 < changed < n // 2, changed


# A z-face of the 256^3 turbulence benchmark run with guard 1e-2 (mixed precision, step 294; FP64 gives the
# same flux on these states), in the x-normal frame of hlld_x: a strong normal field (Bn = 1.61) and a
# transverse velocity jump of ~4.  d_L / max(Bn^2, |rho s (S - S_M)|) = 0.0103 passes a 1e-2 guard, the star
# states amplify By by ~42 and the energy flux is 1.35e3 from O(1) states: the cell behind the face went to
# P ~ 1e10 in the next step.  Faces of this kind left ~50 single-cell pressure spikes in the FP64 run as well.
TURB_WL = np.array([0.33268, -0.17699, -1.73912, -3.19886, 0.27983, 1.6144, 0.50274, 0.75518])
TURB_WR = np.array([0.53041, -0.17699, 0.89125, 0.7546, 3.56277, 1.6144, -0.12061, -0.05063])
TURB_BN = 1.6144


def _m3_hlld_ideal(guard, WL, WR, Bn):
    op = OB.m3_params(ideal_eos=1, gamma_ideal=G53, hlld_guard=guard)
    out = np.zeros(7)
    oracle().or_m3_hlld_x(C.byref(op), OB.ptr(np.ascontiguousarray(WL, dtype=np.float64)),
                          OB.ptr(np.ascontiguousarray(WR, dtype=np.float64)), C.c_double(Bn), OB.ptr(out))
    return out


def test_turbulence_face_amplified_at_guard_1e2_takes_hll_at_the_benchmark_guard():
    ref = _m3_hlld_ideal(0.0, TURB_WL, TURB_WR, TURB_BN)
    assert abs(ref[4]) > 1e3 and abs(ref[5]) > 40, ref                       # the reference's HLLD
    assert np.array_equal(_m3_hlld_ideal(1e-2, TURB_WL, TURB_WR, TURB_BN), ref)  # a 1e-2 guard lets it through
    guard = ic.turbulence3d_params(16).hlld_guard
    assert guard >= 3e-2, guard
    got = _m3_hlld_ideal(guard, TURB_WL, TURB_WR, TURB_BN)
    want = _hll_textbook(TURB_WL, TURB_WR, TURB_BN)
    assert np.allclose(got, want, rtol=1e-12, atol=1e-14 * np.abs(want).max()), (got, want)
    assert np.abs(got).max() < 5, got


def test_turbulence_problem_blows_up_in_reference_arithmetic_and_not_with_the_guard():
    """The 3D turbulence IC (beta 1 field along z, rms Mach 1) at 16^3 for 12 steps: with the reference's
    HLLD the total energy density reaches ~1e11 and dt falls to ~1e-8; with the guard it stays O(1)."""
    o = oracle()
    res = {}
    for guard in (0.0, 1e-2):
        p = ic.turbulence3d_params(16, hlld_guard=guard)
        S = m3_arrays(p, ic.turbulence3d_ic(p))
        dts = np.zeros(12)
        o.or_m3_run(C.byref(m3_oparams(p)), C.byref(S.cstate()), C.c_double(p.dx), C.c_double(p.dy), C.c_double(p.dz),
                    12, C.c_double(0.0), C.c_double(1e30), OB.ptr(dts))
        res[guard] = (S.a["E"].max(), dts.min())
    assert res[0.0][0] > 1e6 and res[0.0][1] < 1e-5, res
    assert res[1e-2][0] < 10 and res[1e-2][1] > 1e-3, res


def _divb2(p, A, exact=False):
    Nx, Ny = p.Nx, p.Ny
    bx, by = A.a["Bx_f"], A.a["By_f"]
    if exact:  # hi + lo in extended precision
        bx = bx.astype(np.longdouble) + A.lo["Bx_lo"].astype(np.longdouble)
        by = by.astype(np.longdouble) + A.lo["By_lo"].astype(np.longdouble)
    bx, by = bx.reshape(Ny, Nx + 1), by.reshape(Ny + 1, Nx)
    return float(np.abs((bx[:, 1:] - bx[:, :-1]) / p.dx + (by[1:] - by[:-1]) / p.dy).max())


def test_compensated_face_update_keeps_divb_at_one_rounding():
    """Periodic 64^2 random state, resistive, no gravity, 400 steps: the plain update's max|div B| grows
    with the step count (7e-14 -> 1.4e-13); the compensated one stays at one rounding of the face values,
    <= 2 ulp(max|B|) (1/dx + 1/dy), and hi + lo stays divergence-free to the increments' rounding."""
    o = oracle()
    out = {}
    for comp in (False, True):
        p = Params(Nx=64, Ny=64, gravity_on=False, eta=2e-3, ct_compensated=comp)
        A = m2_arrays(p, random_mhd2d_state(p, 3))
        op = m2_oparams(p)
        t, hist = 0.0, []
        for _ in range(4):
            t = o.or_m2_run(C.byref(op), C.byref(A.cstate()), C.c_double(p.dx), C.c_double(p.dy), 100, C.c_double(t),
                            C.c_double(1e30), None)
            hist.append(_divb2(p, A))
        out[comp] = (hist, A, p)
    plain, comp = out[False][0], out[True][0]
    A, p = out[True][1], out[True][2]
    bmax = max(np.abs(A.a["Bx_f"]).max(), np.abs(A.a["By_f"]).max())
    bound = 2 * np.spacing(bmax) * (1 / p.dx + 1 / p.dy)
    assert max(comp) <= bound, (comp, bound)
    assert plain[-1] > 5 * max(comp) and plain[-1] > plain[0], (plain, comp)
    assert _divb2(p, A, exact=True) < 0.3 * bound
    assert np.abs(A.lo["Bx_lo"]).max() <= 0.5 * np.spacing(np.abs(A.a["Bx_f"]).max())


def test_compensated_run_agrees_with_plain_to_rounding():
    """The compensated update changes only roundings: 5 steps of the collapse problem at 32^2 (gravity on)
    agree with the plain update to 1e-12 relative per field."""
    o = oracle()
    res = {}
    for comp in (False, True):
        p = ic.collapse2d_params(32, ct_compensated=comp, sor_iters=100)
        A = m2_arrays(p, ic.collapse2d_ic(p))
        o.or_m2_run(C.byref(m2_oparams(p)), C.byref(A.cstate()), C.c_double(p.dx), C.c_double(p.dy), 5,
                    C.c_double(0.0), C.c_double(1e30), None)
        res[comp] = A
    for f in OB.M2Arrays.FIELDS:
        a, b = res[True].a[f], res[False].a[f]
        assert np.abs(a - b).max() <= 1e-12 * max(np.abs(b).max(), 1e-300), f


def test_compensated_3d_face_update_keeps_divb_at_one_rounding():
    """3D turbulence (guarded) at 16^3, 60 steps: max|div B| of the compensated field stays at one
    rounding of |B| ~ 1.1 (<= 2 ulp (1/dx + 1/dy + 1/dz)); the plain update's is larger."""
    o = oracle()
    out = {}
    for comp in (False, True):
        p = ic.turbulence3d_params(16, ct_compensated=comp)
        S = m3_arrays(p, ic.turbulence3d_ic(p))
        o.or_m3_run(C.byref(m3_oparams(p)), C.byref(S.cstate()), C.c_double(p.dx), C.c_double(p.dy), C.c_double(p.dz),
                    60, C.c_double(0.0), C.c_double(1e30), None)
        d = np.zeros(16 ** 3)
        o.or_m3_divb(C.byref(m3_oparams(p)), C.byref(S.cstate()), C.c_double(p.dx), C.c_double(p.dy),
                     C.c_double(p.dz), OB.ptr(d))
        bmax = max(np.abs(S.a[f]).max() for f in ("Bx_f", "By_f", "Bz_f"))
        out[comp] = (np.abs(d).max(), 2 * np.spacing(bmax) * 3 * 16)
    assert out[True][0] <= out[True][1], out
    assert out[False][0] > out[True][0], out
