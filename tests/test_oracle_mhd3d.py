"""Pinning the 3D MHD oracle (oracle/oracle_mhd3d.c), CPU only: golden vectors from the reference's
compiled functions (bit for bit), closed forms of the hybrid EOS (continuity of the cold curve at
rho_nuc, the ideal-gas limit K1 = Kbounce = 0), uniform-state fluxes, div B, mass conservation and the drift of duplicate periodic faces, and
(where oracle/_ref is built) the ideal-EOS mode against the reference's hybrid EOS with K1 = 0."""
import ctypes as C
import os

import numpy as np
import pytest

from oracle import bind as OB
from tests.helpers import oracle, random_mhd3d_state, ref

G = os.path.join(os.path.dirname(__file__), "golden")


def test_hlld_eos_prims_golden_bitwise():
    d = np.load(os.path.join(G, "m3_step.npz"))
    o = oracle()
    p = OB.m3_params()
    F = np.zeros_like(d["F"])
    for k in range(len(F)):
        o.or_m3_hlld_x(C.byref(p), OB.ptr(np.ascontiguousarray(d["WL"][k])), OB.ptr(np.ascontiguousarray(d["WR"][k])),
                       C.c_double(d["Bn"][k]), OB.ptr(F[k]))
    assert np.array_equal(F, d["F"])
    cold = np.zeros_like(d["cold"])
    for k, r in enumerate(d["rho"]):
        o.or_m3_eos_cold(C.byref(p), C.c_double(r), OB.ptr(cold[k]))
    assert np.array_equal(cold, d["cold"])
    W8 = np.zeros_like(d["c8"])
    for k in range(len(W8)):
        o.or_m3_cons_to_prim(C.byref(p), OB.ptr(np.ascontiguousarray(d["c8"][k])), OB.ptr(W8[k]))
    assert np.array_equal(W8, d["W8"])


@pytest.mark.parametrize("per", [1, 0])
def test_reference_ic_run_golden_bitwise(per):
    d = np.load(os.path.join(G, "m3_step.npz"))
    op = OB.m3_params(Nx=10, Ny=9, Nz=8, periodic_x=per, periodic_y=per, periodic_z=per)
    S = OB.M3Arrays(10, 9, 8, {f: d["per%d_in_%s" % (per, f)] for f in OB.M3Arrays.FIELDS})
    dts = np.zeros(3)
    oracle().or_m3_run(C.byref(op), C.byref(S.cstate()), C.c_double(0.1), C.c_double(1 / 9), C.c_double(0.125), 3,
                       C.c_double(0.0), C.c_double(0.15), OB.ptr(dts))
    assert np.array_equal(dts, d["per%d_dts" % per])
    for f in S.FIELDS:
        assert np.array_equal(S.a[f], d["per%d_out_%s" % (per, f)]), f


def test_cold_curve_continuous_at_rho_nuc_and_thermal_split():
    o = oracle()
    p = OB.m3_params()
    lo, hi = np.zeros(3), np.zeros(3)
    o.or_m3_eos_cold(C.byref(p), C.c_double(p.rho_nuc * (1 - 1e-12)), OB.ptr(lo))
    o.or_m3_eos_cold(C.byref(p), C.c_double(p.rho_nuc), OB.ptr(hi))
    assert abs(lo[0] - hi[0]) < 1e-9 * hi[0] and abs(lo[1] - hi[1]) < 1e-9 * hi[1]
    # thermal part: E = kin + mag + rho (e_cold + P_th / ((gamma_th - 1) rho)) (M3:154-163)
    W = np.array([50.0, 0.1, -0.2, 0.3, 400.0, 0.1, 0.2, 0.3])
    c = np.zeros(3)
    o.or_m3_eos_cold(C.byref(p), C.c_double(50.0), OB.ptr(c))
    E = o.or_m3_total_energy(C.byref(p), OB.ptr(W))
    kin = 0.5 * 50 * (0.01 + 0.04 + 0.09)
    mag = 0.5 * (0.01 + 0.04 + 0.09)
    assert abs(E - (kin + mag + 50 * (c[1] + (400 - c[0]) / (0.5 * 50)))) < 1e-10 * E


def test_hybrid_eos_with_zero_cold_curve_is_the_ideal_gas():
    o = oracle()
    ph = OB.m3_params(K1=0.0, Kbounce=0.0, gamma_th=5.0 / 3.0)
    pi = OB.m3_params(ideal_eos=1, gamma_ideal=5.0 / 3.0)
    rng = np.random.default_rng(12)
    for _ in range(200):
        W = np.array([rng.uniform(0.5, 3), *rng.normal(0, 1, 3), rng.uniform(0.1, 2), *rng.normal(0, 1, 3)])
        assert abs(o.or_m3_sound_speed2(C.byref(ph), OB.ptr(W)) - o.or_m3_sound_speed2(C.byref(pi), OB.ptr(W))) < \
            1e-13 * o.or_m3_sound_speed2(C.byref(pi), OB.ptr(W))
        a, b = np.zeros(7), np.zeros(7)
        WR = W.copy(); WR[0] *= 1.2; WR[1] -= 0.3
        Bn = rng.normal(0, 1)
        o.or_m3_hlld_x(C.byref(ph), OB.ptr(W), OB.ptr(WR), C.c_double(Bn), OB.ptr(a))
        o.or_m3_hlld_x(C.byref(pi), OB.ptr(W), OB.ptr(WR), C.c_double(Bn), OB.ptr(b))
        assert np.allclose(a, b, rtol=1e-10, atol=1e-12)


def test_divb_and_conservation_periodic():
    from core_collapse_amd.mhd3d import Params
    pp = Params(Nx=8, Ny=7, Nz=6)
    st = random_mhd3d_state(pp, 4)
    op = OB.m3_params(Nx=8, Ny=7, Nz=6)
    S = OB.M3Arrays(8, 7, 6, {k: v.reshape(-1) for k, v in st.items()})
    m0 = S.a["rho"].sum()
    oracle().or_m3_run(C.byref(op), C.byref(S.cstate()), C.c_double(pp.dx), C.c_double(pp.dy), C.c_double(pp.dz), 3,
                       C.c_double(0.0), C.c_double(1e300), None)
    div = np.zeros(8 * 7 * 6)
    oracle().or_m3_divb(C.byref(op), C.byref(S.cstate()), C.c_double(pp.dx), C.c_double(pp.dy), C.c_double(pp.dz),
                        OB.ptr(div))
    assert np.abs(div).max() < 1e-11
    # Mass telescopes exactly.  Momentum does not quite: the periodic state stores both duplicate
    # faces (i_f = 0 and i_f = Nx), the reference's edge mapping updates them differently
    # (M3:941-1000), so they drift apart and the two boundary fluxes no longer cancel.
    assert abs(S.a["rho"].sum() - m0) < 1e-12 * m0
    bx = S.a["Bx_f"].reshape(6, 7, 9)
    assert 0 < np.abs(bx[:, :, 0] - bx[:, :, 8]).max() < 1e-2


def test_ideal_mode_tracks_reference_hybrid_with_zero_cold_curve():
    r = ref("ref_m3")
    if r is None:
        pytest.skip("oracle/_ref/libref_m3.so is not built here (reference source absent)")
    from core_collapse_amd.mhd3d import Params
    pp = Params(Nx=8, Ny=8, Nz=6, eos=1, gamma=5.0 / 3.0)
    st = random_mhd3d_state(pp, 2)
    A = OB.M3Arrays(8, 8, 6, {k: v.reshape(-1) for k, v in st.items()})
    B = A.copy()
    oracle().or_m3_run(C.byref(OB.m3_params(Nx=8, Ny=8, Nz=6, ideal_eos=1, gamma_ideal=5.0 / 3.0)), C.byref(A.cstate()),
                       C.c_double(pp.dx), C.c_double(pp.dy), C.c_double(pp.dz), 3, C.c_double(0.0), C.c_double(1e300),
                       None)
    r.or_m3_run(C.byref(OB.m3_params(Nx=8, Ny=8, Nz=6, K1=0.0, Kbounce=0.0, gamma_th=5.0 / 3.0)), C.byref(B.cstate()),
                C.c_double(pp.dx), C.c_double(pp.dy), C.c_double(pp.dz), 3, C.c_double(0.0), C.c_double(1e300), None)
    for f in A.FIELDS:
        assert np.allclose(A.a[f], B.a[f], rtol=1e-10, atol=1e-12), f
