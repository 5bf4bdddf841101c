"""Pinning the 2D MHD oracle (oracle/oracle_mhd2d.c), CPU only.

1. Golden vectors from the reference's own compiled step functions (tests/golden/m2_*.npz, made by
   tests/golden/gen_golden.py): the restatement must reproduce them bit for bit.
2. Closed forms the reference's formulas must satisfy, chosen so a dropped term, a wrong sign or a
   transposed operand fails: uniform-state and supersonic HLLD fluxes equal the physical MHD flux;
   |Bn| < 1e-12 gives the textbook HLL flux; hlld_y is the component permutation of hlld_x; SOR
   equals a matrix-splitting iteration built independently; CT keeps div B at roundoff; a periodic
   run without gravity conserves mass and momentum.
3. Where oracle/_ref is built (this container), direct randomized comparison with the reference.
"""
import ctypes as C
import os

import numpy as np
import pytest

from oracle import bind as OB
from tests.helpers import oracle, random_mhd2d_state, ref

G = os.path.join(os.path.dirname(__file__), "golden")


def _hlld(lib, axis, WL, WR, Bn, p=None):
    p = p or OB.m2_params()
    out = np.zeros((len(Bn), 8))
    fn = lib.or_m2_hlld_x if axis == 0 else lib.or_m2_hlld_y
    for k in range(len(Bn)):
        fn(C.byref(p), OB.ptr(np.ascontiguousarray(WL[k])), OB.ptr(np.ascontiguousarray(WR[k])), C.c_double(Bn[k]),
           OB.ptr(out[k]))
    return out


def test_hlld_golden_bitwise_all_branches():
    d = np.load(os.path.join(G, "m2_hlld.npz"))
    assert set(np.unique(d["branch"])) == set(range(7))
    o = oracle()
    assert np.array_equal(_hlld(o, 0, d["WL"], d["WR"], d["Bn"]), d["Fx"])
    assert np.array_equal(_hlld(o, 1, d["WL"], d["WR"], d["Bn"]), d["Fy"])


def test_prims_limiter_speeds_golden_bitwise():
    d = np.load(os.path.join(G, "m2_prims.npz"))
    o = oracle()
    p = OB.m2_params()
    W8 = np.zeros_like(d["c8"])
    for k in range(len(W8)):
        o.or_m2_cons_to_prim(C.byref(p), OB.ptr(np.ascontiguousarray(d["c8"][k])), OB.ptr(W8[k]))
    assert np.array_equal(W8, d["W8"])
    lim = np.array([o.or_m2_mc_limiter(C.c_double(a), C.c_double(b)) for a, b in d["lim_in"]])
    assert np.array_equal(lim, d["lim"])
    cfx = np.array([o.or_m2_fast_speed_x(C.byref(p), OB.ptr(np.ascontiguousarray(w))) for w in d["Wf"]])
    cfy = np.array([o.or_m2_fast_speed_y(C.byref(p), OB.ptr(np.ascontiguousarray(w))) for w in d["Wf"]])
    assert np.array_equal(cfx, d["cfx"]) and np.array_equal(cfy, d["cfy"])


@pytest.mark.parametrize("name,Nx,Ny,per", [("p16", 16, 16, 1), ("c16", 16, 16, 0), ("c24x20", 24, 20, 0),
                                             ("p20x12", 20, 12, 1)])
def test_poisson_golden_bitwise(name, Nx, Ny, per):
    d = np.load(os.path.join(G, "m2_poisson.npz"))
    o = oracle()
    p = OB.m2_params(Nx=Nx, Ny=Ny, periodic_x=per, periodic_y=per, G=1.0)
    gx, gy = np.zeros(Nx * Ny), np.zeros(Nx * Ny)
    o.or_m2_poisson(C.byref(p), OB.ptr(d[name + "_rho"]), C.c_double(1.0 / Nx), C.c_double(1.0 / Ny), OB.ptr(gx),
                    OB.ptr(gy))
    assert np.array_equal(gx, d[name + "_gx"]) and np.array_equal(gy, d[name + "_gy"])


@pytest.mark.parametrize("per", [1, 0])
def test_stage_and_rk2_golden_bitwise(per):
    d = np.load(os.path.join(G, "m2_step.npz"))
    o = oracle()
    op = OB.m2_params(Nx=16, Ny=12, periodic_x=per, periodic_y=per, eta=2e-3, G=1.0, sor_iters=60)
    S0 = OB.M2Arrays(16, 12, {f: d["per%d_in_%s" % (per, f)] for f in OB.M2Arrays.FIELDS})
    S1 = S0.copy()
    gx, gy = d["per%d_gx" % per], d["per%d_gy" % per]
    o.or_m2_advance_stage(C.byref(op), C.byref(S0.cstate()), C.byref(S1.cstate()), OB.ptr(gx), OB.ptr(gy),
                          C.c_double(1 / 16), C.c_double(1 / 12), C.c_double(2e-3))
    for f in S0.FIELDS:
        assert np.array_equal(S1.a[f], d["per%d_stage_%s" % (per, f)]), f
    S2 = S0.copy()
    o.or_m2_rk2_step(C.byref(op), C.byref(S2.cstate()), OB.ptr(gx), OB.ptr(gy), C.c_double(1 / 16), C.c_double(1 / 12),
                     C.c_double(2e-3))
    for f in S0.FIELDS:
        assert np.array_equal(S2.a[f], d["per%d_rk2_%s" % (per, f)]), f


def test_collapse_run_golden_bitwise():
    d = np.load(os.path.join(G, "m2_step.npz"))
    from core_collapse_amd import ic
    pc = ic.collapse2d_params(32)
    op = OB.m2_params(gamma=pc.gamma, cfl=pc.cfl, x0=pc.x0, x1=pc.x1, y0=pc.y0, y1=pc.y1, Nx=32, Ny=32, periodic_x=0,
                      periodic_y=0, eta=pc.eta, G=pc.G)
    S = OB.M2Arrays(32, 32, {f: d["run_in_" + f] for f in OB.M2Arrays.FIELDS})
    dts = np.zeros(12)
    t = oracle().or_m2_run(C.byref(op), C.byref(S.cstate()), C.c_double(pc.dx), C.c_double(pc.dy), 12,
                           C.c_double(0.0), C.c_double(1e300), OB.ptr(dts))
    assert np.array_equal(dts, d["run_dts"]) and t == d["run_t"][0]
    for f in S.FIELDS:
        assert np.array_equal(S.a[f], d["run_out_" + f]), f


# ------------------------------------------------------------------ closed forms
def _mhd_flux_x(W, gamma):
    rho, u, v, w, P, Bx, By, Bz = W
    B2 = Bx * Bx + By * By + Bz * Bz
    pt = P + 0.5 * B2
    E = P / (gamma - 1) + 0.5 * rho * (u * u + v * v + w * w) + 0.5 * B2
    return np.array([rho * u, rho * u * u + pt - Bx * Bx, rho * u * v - Bx * By, rho * u * w - Bx * Bz,
                     (E + pt) * u - Bx * (u * Bx + v * By + w * Bz), u * By - v * Bx, u * Bz - w * Bx])


def _cons(W, gamma):
    rho, u, v, w, P, Bx, By, Bz = W
    return np.array([rho, rho * u, rho * v, rho * w,
                     P / (gamma - 1) + 0.5 * rho * (u * u + v * v + w * w) + 0.5 * (Bx * Bx + By * By + Bz * Bz),
                     By, Bz])


def test_hlld_uniform_state_gives_physical_flux_in_every_branch():
    rng = np.random.default_rng(3)
    o = oracle()
    for k in range(400):
        W = np.array([rng.uniform(0.2, 3), *rng.normal(0, 1.5, 3), rng.uniform(0.05, 2), *rng.normal(0, 1, 3)])
        Bn = 0.0 if k % 4 == 0 else rng.normal(0, 1)
        W[5] = Bn
        got = _hlld(o, 0, W[None], W[None], np.array([Bn]))[0]
        want = _mhd_flux_x(W, 5.0 / 3.0)
        assert np.allclose(got[:7], want, rtol=1e-12, atol=1e-12), (k, got, want)
        assert got[7] == -got[5]


def test_hlld_supersonic_upwinds_and_hll_fallback():
    rng = np.random.default_rng(4)
    o = oracle()
    g = 5.0 / 3.0
    for k in range(200):
        WL = np.array([rng.uniform(0.5, 2), 25.0, *rng.normal(0, 1, 2), rng.uniform(0.1, 1), *rng.normal(0, 0.3, 3)])
        WR = WL.copy(); WR[0] *= 1.3; WR[4] *= 0.7
        Bn = rng.normal(0, 0.3)
        WL[5] = WR[5] = Bn
        got = _hlld(o, 0, WL[None], WR[None], np.array([Bn]))[0]
        assert np.allclose(got[:7], _mhd_flux_x(WL, g), rtol=1e-13, atol=1e-13)
    for k in range(200):  # |Bn| < 1e-12: the HLL average of M2:449-467
        WL = np.array([rng.uniform(0.5, 2), *rng.normal(0, 0.3, 3), rng.uniform(0.1, 1), 0.0, *rng.normal(0, 1, 2)])
        WR = np.array([rng.uniform(0.5, 2), *rng.normal(0, 0.3, 3), rng.uniform(0.1, 1), 0.0, *rng.normal(0, 1, 2)])
        got = _hlld(o, 0, WL[None], WR[None], np.array([0.0]))[0]

        def cf(W):
            a2 = g * W[4] / W[0]
            b2 = (W[5:8] ** 2).sum()
            vA2, vAx2 = b2 / W[0], W[5] ** 2 / W[0]
            return np.sqrt(0.5 * ((a2 + vA2) + np.sqrt(max(0.0, (a2 + vA2) ** 2 - 4 * a2 * vAx2))))
        SL = min(WL[1] - cf(WL), WR[1] - cf(WR))
        SR = max(WL[1] + cf(WL), WR[1] + cf(WR))
        FL, FR, UL, UR = _mhd_flux_x(WL, g), _mhd_flux_x(WR, g), _cons(WL, g), _cons(WR, g)
        if SL >= 0:
            want = FL
        elif SR <= 0:
            want = FR
        else:
            want = (SR * FL - SL * FR + SL * SR * (UR - UL)) / (SR - SL)
        assert np.allclose(got[:7], want, rtol=1e-11, atol=1e-12)


def test_hlld_y_is_the_permuted_hlld_x():
    rng = np.random.default_rng(5)
    o = oracle()
    for k in range(300):
        WL = np.array([rng.uniform(0.2, 3), *rng.normal(0, 1, 3), rng.uniform(0.05, 2), *rng.normal(0, 1, 3)])
        WR = np.array([rng.uniform(0.2, 3), *rng.normal(0, 1, 3), rng.uniform(0.05, 2), *rng.normal(0, 1, 3)])
        Bn = rng.normal(0, 1)
        perm = [0, 2, 1, 3, 4, 6, 5, 7]  # (u,v,w)->(v,u,w), (Bx,By,Bz)->(By,Bx,Bz)
        gy = _hlld(o, 1, WL[None], WR[None], np.array([Bn]))[0]
        fx = _hlld(o, 0, WL[perm][None], WR[perm][None], np.array([Bn]))[0]
        assert np.array_equal(gy[[0, 2, 1, 3, 4, 5, 6]], fx[[0, 1, 2, 3, 4, 5, 6]])


def test_sor_equals_independent_matrix_splitting():
    """600-sweep lexicographic SOR with per-sweep mean removal (M2:207-223) on a clamped 7x5 grid,
    against phi_{k+1} = (D/w + L)^-1 (((1/w - 1) D - U) phi_k + b) built from the 5-point matrix."""
    Nx, Ny, it, w = 7, 5, 600, 1.8
    rng = np.random.default_rng(6)
    rho = 1 + rng.uniform(0, 1, Nx * Ny)
    p = OB.m2_params(Nx=Nx, Ny=Ny, periodic_x=0, periodic_y=0, G=1.0, sor_iters=it, sor_omega=w)
    phi = np.zeros(Nx * Ny)
    dx, dy = 1.0 / Nx, 1.0 / Ny
    oracle().or_m2_sor_phi(C.byref(p), OB.ptr(rho), C.c_double(dx), C.c_double(dy), OB.ptr(phi))
    # matrix of the update phi_new(q) = (sum_nb c_nb phi_nb - rhs)/denom (neighbours clamped)
    n = Nx * Ny
    idx2, idy2 = 1 / dx ** 2, 1 / dy ** 2
    denom = 2 * (idx2 + idy2)
    A = np.zeros((n, n))  # A phi = sum of weighted neighbours
    for j in range(Ny):
        for i in range(Nx):
            q = j * Nx + i
            for (ii, jj, c) in ((min(i + 1, Nx - 1), j, idx2), (max(i - 1, 0), j, idx2), (i, min(j + 1, Ny - 1), idy2),
                                (i, max(j - 1, 0), idy2)):
                A[q, jj * Nx + ii] += c
    rhs = 4 * np.pi * 1.0 * (rho - rho.mean())
    Lm = np.tril(A, -1)          # already-updated neighbours (earlier in lexicographic order)
    Um = A - Lm                  # not-yet-updated neighbours and the cell itself (clamped self-reference)
    M1 = np.eye(n) * denom / w - Lm
    x = np.zeros(n)
    for _ in range(it):
        x = np.linalg.solve(M1, (denom / w - denom) * x + Um @ x - rhs)
        x -= x.mean()
    assert np.allclose(phi, x, rtol=0, atol=1e-10 * np.abs(x).max())


def test_ct_keeps_divb_and_periodic_run_conserves():
    from core_collapse_amd.mhd2d import Params
    pp = Params(Nx=24, Ny=20, eta=1e-3, gravity_on=False)
    st = random_mhd2d_state(pp, 9)
    op = OB.m2_params(Nx=24, Ny=20, eta=1e-3, gravity_on=0)
    S = OB.M2Arrays(24, 20, {k: v.reshape(-1) for k, v in st.items()})
    m0 = [S.a[f].sum() for f in ("rho", "mx", "my", "mz")]
    t = oracle().or_m2_run(C.byref(op), C.byref(S.cstate()), C.c_double(pp.dx), C.c_double(pp.dy), 6, C.c_double(0.0),
                           C.c_double(1e300), None)
    assert t > 0
    div = np.zeros(24 * 20)
    oracle().or_m2_divb(C.byref(op), OB.ptr(S.a["Bx_f"]), OB.ptr(S.a["By_f"]), C.c_double(pp.dx), C.c_double(pp.dy),
                        OB.ptr(div))
    assert np.abs(div).max() < 1e-11
    m1 = [S.a[f].sum() for f in ("rho", "mx", "my", "mz")]
    for a, b in zip(m0, m1):
        assert abs(a - b) < 1e-10 * max(1.0, abs(a))


def test_oracle_equals_reference_randomized():
    r = ref("ref_m2")
    if r is None:
        pytest.skip("oracle/_ref/libref_m2.so is not built here (reference source absent)")
    o = oracle()
    rng = np.random.default_rng(8)
    WL = np.array([[rng.uniform(0.1, 4), *rng.normal(0, 1, 3), rng.uniform(0.01, 2), *rng.normal(0, 1, 3)]
                   for _ in range(2000)])
    WR = np.array([[rng.uniform(0.1, 4), *rng.normal(0, 1, 3), rng.uniform(0.01, 2), *rng.normal(0, 1, 3)]
                   for _ in range(2000)])
    Bn = rng.normal(0, 1, 2000); Bn[::5] = 0
    for axis in (0, 1):
        assert np.array_equal(_hlld(o, axis, WL, WR, Bn), _hlld(r, axis, WL, WR, Bn))
    for per in (1, 0):
        from core_collapse_amd.mhd2d import Params
        pp = Params(Nx=20, Ny=14, periodic_x=bool(per), periodic_y=bool(per), eta=1e-3, G=3.0, sor_iters=80)
        st = random_mhd2d_state(pp, 20 + per)
        op = OB.m2_params(Nx=20, Ny=14, periodic_x=per, periodic_y=per, eta=1e-3, G=3.0, sor_iters=80)
        A = OB.M2Arrays(20, 14, {k: v.reshape(-1) for k, v in st.items()})
        B = A.copy()
        ta = o.or_m2_run(C.byref(op), C.byref(A.cstate()), C.c_double(pp.dx), C.c_double(pp.dy), 4, C.c_double(0.0),
                         C.c_double(1e300), None)
        tb = r.or_m2_run(C.byref(op), C.byref(B.cstate()), C.c_double(pp.dx), C.c_double(pp.dy), 4, C.c_double(0.0),
                         C.c_double(1e300), None)
        assert ta == tb
        for f in A.FIELDS:
            assert np.array_equal(A.a[f], B.a[f]), f
