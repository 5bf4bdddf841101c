"""GPU parity of the 2D MHD step (libccamd.so through its C ABI) against the CPU oracle.

The oracle restates mhd2d_resistive_hlld_ct_gravity.cpp bit for bit (pinned in
test_oracle_mhd2d.py).  GPU FP64 differs from it only by FMA contraction and the order of a few
reductions, so FP64 comparisons use relative tolerances of 1e-9..1e-12 (stated per test); the
mixed-precision path is held to the 1e-4 relative L1 density bar of the benchmark.
"""
import ctypes as C

import numpy as np
import pytest

from core_collapse_amd import CC_FP64, CC_MIXED
from core_collapse_amd import ic
from core_collapse_amd import mhd2d as M
from oracle import bind as OB
from tests.helpers import m2_arrays, m2_dict, m2_oparams, oracle, random_mhd2d_state, rel_err, rel_l1

pytestmark = pytest.mark.gpu


def _random_prims(rng, n):
    W = np.empty((n, 8))
    W[:, 0] = rng.uniform(0.1, 5.0, n)
    W[:, 1:4] = rng.normal(0, 1.0, (n, 3))
    W[:, 4] = rng.uniform(0.01, 3.0, n)
    W[:, 5:8] = rng.normal(0, 1.0, (n, 3))
    return W


def _oracle_hlld(axis, WL, WR, Bn, gamma=5.0 / 3.0):
    o = oracle()
    p = OB.m2_params(gamma=gamma)
    out = np.zeros((len(Bn), 8))
    fn = o.or_m2_hlld_x if axis == 0 else o.or_m2_hlld_y
    for k in range(len(Bn)):
        r = np.zeros(8)
        fn(C.byref(p), OB.ptr(np.ascontiguousarray(WL[k])), OB.ptr(np.ascontiguousarray(WR[k])),
           C.c_double(Bn[k]), OB.ptr(r))
        out[k] = r
    return out


@pytest.mark.parametrize("axis", [0, 1])
def test_hlld_batch_fp64_matches_oracle(axis):
    rng = np.random.default_rng(11 + axis)
    n = 4000
    WL, WR = _random_prims(rng, n), _random_prims(rng, n)
    Bn = rng.normal(0, 1.0, n)
    Bn[::7] = 0.0          # HLL fallback branch (|Bn| < 1e-12)
    WR[::11] = WL[::11]    # uniform states: exact physical flux in every branch
    got = M.hlld_batch(axis, WL, WR, Bn)
    want = _oracle_hlld(axis, WL, WR, Bn)
    scale = np.abs(want).max(axis=1, keepdims=True) + 1e-12
    err = np.abs(got - want) / scale
    # FP64 with FMA contraction vs the unfused oracle: allow 1e-10 relative per face
    assert np.quantile(err, 0.999) < 1e-10, np.quantile(err, 0.999)
    assert err.max() < 1e-6, err.max()


@pytest.mark.parametrize("axis", [0, 1])
def test_hlld_batch_mixed_close_to_oracle(axis):
    rng = np.random.default_rng(21 + axis)
    n = 4000
    WL, WR = _random_prims(rng, n), _random_prims(rng, n)
    Bn = rng.normal(0, 1.0, n)
    # faces placed at the HLL switch (|Bn| < 1e-12) and at the Alfven degeneracy of the left star state
    # (B along the normal, c_s < v_A: rho (S_L - u)(S_L - S_M) - Bn^2 ~ 0), where the mixed path recomputes in FP64
    Bn[::13] = 0.0
    deg = np.arange(5, n, 17)
    WL[deg, 6:8] = 1e-3 * rng.normal(0, 1.0, (len(deg), 2))
    WR[deg] = WL[deg]
    WR[deg, 1] += 1e-3
    WL[deg, 4] = 0.3 * WL[deg, 0]
    WR[deg, 4] = WL[deg, 4]
    Bn[deg] = 1.5 * np.sqrt(WL[deg, 0])
    got = M.hlld_batch(axis, WL, WR, Bn, precision=CC_MIXED)
    want = _oracle_hlld(axis, WL, WR, Bn)
    scale = np.abs(want).max(axis=1, keepdims=True) + 1e-12
    err = np.abs(got - want) / scale
    # FP32 Riemann solver on FP32-rounded states: ~1e-7 typical, a few hundred ulps where the HLLD's
    # subtractions cancel; the bound on the worst face is what catches a wrong branch on a minority of faces
    print("mixed HLLD rel err: median %.2e, 99.9%% %.2e, max %.2e" % (np.median(err), np.quantile(err, 0.999), err.max()))
    assert np.median(err) < 1e-6, np.median(err)
    assert np.quantile(err, 0.999) < 1e-4, np.quantile(err, 0.999)
    assert err.max() < 1e-2, err.max()
    assert err[::13].max() < 1e-4, err[::13].max()  # HLL faces: no cancellation-prone star states


def _oracle_poisson(p, rho):
    o = oracle()
    op = m2_oparams(p)
    gx, gy = np.zeros(rho.size), np.zeros(rho.size)
    o.or_m2_poisson(C.byref(op), OB.ptr(np.ascontiguousarray(rho.reshape(-1))), C.c_double(p.dx), C.c_double(p.dy),
                    OB.ptr(gx), OB.ptr(gy))
    return gx.reshape(rho.shape), gy.reshape(rho.shape)


@pytest.mark.parametrize("Nx,Ny,periodic,iters", [(32, 32, False, 600), (72, 40, False, 600), (130, 67, False, 37),
                                                   (24, 24, True, 200), (20, 28, True, 600), (2, 2, False, 7),
                                                   (3, 2, False, 40), (2, 97, False, 33), (97, 2, False, 33),
                                                   (5, 17, False, 600), (65, 3, False, 1), (129, 130, False, 17)])
def test_poisson_sor_fp64_matches_oracle(Nx, Ny, periodic, iters):
    rng = np.random.default_rng(Nx * 1000 + Ny)
    p = M.Params(Nx=Nx, Ny=Ny, periodic_x=periodic, periodic_y=periodic, sor_iters=iters, G=1.0)
    rho = 1.0 + rng.uniform(0, 1, (Ny, Nx)) + 5.0 * np.exp(-((np.arange(Nx)[None, :] - Nx / 2) ** 2 +
                                                           (np.arange(Ny)[:, None] - Ny / 3) ** 2) / 20.0)
    gx, gy = M.poisson_sor(rho, p.dx, p.dy, p.G, iters, p.sor_omega, periodic, periodic)
    ox, oy = _oracle_poisson(p, rho)
    # same lexicographic iterate; differences are rounding only (FMA, deferred mean removal)
    assert rel_err(gx, ox) < 1e-9 and rel_err(gy, oy) < 1e-9, (rel_err(gx, ox), rel_err(gy, oy))


@pytest.mark.parametrize("Nx,Ny,px,py,iters", [(18, 18, True, True, 150), (64, 64, True, True, 100),
                                                (65, 65, True, False, 80), (100, 128, False, True, 60),
                                                (129, 130, True, True, 40), (192, 192, True, True, 30),
                                                (40, 256, True, True, 20), (300, 250, True, True, 10),
                                                (16, 16, True, True, 100), (200, 300, True, True, 5)])
def test_poisson_sor_periodic_schedules_match_oracle(Nx, Ny, px, py, iters):
    """Periodic and mixed boundaries: the one-wave schedule (Ny <= 256 rows over 1-4 row slots per lane,
    Nx, Ny >= 18) and the one-workgroup schedule (other sizes) give the lexicographic iterate."""
    rng = np.random.default_rng(Nx * 7 + Ny)
    p = M.Params(Nx=Nx, Ny=Ny, periodic_x=px, periodic_y=py, sor_iters=iters, G=1.0)
    rho = 1.0 + rng.uniform(0, 1, (Ny, Nx)) + 5.0 * np.exp(-((np.arange(Nx)[None, :] - Nx / 3) ** 2 +
                                                           (np.arange(Ny)[:, None] - Ny / 2) ** 2) / 30.0)
    gx, gy = M.poisson_sor(rho, p.dx, p.dy, p.G, iters, p.sor_omega, px, py)
    ox, oy = _oracle_poisson(p, rho)
    assert rel_err(gx, ox) < 1e-9 and rel_err(gy, oy) < 1e-9, (rel_err(gx, ox), rel_err(gy, oy))


@pytest.mark.parametrize("Nx,Ny,iters", [(96, 80, 600), (2, 2, 7), (5, 17, 600), (97, 2, 33), (130, 67, 37)])
def test_poisson_sor_mixed_close_to_oracle(Nx, Ny, iters):
    rng = np.random.default_rng(5)
    p = M.Params(Nx=Nx, Ny=Ny, periodic_x=False, periodic_y=False, sor_iters=iters, G=1.0)
    rho = 1.0 + rng.uniform(0, 1, (Ny, Nx))
    gx, gy = M.poisson_sor(rho, p.dx, p.dy, p.G, iters, 1.8, False, False, precision=CC_MIXED)
    ox, oy = _oracle_poisson(p, rho)
    assert rel_err(gx, ox) < 1e-5 and rel_err(gy, oy) < 1e-5, (rel_err(gx, ox), rel_err(gy, oy))


@pytest.mark.parametrize("Nx,Ny,iters", [(96, 80, 600), (130, 67, 37), (5, 17, 600)])
@pytest.mark.parametrize("tile", ["16,64,s", "16,64"])
def test_mixed_solve_on_the_fp64_kernels_is_the_fp64_solve(monkeypatch, Nx, Ny, iters, tile):
    """A mixed-precision grid may run the FP64 SOR kernels (CC_SOR_TILE flag d; the autotuner's choice where
    they are faster): its solve is then the FP64 solve, bit for bit."""
    rng = np.random.default_rng(8)
    p = M.Params(Nx=Nx, Ny=Ny, periodic_x=False, periodic_y=False, sor_iters=iters, G=1.0)
    rho = 1.0 + rng.uniform(0, 1, (Ny, Nx))
    monkeypatch.setenv("CC_SOR_WS", "0")
    monkeypatch.setenv("CC_SOR_TILE", tile + ("d" if "," in tile[tile.index(",") + 1:] else ",d"))
    gm = M.poisson_sor(rho, p.dx, p.dy, p.G, iters, 1.8, False, False, precision=CC_MIXED)
    monkeypatch.setenv("CC_SOR_TILE", tile)
    gd = M.poisson_sor(rho, p.dx, p.dy, p.G, iters, 1.8, False, False, precision=CC_FP64)
    gr = M.poisson_sor(rho, p.dx, p.dy, p.G, iters, 1.8, False, False, precision=CC_MIXED)  # mixed kernels
    for a, b in zip(gm, gd):
        assert np.array_equal(a, b)
    assert not all(np.array_equal(a, b) for a, b in zip(gr, gd))  # the flag does switch the kernels


@pytest.mark.parametrize("periodic", [True, False])
def test_advance_stage_matches_oracle(periodic):
    p = M.Params(Nx=40, Ny=24, periodic_x=periodic, periodic_y=periodic, eta=2e-3, G=1.0, sor_iters=50)
    st = random_mhd2d_state(p, 3)
    rng = np.random.default_rng(9)
    gx, gy = rng.normal(0, 0.5, (p.Ny, p.Nx)), rng.normal(0, 0.5, (p.Ny, p.Nx))
    dt = 2e-3
    with M.Mhd2D(p) as h:
        h.upload(st)
        h.set_gravity(gx, gy)
        h.advance_stage(dt)
        got = h.download()
    o = oracle()
    S0 = m2_arrays(p, st)
    S1 = S0.copy()
    o.or_m2_advance_stage(C.byref(m2_oparams(p)), C.byref(S0.cstate()), C.byref(S1.cstate()),
                          OB.ptr(np.ascontiguousarray(gx.reshape(-1))), OB.ptr(np.ascontiguousarray(gy.reshape(-1))),
                          C.c_double(p.dx), C.c_double(p.dy), C.c_double(dt))
    want = m2_dict(p, S1)
    for f in M.FIELDS:
        assert rel_err(got[f], want[f]) < 1e-11, (f, rel_err(got[f], want[f]))


@pytest.mark.parametrize("periodic", [True, False])
def test_rk2_step_matches_oracle(periodic):
    p = M.Params(Nx=32, Ny=36, periodic_x=periodic, periodic_y=periodic, eta=1e-3, G=2.0, sor_iters=120)
    st = random_mhd2d_state(p, 4)
    dt = 1.5e-3
    with M.Mhd2D(p) as h:
        h.upload(st)
        h.compute_poisson_gravity_periodic()
        h.rk2_step(dt)
        got = h.download()
    o = oracle()
    op = m2_oparams(p)
    S = m2_arrays(p, st)
    gx, gy = np.zeros(p.Nx * p.Ny), np.zeros(p.Nx * p.Ny)
    o.or_m2_poisson(C.byref(op), OB.ptr(S.a["rho"]), C.c_double(p.dx), C.c_double(p.dy), OB.ptr(gx), OB.ptr(gy))
    o.or_m2_rk2_step(C.byref(op), C.byref(S.cstate()), OB.ptr(gx), OB.ptr(gy), C.c_double(p.dx), C.c_double(p.dy),
                     C.c_double(dt))
    want = m2_dict(p, S)
    for f in M.FIELDS:
        assert rel_err(got[f], want[f]) < 1e-10, (f, rel_err(got[f], want[f]))


def test_compute_dt_matches_oracle():
    p = M.Params(Nx=48, Ny=40)
    st = random_mhd2d_state(p, 6)
    with M.Mhd2D(p) as h:
        h.upload(st)
        dt = h.compute_dt()
    o = oracle()
    S = m2_arrays(p, st)
    want = o.or_m2_compute_dt(C.byref(m2_oparams(p)), C.byref(S.cstate()), C.c_double(p.dx), C.c_double(p.dy))
    assert abs(dt - want) <= 1e-14 * want, (dt, want)


@pytest.mark.parametrize("comp", [False, True])
@pytest.mark.parametrize("N,nsteps", [(48, 8)])
def test_collapse_run_matches_oracle(N, nsteps, comp):
    """The benchmark's Gaussian-collapse problem (clamped boundaries, gravity, resistivity),
    nsteps of the driver loop with device-side dt, against the oracle's loop; with the reference's face
    update and with the compensated one."""
    p = ic.collapse2d_params(N, ct_compensated=comp)
    st = ic.collapse2d_ic(p)
    with M.Mhd2D(p) as h:
        h.upload(st)
        t, dts = h.run(nsteps, 0.0, 1e300, return_dts=True)
        got = h.download()
    o = oracle()
    S = m2_arrays(p, st)
    odts = np.zeros(nsteps)
    ot = o.or_m2_run(C.byref(m2_oparams(p)), C.byref(S.cstate()), C.c_double(p.dx), C.c_double(p.dy), nsteps,
                     C.c_double(0.0), C.c_double(1e300), OB.ptr(odts))
    want = m2_dict(p, S)
    assert rel_err(dts, odts) < 1e-11, (dts, odts)
    assert abs(t - ot) < 1e-11 * ot
    for f in M.FIELDS:
        assert rel_err(got[f], want[f]) < 1e-9, (f, rel_err(got[f], want[f]))


def test_config1_full_run_matches_the_reference_step_code():
    """Config 1 itself: the 2D collapse at 128^2 in plain FP64 (the reference's face update) for its 200
    steps, against the same run through the reference program's own step functions (compiled from its
    source; tests/golden/m2_config1.npz): every dt to 1e-10 relative, every field to 1e-8 relative to its
    maximum.  The GPU differs from the reference only by FMA contraction and its reduction order."""
    import os
    d = np.load(os.path.join(os.path.dirname(__file__), "golden", "m2_config1.npz"))
    p = ic.collapse2d_params(128, ct_compensated=False)
    st = ic.collapse2d_ic(p)
    assert st["rho"].sum() == d["in_rho_sum"][0]
    with M.Mhd2D(p) as h:
        h.upload(st)
        t, dts = h.run(200, 0.0, 1e300, return_dts=True)
        got = h.download()
        divb = h.divB_max()
    e_dt = rel_err(dts, d["dts"])
    errs = {f: rel_err(got[f].reshape(-1), d["out_" + f]) for f in M.FIELDS}
    print("config 1 (128^2, 200 steps): t %.9g vs %.9g, dt err %.2e, fields %s, max|divB| %.2e"
          % (t, d["t"][0], e_dt, {k: "%.1e" % v for k, v in errs.items()}, divb), flush=True)
    assert e_dt < 1e-10, e_dt
    assert abs(t - d["t"][0]) <= 1e-10 * d["t"][0]
    for f, e in errs.items():
        assert e < 1e-8, (f, e)


@pytest.mark.parametrize("prec", [CC_FP64, CC_MIXED])
def test_compensated_face_update_keeps_divb_at_one_rounding(prec):
    """The collapse at 256^2 for 300 steps: with the compensated face update max|div B| stays at one
    rounding of the face values (<= 2 ulp(max|B|) (1/dx + 1/dy), far below 1e-12); the reference's update,
    the same run otherwise, accumulates roundings."""
    out = {}
    for comp in (False, True):
        p = ic.collapse2d_params(256, precision=prec, ct_compensated=comp)
        with M.Mhd2D(p) as h:
            h.upload(ic.collapse2d_ic(p))
            h.run(300)
            s = h.download()
            out[comp] = (h.divB_max(), max(np.abs(s["Bx_f"]).max(), np.abs(s["By_f"]).max()), p)
    dc, bmax, p = out[True]
    bound = 2 * np.spacing(bmax) * (1 / p.dx + 1 / p.dy)
    print("max|divB| after 300 steps at 256^2: compensated %.2e (bound %.2e), reference update %.2e"
          % (dc, bound, out[False][0]), flush=True)
    assert dc <= bound and dc < 1e-12, (dc, bound)
    assert out[False][0] > dc


def test_divb_stays_at_roundoff():
    p = M.Params(Nx=64, Ny=64, eta=1e-3, G=4.0, sor_iters=100)
    st = ic.mhd2d_reference_ic(p)
    with M.Mhd2D(p) as h:
        h.upload(st)
        d0 = h.divB_max()
        h.run(10)
        d1 = h.divB_max()
    assert d0 < 1e-11 and d1 < 1e-11, (d0, d1)


def test_mixed_vs_fp64_l1_density():
    N, nsteps = 96, 20
    out = {}
    for prec in (CC_FP64, CC_MIXED):
        p = ic.collapse2d_params(N, precision=prec)
        with M.Mhd2D(p) as h:
            h.upload(ic.collapse2d_ic(p))
            h.run(nsteps)
            out[prec] = h.download()
    e = rel_l1(out[CC_MIXED]["rho"], out[CC_FP64]["rho"])
    assert e <= 1e-4, e


SHEARED_SHAPES = [(CC_FP64, 16, 64), (CC_FP64, 16, 32), (CC_FP64, 32, 32), (CC_FP64, 16, 80), (CC_MIXED, 16, 64),
                  (CC_MIXED, 16, 48), (CC_MIXED, 16, 32), (CC_MIXED, 16, 128), (CC_MIXED, 32, 64), (CC_MIXED, 32, 32),
                  (CC_MIXED, 16, 80)]


@pytest.mark.parametrize("prec,S,K", SHEARED_SHAPES)
def test_sheared_tiles_match_oracle_and_rectangular_tiles(prec, S, K, monkeypatch):
    """Sheared SOR tiles (CC_SOR_TILE="S,K,s"): the lexicographic iterate on grids with interior and edge
    tiles and a partial last sweep group -- FP64 against the oracle at 1e-9 and against the rectangular
    tiles at 1e-12, mixed at the 1e-5 bar of test_poisson_sor_mixed_close_to_oracle; one launch per level
    and the dataflow launch give the same bits."""
    monkeypatch.setenv("CC_SOR_WS", "0")
    for Nx, Ny, iters in ((300, 170, 600), (130, 67, 37), (64, 64, 600), (200, 90, 45)):
        rng = np.random.default_rng(Nx + Ny + iters)
        p = M.Params(Nx=Nx, Ny=Ny, periodic_x=False, periodic_y=False, sor_iters=iters, G=1.0)
        rho = 1.0 + rng.uniform(0, 1, (Ny, Nx)) + 5.0 * np.exp(-((np.arange(Nx)[None, :] - Nx / 2) ** 2 +
                                                               (np.arange(Ny)[:, None] - Ny / 3) ** 2) / 20.0)
        out = {}
        for form in ("flow", "levels", "rect"):
            monkeypatch.setenv("CC_SOR_TILE", "16,64" if form == "rect" else "%d,%d,s" % (S, K))
            if form == "levels":
                monkeypatch.setenv("CC_SOR_LEVELS", "1")
            else:
                monkeypatch.delenv("CC_SOR_LEVELS", raising=False)
            out[form] = M.poisson_sor(rho, p.dx, p.dy, p.G, iters, p.sor_omega, False, False, precision=prec)
        ox, oy = _oracle_poisson(p, rho)
        gx, gy = out["flow"]
        assert np.array_equal(gx, out["levels"][0]) and np.array_equal(gy, out["levels"][1])
        tol = 1e-9 if prec == CC_FP64 else 1e-5
        assert rel_err(gx, ox) < tol and rel_err(gy, oy) < tol, (Nx, Ny, iters, rel_err(gx, ox), rel_err(gy, oy))
        if prec == CC_FP64:
            rx, ry = out["rect"]
            assert rel_err(gx, rx) < 1e-12 and rel_err(gy, ry) < 1e-12, (rel_err(gx, rx), rel_err(gy, ry))


@pytest.mark.parametrize("prec", [CC_FP64, CC_MIXED])
def test_sheared_collapse_steps_match_oracle(prec, monkeypatch):
    """Collapse steps through the captured step graph with sheared SOR tiles (the dataflow launch)."""
    monkeypatch.setenv("CC_SOR_WS", "0")
    monkeypatch.setenv("CC_SOR_TILE", "16,64,s")
    monkeypatch.delenv("CC_SOR_LEVELS", raising=False)
    N, nsteps = 96, 3
    p = ic.collapse2d_params(N, precision=prec)
    st = ic.collapse2d_ic(p)
    with M.Mhd2D(p) as h:
        h.upload(st)
        t = h.run(nsteps)
        got = h.download()
    o = oracle()
    S_ = m2_arrays(p, st)
    ot = o.or_m2_run(C.byref(m2_oparams(p)), C.byref(S_.cstate()), C.c_double(p.dx), C.c_double(p.dy), nsteps,
                     C.c_double(0.0), C.c_double(1e300), None)
    want = m2_dict(p, S_)
    if prec == CC_FP64:
        assert abs(t - ot) < 1e-11 * ot
        for f in M.FIELDS:
            assert rel_err(got[f], want[f]) < 1e-9, f
    else:
        assert rel_l1(got["rho"], want["rho"]) <= 1e-4


@pytest.mark.parametrize("prec", [CC_FP64, CC_MIXED])
def test_sor_dataflow_equals_level_launches_bitwise(prec, monkeypatch):
    """The default tiled solve (one persistent launch, tiles started as their waits clear) and the
    one-launch-per-level solve (CC_SOR_LEVELS) give the same bits: a standalone solve on an uneven grid,
    and collapse steps run through the captured step graph."""
    monkeypatch.setenv("CC_SOR_WS", "0")
    rng = np.random.default_rng(17)
    Nx, Ny = 300, 170
    rho = 1.0 + rng.uniform(0, 1, (Ny, Nx))
    p = ic.collapse2d_params(160, precision=prec)
    st = ic.collapse2d_ic(p)
    out = {}
    for levels in (False, True):
        if levels:
            monkeypatch.setenv("CC_SOR_LEVELS", "1")
        else:
            monkeypatch.delenv("CC_SOR_LEVELS", raising=False)
        gx, gy = M.poisson_sor(rho, 1.0 / Nx, 1.0 / Ny, 1.0, 600, 1.8, False, False, precision=prec)
        with M.Mhd2D(p) as h:
            h.upload(st)
            t = h.run(3)
            out[levels] = (gx, gy, t, h.download())
    a, b = out[False], out[True]
    assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1]) and a[2] == b[2]
    for f in M.FIELDS:
        assert np.array_equal(a[3][f], b[3][f]), f


@pytest.mark.parametrize("mode", ["merged", "separate", "merged_sheared"])
@pytest.mark.parametrize("prec", [CC_FP64, CC_MIXED])
def test_batch_equals_instances_run_alone(prec, mode, monkeypatch):
    """cc_mhd2d_batch_run: independent collapse instances of uneven sizes, stepped together in one
    graph with per-instance dt -- their SOR tile levels merged into one launch per level (default) or
    each instance's step a separate branch (CC_BATCH_SEPARATE_SOR) -- give bit for bit the times, dts
    and states of the same instances run one by one, over two calls (the second continuing the
    first), with one instance stopping at its own t_end inside the run.  merged_sheared: the instances
    and the merged launches use sheared SOR tiles."""
    from core_collapse_amd.mhd2d import Mhd2D, Mhd2DBatch
    separate_sor = mode == "separate"
    monkeypatch.delenv("CC_SOR_WS", raising=False)  # LDS tiles (wave-streamed solves are never merged)
    if mode == "merged_sheared":
        monkeypatch.setenv("CC_SOR_TILE", "16,64,s")
    elif mode == "merged":
        monkeypatch.setenv("CC_SOR_TILE", "16,64")  # rectangular tiles
    else:
        monkeypatch.delenv("CC_SOR_TILE", raising=False)
    if separate_sor:
        monkeypatch.setenv("CC_BATCH_SEPARATE_SOR", "1")
    else:
        monkeypatch.delenv("CC_BATCH_SEPARATE_SOR", raising=False)
    sample = ic.collapse2d_batch_sample(5, seed=9, nmin=24, nmax=80)

    def make():
        sims = []
        for e in sample:
            p = ic.collapse2d_params(e["N"], eta=e["eta"], precision=prec, sor_iters=80)
            h = Mhd2D(p)
            h.upload(ic.collapse2d_ic(p, **e["ic"]))
            sims.append(h)
        return sims

    alone = make()
    p2 = alone[2].params
    dt0 = oracle().or_m2_compute_dt(C.byref(m2_oparams(p2)), C.byref(m2_arrays(p2, alone[2].download()).cstate()),
                                    C.c_double(p2.dx), C.c_double(p2.dy))
    t_end = np.full(len(sample), 1e300)
    t_end[2] = 8.5 * dt0   # reached within the 19 steps
    t_alone = np.array([h.run(12, 0.0, t_end[i]) for i, h in enumerate(alone)])
    t_alone = np.array([h.run(7, t_alone[i], t_end[i]) for i, h in enumerate(alone)])
    batch_sims = make()
    b = Mhd2DBatch(batch_sims)
    print("batch sizes", [e["N"] for e in sample], b.info(), flush=True)
    assert b.info()["merged"] == (not separate_sor)
    assert b.info()["tile"][2] == (mode != "merged") or separate_sor
    t = b.run(12, None, t_end)
    t = b.run(7, t, t_end)
    b.close()
    assert np.array_equal(t, t_alone), (t, t_alone)
    assert abs(t[2] - t_end[2]) <= 1e-15 * t_end[2]
    for h1, h2 in zip(alone, batch_sims):
        s1, s2 = h1.download(), h2.download()
        for f in M.FIELDS:
            assert np.array_equal(s1[f], s2[f]), f
    # every instance against the oracle's run of the same loop (M2:1111-1130), both calls, its own t_end:
    # FP64 to 1e-9 relative per field, the mixed form at the 1e-4 relative L1 density bar
    o = oracle()
    for i, (e, h) in enumerate(zip(sample, batch_sims)):
        p = h.params
        S_ = m2_arrays(p, ic.collapse2d_ic(p, **e["ic"]))
        op = m2_oparams(p)
        ot = o.or_m2_run(C.byref(op), C.byref(S_.cstate()), C.c_double(p.dx), C.c_double(p.dy), 12, C.c_double(0.0),
                         C.c_double(t_end[i]), None)
        ot = o.or_m2_run(C.byref(op), C.byref(S_.cstate()), C.c_double(p.dx), C.c_double(p.dy), 7, C.c_double(ot),
                         C.c_double(t_end[i]), None)
        got, want = h.download(), m2_dict(p, S_)
        if prec == CC_FP64:
            assert abs(t[i] - ot) <= 1e-11 * ot, (i, t[i], ot)
            for f in M.FIELDS:
                assert rel_err(got[f], want[f]) < 1e-9, (i, f, rel_err(got[f], want[f]))
        else:
            assert abs(t[i] - ot) <= 1e-6 * ot, (i, t[i], ot)
            assert rel_l1(got["rho"], want["rho"]) <= 1e-4, (i, rel_l1(got["rho"], want["rho"]))
    for h in alone + batch_sims:
        h.close()


def test_batch_of_64_sampled_collapses_equals_instances_run_alone():
    """The 64-instance batch itself (collapse2d_batch_sample(64, seed 42, N in [128, 1024]), mixed
    precision, 600 SOR sweeps): 3 batch steps give bit for bit the times and states of the instances
    stepped alone, and the face field stays divergence-free at round-off (max|div B| <= 1e-12)."""
    from core_collapse_amd.mhd2d import Mhd2D, Mhd2DBatch
    sample = ic.collapse2d_batch_sample(64, seed=42, nmin=128, nmax=1024)
    assert len(sample) == 64 and min(e["N"] for e in sample) >= 128 and max(e["N"] for e in sample) <= 1024

    def make():
        sims = []
        for e in sample:
            p = ic.collapse2d_params(e["N"], eta=e["eta"], precision=CC_MIXED)
            h = Mhd2D(p)
            h.upload(ic.collapse2d_ic(p, **e["ic"]))
            sims.append(h)
        return sims

    batch_sims = make()
    b = Mhd2DBatch(batch_sims)
    print("batch of 64: cells %d, %s" % (sum(e["N"] ** 2 for e in sample), b.info()), flush=True)
    t = b.run(3, None, np.full(64, 1e300))
    b.close()
    divb = max(h.divB_max() for h in batch_sims)
    print("batch max|divB| after 3 steps: %.3e" % divb, flush=True)
    assert divb <= 1e-12, divb
    for i, h in enumerate(batch_sims):
        p = h.params
        with Mhd2D(p) as a:
            a.upload(ic.collapse2d_ic(p, **sample[i]["ic"]))
            ta = a.run(3)
            sa = a.download()
        assert ta == t[i], (i, ta, t[i])
        sb = h.download()
        for f in M.FIELDS:
            assert np.array_equal(sa[f], sb[f]), (i, f)
        h.close()


@pytest.mark.parametrize("prec", [CC_FP64, CC_MIXED])
def test_sor_autotune_picks_a_compiled_shape_and_keeps_parity(prec, monkeypatch):
    """cc_sor_autotune: the handle times every compiled tile shape once for its grid and runs with the
    fastest; the solve is the same lexicographic iterate whatever the shape (FP64 against the oracle at
    the tolerance of test_collapse_run_matches_oracle, mixed at the 1e-4 L1 bar)."""
    monkeypatch.setenv("CC_SOR_WS", "0")
    N, nsteps = 80, 4
    p = ic.collapse2d_params(N, precision=prec)
    st = ic.collapse2d_ic(p)
    M.sor_autotune(True)
    try:
        with M.Mhd2D(p) as h:
            h.upload(st)
            t = h.run(nsteps)
            got = h.download()
        S, WJ, sheared, fp64k = M.sor_tile_choice(N, N, p.sor_iters, prec, with_kernels=True)
        print("autotuned SOR tile", S, WJ, sheared, "fp64 kernels" if fp64k else "", flush=True)
        assert not (fp64k and prec == CC_FP64)
        if sheared:
            assert (S, WJ) in ({(16, 64), (16, 32), (32, 32), (16, 80)} if prec == CC_FP64 or fp64k else
                               {(16, 64), (16, 48), (16, 32), (16, 128), (32, 64), (32, 32), (16, 80)})
        else:
            assert (S, WJ) in {(16, 64), (32, 64), (16, 128), (32, 128)}
    finally:
        M.sor_autotune(False)
    o = oracle()
    S_ = m2_arrays(p, st)
    ot = o.or_m2_run(C.byref(m2_oparams(p)), C.byref(S_.cstate()), C.c_double(p.dx), C.c_double(p.dy), nsteps,
                     C.c_double(0.0), C.c_double(1e300), None)
    want = m2_dict(p, S_)
    if prec == CC_FP64:
        assert abs(t - ot) < 1e-11 * ot
        for f in M.FIELDS:
            assert rel_err(got[f], want[f]) < 1e-9, f
    else:
        assert rel_l1(got["rho"], want["rho"]) <= 1e-4
