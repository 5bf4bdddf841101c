"""GPU parity of the 3D Euler HLLC + FFT gravity step (libccamd.so through its C ABI) against the
oracle (oracle/oracle_e3.c, pinned in test_oracle_euler3d.py).

* Hydro without gravity, FP64: the kernels are compiled without FP contraction and repeat the
  oracle's operation order, so rhs, Heun steps and dt are bit-identical.
* Gravity: the FFT solve refined in FP64 to max|rhs - L_h phi| <= 1e-10 max|rhs| against the
  oracle's exact spectral solve: phi and g agree to 1e-9 relative (the residual bound divided by
  the smallest eigenvalue of L_h gives the error bound); the refinement converges in FP64 and in
  mixed precision (FP32 FFT), where it needs more corrections.
* Full steps with gravity: 1e-9 (FP64); mixed precision vs FP64: relative L1 density <= 1e-4.
"""
import ctypes as C

import numpy as np
import pytest

from core_collapse_amd import CC_FP64, CC_MIXED
from core_collapse_amd import euler3d as E
from core_collapse_amd import ic
from oracle import bind as OB
from tests.helpers import oracle

pytestmark = pytest.mark.gpu


def _op(p):
    return OB.e3_params(gamma=p.gamma, cfl=p.cfl, x0=p.x0, x1=p.x1, y0=p.y0, y1=p.y1, z0=p.z0, z1=p.z1, Nx=p.Nx,
                        Ny=p.Ny, Nz=p.Nz, floor_rho=p.floor_rho, floor_P=p.floor_P, gravity_on=int(p.gravity_on), G=p.G)


def _state(rng, p):
    Nx, Ny, Nz = p.Nx, p.Ny, p.Nz
    x = (np.arange(Nx) + 0.5) / Nx
    y = (np.arange(Ny) + 0.5) / Ny
    z = (np.arange(Nz) + 0.5) / Nz
    Z, Y, X = np.meshgrid(z, y, x, indexing="ij")

    def smooth(a):
        ph = rng.uniform(0, 2 * np.pi, 3)
        return a * np.sin(2 * np.pi * X + ph[0]) * np.cos(2 * np.pi * Y + ph[1]) * np.cos(2 * np.pi * Z + ph[2])
    rho = 1.0 + 0.5 * smooth(1.0) + 0.05 * rng.standard_normal(X.shape)
    u, v, w = smooth(0.8) + 0.1 * rng.standard_normal(X.shape), smooth(0.8), smooth(0.8)
    P = 0.8 + 0.3 * smooth(1.0)
    return {"rho": rho, "mx": rho * u, "my": rho * v, "mz": rho * w,
            "E": P / (p.gamma - 1.0) + 0.5 * rho * ((u * u + v * v) + w * w)}


def _rel(a, b):
    return np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)


@pytest.mark.parametrize("shape", [(16, 8, 32), (32, 32, 32)])
def test_hydro_rhs_and_steps_bitwise_without_gravity(shape):
    Nx, Ny, Nz = shape
    p = E.Params(Nx=Nx, Ny=Ny, Nz=Nz, gravity_on=False)
    st = _state(np.random.default_rng(1), p)
    o = oracle()
    op = _op(p)
    with E.Euler3D(p) as h:
        h.upload(st)
        d = h.rhs()
        U = OB.E3Arrays(Nx, Ny, Nz, st)
        dU = OB.E3Arrays(Nx, Ny, Nz)
        z = np.zeros(Nx * Ny * Nz)
        o.or_e3_rhs_g(C.byref(op), C.byref(U.cstate()), C.byref(dU.cstate()), OB.ptr(z), OB.ptr(z), OB.ptr(z))
        for f in E.FIELDS:
            assert np.array_equal(d[f].reshape(-1), dU[f]), f
        assert h.compute_dt() == o.or_e3_compute_dt(C.byref(op), C.byref(U.cstate()))
        t, dts = h.run(4, 0.0, 1e300, return_dts=True)
        odts = np.zeros(4)
        ot = o.or_e3_run(C.byref(op), C.byref(U.cstate()), 4, C.c_double(0.0), C.c_double(1e300), OB.ptr(odts))
        got = h.download()
    assert np.array_equal(dts, odts) and t == ot
    for f in E.FIELDS:
        assert np.array_equal(got[f].reshape(-1), U[f]), f


@pytest.mark.parametrize("prec", [CC_FP64, CC_MIXED])
@pytest.mark.parametrize("shape", [(16, 32, 8), (32, 32, 32), (64, 16, 32), (1024, 16, 8), (16, 1024, 8), (8, 16, 1024),
                                   (128, 8, 256)])
def test_fft_gravity_matches_oracle(monkeypatch, prec, shape):
    """Every radix-stage combination of the Stockham passes: N = 8 (8), 16 (8, 2), 32 (8, 4), 64 (8, 8),
    128 (8, 8, 2), 256 (8, 8, 4), 1024 (8, 8, 8, 2), on each axis.  The mixed case runs the FP32 FFT
    (CC_E3_FFT=fp32; at this tolerance a mixed handle otherwise takes the FP64 one)."""
    monkeypatch.setenv("CC_E3_FFT", "fp32")
    Nx, Ny, Nz = shape
    p = E.Params(Nx=Nx, Ny=Ny, Nz=Nz, x1=2.0, G=0.7, precision=prec, poisson_tol=1e-10, poisson_max_iter=12)
    rng = np.random.default_rng(2)
    st = _state(rng, p)
    st["rho"] = np.exp(rng.normal(0, 0.7, (Nz, Ny, Nx)))
    with E.Euler3D(p) as h:
        h.upload(st)
        gx, gy, gz, phi = h.gravity()
        stats = h.poisson_stats()
    o = oracle()
    n = Nx * Ny * Nz
    ogx, ogy, ogz, ophi = np.zeros(n), np.zeros(n), np.zeros(n), np.zeros(n)
    o.or_e3_poisson(C.byref(_op(p)), OB.ptr(np.ascontiguousarray(st["rho"].reshape(-1))), OB.ptr(ogx), OB.ptr(ogy),
                    OB.ptr(ogz), OB.ptr(ophi))
    print(shape, prec, stats, flush=True)
    assert stats["solves"] == 1 and stats["unconverged"] == 0 and stats["resid_last"] <= 1e-10
    assert stats["iters_last"] >= 1
    assert _rel(phi.reshape(-1) - phi.mean(), ophi) < 1e-9
    for a, b in ((gx, ogx), (gy, ogy), (gz, ogz)):
        assert _rel(a.reshape(-1), b) < 1e-9


@pytest.mark.parametrize("fft", ["fp32", "fp64"])
def test_refinement_warm_start_and_uniform_density(monkeypatch, fft):
    monkeypatch.setenv("CC_E3_FFT", fft)
    p = E.Params(Nx=16, Ny=16, Nz=16, precision=CC_MIXED)
    rng = np.random.default_rng(3)
    st = _state(rng, p)
    with E.Euler3D(p) as h:
        h.upload(st)
        h.gravity()
        first = h.poisson_stats()["iters_last"]
        h.gravity()   # the same density again: the kept phi already satisfies the tolerance
        s = h.poisson_stats()
        assert s["solves"] == 2 and s["iters_last"] == 0 and first >= 1
        st["rho"] = np.full_like(st["rho"], 3.0)
        h.upload(st)
        gx, gy, gz, phi = h.gravity()
        assert np.all(phi == 0.0) and np.all(gx == 0.0) and h.poisson_stats()["unconverged"] == 0


@pytest.mark.parametrize("prec", [CC_FP64, CC_MIXED])
def test_sphere_steps_match_oracle(prec):
    p = ic.sphere3d_params(32, precision=prec)
    st = ic.sphere3d_ic(p)
    nsteps = 6
    with E.Euler3D(p) as h:
        h.upload(st)
        m0 = h.totals()["mass"]
        t, dts = h.run(nsteps, 0.0, 1e300, return_dts=True)
        got = h.download()
        m1 = h.totals()["mass"]
        stats = h.poisson_stats()
    U = OB.E3Arrays(p.Nx, p.Ny, p.Nz, st)
    odts = np.zeros(nsteps)
    ot = oracle().or_e3_run(C.byref(_op(p)), C.byref(U.cstate()), nsteps, C.c_double(0.0), C.c_double(1e300),
                            OB.ptr(odts))
    print(prec, stats, flush=True)
    assert stats["solves"] == 2 * nsteps and stats["unconverged"] == 0
    assert abs(m1 - m0) <= 1e-13 * m0
    if prec == CC_FP64:
        assert _rel(dts, odts) < 1e-10 and abs(t - ot) < 1e-10 * ot
        for f in E.FIELDS:
            assert _rel(got[f].reshape(-1), U[f]) < 1e-9, (f, _rel(got[f].reshape(-1), U[f]))
    else:
        rho = U["rho"]
        l1 = np.abs(got["rho"].reshape(-1) - rho).sum() / np.abs(rho).sum()
        assert l1 <= 1e-4, l1


def test_mixed_vs_fp64_density_error():
    out = {}
    for prec in (CC_FP64, CC_MIXED):
        p = ic.sphere3d_params(64, precision=prec)
        with E.Euler3D(p) as h:
            h.upload(ic.sphere3d_ic(p))
            h.run(20)
            out[prec] = h.download()["rho"]
    l1 = np.abs(out[CC_MIXED] - out[CC_FP64]).sum() / np.abs(out[CC_FP64]).sum()
    print("mixed vs fp64 L1(rho)", l1, flush=True)
    assert l1 <= 1e-4, l1


def test_bad_grid_is_rejected():
    from core_collapse_amd._lib import CCError
    with pytest.raises(CCError):
        E.Euler3D(E.Params(Nx=24, Ny=16, Nz=16))


@pytest.mark.parametrize("prec", [CC_FP64, CC_MIXED])
@pytest.mark.parametrize("shape", [(8, 8, 8), (16, 8, 32), (64, 16, 16), (32, 32, 8)])
def test_fused_pass_equals_split_passes(monkeypatch, prec, shape):
    """The faces + update pass (k_e3_fused: one z march per 32 x 7 column of cells, x/y fluxes through LDS, z fluxes
    in registers) against the face kernels + k_e3_update, three steps with gravity, z chunks of 3 planes (seams,
    the one-plane warm-up, tiles wider than an 8-cell axis): the same arithmetic in the same order without FP
    contraction, so the states, dts and t are bit-identical in both precisions."""
    Nx, Ny, Nz = shape
    monkeypatch.setenv("CC_E3_ZCHUNK", "3")
    p = E.Params(Nx=Nx, Ny=Ny, Nz=Nz, gravity_on=True, precision=prec)
    st = _state(np.random.default_rng(4), p)
    out = {}
    for fused in ("1", "0"):
        monkeypatch.setenv("CC_E3_FUSED", fused)
        with E.Euler3D(p) as h:
            h.upload(st)
            t, dts = h.run(3, 0.0, 1e300, return_dts=True)
            out[fused] = (t, dts, h.download())
    (ta, da, a), (tb, db, b) = out["1"], out["0"]
    assert ta == tb and np.array_equal(da, db), (da, db)
    for f in E.FIELDS:
        assert np.array_equal(a[f], b[f]), (f, _rel(a[f], b[f]))


def test_mixed_handle_runs_the_fp64_fft_at_tight_tolerance(monkeypatch):
    """At poisson_tol < 1e-6 two FP32-FFT corrections cost more than the one FP64-FFT correction that reaches
    the tolerance, so a mixed handle solves gravity with the FP64 FFT: phi and g equal an FP64 handle's bit for
    bit, after one correction."""
    monkeypatch.delenv("CC_E3_FFT", raising=False)
    rng = np.random.default_rng(6)
    out = {}
    for prec in (CC_MIXED, CC_FP64):
        p = E.Params(Nx=32, Ny=16, Nz=32, G=0.7, precision=prec, poisson_tol=1e-10, poisson_max_iter=12)
        st = _state(rng if prec == CC_MIXED else np.random.default_rng(6), p)
        with E.Euler3D(p) as h:
            h.upload(st)
            out[prec] = (h.gravity(), h.poisson_stats())
    (gm, sm), (gd, sd) = out[CC_MIXED], out[CC_FP64]
    assert sm["iters_last"] == sd["iters_last"] == 1, (sm, sd)
    for a, b in zip(gm, gd):
        assert np.array_equal(a, b)
