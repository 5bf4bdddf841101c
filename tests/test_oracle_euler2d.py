"""Pinning the 2D Euler oracle (oracle/oracle_eh.c), CPU only: golden vectors from the reference's own
functions (tests/golden/eh_step.npz, bit for bit), the program's default IC as libccamd builds it
(core_collapse_amd/ic.py), and closed forms -- uniform, supersonic and stationary-contact HLLC fluxes,
the converged SOR against a direct spectral solve of the same 5-point periodic Poisson problem, and
conservation of a periodic run."""
import ctypes as C
import os

import numpy as np
import pytest

from oracle import bind as OB
from tests.helpers import oracle, ref

G = os.path.join(os.path.dirname(__file__), "golden")
SIZES = {"a": (16, 16, 1), "b": (24, 20, 0)}


def _flux(lib, fn, p, WL, WR):
    out = np.zeros((len(WL), 4))
    for k in range(len(WL)):
        getattr(lib, fn)(C.byref(p), OB.ptr(np.ascontiguousarray(WL[k])), OB.ptr(np.ascontiguousarray(WR[k])),
                         OB.ptr(out[k]))
    return out


def test_hllc_golden_bitwise():
    d = np.load(os.path.join(G, "eh_step.npz"))
    p = OB.eh_params()
    assert np.array_equal(_flux(oracle(), "or_eh_hllc_x", p, d["hllc_WL"], d["hllc_WR"]), d["hllc_Fx"])
    assert np.array_equal(_flux(oracle(), "or_eh_hllc_y", p, d["hllc_WL"], d["hllc_WR"]), d["hllc_Fy"])


@pytest.mark.parametrize("tag", ["a", "b"])
def test_poisson_dt_rhs_rk2_run_golden_bitwise(tag):
    d = np.load(os.path.join(G, "eh_step.npz"))
    Nx, Ny, per = SIZES[tag]
    o = oracle()
    p = OB.eh_params(Nx=Nx, Ny=Ny, periodic_x=per, periodic_y=per)
    S = OB.EHArrays(Nx, Ny, {f: d[tag + "_ic_" + f] for f in OB.EHArrays.FIELDS})
    gx, gy = np.zeros(Nx * Ny), np.zeros(Nx * Ny)
    o.or_eh_poisson(C.byref(p), C.byref(S.cstate()), OB.ptr(gx), OB.ptr(gy), None)
    assert np.array_equal(gx, d[tag + "_gx"]) and np.array_equal(gy, d[tag + "_gy"])
    assert o.or_eh_compute_dt(C.byref(p), C.byref(S.cstate())) == d[tag + "_dt"][0]
    dU = OB.EHArrays(Nx, Ny)
    o.or_eh_rhs(C.byref(p), C.byref(S.cstate()), C.byref(dU.cstate()))
    for f in S.FIELDS:
        assert np.array_equal(dU.a[f], d[tag + "_rhs_" + f]), f
    T = S.copy()
    o.or_eh_rk2_step(C.byref(p), C.byref(T.cstate()), C.c_double(2e-3))
    for f in S.FIELDS:
        assert np.array_equal(T.a[f], d[tag + "_rk2_" + f]), f
    R = S.copy()
    dts = np.zeros(10)
    t = o.or_eh_run(C.byref(p), C.byref(R.cstate()), 10, C.c_double(0.0), C.c_double(5.0), OB.ptr(dts))
    assert np.array_equal(dts, d[tag + "_run_dts"]) and t == d[tag + "_run_t"][0]
    for f in S.FIELDS:
        assert np.array_equal(R.a[f], d[tag + "_run_" + f]), f


@pytest.mark.parametrize("tag", ["a", "b"])
def test_reference_ic_restated_bitwise(tag):
    from core_collapse_amd import ic
    from core_collapse_amd.euler2d import Params
    d = np.load(os.path.join(G, "eh_step.npz"))
    Nx, Ny, per = SIZES[tag]
    U = ic.euler2d_reference_ic(Params(Nx=Nx, Ny=Ny))
    for f in ("rho", "mx", "my", "E"):
        assert np.array_equal(U[f].reshape(-1), d[tag + "_ic_" + f]), f


def _phys(W, gamma, axis):
    rho, u, v, P = W
    E = P / (gamma - 1) + 0.5 * rho * (u * u + v * v)
    un = u if axis == 0 else v
    return np.array([rho * un, rho * u * un + (P if axis == 0 else 0), rho * v * un + (P if axis == 1 else 0),
                     (E + P) * un])


def test_hllc_closed_forms():
    o = oracle()
    p = OB.eh_params(gamma=1.4)
    rng = np.random.default_rng(3)
    for k in range(300):
        W = np.array([rng.uniform(0.2, 3), *rng.normal(0, 1.5, 2), rng.uniform(0.05, 2)])
        for axis, fn in ((0, "or_eh_hllc_x"), (1, "or_eh_hllc_y")):
            got = _flux(o, fn, p, W[None], W[None])[0]
            assert np.allclose(got, _phys(W, 1.4, axis), rtol=1e-12, atol=1e-12)
            Wf = W.copy()
            Wf[1 + axis] = 30.0  # supersonic to the right: upwind F_L
            V = Wf.copy(); V[0] *= 1.5; V[3] *= 0.5
            assert np.allclose(_flux(o, fn, p, Wf[None], V[None])[0], _phys(Wf, 1.4, axis), rtol=1e-13, atol=1e-13)
    # stationary contact: equal pressure, zero velocity, different density -> flux (0, P, 0, 0)
    for k in range(50):
        P = rng.uniform(0.1, 2)
        L = np.array([rng.uniform(0.1, 1), 0.0, 0.0, P]); R = np.array([rng.uniform(1, 5), 0.0, 0.0, P])
        assert np.allclose(_flux(o, "or_eh_hllc_x", p, L[None], R[None])[0], [0, P, 0, 0], atol=1e-13)
        assert np.allclose(_flux(o, "or_eh_hllc_y", p, L[None], R[None])[0], [0, 0, P, 0], atol=1e-13)


def test_converged_sor_equals_spectral_solve_of_the_5point_problem():
    """The SOR solution at tolerance 1e-10 against the exact solution of the same periodic 5-point
    Poisson equation (FFT with the discrete eigenvalues (2 cos(2 pi k / N) - 2) / h^2)."""
    Nx, Ny = 24, 20
    p = OB.eh_params(Nx=Nx, Ny=Ny, G=1.3)
    rng = np.random.default_rng(9)
    rho = 1 + 0.5 * rng.random((Ny, Nx))
    S = OB.EHArrays(Nx, Ny, {"rho": rho, "mx": np.zeros(Nx * Ny), "my": np.zeros(Nx * Ny), "E": np.ones(Nx * Ny)})
    gx, gy, phi = np.zeros(Nx * Ny), np.zeros(Nx * Ny), np.zeros(Nx * Ny)
    it = oracle().or_eh_poisson(C.byref(p), C.byref(S.cstate()), OB.ptr(gx), OB.ptr(gy), OB.ptr(phi))
    assert 0 < it < 50000
    dx, dy = 1.0 / Nx, 1.0 / Ny
    rhs = 4 * np.pi * 1.3 * (rho - rho.mean())
    kx = (2 * np.cos(2 * np.pi * np.arange(Nx) / Nx) - 2) / dx ** 2
    ky = (2 * np.cos(2 * np.pi * np.arange(Ny) / Ny) - 2) / dy ** 2
    lam = ky[:, None] + kx[None, :]
    lam[0, 0] = 1.0
    ph = np.fft.fft2(rhs) / lam
    ph[0, 0] = 0.0
    exact = np.real(np.fft.ifft2(ph))
    assert np.abs(phi.reshape(Ny, Nx) - exact).max() < 1e-9 * np.abs(exact).max()
    gxe = -(np.roll(exact, -1, 1) - np.roll(exact, 1, 1)) / (2 * dx)
    assert np.abs(gx.reshape(Ny, Nx) - gxe).max() < 1e-8 * np.abs(gxe).max()


def test_periodic_run_conserves():
    from core_collapse_amd import ic
    from core_collapse_amd.euler2d import Params
    for grav in (0, 1):
        p = OB.eh_params(Nx=20, Ny=16, gravity_on=grav)
        U = ic.euler2d_reference_ic(Params(Nx=20, Ny=16))
        S = OB.EHArrays(20, 16, {f: U[f] for f in U})
        s0 = {f: S.a[f].sum() for f in S.FIELDS}
        oracle().or_eh_run(C.byref(p), C.byref(S.cstate()), 5, C.c_double(0.0), C.c_double(5.0), None)
        assert abs(S.a["rho"].sum() - s0["rho"]) < 1e-12 * s0["rho"]
        if not grav:  # gravity is a momentum and energy source
            for f in ("mx", "my", "E"):
                assert abs(S.a[f].sum() - s0[f]) < 1e-11 * max(1.0, abs(s0[f])), f


def test_oracle_equals_reference_randomized():
    r = ref("ref_eh")
    if r is None:
        pytest.skip("oracle/_ref/libref_eh.so is not built here (reference source absent)")
    o = oracle()
    rng = np.random.default_rng(11)
    WL = np.array([[rng.uniform(0.05, 4), *rng.normal(0, 2, 2), rng.uniform(0.001, 2)] for _ in range(3000)])
    WR = np.array([[rng.uniform(0.05, 4), *rng.normal(0, 2, 2), rng.uniform(0.001, 2)] for _ in range(3000)])
    p = OB.eh_params()
    for fn in ("or_eh_hllc_x", "or_eh_hllc_y"):
        assert np.array_equal(_flux(o, fn, p, WL, WR), _flux(r, fn, p, WL, WR))
    p = OB.eh_params(Nx=18, Ny=14, periodic_x=1, periodic_y=0, G=2.0)
    A = OB.EHArrays(18, 14)
    r.or_eh_init_random_blob(C.byref(p), C.byref(A.cstate()))
    B = A.copy()
    ta = o.or_eh_run(C.byref(p), C.byref(A.cstate()), 4, C.c_double(0.0), C.c_double(5.0), None)
    tb = r.or_eh_run(C.byref(p), C.byref(B.cstate()), 4, C.c_double(0.0), C.c_double(5.0), None)
    assert ta == tb
    for f in A.FIELDS:
        assert np.array_equal(A.a[f], B.a[f]), f
