"""GPU parity of the 2D Navier-Stokes HLLC + FFT step (libccamd.so through its C ABI) against the
oracle (oracle/oracle_ns.c, pinned bit for bit to the reference in test_oracle_ns2d.py).

The NS kernels and the FFT are compiled without FP contraction and the FFT uses the reference's
own twiddle sequence (built on the host by repeated multiplication), so line transforms round like
the reference; the mean of rho inside the Poisson solve is a fixed-order tree instead of the
sequential loop.  Tolerances: FFT and HLLC bit for bit, Poisson / rhs / steps / runs
<= 1e-10 relative (stated per test); mixed precision <= 1e-4 L1.
"""
import ctypes as C
import os

import numpy as np
import pytest

from core_collapse_amd import CC_FP64, CC_MIXED, _lib
from core_collapse_amd import ic
from core_collapse_amd import ns2d as NSM
from oracle import bind as OB
from tests.helpers import oracle

pytestmark = pytest.mark.gpu
G = os.path.join(os.path.dirname(__file__), "golden")
SIZES = {"a": (32, 32, 1), "b": (16, 32, 0)}


def _rel(a, b):
    a, b = np.asarray(a, dtype=np.float64).reshape(-1), np.asarray(b, dtype=np.float64).reshape(-1)
    return np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)


@pytest.mark.parametrize("shape", [(16, 16), (32, 8)])
def test_fft2d_matches_reference_vectors(shape):
    Nx, Ny = shape
    d = np.load(os.path.join(G, "ns_step.npz"))
    x = d["fft_in_%dx%d" % (Nx, Ny)]
    z = x[..., 0] + 1j * x[..., 1]
    for inv in (0, 1):
        got = NSM.fft2d(z, inverse=bool(inv))
        w = d["fft_out%d_%dx%d" % (inv, Nx, Ny)]
        want = w[..., 0] + 1j * w[..., 1]
        assert np.array_equal(got, want), (inv, int((got != want).sum()))  # bit for bit


@pytest.mark.parametrize("tag", ["a", "b"])
def test_poisson_dt_rhs_heun_match_reference_vectors(tag):
    d = np.load(os.path.join(G, "ns_step.npz"))
    Nx, Ny, vis = SIZES[tag]
    p = NSM.Params(Nx=Nx, Ny=Ny, viscosity_on=bool(vis), mu=1e-3)
    U = d[tag + "_ic"]
    with NSM.NS2D(p) as s:
        s.upload(U)
        phi = s.poisson(np.maximum(U[0::4], p.floor_rho))
        assert _rel(phi, d[tag + "_phi"]) <= 1e-10
        assert abs(s.compute_dt() - d[tag + "_dt"][0]) <= 1e-14 * d[tag + "_dt"][0]
        assert _rel(s.compute_rhs(), d[tag + "_rhs"]) <= 1e-10
        s.heun_step(2e-3)
        assert _rel(s.download(), d[tag + "_heun"]) <= 1e-12


@pytest.mark.parametrize("tag", ["a", "b"])
def test_run_matches_reference_vectors(tag):
    d = np.load(os.path.join(G, "ns_step.npz"))
    Nx, Ny, vis = SIZES[tag]
    with NSM.NS2D(NSM.Params(Nx=Nx, Ny=Ny, viscosity_on=bool(vis), mu=1e-3)) as s:
        s.upload(d[tag + "_ic"])
        t, dts = s.run(20, t_end=10.0, return_dts=True)
        assert _rel(dts, d[tag + "_run_dts"]) <= 1e-10
        assert _rel(s.download(), d[tag + "_run"]) <= 1e-10


def test_default_problem_matches_oracle():
    """The reference default (128^2, mu 1e-6, G 10) for 3 steps."""
    p = NSM.Params()
    U = ic.ns2d_reference_ic(p).reshape(-1).copy()
    R = U.copy()
    dts_o = np.zeros(3)
    oracle().or_ns_run(C.byref(OB.ns_params()), OB.ptr(R), 3, C.c_double(0.0), C.c_double(10.0), OB.ptr(dts_o))
    with NSM.NS2D(p) as s:
        s.upload(U)
        t, dts = s.run(3, t_end=10.0, return_dts=True)
        assert _rel(dts, dts_o) <= 1e-12
        assert _rel(s.download(), R) <= 1e-10


def test_non_power_of_two_is_an_error():
    with pytest.raises(_lib.CCError) as e:
        NSM.NS2D(NSM.Params(Nx=96, Ny=64))
    assert e.value.status == 2


def test_mixed_hllc_close_to_fp64():
    p = NSM.Params(Nx=64, Ny=64)
    U = ic.ns2d_reference_ic(p)
    res = {}
    for prec in (CC_FP64, CC_MIXED):
        with NSM.NS2D(NSM.Params(Nx=64, Ny=64, precision=prec)) as s:
            s.upload(U)
            s.run(10, t_end=10.0)
            res[prec] = s.download()
    a, b = res[CC_MIXED][..., 0], res[CC_FP64][..., 0]
    assert np.abs(a - b).sum() / np.abs(b).sum() <= 1e-4
