"""GPU parity of the 2D Euler HLLC + SOR step (libccamd.so through its C ABI) against the oracle
(oracle/oracle_eh.c, itself pinned bit for bit to the reference in test_oracle_euler2d.py).

The EH kernels are compiled without FP contraction, so the HLLC, the MUSCL states and the stage
update round like the reference; what differs is the order of the two sums inside the SOR solve
(mean of rho, mean of phi: fixed-order trees instead of the sequential loop).  The SOR stops on an
in-sweep residual below 1e-10, so the GPU may stop one sweep earlier or later than the oracle; the
resulting g differs by at most the last sweep's correction.  Tolerances: HLLC bit for bit;
gravity, rhs, steps and runs <= 1e-8 relative (stated per test); mixed precision <= 1e-4 L1.
"""
import os

import numpy as np
import pytest

from core_collapse_amd import CC_FP64, CC_MIXED
from core_collapse_amd import euler2d as E
from core_collapse_amd import ic
from oracle import bind as OB
from tests.helpers import oracle

pytestmark = pytest.mark.gpu
G = os.path.join(os.path.dirname(__file__), "golden")
SIZES = {"a": (16, 16, 1), "b": (24, 20, 0)}
FIELDS = ("rho", "mx", "my", "E")


def _rel(a, b):
    a, b = np.asarray(a, dtype=np.float64).reshape(-1), np.asarray(b, dtype=np.float64).reshape(-1)
    return np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)


def _sim(tag, **kw):
    Nx, Ny, per = SIZES[tag]
    return E.Euler2D(E.Params(Nx=Nx, Ny=Ny, periodic_x=bool(per), periodic_y=bool(per), **kw))


@pytest.mark.parametrize("variant", ["eh", "ns"])
def test_hllc_batch_matches_reference_vectors(variant):
    d = np.load(os.path.join(G, "%s_step.npz" % variant))
    gamma, floor = (1.2, 1e-12) if variant == "eh" else (1.4, 1e-10)
    for axis, key in ((0, "hllc_Fx"), (1, "hllc_Fy")):
        got = E.hllc_batch(d["hllc_WL"], d["hllc_WR"], axis=axis, variant=variant, gamma=gamma, floor_rho=floor)
        want = d[key]
        # compiled without contraction, FP64 division and sqrt correctly rounded: bit for bit
        assert np.array_equal(got, want), (variant, axis, int((got != want).sum()))


@pytest.mark.parametrize("tag", ["a", "b"])
def test_gravity_dt_rhs_rk2_match_reference_vectors(tag):
    d = np.load(os.path.join(G, "eh_step.npz"))
    with _sim(tag) as s:
        s.upload({f: d[tag + "_ic_" + f] for f in FIELDS})
        gx, gy, phi, sweeps = s.gravity()
        assert sweeps > 0
        assert _rel(gx, d[tag + "_gx"]) <= 1e-8 and _rel(gy, d[tag + "_gy"]) <= 1e-8
        assert abs(s.compute_dt() - d[tag + "_dt"][0]) <= 1e-14 * d[tag + "_dt"][0]
        r = s.rhs()
        for f in FIELDS:
            assert _rel(r[f], d[tag + "_rhs_" + f]) <= 1e-8, f
        s.rk2_step(2e-3)
        out = s.download()
        for f in FIELDS:
            assert _rel(out[f], d[tag + "_rk2_" + f]) <= 1e-10, f


@pytest.mark.parametrize("tag", ["a", "b"])
def test_run_matches_reference_vectors(tag):
    d = np.load(os.path.join(G, "eh_step.npz"))
    with _sim(tag) as s:
        s.upload({f: d[tag + "_ic_" + f] for f in FIELDS})
        t, dts = s.run(10, t=0.0, t_end=5.0, return_dts=True)
        assert _rel(dts, d[tag + "_run_dts"]) <= 1e-9
        assert abs(t - d[tag + "_run_t"][0]) <= 1e-9 * t
        out = s.download()
        for f in FIELDS:
            assert _rel(out[f], d[tag + "_run_" + f]) <= 1e-8, f


@pytest.mark.parametrize("Nx,Ny,per", [(50, 50, True), (96, 80, False), (130, 70, True)])
def test_run_matches_oracle(Nx, Ny, per):
    """The program's default IC (50^2 is the reference default); 96x80 keeps phi in LDS, 130x70
    (9100 cells) runs the SOR on global memory."""
    p = E.Params(Nx=Nx, Ny=Ny, periodic_x=per, periodic_y=per)
    U = ic.euler2d_reference_ic(p)
    op = OB.eh_params(Nx=Nx, Ny=Ny, periodic_x=int(per), periodic_y=int(per))
    S = OB.EHArrays(Nx, Ny, U)
    dts_o = np.zeros(3)
    import ctypes as C
    t_o = oracle().or_eh_run(C.byref(op), C.byref(S.cstate()), 3, C.c_double(0.0), C.c_double(5.0),
                             OB.ptr(dts_o))
    with E.Euler2D(p) as s:
        s.upload(U)
        t, dts = s.run(3, t_end=5.0, return_dts=True)
        out = s.download()
        print("sweeps (last, total, solves):", s.sweep_stats())
    assert _rel(dts, dts_o) <= 1e-9 and abs(t - t_o) <= 1e-9 * t_o
    for f in FIELDS:
        assert _rel(out[f], S.a[f]) <= 1e-8, f


def test_bad_dt_is_reported_not_fatal():
    p = E.Params(Nx=16, Ny=16)
    U = ic.euler2d_reference_ic(p)
    U["E"][3, 4] = np.inf  # infinite sound speed -> dt = 0 -> "Bad dt" (EH:699-702)
    with E.Euler2D(p) as s:
        s.upload(U)
        with pytest.raises(E.BadDt):
            s.run(2, t_end=5.0)


def test_mixed_hllc_close_to_fp64():
    p64 = E.Params(Nx=64, Ny=64)
    U = ic.euler2d_reference_ic(p64)
    res = {}
    for prec in (CC_FP64, CC_MIXED):
        with E.Euler2D(E.Params(Nx=64, Ny=64, precision=prec)) as s:
            s.upload(U)
            s.run(10, t_end=5.0)
            res[prec] = s.download()
    a, b = res[CC_MIXED]["rho"], res[CC_FP64]["rho"]
    assert np.abs(a - b).sum() / np.abs(b).sum() <= 1e-4
