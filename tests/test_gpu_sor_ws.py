"""GPU parity of the wave-streamed SOR (core_collapse_amd/csrc/sorws.hip) on clamped grids.

The solve is the lexicographic iterate of compute_poisson_gravity_periodic (M2:178-237).  FP64: against the
oracle at 1e-9 relative (FMA contraction and the deferred mean removal are the only differences) and
against the LDS-tiled schedule (CC_SOR_WS=0) at 1e-12; mixed (FP32 sweeps of a per-tile correction on an
FP64 base): against the oracle at the 1e-5 bar of the other mixed SOR schedules.  Both compiled tile
heights (4 and 8 sweeps per tile) run; grids cover interior and edge tiles, partial sweep blocks, grids
narrower than a tile and one-sweep solves.
"""
import ctypes as C

import numpy as np
import pytest

from core_collapse_amd import CC_FP64, CC_MIXED
from core_collapse_amd import ic
from core_collapse_amd import mhd2d as M
from oracle import bind as OB
from tests.helpers import m2_arrays, m2_dict, m2_oparams, oracle, rel_err, rel_l1

pytestmark = pytest.mark.gpu

GRIDS = [(32, 32, 600), (72, 40, 600), (130, 67, 37), (2, 2, 7), (3, 2, 40), (2, 97, 33), (97, 2, 33),
         (5, 17, 600), (65, 3, 1), (129, 130, 17), (300, 170, 61), (64, 64, 9)]


def _rho(Nx, Ny, seed):
    rng = np.random.default_rng(seed)
    return 1.0 + rng.uniform(0, 1, (Ny, Nx)) + 5.0 * np.exp(-((np.arange(Nx)[None, :] - Nx / 2) ** 2 +
                                                           (np.arange(Ny)[:, None] - Ny / 3) ** 2) / 20.0)


def _oracle(Nx, Ny, iters, rho):
    p = M.Params(Nx=Nx, Ny=Ny, periodic_x=False, periodic_y=False, sor_iters=iters, G=1.0)
    gx, gy = np.zeros(rho.size), np.zeros(rho.size)
    oracle().or_m2_poisson(C.byref(m2_oparams(p)), OB.ptr(np.ascontiguousarray(rho.reshape(-1))), C.c_double(p.dx),
                           C.c_double(p.dy), OB.ptr(gx), OB.ptr(gy))
    return p, gx.reshape(rho.shape), gy.reshape(rho.shape)


@pytest.mark.parametrize("SW", [4, 8])
@pytest.mark.parametrize("prec", [CC_FP64, CC_MIXED])
def test_wave_streamed_sor_matches_oracle(prec, SW, monkeypatch):
    monkeypatch.setenv("CC_SOR_WS", "1")
    monkeypatch.setenv("CC_SOR_WS_SW", str(SW))
    tol = 1e-9 if prec == CC_FP64 else 1e-5
    for Nx, Ny, iters in GRIDS:
        print("grid", Nx, Ny, iters, "precision", prec, "SW", SW, flush=True)
        rho = _rho(Nx, Ny, Nx * 1000 + Ny)
        p, ox, oy = _oracle(Nx, Ny, iters, rho)
        gx, gy = M.poisson_sor(rho, p.dx, p.dy, p.G, iters, p.sor_omega, False, False, precision=prec)
        assert rel_err(gx, ox) < tol and rel_err(gy, oy) < tol, (Nx, Ny, iters, rel_err(gx, ox), rel_err(gy, oy))


def test_wave_streamed_fp64_equals_lds_tiles(monkeypatch):
    for Nx, Ny, iters in ((300, 170, 600), (130, 67, 37), (257, 300, 100)):
        print("grid", Nx, Ny, iters, flush=True)
        rho = _rho(Nx, Ny, 3 * Nx + Ny)
        out = {}
        for ws in ("1", "0"):
            monkeypatch.setenv("CC_SOR_WS", ws)
            out[ws] = M.poisson_sor(rho, 1.0 / Nx, 1.0 / Ny, 1.0, iters, 1.8, False, False)
        for a, b in zip(out["1"], out["0"]):
            assert rel_err(a, b) < 1e-12, rel_err(a, b)


@pytest.mark.parametrize("prec", [CC_FP64, CC_MIXED])
def test_wave_streamed_solves_repeat_bit_for_bit(prec, monkeypatch):
    """The schedule's result does not depend on which waves run which tiles or when: repeated solves,
    and solves on a grid of 1 and 2 waves per CU (CC_SOR_WS_WAVES), give the same bits."""
    monkeypatch.setenv("CC_SOR_WS", "1")
    Nx, Ny, iters = 520, 390, 200
    rho = _rho(Nx, Ny, 11)
    ref = M.poisson_sor(rho, 1.0 / Nx, 1.0 / Ny, 1.0, iters, 1.8, False, False, precision=prec)
    for waves in ("8", "1", "2"):
        monkeypatch.setenv("CC_SOR_WS_WAVES", waves)
        got = M.poisson_sor(rho, 1.0 / Nx, 1.0 / Ny, 1.0, iters, 1.8, False, False, precision=prec)
        assert np.array_equal(got[0], ref[0]) and np.array_equal(got[1], ref[1]), waves


@pytest.mark.parametrize("prec", [CC_FP64, CC_MIXED])
def test_collapse_steps_with_wave_streamed_sor_match_oracle(prec, monkeypatch):
    """Collapse steps through the captured step graph (two wave-streamed solves per Heun step)."""
    monkeypatch.setenv("CC_SOR_WS", "1")
    N, nsteps = 96, 3
    p = ic.collapse2d_params(N, precision=prec)
    st = ic.collapse2d_ic(p)
    with M.Mhd2D(p) as h:
        h.upload(st)
        t = h.run(nsteps)
        got = h.download()
    S_ = m2_arrays(p, st)
    ot = oracle().or_m2_run(C.byref(m2_oparams(p)), C.byref(S_.cstate()), C.c_double(p.dx), C.c_double(p.dy), nsteps,
                            C.c_double(0.0), C.c_double(1e300), None)
    want = m2_dict(p, S_)
    if prec == CC_FP64:
        assert abs(t - ot) < 1e-11 * ot
        for f in M.FIELDS:
            assert rel_err(got[f], want[f]) < 1e-9, f
    else:
        assert rel_l1(got["rho"], want["rho"]) <= 1e-4


def test_batch_with_wave_streamed_sor_equals_instances_run_alone(monkeypatch):
    """cc_mhd2d_batch_run with wave-streamed solves: the instances' SOR launches stay separate (one branch
    of the batch graph each, never merged), and three batch steps give bit for bit the times and states
    of the same instances stepped alone."""
    from core_collapse_amd.mhd2d import Mhd2D, Mhd2DBatch
    monkeypatch.setenv("CC_SOR_WS", "1")
    sample = ic.collapse2d_batch_sample(4, seed=5, nmin=40, nmax=140)

    def make():
        sims = []
        for e in sample:
            p = ic.collapse2d_params(e["N"], eta=e["eta"], precision=CC_MIXED, sor_iters=120)
            h = Mhd2D(p)
            h.upload(ic.collapse2d_ic(p, **e["ic"]))
            sims.append(h)
        return sims

    alone = make()
    t_alone = [h.run(3) for h in alone]
    sims = make()
    b = Mhd2DBatch(sims)
    assert not b.info()["merged"]
    t = b.run(3, None, np.full(len(sample), 1e300))
    b.close()
    assert list(t) == t_alone, (t, t_alone)
    for h1, h2 in zip(alone, sims):
        s1, s2 = h1.download(), h2.download()
        for f in M.FIELDS:
            assert np.array_equal(s1[f], s2[f]), f
        h1.close()
        h2.close()
