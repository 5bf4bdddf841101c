"""GPU parity of the one-pass edge EMFs + CT of the 3D MHD stage (k_ect3d, z-marching tiles) on grids that
span several tiles and z chunks, against the CPU oracle (mhd3d_resistive_hlld_ct.cpp restated,
M3:694-1066) and against the split passes k_edge3d + k_ct3d of the same library; and of the column pass that
solves the z faces and updates the cells (k_zupd3d, CC_M3_ZUPD) against the split k_face3d<2> + k_update3d.

Tiles are 32 x 28, 32 x 12 or 64 x 5 cells (CC_M3_ECT_SHAPE 0, 1 and 3, 2), so a 70 x 53 x 11 grid has partial
tiles on both axes; CC_M3_ZCHUNK = 3 splits z into four chunks (and CC_M3_JCHUNK the column-marching Joule
pass k_joule3d likewise), so chunk seams, the index maps at both z ends and the faces of index Nx, Ny, Nz are
all exercised, with periodic and clamped boundaries.  FP64 bars allow
FMA-contraction rounding (1e-11 for a stage, 1e-10 for a Heun step); one pass against split passes on the
same input: 1e-12."""
import ctypes as C

import numpy as np
import pytest

from core_collapse_amd import CC_FP64, CC_MIXED
from core_collapse_amd import mhd3d as M
from tests.helpers import m3_arrays, m3_dict, m3_oparams, oracle, random_mhd3d_state, rel_err

pytestmark = pytest.mark.gpu

GRID = (70, 53, 11)
BCS = {"periodic": (True, True, True), "clamped": (False, False, False), "mixed": (True, False, True)}


def _params(eos, bc, comp, precision=CC_FP64, eta=2e-3):
    px, py, pz = BCS[bc]
    return M.Params(Nx=GRID[0], Ny=GRID[1], Nz=GRID[2], eos=eos, periodic_x=px, periodic_y=py, periodic_z=pz,
                    eta=eta, ct_compensated=comp, precision=precision)


@pytest.mark.parametrize("shape,zchunk", [(0, 3), (0, 64), (1, 3), (2, 5), (3, 4)])
@pytest.mark.parametrize("eos,bc,comp", [(1, "periodic", False), (1, "periodic", True), (0, "clamped", False),
                                         (1, "mixed", True), (0, "periodic", True)])
def test_fused_stage_and_step_match_oracle(monkeypatch, eos, bc, comp, shape, zchunk):
    monkeypatch.setenv("CC_M3_FUSED", "1")
    monkeypatch.setenv("CC_M3_ECT_SHAPE", str(shape))
    monkeypatch.setenv("CC_M3_ZCHUNK", str(zchunk))
    monkeypatch.setenv("CC_M3_JCHUNK", str(zchunk))  # the column-marching Joule pass, same chunks
    p = _params(eos, bc, comp)
    st = random_mhd3d_state(p, 11 + eos)
    dt = 1e-3
    o = oracle()
    op = m3_oparams(p)
    with M.Mhd3D(p) as h:  # one stage (advance_stage) ...
        h.upload(st)
        h.advance_stage(dt)
        got = h.download()
    S0 = m3_arrays(p, st)
    S1 = S0.copy()
    o.or_m3_advance_stage(C.byref(op), C.byref(S0.cstate()), C.byref(S1.cstate()), C.c_double(p.dx),
                          C.c_double(p.dy), C.c_double(p.dz), C.c_double(dt))
    want = m3_dict(p, S1)
    for f in M.FIELDS:
        assert rel_err(got[f], want[f]) < 1e-11, ("stage", f, rel_err(got[f], want[f]))
    with M.Mhd3D(p) as h:  # ... and one Heun step (rk2_step), whose second stage averages
        h.upload(st)
        h.rk2_step(dt)
        got = h.download()
    S = m3_arrays(p, st)
    o.or_m3_rk2_step(C.byref(op), C.byref(S.cstate()), C.c_double(p.dx), C.c_double(p.dy), C.c_double(p.dz),
                     C.c_double(dt))
    want = m3_dict(p, S)
    for f in M.FIELDS:
        assert rel_err(got[f], want[f]) < 1e-10, ("rk2", f, rel_err(got[f], want[f]))


@pytest.mark.parametrize("precision", [CC_FP64, CC_MIXED])
def test_fused_run_equals_split_passes(monkeypatch, precision):
    """Four driver steps (device dt) with the fused pass and with the split passes from the same state."""
    p = _params(1, "periodic", True, precision=precision, eta=1e-3)
    st = random_mhd3d_state(p, 5)
    out = {}
    for fused in (1, 0):
        monkeypatch.setenv("CC_M3_FUSED", str(fused))
        monkeypatch.setenv("CC_M3_ZCHUNK", "4")
        with M.Mhd3D(p) as h:
            h.upload(st)
            t, dts = h.run(4, return_dts=True)
            out[fused] = (h.download(), dts, h.divB_max())
    (a, da, ba), (b, db, bb) = out[1], out[0]
    assert rel_err(da, db) < 1e-12, (da, db)
    for f in M.FIELDS:
        assert rel_err(a[f], b[f]) < 1e-12, (f, rel_err(a[f], b[f]))
    assert ba < 1e-12 and bb < 1e-12, (ba, bb)


@pytest.mark.parametrize("grid", [(2, 3, 2), (5, 2, 7), (3, 33, 4)])
@pytest.mark.parametrize("bc", ["periodic", "clamped"])
def test_fused_passes_on_tiny_grids_match_oracle(monkeypatch, grid, bc):
    """Grids of 2..7 cells on an axis: every tile is mostly halo, and the index maps wrap several times."""
    monkeypatch.setenv("CC_M3_FUSED", "1")
    monkeypatch.setenv("CC_M3_ZCHUNK", "3")
    monkeypatch.setenv("CC_M3_JCHUNK", "2")
    px, py, pz = BCS[bc]
    p = M.Params(Nx=grid[0], Ny=grid[1], Nz=grid[2], eos=1, periodic_x=px, periodic_y=py, periodic_z=pz, eta=2e-3,
                 ct_compensated=True)
    st = random_mhd3d_state(p, 3)
    dt = 1e-3
    with M.Mhd3D(p) as h:
        h.upload(st)
        h.rk2_step(dt)
        got = h.download()
    S = m3_arrays(p, st)
    oracle().or_m3_rk2_step(C.byref(m3_oparams(p)), C.byref(S.cstate()), C.c_double(p.dx), C.c_double(p.dy),
                            C.c_double(p.dz), C.c_double(dt))
    want = m3_dict(p, S)
    for f in M.FIELDS:
        assert rel_err(got[f], want[f]) < 1e-10, (f, rel_err(got[f], want[f]))


ZUPD_GRIDS = [(70, 53, 11), (2, 3, 2), (5, 2, 7), (3, 33, 4)]


@pytest.mark.parametrize("grid", ZUPD_GRIDS)
@pytest.mark.parametrize("bc", ["periodic", "clamped", "mixed"])
@pytest.mark.parametrize("eos", [1, 0])
@pytest.mark.parametrize("precision", [CC_FP64, CC_MIXED])
def test_zface_update_pass_matches_split_passes(monkeypatch, grid, bc, eos, precision):
    """One stage and one Heun step with the z faces + update in one column march (z chunks of 3 planes, so the
    chunk seams and the two-plane start of the march are exercised) against the split passes.  FP64: the same
    arithmetic, up to FMA contraction (1e-13).  Mixed: the FP32 solver of the two kernels may round a face flux
    one ulp apart, which moves a stage by ~1e-9 of a field (measured <= 7.2e-9 after a Heun step); 2e-8 bounds
    it.  An index or chunk-seam error moves a field by ~1e-3."""
    monkeypatch.setenv("CC_M3_ZUCHUNK", "3")
    px, py, pz = BCS[bc]
    p = M.Params(Nx=grid[0], Ny=grid[1], Nz=grid[2], eos=eos, periodic_x=px, periodic_y=py, periodic_z=pz, eta=2e-3,
                 ct_compensated=True, precision=precision, hlld_guard=1e-2)
    st = random_mhd3d_state(p, 3 + eos)
    tol = 1e-13 if precision == CC_FP64 else 2e-8
    for what in ("stage", "rk2"):
        out = {}
        for zupd in (1, 0):
            monkeypatch.setenv("CC_M3_ZUPD", str(zupd))
            with M.Mhd3D(p) as h:
                h.upload(st)
                h.advance_stage(1e-3) if what == "stage" else h.rk2_step(1e-3)
                out[zupd] = h.download()
        for f in M.FIELDS:
            e = rel_err(out[1][f], out[0][f])
            assert e <= tol, (what, f, e)


@pytest.mark.parametrize("eos,periodic", [(0, True), (1, False)])
def test_zface_update_pass_deferred_cells_match_oracle(monkeypatch, eos, periodic):
    """Mixed, with degen_tau = 2 and no HLLD guard: the FP32 solver flags every subsonic face, so nearly every
    cell of the column pass is deferred (its two z-face fluxes stored, the cell queued) and redone after the
    FP64 pass over the queued z faces; the stage then equals the FP64 oracle stage up to the FP32 rounding of
    the stored fluxes (~1e-9 here, bar 1e-7).  z chunks of 3 planes, 12 x 10 x 9 cells."""
    monkeypatch.setenv("CC_M3_ZUPD", "1")
    monkeypatch.setenv("CC_M3_ZUCHUNK", "3")
    p = M.Params(Nx=12, Ny=10, Nz=9, eos=eos, periodic_x=periodic, periodic_y=periodic, periodic_z=periodic,
                 eta=2e-3, precision=CC_MIXED, degen_tau=2.0)
    st = random_mhd3d_state(p, 5)
    dt = 1e-3
    with M.Mhd3D(p) as h:
        h.upload(st)
        h.advance_stage(dt)
        got = h.download()
    S0 = m3_arrays(p, st)
    S1 = S0.copy()
    oracle().or_m3_advance_stage(C.byref(m3_oparams(p)), C.byref(S0.cstate()), C.byref(S1.cstate()),
                                 C.c_double(p.dx), C.c_double(p.dy), C.c_double(p.dz), C.c_double(dt))
    want = m3_dict(p, S1)
    for f in M.FIELDS:
        e = rel_err(got[f], want[f])
        assert e < 1e-7, (f, e)
