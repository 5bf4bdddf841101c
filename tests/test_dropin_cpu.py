"""The drop-in programs on the CPU (no device needed): their command lines default to the reference
Params, and the initial conditions libccamd builds for them reproduce the reference runs' first
stored frames bit for bit.

The first-frame pins compare SHA-256 digests of the reference programs' own output (their default
runs, tests/golden/dropin.json, made by tests/golden/gen_dropin_golden.py) with the oracle's frame
restatement applied to the C++ IC: EH and NS store the IC itself at t = 0; M3 stores no t = 0
frame, so its IC is advanced one step by the oracle (bit-identical to the reference step) first.
M2's default run is recorded by its log only; its IC is pinned through the reference's log lines.
"""
import ctypes as C
import os
import re
import subprocess

import numpy as np
import pytest

from core_collapse_amd import ic
from oracle import bind as OB
from tests import dropin_ref as D
from tests.helpers import m2_arrays, m2_oparams, m3_arrays, m3_oparams, oracle

# Reference Params defaults (and IC constants) each program must expose as a flag.
REFERENCE_DEFAULTS = {
    "m2": {  # M2:44-86, gamma_eff M2:1011
        "gamma": 5.0 / 3.0, "cfl": 0.45, "x0": 0.0, "x1": 1.0, "y0": 0.0, "y1": 1.0, "Nx": 192, "Ny": 192,
        "t_end": 0.35, "dt_store": 1e-3, "floor_rho": 1e-10, "floor_P": 1e-10, "periodic_x": True,
        "periodic_y": True, "eta": 5e-4, "gravity_on": True, "G": 8.0, "sor_iters": 600, "sor_omega": 1.8,
        "rho_bg": 1.0, "rho_core": 20.0, "r_core": 0.12, "r_edge": 0.02, "P0": 0.15, "vin": 1.25, "rot_omega": 1.0,
        "B0x": 0.15, "Bshear": 0.05, "gamma_eff": 1.05},
    "m3": {  # M3:43-89
        "gamma1": 1.30, "gamma2": 2.50, "K1": 0.05, "rho_nuc": 80.0, "gamma_th": 1.50, "rho_bounce": 120.0,
        "gamma_bounce": 3.00, "Kbounce": 0.02, "gamma_init": 5.0 / 3.0, "cfl": 0.4, "x0": 0.0, "x1": 1.0,
        "y0": 0.0, "y1": 1.0, "z0": 0.0, "z1": 1.0, "Nx": 48, "Ny": 48, "Nz": 48, "t_end": 0.15, "dt_store": 1e-3,
        "floor_rho": 1e-10, "floor_P": 1e-10, "periodic_x": True, "periodic_y": True, "periodic_z": True,
        "eta": 5e-4, "rho_bg": 1.0, "rho_core": 20.0, "r_core": 0.12, "r_edge": 0.02, "P0": 0.15, "vin": 1.25,
        "rot_omega": 1.0, "B0x": 0.15, "B0z": 0.05, "Bshear": 0.05},
    "eh": {  # EH:29-56, SOR constants EH:329-331, IC constants EH:585-603
        "gamma": 1.2, "cfl": 0.45, "x0": 0.0, "x1": 1.0, "y0": 0.0, "y1": 1.0, "Nx": 50, "Ny": 50, "t_end": 5.0,
        "dt_store": 1e-3, "floor_rho": 1e-12, "floor_P": 1e-12, "periodic_x": True, "periodic_y": True,
        "gravity_on": True, "G": 1.0, "max_iter": 50000, "tol": 1e-10, "omega": 1.7, "seed": 42, "rho0": 1.0,
        "P0": 1.0, "amp_rho": 0.10, "amp_P": 0.10, "vmax": 0.05, "sigma": 0.08, "bump_amp": 0.5},
    "ns": {  # NS:54-86, IC constants NS:642-660
        "gamma": 1.4, "cfl": 0.35, "x0": 0.0, "x1": 1.0, "y0": 0.0, "y1": 1.0, "Nx": 128, "Ny": 128,
        "t_end": 10.0, "dt_store": 5e-3, "seed": 42, "floor_rho": 1e-10, "floor_P": 1e-10, "viscosity_on": True,
        "mu": 1e-6, "cfl_visc": 0.05, "gravity_on": True, "G": 10.0, "out_h5": "out_ns2d_poisson.h5",
        "rho0": 1.0, "P0": 0.05, "A": 1.0, "sigma_frac": 0.10, "Arho": 0.01, "Mach0": 0.05},
}


def _program(tag):
    path = os.path.join(D.BIN, D.PROGRAM[tag])
    if not os.path.exists(path):
        pytest.skip("drop-in programs not built (make -C core_collapse_amd/programs)")
    return path


def _help_defaults(tag):
    out = subprocess.run([_program(tag), "--help"], capture_output=True, text=True, timeout=60)
    assert out.returncode == 0, out.stderr
    got = {}
    for line in out.stdout.splitlines():
        m = re.match(r"\s+--(\S+)\s.*\[(.*)\]$", line)
        if m:
            got[m.group(1)] = m.group(2)
    return got


@pytest.mark.parametrize("tag", ["m2", "m3", "eh", "ns"])
def test_flags_default_to_reference_params(tag):
    got = _help_defaults(tag)
    for name, want in REFERENCE_DEFAULTS[tag].items():
        assert name in got, "missing flag --%s" % name
        v = got[name]
        if isinstance(want, bool):
            assert v == ("true" if want else "false"), (name, v)
        elif isinstance(want, str):
            assert v == want, (name, v)
        elif isinstance(want, int):
            assert int(v) == want, (name, v)
        else:
            assert float(v) == want, (name, v)


@pytest.mark.parametrize("tag", ["m2", "m3", "eh", "ns"])
def test_bad_flags_are_rejected(tag):
    prog = _program(tag)
    for args in (["--no_such_flag", "1"], ["--Nx"], ["--Nx", "abc"], ["--precision", "fp16"]):
        r = subprocess.run([prog] + args, capture_output=True, text=True, timeout=60)
        assert r.returncode == 2, (args, r.returncode, r.stdout, r.stderr)


def test_ns_rejects_non_power_of_two_like_the_reference():
    r = subprocess.run([_program("ns"), "--Nx", "48"], capture_output=True, text=True, timeout=60)
    assert r.returncode == 1
    assert r.stderr.strip() == "Nx and Ny must be powers of two for the FFT Poisson solver."  # NS:614-617


def _check_first_frame(tag, frame):
    meta, _ = D.fixture(tag)
    for name in D.FIELDS[tag]:
        assert D.digest(frame[name]) == meta["first_sha256"][name], name


def test_m2_reference_ic_reproduces_the_reference_log():
    """M2's default run is recorded by its log only (tests/golden/gen_dropin_golden.py says why),
    so the IC is pinned by the reference's log instead: the
    oracle (bit-identical to the reference step) started from the C++ IC prints the same first
    lines, digit for digit -- including max|divB| at 12 digits, round-off noise that depends on
    every bit of the face field."""
    from core_collapse_amd.mhd2d import Params
    meta, _ = D.fixture("m2")
    ref = meta["stdout"]
    p = Params()
    op = m2_oparams(p)
    S = m2_arrays(p, ic.mhd2d_reference_ic(p))
    fr0 = D.frame_m2(op, S, p.dx, p.dy)
    assert ref[0] == "init max|divB|=%s" % D._fmt(np.abs(fr0["/divB"]).max(), 12)
    times, frames, log = D.run_m2(op, S, 0.0035, 1e-3)  # the first two stored frames (and one at t_end)
    assert log[:2] == ref[1:3]


def test_eh_reference_ic_first_frame_bitwise():
    from core_collapse_amd.euler2d import Params
    p = Params()
    op = OB.eh_params()
    S = OB.EHArrays(p.Nx, p.Ny, {k: v.reshape(-1) for k, v in ic.euler2d_reference_ic(p).items()})
    _check_first_frame("eh", D.frame_eh(op, S))


def test_ns_reference_ic_first_frame_bitwise():
    from core_collapse_amd.ns2d import Params
    p = Params()
    U = ic.ns2d_reference_ic(p).reshape(-1).copy()
    _check_first_frame("ns", D.frame_ns(OB.ns_params(), U))


def test_m3_reference_ic_first_stored_frame_bitwise():
    from core_collapse_amd.mhd3d import Params
    p = Params()
    op = m3_oparams(p)
    S = m3_arrays(p, ic.mhd3d_reference_state(p))
    _, arrays = D.fixture("m3")
    dts = np.zeros(1)
    t = oracle().or_m3_run(C.byref(op), C.byref(S.cstate()), C.c_double(p.dx), C.c_double(p.dy),
                           C.c_double(p.dz), 1, C.c_double(0.0), C.c_double(0.15), OB.ptr(dts))
    assert t == arrays["t"][0]
    _check_first_frame("m3", D.frame_m3(op, S, p.dx, p.dy, p.dz))


@pytest.mark.parametrize("tag", ["m3", "eh", "ns"])
def test_fixture_layout_is_the_reference_layout(tag):
    """The recorded reference files: F64LE contiguous datasets, coordinates, (Nt, ...) fields."""
    meta, arrays = D.fixture(tag)
    d = REFERENCE_DEFAULTS[tag]
    nt = len(arrays["t"])
    frame = [d["Nz"], d["Ny"], d["Nx"]] if tag == "m3" else [d["Ny"], d["Nx"]]
    want = {"/t": [nt], "/x": [d["Nx"]], "/y": [d["Ny"]]}
    if tag == "m3":
        want["/z"] = [d["Nz"]]
    for f in D.FIELDS[tag]:
        want[f] = [nt] + frame
    assert {k: v["shape"] for k, v in meta["layout"].items()} == want
    assert all(v["f64le"] and v["layout"] == "CONTIGUOUS" for v in meta["layout"].values())
    x = arrays["x"]
    assert np.array_equal(x, d["x0"] + (np.arange(d["Nx"]) + 0.5) * ((d["x1"] - d["x0"]) / d["Nx"]))
