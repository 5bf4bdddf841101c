"""The C ABI library loads and exports every entry point include/ccamd.h declares (no GPU calls)."""
import os
import re
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared():
    src = open(os.path.join(ROOT, "include", "ccamd.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(cc_[a-z0-9_]+)\s*\(", src)))


def test_header_declares_the_step_api():
    names = declared()
    for must in ("cc_mhd2d_create", "cc_mhd2d_run", "cc_poisson2d_sor", "cc_mhd3d_create", "cc_mhd3d_run",
                 "cc_last_error"):
        assert must in names


def test_library_exports_every_declared_symbol():
    lib = os.path.join(ROOT, "core_collapse_amd", "libccamd.so")
    if not os.path.exists(lib):
        subprocess.check_call(["make", "-C", os.path.join(ROOT, "core_collapse_amd", "csrc"), "-j8"])
    out = subprocess.check_output(["nm", "-D", "--defined-only", lib]).decode()
    exported = {line.split()[-1] for line in out.splitlines() if " T " in line}
    missing = [n for n in declared() if n not in exported]
    assert not missing, missing


def test_python_binding_prototypes_cover_the_header():
    from core_collapse_amd import _lib
    missing = [n for n in declared() if n not in _lib.exported_symbols()]
    assert not missing, missing


def test_library_loads_and_reports_errors_without_a_device():
    import ctypes as C
    from core_collapse_amd import _lib
    L = _lib.lib()
    assert L.cc_version() >= 10000
    assert isinstance(L.cc_last_error(), bytes)
    h = C.c_void_p()
    assert L.cc_mhd2d_create(None, C.byref(h)) == 1  # CC_ERR_INVALID_ARG before any device call


def test_new_entry_points_reject_bad_arguments_before_touching_a_device():
    """Argument errors of the batch, 3D Euler, SOR-autotune and sampler entry points come back as status
    codes with a message, before any HIP call (so they hold on a machine without a GPU)."""
    import ctypes as C
    import numpy as np
    from core_collapse_amd import _lib
    L = _lib.lib()
    b = C.c_void_p()
    assert L.cc_mhd2d_batch_create(None, 4, C.byref(b)) == 1
    arr = (C.c_void_p * 2)(None, None)
    assert L.cc_mhd2d_batch_create(arr, 2, C.byref(b)) == 1 and b"null handle" in L.cc_last_error()
    assert L.cc_mhd2d_batch_create(arr, 0, C.byref(b)) == 1
    t = np.zeros(1)
    assert L.cc_mhd2d_batch_run(None, 1, t.ctypes.data_as(_lib.DP), t.ctypes.data_as(_lib.DP)) == 1
    assert L.cc_mhd2d_batch_destroy(None) == 0
    h = C.c_void_p()
    for Nx, Ny, Nz in ((24, 16, 16), (16, 4, 16), (16, 16, 2048)):
        p = _lib.Euler3dParams(Nx=Nx, Ny=Ny, Nz=Nz, x1=1.0, y1=1.0, z1=1.0, gamma=5.0 / 3.0, cfl=0.4,
                               poisson_tol=1e-10, poisson_max_iter=8)
        assert L.cc_euler3d_create(C.byref(p), C.byref(h)) == 2, (Nx, Ny, Nz)  # CC_ERR_NON_POW2
    p = _lib.Euler3dParams(Nx=16, Ny=16, Nz=16, x1=1.0, y1=1.0, z1=1.0, gamma=5.0 / 3.0, cfl=0.4,
                           poisson_tol=1e-10, poisson_max_iter=99)
    assert L.cc_euler3d_create(C.byref(p), C.byref(h)) == 1  # max_iter out of [0, 64]
    p.poisson_max_iter, p.gamma = 8, 1.0
    assert L.cc_euler3d_create(C.byref(p), C.byref(h)) == 1  # gamma <= 1
    assert L.cc_euler3d_destroy(None) == 0
    S, WJ, sh = C.c_int(), C.c_int(), C.c_int()
    assert L.cc_sor_tile_choice(128, 128, 600, 1, C.byref(S), C.byref(WJ), C.byref(sh), None) == 0
    assert L.cc_sor_tile_choice(128, 128, 600, 1, None, C.byref(WJ), None, None) == 1
    N = (C.c_int * 1)()
    k = (_lib.Collapse2dIc * 1)()
    e = np.zeros(1)
    assert L.cc_collapse2d_batch_sample(1, 1, 256, 128, N, k, e.ctypes.data_as(_lib.DP)) == 1  # nmax < nmin
