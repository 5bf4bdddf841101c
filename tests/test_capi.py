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
