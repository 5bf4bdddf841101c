"""Sanitizer builds of the host code, run on the CPU (tests/san/Makefile builds into a temporary folder):

* the oracle (oracle/*.c) with AddressSanitizer + UndefinedBehaviorSanitizer, errors aborting, under the
  oracle's own test files (the oracle is loaded through CC_ORACLE_LIB, the ASan runtime preloaded);
* the host-only sources of libccamd -- initial conditions (ic.cpp), the SOR tile schedules and their
  dataflow dependencies (sorsched.cpp), error handling (ccamd.cpp) -- with ASan + UBSan, driven over grid
  sizes and tile shapes by tests/san/host_driver.cpp;
* the drop-in programs' snapshot writer (a producer thread and its writer thread) under ThreadSanitizer.
"""
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SAN = os.path.join(ROOT, "tests", "san")


def _lib(name):
    return subprocess.check_output(["gcc", "-print-file-name=" + name], text=True).strip()


@pytest.fixture(scope="module")
def out(tmp_path_factory):
    if not os.path.exists("/opt/rocm/bin/hipcc") or not os.path.exists("/opt/conda/include/hdf5.h"):
        pytest.skip("hipcc or the HDF5 headers are not available")
    d = str(tmp_path_factory.mktemp("san"))
    r = subprocess.run(["make", "-j8", "-C", SAN, "OUT=" + d], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-3000:]
    return d


def test_host_code_under_asan_ubsan(out):
    r = subprocess.run([os.path.join(out, "host_san")], capture_output=True, text=True, timeout=600,
                       env=dict(os.environ, ASAN_OPTIONS="detect_leaks=0"))
    assert r.returncode == 0 and "ok" in r.stdout, (r.stdout[-2000:], r.stderr[-4000:])


def test_snapshot_writer_under_tsan(out, tmp_path):
    r = subprocess.run([os.path.join(out, "snapshot_tsan"), str(tmp_path)], capture_output=True, text=True,
                       timeout=300)
    assert r.returncode == 0 and "ok" in r.stdout and "WARNING: ThreadSanitizer" not in r.stderr, r.stderr[-4000:]


def test_oracle_under_asan_ubsan(out):
    files = [os.path.join(ROOT, "tests", f) for f in ("test_oracle_mhd2d.py", "test_oracle_mhd3d.py",
                                                      "test_oracle_euler2d.py", "test_oracle_euler3d.py",
                                                      "test_oracle_ns2d.py", "test_oracle_robust.py")]
    env = dict(os.environ, LD_PRELOAD=_lib("libasan.so"), ASAN_OPTIONS="detect_leaks=0",
               CC_ORACLE_LIB=os.path.join(out, "liboracle_san.so"))
    r = subprocess.run([sys.executable, "-m", "pytest", "-q", "-x", "-p", "no:cacheprovider", "-m", "not gpu"] + files,
                       capture_output=True, text=True, timeout=1200, env=env, cwd=ROOT)
    assert r.returncode == 0, (r.stdout[-3000:], r.stderr[-3000:])
