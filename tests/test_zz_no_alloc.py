"""cc_fft2d_noalloc (the allocation-free variant of cc_fft2d) through the C ABI on device memory.

The entry point keeps its twiddle tables in a fixed device array and allocates nothing; the caller
passes device buffers.  Its output must equal cc_fft2d's (host buffers, library-allocated device
memory) bit for bit, on two calls in a row -- the second on a smaller grid, so the tables left in
the array by the first call must not leak into it -- forward and inverse.  Device memory for the
test comes from the HIP runtime that libccamd.so itself loads (hipMalloc / hipMemcpy over ctypes).
"""
import ctypes as C
import os

import numpy as np
import pytest

from core_collapse_amd import _lib

pytestmark = pytest.mark.gpu

H2D, D2H = 1, 2


def _hip():
    _lib.lib()  # loads libccamd.so and with it the HIP runtime it links
    for name in ("libamdhip64.so.7", "libamdhip64.so"):
        try:
            return C.CDLL(os.path.join("/opt/rocm/lib", name))
        except OSError:
            continue
    pytest.fail("HIP runtime library not found next to libccamd.so's")


def _noalloc(Nx, Ny, inverse, dev, stream=None):
    f = _lib.lib().cc_fft2d_noalloc
    f.restype = C.c_int
    f.argtypes = [C.c_int, C.c_int, C.c_int, C.c_void_p, C.c_void_p]
    return f(Nx, Ny, int(inverse), dev, stream)


def _run_noalloc(hip, Nx, Ny, inverse, a):
    """a: complex128 (Ny, Nx); returns the in-place result of cc_fft2d_noalloc on a device copy."""
    nbytes = a.nbytes
    dev = C.c_void_p()
    assert hip.hipMalloc(C.byref(dev), C.c_size_t(nbytes)) == 0
    try:
        host = np.ascontiguousarray(a)
        assert hip.hipMemcpy(dev, host.ctypes.data_as(C.c_void_p), C.c_size_t(nbytes), H2D) == 0
        st = _noalloc(Nx, Ny, inverse, dev)
        assert st == 0, _lib.lib().cc_last_error()
        assert hip.hipDeviceSynchronize() == 0
        out = np.empty_like(host)
        assert hip.hipMemcpy(out.ctypes.data_as(C.c_void_p), dev, C.c_size_t(nbytes), D2H) == 0
        return out
    finally:
        hip.hipFree(dev)


def _run_reference_entry(Nx, Ny, inverse, a):
    out = np.ascontiguousarray(a).copy()
    st = _lib.lib().cc_fft2d(Nx, Ny, int(inverse), out.view(np.float64).ctypes.data_as(_lib.DP))
    assert st == 0, _lib.lib().cc_last_error()
    return out


def test_noalloc_fft_equals_cc_fft2d_on_two_calls_the_second_smaller():
    hip = _hip()
    rng = np.random.default_rng(5)
    for (Nx, Ny) in ((256, 128), (64, 32)):  # the second call on a smaller grid
        for inverse in (False, True):
            a = rng.standard_normal((Ny, Nx)) + 1j * rng.standard_normal((Ny, Nx))
            print("cc_fft2d_noalloc %dx%d inverse=%d" % (Nx, Ny, inverse), flush=True)
            got = _run_noalloc(hip, Nx, Ny, inverse, a)
            want = _run_reference_entry(Nx, Ny, inverse, a)
            assert np.array_equal(got.view(np.float64), want.view(np.float64)), (Nx, Ny, inverse)
    # numpy's FFT as an independent check of the transform itself (the reference's sign and 1/n)
    a = rng.standard_normal((32, 64)) + 1j * rng.standard_normal((32, 64))
    got = _run_noalloc(hip, 64, 32, False, a)
    assert np.abs(got - np.fft.fft2(a)).max() < 1e-10 * np.abs(a).sum()


def test_noalloc_fft_rejects_sizes_beyond_its_scratch_before_device_work():
    _hip()
    assert _noalloc(8192, 8, False, C.c_void_p(16)) == 1   # CC_ERR_INVALID_ARG: line longer than 4096
    assert _noalloc(48, 8, False, C.c_void_p(16)) == 2     # CC_ERR_NON_POW2
    assert _noalloc(64, 8, False, None) == 1               # null data
