"""2D Navier-Stokes HLLC + FFT self-gravity step (navier_stokes_2d_hllc_poisson.cpp, cited NS:<line>).

``Params`` mirrors struct Params (NS:54-86, hot-path subset) with the same defaults plus the
precision.  ``NS2D`` exposes solve_poisson_periodic_fft (NS:285) as poisson(), compute_rhs (NS:363),
compute_dt (NS:521), the Heun update of the driver loop (NS:758-781) as heun_step() and the loop as
run().  The state is the reference's AoS ``std::vector<U4>`` (NS:88-93): float64 array (Ny, Nx, 4)
with components {rho, mx, my, E}.  Grids must be powers of two (NS:614-617).
"""
import ctypes as C
from dataclasses import dataclass, fields

import numpy as np

from . import _lib
from ._lib import CC_FP64, DP, check


@dataclass
class Params:
    gamma: float = 1.4
    cfl: float = 0.35
    x0: float = 0.0
    x1: float = 1.0
    y0: float = 0.0
    y1: float = 1.0
    Nx: int = 128
    Ny: int = 128
    floor_rho: float = 1e-10
    floor_P: float = 1e-10
    viscosity_on: bool = True
    mu: float = 1e-6
    cfl_visc: float = 0.05
    gravity_on: bool = True
    G: float = 10.0
    precision: int = CC_FP64

    @property
    def dx(self):
        return (self.x1 - self.x0) / self.Nx

    @property
    def dy(self):
        return (self.y1 - self.y0) / self.Ny

    def c_struct(self):
        return _lib.Ns2dParams(**{f.name: (int(getattr(self, f.name)) if isinstance(getattr(self, f.name), bool)
                                           else getattr(self, f.name)) for f in fields(self)})


def _p(a):
    return a.ctypes.data_as(DP)


class NS2D:
    def __init__(self, params: Params):
        self.params = params
        self._h = C.c_void_p()
        self._ps = params.c_struct()
        check(_lib.lib().cc_ns2d_create(C.byref(self._ps), C.byref(self._h)))

    def close(self):
        if self._h:
            _lib.lib().cc_ns2d_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    @property
    def shape(self):
        return (self.params.Ny, self.params.Nx, 4)

    def upload(self, U):
        U = np.ascontiguousarray(U, dtype=np.float64)
        if U.size != 4 * self.params.Nx * self.params.Ny:
            raise ValueError("state has %d values, expected %s" % (U.size, self.shape))
        check(_lib.lib().cc_ns2d_upload(self._h, _p(U)))

    def download(self):
        U = np.empty(self.shape)
        check(_lib.lib().cc_ns2d_download(self._h, _p(U)))
        return U

    def poisson(self, rho):
        rho = np.ascontiguousarray(rho, dtype=np.float64)
        phi = np.empty((self.params.Ny, self.params.Nx))
        check(_lib.lib().cc_ns2d_poisson(self._h, _p(rho), _p(phi)))
        return phi

    def compute_rhs(self):
        dU = np.empty(self.shape)
        check(_lib.lib().cc_ns2d_rhs(self._h, _p(dU)))
        return dU

    def compute_dt(self):
        dt = C.c_double()
        check(_lib.lib().cc_ns2d_compute_dt(self._h, C.byref(dt)))
        return dt.value

    def heun_step(self, dt):
        check(_lib.lib().cc_ns2d_heun_step(self._h, float(dt)))

    def run(self, nsteps, t=0.0, t_end=1e300, return_dts=False):
        tt = C.c_double(t)
        dts = np.zeros(max(nsteps, 1))
        check(_lib.lib().cc_ns2d_run(self._h, int(nsteps), C.byref(tt), float(t_end), _p(dts)))
        return (tt.value, dts[:nsteps]) if return_dts else tt.value

    def sync(self):
        check(_lib.lib().cc_ns2d_sync(self._h))

    # -- snapshot frame and in-situ diagnostics ------------------------------------------------
    def frame(self):
        """The fields one snapshot stores, computed on the device: {name: array}."""
        return _lib.frame_of("ns2d", self._h, (self.params.Ny, self.params.Nx))

    def totals(self):
        """Cell-volume-weighted totals of the conserved variables (and max|divB| for MHD)."""
        return _lib.totals_of("ns2d", self._h)

    def compare(self, other):
        """Per frame field (mean|a-b|, sum|a-b|/sum|b|, max|a-b|) of this handle against `other`."""
        return _lib.compare_of("ns2d", self._h, other._h)

    def set_profiling(self, on=True):
        check(_lib.lib().cc_ns2d_set_profiling(self._h, int(bool(on))))

    def reset_profiling(self):
        check(_lib.lib().cc_ns2d_reset_profiling(self._h))

    def kernel_time(self, name):
        n = C.c_int()
        ms = C.c_double()
        check(_lib.lib().cc_ns2d_kernel_time(self._h, name.encode(), C.byref(n), C.byref(ms)))
        return n.value, ms.value


def fft2d(data, inverse=False):
    """fft2d (NS:182-195) on the device; data is complex128 (Ny, Nx); returns a new array."""
    a = np.ascontiguousarray(data, dtype=np.complex128).copy()
    Ny, Nx = a.shape
    check(_lib.lib().cc_fft2d(Nx, Ny, int(bool(inverse)), a.view(np.float64).ctypes.data_as(DP)))
    return a
