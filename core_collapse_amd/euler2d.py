"""2D Euler HLLC + SOR self-gravity step (euler2d_gravity_hllc.cpp, cited EH:<line>).

``Params`` mirrors struct Params (EH:29-56, hot-path subset) with the same defaults, plus the SOR
constants of EH:329-331 and the precision.  ``Euler2D`` exposes compute_poisson_gravity_periodic
(EH:288) as gravity(), compute_dt (EH:390), rhs (EH:406), rk2_step (EH:529) and the driver loop
(EH:697-708) as run(); a non-finite or non-positive dt raises ``BadDt`` where the program prints
"Bad dt" and returns 1.  State: SoA float64 arrays of shape (Ny, Nx).
"""
import ctypes as C
from dataclasses import dataclass, fields

import numpy as np

from . import _lib
from ._lib import CC_FP64, DP, CCError, check

FIELDS = ("rho", "mx", "my", "E")


class BadDt(CCError):
    pass


@dataclass
class Params:
    gamma: float = 1.2
    cfl: float = 0.45
    x0: float = 0.0
    x1: float = 1.0
    y0: float = 0.0
    y1: float = 1.0
    Nx: int = 50
    Ny: int = 50
    floor_rho: float = 1e-12
    floor_P: float = 1e-12
    periodic_x: bool = True
    periodic_y: bool = True
    gravity_on: bool = True
    G: float = 1.0
    max_iter: int = 50000
    tol: float = 1e-10
    omega: float = 1.7
    precision: int = CC_FP64

    @property
    def dx(self):
        return (self.x1 - self.x0) / self.Nx

    @property
    def dy(self):
        return (self.y1 - self.y0) / self.Ny

    def c_struct(self):
        return _lib.Euler2dParams(**{f.name: (int(getattr(self, f.name)) if isinstance(getattr(self, f.name), bool)
                                              else getattr(self, f.name)) for f in fields(self)})


def _p(a):
    return a.ctypes.data_as(DP)


class Euler2D:
    def __init__(self, params: Params):
        self.params = params
        self._h = C.c_void_p()
        self._ps = params.c_struct()
        check(_lib.lib().cc_euler2d_create(C.byref(self._ps), C.byref(self._h)))

    def close(self):
        if self._h:
            _lib.lib().cc_euler2d_destroy(self._h)
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
        return (self.params.Ny, self.params.Nx)

    def upload(self, state):
        arrs = []
        for f in FIELDS:
            a = np.ascontiguousarray(state[f], dtype=np.float64)
            if a.size != self.params.Nx * self.params.Ny:
                raise ValueError("field %s has %d values, expected %s" % (f, a.size, self.shape))
            arrs.append(a)
        check(_lib.lib().cc_euler2d_upload(self._h, *[_p(a) for a in arrs]))

    def download(self):
        out = {f: np.empty(self.shape) for f in FIELDS}
        check(_lib.lib().cc_euler2d_download(self._h, *[_p(out[f]) for f in FIELDS]))
        return out

    def gravity(self):
        """(gx, gy, phi, sweeps) of the current state (EH:288-386)."""
        gx, gy, phi = (np.empty(self.shape) for _ in range(3))
        n = C.c_int()
        check(_lib.lib().cc_euler2d_gravity(self._h, _p(gx), _p(gy), _p(phi), C.byref(n)))
        return gx, gy, phi, n.value

    def compute_dt(self):
        dt = C.c_double()
        check(_lib.lib().cc_euler2d_compute_dt(self._h, C.byref(dt)))
        return dt.value

    def rhs(self):
        out = {f: np.empty(self.shape) for f in FIELDS}
        check(_lib.lib().cc_euler2d_rhs(self._h, *[_p(out[f]) for f in FIELDS]))
        return out

    def rk2_step(self, dt):
        check(_lib.lib().cc_euler2d_rk2_step(self._h, float(dt)))

    def run(self, nsteps, t=0.0, t_end=1e300, return_dts=False):
        tt = C.c_double(t)
        dts = np.zeros(max(nsteps, 1))
        st = _lib.lib().cc_euler2d_run(self._h, int(nsteps), C.byref(tt), float(t_end), _p(dts))
        if st == 3:
            raise BadDt(st, _lib.lib().cc_last_error().decode(errors="replace"))
        check(st)
        return (tt.value, dts[:nsteps]) if return_dts else tt.value

    def sweep_stats(self):
        last, solves = C.c_int(), C.c_int()
        total = C.c_longlong()
        check(_lib.lib().cc_euler2d_sweep_stats(self._h, C.byref(last), C.byref(total), C.byref(solves)))
        return last.value, total.value, solves.value

    def sync(self):
        check(_lib.lib().cc_euler2d_sync(self._h))

    # -- snapshot frame and in-situ diagnostics ------------------------------------------------
    def frame(self):
        """The fields one snapshot stores, computed on the device: {name: array}."""
        return _lib.frame_of("euler2d", self._h, (self.params.Ny, self.params.Nx))

    def totals(self):
        """Cell-volume-weighted totals of the conserved variables (and max|divB| for MHD)."""
        return _lib.totals_of("euler2d", self._h)

    def compare(self, other):
        """Per frame field (mean|a-b|, sum|a-b|/sum|b|, max|a-b|) of this handle against `other`."""
        return _lib.compare_of("euler2d", self._h, other._h)

    def set_profiling(self, on=True):
        check(_lib.lib().cc_euler2d_set_profiling(self._h, int(bool(on))))

    def reset_profiling(self):
        check(_lib.lib().cc_euler2d_reset_profiling(self._h))

    def kernel_time(self, name):
        n = C.c_int()
        ms = C.c_double()
        check(_lib.lib().cc_euler2d_kernel_time(self._h, name.encode(), C.byref(n), C.byref(ms)))
        return n.value, ms.value


def hllc_batch(WL, WR, axis=0, variant="eh", gamma=1.2, floor_rho=1e-12, precision=CC_FP64):
    """HLLC of EH (EH:125-219, variant 'eh') or NS (NS:212-282, variant 'ns') on n x 4 {rho,u,v,P}
    states; returns n x 4 {f_rho, f_mx, f_my, f_E}."""
    WL = np.ascontiguousarray(WL, dtype=np.float64)
    WR = np.ascontiguousarray(WR, dtype=np.float64)
    n = WL.shape[0]
    out = np.empty((n, 4))
    check(_lib.lib().cc_hllc2d_batch(0 if variant == "eh" else 1, int(axis), n, _p(WL), _p(WR), float(gamma),
                                     float(floor_rho), int(precision), _p(out)))
    return out
