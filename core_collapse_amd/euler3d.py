"""3D Euler HLLC + FFT self-gravity step (the periodic uniform-sphere collapse benchmark).

No reference program has this step: it is the step of euler2d_gravity_hllc.cpp (EH:406-550:
floored primitives, 4-point MUSCL-MC, the EH HLLC, flux divergence + rho g source, Heun with floors)
extended to three axes, with gravity from the 7-point discrete Poisson equation solved by FFT with
FP64 iterative refinement to a relative residual ``poisson_tol``.  State: SoA float64 arrays of
shape (Nz, Ny, Nx); grids are powers of two in [8, 1024] on every axis (periodic).
"""
import ctypes as C
from dataclasses import dataclass, fields

import numpy as np

from . import _lib
from ._lib import CC_FP64, DP, CCError, check

FIELDS = ("rho", "mx", "my", "mz", "E")


class BadDt(CCError):
    pass


@dataclass
class Params:
    Nx: int = 64
    Ny: int = 64
    Nz: int = 64
    x0: float = 0.0
    x1: float = 1.0
    y0: float = 0.0
    y1: float = 1.0
    z0: float = 0.0
    z1: float = 1.0
    gamma: float = 5.0 / 3.0
    cfl: float = 0.4
    floor_rho: float = 1e-10
    floor_P: float = 1e-10
    gravity_on: bool = True
    G: float = 1.0
    poisson_tol: float = 1e-10
    poisson_max_iter: int = 8
    precision: int = CC_FP64

    @property
    def dx(self):
        return (self.x1 - self.x0) / self.Nx

    @property
    def dy(self):
        return (self.y1 - self.y0) / self.Ny

    @property
    def dz(self):
        return (self.z1 - self.z0) / self.Nz

    def c_struct(self):
        return _lib.Euler3dParams(**{f.name: (int(getattr(self, f.name)) if isinstance(getattr(self, f.name), bool)
                                              else getattr(self, f.name)) for f in fields(self)})


def _p(a):
    return a.ctypes.data_as(DP)


class Euler3D:
    def __init__(self, params: Params):
        self.params = params
        self._h = C.c_void_p()
        self._ps = params.c_struct()
        check(_lib.lib().cc_euler3d_create(C.byref(self._ps), C.byref(self._h)))

    def close(self):
        if self._h:
            _lib.lib().cc_euler3d_destroy(self._h)
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
        return (self.params.Nz, self.params.Ny, self.params.Nx)

    def upload(self, state):
        arrs = []
        for f in FIELDS:
            a = np.ascontiguousarray(state[f], dtype=np.float64)
            if a.size != int(np.prod(self.shape)):
                raise ValueError("field %s has %d values, expected %s" % (f, a.size, self.shape))
            arrs.append(a)
        check(_lib.lib().cc_euler3d_upload(self._h, *[_p(a) for a in arrs]))

    def download(self):
        out = {f: np.empty(self.shape) for f in FIELDS}
        check(_lib.lib().cc_euler3d_download(self._h, *[_p(out[f]) for f in FIELDS]))
        return out

    def gravity(self):
        """(gx, gy, gz, phi) of the current state."""
        out = [np.empty(self.shape) for _ in range(4)]
        check(_lib.lib().cc_euler3d_gravity(self._h, *[_p(a) for a in out]))
        return tuple(out)

    def poisson_stats(self):
        s, il, im, it, u = C.c_int(), C.c_int(), C.c_int(), C.c_int(), C.c_int()
        rl, rw = C.c_double(), C.c_double()
        check(_lib.lib().cc_euler3d_poisson_stats(self._h, C.byref(s), C.byref(il), C.byref(im), C.byref(it),
                                                  C.byref(rl), C.byref(rw), C.byref(u)))
        return {"solves": s.value, "iters_last": il.value, "iters_max": im.value, "iters_total": it.value,
                "resid_last": rl.value,
                "resid_worst": rw.value, "unconverged": u.value}

    def rhs(self):
        out = {f: np.empty(self.shape) for f in FIELDS}
        check(_lib.lib().cc_euler3d_rhs(self._h, *[_p(out[f]) for f in FIELDS]))
        return out

    def compute_dt(self):
        dt = C.c_double()
        check(_lib.lib().cc_euler3d_compute_dt(self._h, C.byref(dt)))
        return dt.value

    def rk2_step(self, dt):
        check(_lib.lib().cc_euler3d_rk2_step(self._h, float(dt)))

    def run(self, nsteps, t=0.0, t_end=1e300, return_dts=False):
        tt = C.c_double(t)
        dts = np.zeros(max(nsteps, 1))
        st = _lib.lib().cc_euler3d_run(self._h, int(nsteps), C.byref(tt), float(t_end), _p(dts))
        if st == 3:  # CC_ERR_BAD_DT
            raise BadDt(st, _lib.lib().cc_last_error().decode(errors="replace"))
        check(st)
        return (tt.value, dts[:nsteps]) if return_dts else tt.value

    def totals(self):
        return _lib.totals_of("euler3d", self._h)

    def sync(self):
        check(_lib.lib().cc_euler3d_sync(self._h))

    def set_profiling(self, on=True):
        check(_lib.lib().cc_euler3d_set_profiling(self._h, int(bool(on))))

    def reset_profiling(self):
        check(_lib.lib().cc_euler3d_reset_profiling(self._h))

    def kernel_time(self, name):
        n, ms = C.c_int(), C.c_double()
        check(_lib.lib().cc_euler3d_kernel_time(self._h, name.encode(), C.byref(n), C.byref(ms)))
        return n.value, ms.value
