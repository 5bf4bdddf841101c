"""3D resistive MHD step (mhd3d_resistive_hlld_ct.cpp, cited M3:<line>).

``Params`` mirrors struct Params (M3:43-89, hot-path subset) with the same defaults, plus the EOS
switch (``eos=0`` the reference hybrid EOS, ``eos=1`` ideal gas with ``gamma``) and the precision.
``Mhd3D`` exposes compute_dt (M3:626), advance_stage (M3:694), rk2_step (M3:1034), the driver loop
(M3:1285-1291) as run(), and divB_max (M3:1070).  Layout (M3:91-99): cells (Nz, Ny, Nx);
Bx_f (Nz, Ny, Nx+1); By_f (Nz, Ny+1, Nx); Bz_f (Nz+1, Ny, Nx); float64.
"""
import ctypes as C
from dataclasses import dataclass, fields

import numpy as np

from . import _lib
from ._lib import CC_FP64, DP, check

FIELDS = ("rho", "mx", "my", "mz", "E", "Bx_f", "By_f", "Bz_f")


@dataclass
class Params:
    gamma1: float = 1.30
    gamma2: float = 2.50
    K1: float = 0.05
    rho_nuc: float = 80.0
    gamma_th: float = 1.50
    rho_bounce: float = 120.0
    gamma_bounce: float = 3.00
    Kbounce: float = 0.02
    cfl: float = 0.4
    x0: float = 0.0
    x1: float = 1.0
    y0: float = 0.0
    y1: float = 1.0
    z0: float = 0.0
    z1: float = 1.0
    Nx: int = 48
    Ny: int = 48
    Nz: int = 48
    floor_rho: float = 1e-10
    floor_P: float = 1e-10
    periodic_x: bool = True
    periodic_y: bool = True
    periodic_z: bool = True
    eta: float = 5e-4
    eos: int = 0
    gamma: float = 5.0 / 3.0
    precision: int = CC_FP64
    degen_tau: float = 0.0   # mixed precision: FP64 fallback margin near the Alfven degeneracy (0: default)
    hlld_guard: float = 0.0  # HLL for faces whose HLLD star-state denominator is below this fraction of its scale
    ct_compensated: bool = False  # FP32 low words of the faces carried through CT and the Heun average

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
        return _lib.Mhd3dParams(**{f.name: (int(getattr(self, f.name)) if isinstance(getattr(self, f.name), bool)
                                            else getattr(self, f.name)) for f in fields(self)})


def _p(a):
    return a.ctypes.data_as(DP)


class Mhd3D:
    def __init__(self, params: Params):
        self.params = params
        self._h = C.c_void_p()
        self._ps = params.c_struct()
        check(_lib.lib().cc_mhd3d_create(C.byref(self._ps), C.byref(self._h)))

    def close(self):
        if self._h:
            _lib.lib().cc_mhd3d_destroy(self._h)
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

    def shapes(self):
        Nx, Ny, Nz = self.params.Nx, self.params.Ny, self.params.Nz
        c = (Nz, Ny, Nx)
        return {"rho": c, "mx": c, "my": c, "mz": c, "E": c, "Bx_f": (Nz, Ny, Nx + 1), "By_f": (Nz, Ny + 1, Nx),
                "Bz_f": (Nz + 1, Ny, Nx)}

    def upload(self, state):
        sh = self.shapes()
        arrs = []
        for f in FIELDS:
            a = np.ascontiguousarray(state[f], dtype=np.float64)
            if a.size != int(np.prod(sh[f])):
                raise ValueError("field %s has %d values, expected %s" % (f, a.size, sh[f]))
            arrs.append(a)
        check(_lib.lib().cc_mhd3d_upload(self._h, *[_p(a) for a in arrs]))

    def download(self):
        sh = self.shapes()
        out = {f: np.empty(sh[f]) for f in FIELDS}
        check(_lib.lib().cc_mhd3d_download(self._h, *[_p(out[f]) for f in FIELDS]))
        return out

    def compute_dt(self):
        dt = C.c_double()
        check(_lib.lib().cc_mhd3d_compute_dt(self._h, C.byref(dt)))
        return dt.value

    def advance_stage(self, dt):
        check(_lib.lib().cc_mhd3d_stage(self._h, float(dt)))

    def rk2_step(self, dt):
        check(_lib.lib().cc_mhd3d_rk2_step(self._h, float(dt)))

    def run(self, nsteps, t=0.0, t_end=1e300, return_dts=False):
        tt = C.c_double(t)
        dts = np.zeros(max(nsteps, 1))
        check(_lib.lib().cc_mhd3d_run(self._h, int(nsteps), C.byref(tt), float(t_end), _p(dts)))
        return (tt.value, dts[:nsteps]) if return_dts else tt.value

    def divB_max(self):
        v = C.c_double()
        check(_lib.lib().cc_mhd3d_divb_max(self._h, C.byref(v)))
        return v.value

    def sync(self):
        check(_lib.lib().cc_mhd3d_sync(self._h))

    # -- snapshot frame and in-situ diagnostics ------------------------------------------------
    def frame(self):
        """The fields one snapshot stores, computed on the device: {name: array}."""
        return _lib.frame_of("mhd3d", self._h, (self.params.Nz, self.params.Ny, self.params.Nx))

    def totals(self):
        """Cell-volume-weighted totals of the conserved variables (and max|divB| for MHD)."""
        return _lib.totals_of("mhd3d", self._h)

    def compare(self, other):
        """Per frame field (mean|a-b|, sum|a-b|/sum|b|, max|a-b|) of this handle against `other`."""
        return _lib.compare_of("mhd3d", self._h, other._h)

    def set_profiling(self, on=True):
        check(_lib.lib().cc_mhd3d_set_profiling(self._h, int(bool(on))))

    def reset_profiling(self):
        check(_lib.lib().cc_mhd3d_reset_profiling(self._h))

    def kernel_time(self, name):
        n = C.c_int()
        ms = C.c_double()
        check(_lib.lib().cc_mhd3d_kernel_time(self._h, name.encode(), C.byref(n), C.byref(ms)))
        return n.value, ms.value


def hlld_batch(params: Params, WL, WR, Bn):
    """hlld_x (M3:415-568) with params' EOS and precision on n states (n x 8); returns n x 7."""
    WL = np.ascontiguousarray(WL, dtype=np.float64)
    WR = np.ascontiguousarray(WR, dtype=np.float64)
    Bn = np.ascontiguousarray(Bn, dtype=np.float64)
    out = np.empty((Bn.size, 7))
    ps = params.c_struct()
    check(_lib.lib().cc_hlld3d_batch(C.byref(ps), Bn.size, _p(WL), _p(WR), _p(Bn), _p(out)))
    return out
