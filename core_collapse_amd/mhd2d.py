"""2D resistive MHD + SOR self-gravity step (mhd2d_resistive_hlld_ct_gravity.cpp, cited M2:<line>).

``Params`` mirrors struct Params (M2:44-86) with the same defaults; ``Mhd2D`` holds one grid's
state on the GPU and exposes the reference's step functions under their names:

    compute_poisson_gravity_periodic()   M2:178-237  (result kept on the device as g)
    compute_dt()                         M2:579-598
    advance_stage(dt)                    M2:690-834  (in place)
    rk2_step(dt)                         M2:838-869  (in place, inner Poisson solve included)
    run(nsteps, t, t_end)                the driver loop M2:1111-1117, device-resident dt
    divB_max()                           max |divB_cc| (M2:873-879), as printed at M2:1126

State arrays use the reference layout (M2:88-95): cells (Ny, Nx) i-fastest; Bx_f (Ny, Nx+1);
By_f (Ny+1, Nx); all float64.
"""
import ctypes as C
from dataclasses import dataclass, fields

import numpy as np

from . import _lib
from ._lib import CC_FP64, CC_MIXED, DP, check

FIELDS = ("rho", "mx", "my", "mz", "E", "Bz", "Bx_f", "By_f")


@dataclass
class Params:
    """struct Params of M2:44-86 (hot-path subset) plus the precision switch of this library."""
    gamma: float = 5.0 / 3.0
    cfl: float = 0.45
    x0: float = 0.0
    x1: float = 1.0
    y0: float = 0.0
    y1: float = 1.0
    Nx: int = 192
    Ny: int = 192
    floor_rho: float = 1e-10
    floor_P: float = 1e-10
    periodic_x: bool = True
    periodic_y: bool = True
    eta: float = 5e-4
    gravity_on: bool = True
    G: float = 8.0
    sor_iters: int = 600
    sor_omega: float = 1.8
    precision: int = CC_FP64
    # no reference counterpart (0 keeps the reference's arithmetic): HLL for faces whose HLLD star-state
    # denominator is below hlld_guard of its scale; FP32 low words carried through CT and the Heun average
    hlld_guard: float = 0.0
    ct_compensated: bool = False

    @property
    def dx(self):
        return (self.x1 - self.x0) / self.Nx

    @property
    def dy(self):
        return (self.y1 - self.y0) / self.Ny

    def c_struct(self):
        return _lib.Mhd2dParams(**{f.name: (int(getattr(self, f.name)) if isinstance(getattr(self, f.name), bool)
                                            else getattr(self, f.name)) for f in fields(self)})


def _p(a):
    return a.ctypes.data_as(DP)


class Mhd2D:
    """One 2D grid on the GPU (a cc_mhd2d handle)."""

    def __init__(self, params: Params):
        self.params = params
        self._h = C.c_void_p()
        self._ps = params.c_struct()
        check(_lib.lib().cc_mhd2d_create(C.byref(self._ps), C.byref(self._h)))

    def close(self):
        if self._h:
            _lib.lib().cc_mhd2d_destroy(self._h)
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

    # -- state transfer ------------------------------------------------------------------
    def shapes(self):
        Nx, Ny = self.params.Nx, self.params.Ny
        return {"rho": (Ny, Nx), "mx": (Ny, Nx), "my": (Ny, Nx), "mz": (Ny, Nx), "E": (Ny, Nx), "Bz": (Ny, Nx),
                "Bx_f": (Ny, Nx + 1), "By_f": (Ny + 1, Nx)}

    def upload(self, state):
        sh = self.shapes()
        arrs = []
        for f in FIELDS:
            a = np.ascontiguousarray(state[f], dtype=np.float64)
            if a.size != int(np.prod(sh[f])):
                raise ValueError("field %s has %d values, expected %s" % (f, a.size, sh[f]))
            arrs.append(a)
        check(_lib.lib().cc_mhd2d_upload(self._h, *[_p(a) for a in arrs]))

    def download(self):
        sh = self.shapes()
        out = {f: np.empty(sh[f]) for f in FIELDS}
        check(_lib.lib().cc_mhd2d_download(self._h, *[_p(out[f]) for f in FIELDS]))
        return out

    # -- the reference step functions ------------------------------------------------------
    def compute_poisson_gravity_periodic(self):
        check(_lib.lib().cc_mhd2d_gravity(self._h))

    def gravity(self):
        n = self.params.Nx * self.params.Ny
        gx, gy = np.empty(n), np.empty(n)
        check(_lib.lib().cc_mhd2d_get_gravity(self._h, _p(gx), _p(gy)))
        return gx.reshape(self.params.Ny, self.params.Nx), gy.reshape(self.params.Ny, self.params.Nx)

    def set_gravity(self, gx, gy):
        gx = np.ascontiguousarray(gx, dtype=np.float64)
        gy = np.ascontiguousarray(gy, dtype=np.float64)
        check(_lib.lib().cc_mhd2d_set_gravity(self._h, _p(gx), _p(gy)))

    def compute_dt(self):
        dt = C.c_double()
        check(_lib.lib().cc_mhd2d_compute_dt(self._h, C.byref(dt)))
        return dt.value

    def advance_stage(self, dt):
        check(_lib.lib().cc_mhd2d_stage(self._h, float(dt)))

    def rk2_step(self, dt):
        check(_lib.lib().cc_mhd2d_rk2_step(self._h, float(dt)))

    def run(self, nsteps, t=0.0, t_end=1e300, return_dts=False):
        tt = C.c_double(t)
        dts = np.zeros(max(nsteps, 1))
        check(_lib.lib().cc_mhd2d_run(self._h, int(nsteps), C.byref(tt), float(t_end), _p(dts)))
        return (tt.value, dts[:nsteps]) if return_dts else tt.value

    def divB_max(self):
        v = C.c_double()
        check(_lib.lib().cc_mhd2d_divb_max(self._h, C.byref(v)))
        return v.value

    def sync(self):
        check(_lib.lib().cc_mhd2d_sync(self._h))

    # -- snapshot frame and in-situ diagnostics ------------------------------------------------
    def frame(self):
        """The fields one snapshot stores, computed on the device: {name: array}."""
        return _lib.frame_of("mhd2d", self._h, (self.params.Ny, self.params.Nx))

    def totals(self):
        """Cell-volume-weighted totals of the conserved variables (and max|divB| for MHD)."""
        return _lib.totals_of("mhd2d", self._h)

    def compare(self, other):
        """Per frame field (mean|a-b|, sum|a-b|/sum|b|, max|a-b|) of this handle against `other`."""
        return _lib.compare_of("mhd2d", self._h, other._h)

    # -- profiling ---------------------------------------------------------------------------
    def set_profiling(self, on=True):
        check(_lib.lib().cc_mhd2d_set_profiling(self._h, int(bool(on))))

    def reset_profiling(self):
        check(_lib.lib().cc_mhd2d_reset_profiling(self._h))

    def kernel_time(self, name):
        n = C.c_int()
        ms = C.c_double()
        check(_lib.lib().cc_mhd2d_kernel_time(self._h, name.encode(), C.byref(n), C.byref(ms)))
        return n.value, ms.value


def sor_autotune(enable=True):
    """Time the compiled SOR tile shapes once per grid size for handles created afterwards (cc_sor_autotune)."""
    check(_lib.lib().cc_sor_autotune(int(bool(enable))))


def sor_tile_choice(Nx, Ny, sor_iters=600, precision=CC_FP64, with_kernels=False):
    """(sweeps per tile, rows or k values per tile, sheared) of the SOR schedule a grid gets
    (cc_sor_tile_choice); with_kernels=True appends whether a mixed grid runs the FP64 kernels."""
    S, WJ, sh, d = C.c_int(), C.c_int(), C.c_int(), C.c_int()
    check(_lib.lib().cc_sor_tile_choice(int(Nx), int(Ny), int(sor_iters), int(precision), C.byref(S), C.byref(WJ),
                                        C.byref(sh), C.byref(d)))
    return (S.value, WJ.value, bool(sh.value)) + ((bool(d.value),) if with_kernels else ())


def sor_flow_schedule(Nx, Ny, iters, W=64, WJ=64, S=16, sheared=False):
    """The tiled SOR's dataflow schedule (cc_sor_flow_schedule; host only): a list of tiles (TS, TJ, TI)
    -- (TS, TK, TI) for sheared tiles -- in launch order and, per tile, the indices of the tiles it
    waits for."""
    lib = _lib.lib()
    n, nd = C.c_int(), C.c_int()
    args = [int(v) for v in (Nx, Ny, iters, W, WJ, S, bool(sheared))]
    check(lib.cc_sor_flow_schedule(*args, C.byref(n), C.byref(nd), None, None, None))
    tiles = np.zeros(3 * n.value, dtype=np.int32)
    off = np.zeros(n.value + 1, dtype=np.int32)
    idx = np.zeros(max(nd.value, 1), dtype=np.int32)
    check(lib.cc_sor_flow_schedule(*args, C.byref(n), C.byref(nd), tiles.ctypes.data, off.ctypes.data,
                                   idx.ctypes.data))
    tl = [tuple(int(v) for v in tiles[3 * k:3 * k + 3]) for k in range(n.value)]
    return tl, [[int(v) for v in idx[off[k]:off[k + 1]]] for k in range(n.value)]


class Mhd2DBatch:
    """Independent Mhd2D instances stepped together: one HIP graph launch per step runs every
    instance's Heun step as a parallel branch, each with its own dt (cc_mhd2d_batch_*)."""

    def __init__(self, sims):
        self.sims = list(sims)
        self._b = C.c_void_p()
        arr = (C.c_void_p * len(self.sims))(*[s._h for s in self.sims])
        check(_lib.lib().cc_mhd2d_batch_create(arr, len(self.sims), C.byref(self._b)))

    def run(self, nsteps, t=None, t_end=None):
        n = len(self.sims)
        tt = np.zeros(n) if t is None else np.array(t, dtype=np.float64)
        te = np.full(n, 1e300) if t_end is None else np.broadcast_to(np.asarray(t_end, dtype=np.float64), (n,)).copy()
        check(_lib.lib().cc_mhd2d_batch_run(self._b, int(nsteps), _p(tt), _p(te)))
        return tt

    def info(self):
        """{"merged": SOR tile levels merged across instances, "tile": (S, WJ, sheared) of the merged
        launches, "fp64_kernels": a mixed batch running the FP64 SOR kernels}."""
        m, S, WJ, sh, d = C.c_int(), C.c_int(), C.c_int(), C.c_int(), C.c_int()
        check(_lib.lib().cc_mhd2d_batch_info(self._b, C.byref(m), C.byref(S), C.byref(WJ), C.byref(sh), C.byref(d)))
        return {"merged": bool(m.value), "tile": (S.value, WJ.value, bool(sh.value)), "fp64_kernels": bool(d.value)}

    def close(self):
        if self._b:
            _lib.lib().cc_mhd2d_batch_destroy(self._b)
            self._b = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def hlld_batch(axis, WL, WR, Bn, gamma=5.0 / 3.0, floor_rho=1e-10, precision=CC_FP64):
    """hlld_x (axis 0, M2:372) / hlld_y (axis 1, M2:540) on n independent states (n x 8 arrays)."""
    WL = np.ascontiguousarray(WL, dtype=np.float64)
    WR = np.ascontiguousarray(WR, dtype=np.float64)
    Bn = np.ascontiguousarray(Bn, dtype=np.float64)
    n = Bn.size
    out = np.empty((n, 8))
    check(_lib.lib().cc_hlld2d_batch(int(axis), n, _p(WL), _p(WR), _p(Bn), float(gamma), float(floor_rho),
                                     int(precision), _p(out)))
    return out


def poisson_sor(rho, dx, dy, G, sor_iters=600, omega=1.8, periodic_x=True, periodic_y=True, precision=CC_FP64,
                want_phi=False):
    """compute_poisson_gravity_periodic (M2:178-237) on a host density array (Ny, Nx)."""
    rho = np.ascontiguousarray(rho, dtype=np.float64)
    Ny, Nx = rho.shape
    gx, gy = np.empty_like(rho), np.empty_like(rho)
    phi = np.empty_like(rho) if want_phi else None
    check(_lib.lib().cc_poisson2d_sor(Nx, Ny, int(periodic_x), int(periodic_y), float(dx), float(dy), float(G),
                                      int(sor_iters), float(omega), int(precision), _p(rho), _p(gx), _p(gy),
                                      _p(phi) if want_phi else None))
    return (gx, gy, phi) if want_phi else (gx, gy)
