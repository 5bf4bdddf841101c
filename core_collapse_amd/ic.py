"""Initial conditions (host side), built by libccamd's C++ IC entry points (cc_*_ic_*).

The reference programs build their ICs inline in main(); libccamd restates them with the same libm
calls and libstdc++ <random> distributions, so the arrays returned here equal the reference's bit
for bit.  The benchmark-problem ICs (Gaussian collapse, solenoidal turbulence) are defined in
include/ccamd.h.  Every function returns arrays in the layout of the matching upload call.
"""
import ctypes as C

import numpy as np

from . import _lib
from ._lib import check
from .mhd2d import Params as Mhd2dParams


def _p(a):
    return a.ctypes.data_as(_lib.DP)


def _knobs(cls, defaults, kw):
    k = cls()
    getattr(_lib.lib(), defaults)(C.byref(k))
    for name, v in kw.items():
        if name not in dict(cls._fields_):
            raise TypeError("unknown IC parameter %r" % name)
        setattr(k, name, v)
    return k


def _m2_state(p):
    Nx, Ny = p.Nx, p.Ny
    return {"rho": np.empty((Ny, Nx)), "mx": np.empty((Ny, Nx)), "my": np.empty((Ny, Nx)), "mz": np.empty((Ny, Nx)),
            "E": np.empty((Ny, Nx)), "Bz": np.empty((Ny, Nx)), "Bx_f": np.empty((Ny, Nx + 1)),
            "By_f": np.empty((Ny + 1, Nx))}


_M2_ORDER = ("rho", "mx", "my", "mz", "E", "Bz", "Bx_f", "By_f")
_M3_ORDER = ("rho", "mx", "my", "mz", "E", "Bx_f", "By_f", "Bz_f")


def mhd2d_reference_ic(p: Mhd2dParams = None, **knobs):
    """The default IC of mhd2d_resistive_hlld_ct_gravity.cpp (M2:954-1048): face B from a corner
    vector potential Az, tanh-core overdensity with infall and rotation.  knobs: the IC members of
    Params (M2:74-85) and gamma_eff (M2:1011)."""
    p = p or Mhd2dParams()
    k = _knobs(_lib.Mhd2dIc, "cc_mhd2d_ic_defaults", knobs)
    st = _m2_state(p)
    ps = p.c_struct()
    check(_lib.lib().cc_mhd2d_ic_reference(C.byref(ps), C.byref(k), *[_p(st[f]) for f in _M2_ORDER]))
    return st


def collapse2d_params(N=128, precision=0, **kw):
    """Parameters of the 2D MHD Gaussian-collapse benchmark: box [-0.5,0.5]^2, clamped ("outflow")
    boundaries, gamma 5/3, G 1, eta 1e-4, CFL 0.4, SOR 600 sweeps at omega 1.8, compensated face update
    (max|div B| at one rounding).  The HLLD keeps the reference's arithmetic: at plasma beta >= 20 no face
    comes near its star-state degeneracy."""
    d = dict(gamma=5.0 / 3.0, cfl=0.4, x0=-0.5, x1=0.5, y0=-0.5, y1=0.5, Nx=N, Ny=N, periodic_x=False,
             periodic_y=False, eta=1e-4, gravity_on=True, G=1.0, sor_iters=600, sor_omega=1.8, precision=precision,
             ct_compensated=True)
    d.update(kw)
    return Mhd2dParams(**d)


def collapse2d_ic(p: Mhd2dParams, **knobs):
    """Gaussian collapse (cc_mhd2d_ic_collapse): rho = 1 + (contrast-1) exp(-r^2/width^2), p = P0,
    Box-Muller velocity noise of amplitude vamp from std::mt19937(seed), uniform Bx = B0.  knobs:
    seed (42), contrast (10), width (0.1), P0 (0.1), vamp (1e-3), B0 (0.05)."""
    k = _knobs(_lib.Collapse2dIc, "cc_collapse2d_ic_defaults", knobs)
    st = _m2_state(p)
    ps = p.c_struct()
    check(_lib.lib().cc_mhd2d_ic_collapse(C.byref(ps), C.byref(k), *[_p(st[f]) for f in _M2_ORDER]))
    return st


def collapse2d_batch_sample(n=64, seed=42, nmin=128, nmax=1024):
    """The batch of collapse instances (cc_collapse2d_batch_sample): n dicts {"N", "eta", "ic"} --
    grid N log-uniform in [nmin, nmax], eta U[1e-5, 1e-3], IC knobs with contrast U[5, 50],
    B0 U[0.01, 0.1] and seed 42 + i; the instance's parameters are collapse2d_params(N, eta=eta)."""
    N = (C.c_int * n)()
    knobs = (_lib.Collapse2dIc * n)()
    eta = np.empty(n)
    check(_lib.lib().cc_collapse2d_batch_sample(C.c_uint(seed), n, nmin, nmax, N, knobs, _p(eta)))
    return [dict(N=int(N[i]), eta=float(eta[i]),
                 ic=dict(seed=knobs[i].seed, contrast=knobs[i].contrast, width=knobs[i].width, P0=knobs[i].P0,
                         vamp=knobs[i].vamp, B0=knobs[i].B0)) for i in range(n)]


def _m3_state(p):
    Nx, Ny, Nz = p.Nx, p.Ny, p.Nz
    c = (Nz, Ny, Nx)
    return {"rho": np.empty(c), "mx": np.empty(c), "my": np.empty(c), "mz": np.empty(c), "E": np.empty(c),
            "Bx_f": np.empty((Nz, Ny, Nx + 1)), "By_f": np.empty((Nz, Ny + 1, Nx)), "Bz_f": np.empty((Nz + 1, Ny, Nx))}


def mhd3d_reference_state(p=None, **knobs):
    """Conserved state of the default IC of mhd3d_resistive_hlld_ct.cpp (M3:1159-1272) through the
    EOS of p.  knobs: the IC members of Params (M3:76-88) and gamma_init (M3:53)."""
    from .mhd3d import Params as P3
    p = p or P3()
    k = _knobs(_lib.Mhd3dIc, "cc_mhd3d_ic_defaults", knobs)
    st = _m3_state(p)
    ps = p.c_struct()
    check(_lib.lib().cc_mhd3d_ic_reference(C.byref(ps), C.byref(k), *[_p(st[f]) for f in _M3_ORDER]))
    return st


def turbulence3d_params(N=256, precision=0, **kw):
    """The 3D MHD turbulence benchmark: periodic [0,1)^3, ideal gas gamma 5/3, eta 5e-5, CFL 0.4, with the
    HLLD degeneracy guard at 5e-2 and the compensated face update.  The beta = 1 field along z puts the z
    faces at the HLLD's star-state degeneracy (fast speed = Alfven speed), where the reference's formulas
    divide by a near-zero denominator guarded only by 1e-14: without the guard the run produces energy
    spikes of 1e3..1e11 within a few steps, and two FP64 builds differing only by FMA contraction part by
    L1(rho) ~ 1e-3.  A guard of 1e-2 still lets star states be amplified ~40x (energy fluxes of 1e3 from
    O(1) states): at 256^3 it leaves ~50 single-cell pressure spikes in 500 steps in either precision and
    blew up one mixed run; 3e-2 and above leave none.  hlld_guard=0 gives the reference's arithmetic."""
    from .mhd3d import Params as P3
    d = dict(Nx=N, Ny=N, Nz=N, eos=1, gamma=5.0 / 3.0, eta=5e-5, cfl=0.4, precision=precision, hlld_guard=5e-2,
             ct_compensated=True)
    d.update(kw)
    return P3(**d)


def turbulence3d_ic(p, **knobs):
    """Solenoidal turbulence (cc_mhd3d_ic_turbulence): rho = 1, p = 1/gamma, uniform Bz with plasma
    beta `beta`, random-phase velocity modes 1 <= |k| <= kmax with |v(k)| ~ k^-2 scaled to rms Mach
    `mach`.  knobs: seed (42), kmin (1), kmax (8), mach (1), beta (1)."""
    k = _knobs(_lib.Turbulence3dIc, "cc_turbulence3d_ic_defaults", knobs)
    st = _m3_state(p)
    ps = p.c_struct()
    check(_lib.lib().cc_mhd3d_ic_turbulence(C.byref(ps), C.byref(k), *[_p(st[f]) for f in _M3_ORDER]))
    return st


def euler2d_reference_ic(p=None, **knobs):
    """init_random_blob of euler2d_gravity_hllc.cpp (EH:574-642).  knobs: seed (42) and the constants
    of EH:585-603.  Returns SoA arrays of shape (Ny, Nx)."""
    from .euler2d import Params as EHParams
    p = p or EHParams()
    k = _knobs(_lib.Euler2dIc, "cc_euler2d_ic_defaults", knobs)
    out = {f: np.empty((p.Ny, p.Nx)) for f in ("rho", "mx", "my", "E")}
    ps = p.c_struct()
    check(_lib.lib().cc_euler2d_ic_reference(C.byref(ps), C.byref(k), *[_p(out[f]) for f in ("rho", "mx", "my", "E")]))
    return out


def ns2d_reference_ic(p=None, **knobs):
    """The collapse IC in main() of navier_stokes_2d_hllc_poisson.cpp (NS:638-688).  knobs: seed (42)
    and the constants of NS:642-660.  Returns the AoS state (Ny, Nx, 4)."""
    from .ns2d import Params as NSParams
    p = p or NSParams()
    k = _knobs(_lib.Ns2dIc, "cc_ns2d_ic_defaults", knobs)
    U = np.empty((p.Ny, p.Nx, 4))
    ps = p.c_struct()
    check(_lib.lib().cc_ns2d_ic_reference(C.byref(ps), C.byref(k), _p(U)))
    return U


def sphere3d_params(N=512, precision=0, **kw):
    """The 3D Euler self-gravity benchmark: periodic [0,1)^3, gamma 5/3, G 1, CFL 0.4, FFT gravity
    refined in FP64 to a relative residual of 1e-10."""
    from .euler3d import Params as E3
    d = dict(Nx=N, Ny=N, Nz=N, gamma=5.0 / 3.0, G=1.0, cfl=0.4, poisson_tol=1e-10, poisson_max_iter=8,
             precision=precision)
    d.update(kw)
    return E3(**d)


def sphere3d_ic(p, **knobs):
    """Uniform sphere (cc_euler3d_ic_sphere): rho = rho_in (10) inside radius (0.2) of the box centre,
    rho_out (0.1) outside, times (1 + noise N(0,1)) with noise 0.01 from std::mt19937(seed = 42), p = P0
    (0.01), at rest.  Returns SoA arrays of shape (Nz, Ny, Nx)."""
    k = _knobs(_lib.Sphere3dIc, "cc_sphere3d_ic_defaults", knobs)
    sh = (p.Nz, p.Ny, p.Nx)
    st = {f: np.empty(sh) for f in ("rho", "mx", "my", "mz", "E")}
    ps = p.c_struct()
    check(_lib.lib().cc_euler3d_ic_sphere(C.byref(ps), C.byref(k), *[_p(st[f]) for f in ("rho", "mx", "my", "mz", "E")]))
    return st
