"""ctypes bindings for the CPU parity oracle (test infrastructure only).

The same C API (oracle/oracle.h) is exported by liboracle.so (the restatement)
and by _ref/libref_*.so (the reference programs' own sources, compiled in this
container); ``load(kind)`` returns either, so tests can compare them.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
DP = C.POINTER(C.c_double)


class M2Params(C.Structure):
    _fields_ = [("gamma", C.c_double), ("cfl", C.c_double), ("x0", C.c_double), ("x1", C.c_double),
                ("y0", C.c_double), ("y1", C.c_double), ("Nx", C.c_int), ("Ny", C.c_int),
                ("floor_rho", C.c_double), ("floor_P", C.c_double), ("periodic_x", C.c_int),
                ("periodic_y", C.c_int), ("eta", C.c_double), ("gravity_on", C.c_int), ("G", C.c_double),
                ("sor_iters", C.c_int), ("sor_omega", C.c_double), ("hlld_guard", C.c_double),
                ("ct_comp", C.c_int)]


FP = C.POINTER(C.c_float)


class M2State(C.Structure):
    _fields_ = [(n, DP) for n in ("rho", "mx", "my", "mz", "E", "Bz", "Bx_f", "By_f")] + \
               [(n, FP) for n in ("Bx_lo", "By_lo")]


class M3Params(C.Structure):
    _fields_ = [(n, C.c_double) for n in ("gamma1", "gamma2", "K1", "rho_nuc", "gamma_th", "rho_bounce",
                                           "gamma_bounce", "Kbounce", "gamma_init", "cfl",
                                           "x0", "x1", "y0", "y1", "z0", "z1")] + \
               [("Nx", C.c_int), ("Ny", C.c_int), ("Nz", C.c_int), ("floor_rho", C.c_double),
                ("floor_P", C.c_double), ("periodic_x", C.c_int), ("periodic_y", C.c_int),
                ("periodic_z", C.c_int), ("eta", C.c_double), ("ideal_eos", C.c_int),
                ("gamma_ideal", C.c_double), ("hlld_guard", C.c_double), ("ct_comp", C.c_int)]


class M3State(C.Structure):
    _fields_ = [(n, DP) for n in ("rho", "mx", "my", "mz", "E", "Bx_f", "By_f", "Bz_f")] + \
               [(n, FP) for n in ("Bx_lo", "By_lo", "Bz_lo")]


class EHParams(C.Structure):
    _fields_ = [("gamma", C.c_double), ("cfl", C.c_double), ("x0", C.c_double), ("x1", C.c_double),
                ("y0", C.c_double), ("y1", C.c_double), ("Nx", C.c_int), ("Ny", C.c_int),
                ("floor_rho", C.c_double), ("floor_P", C.c_double), ("periodic_x", C.c_int),
                ("periodic_y", C.c_int), ("gravity_on", C.c_int), ("G", C.c_double), ("max_iter", C.c_int),
                ("tol", C.c_double), ("omega", C.c_double)]


class EHState(C.Structure):
    _fields_ = [(n, DP) for n in ("rho", "mx", "my", "E")]


class NSParams(C.Structure):
    _fields_ = [("gamma", C.c_double), ("cfl", C.c_double), ("x0", C.c_double), ("x1", C.c_double),
                ("y0", C.c_double), ("y1", C.c_double), ("Nx", C.c_int), ("Ny", C.c_int),
                ("floor_rho", C.c_double), ("floor_P", C.c_double), ("viscosity_on", C.c_int),
                ("mu", C.c_double), ("cfl_visc", C.c_double), ("gravity_on", C.c_int), ("G", C.c_double)]


class E3Params(C.Structure):
    _fields_ = [(n, C.c_double) for n in ("gamma", "cfl", "x0", "x1", "y0", "y1", "z0", "z1")] + \
               [("Nx", C.c_int), ("Ny", C.c_int), ("Nz", C.c_int), ("floor_rho", C.c_double),
                ("floor_P", C.c_double), ("gravity_on", C.c_int), ("G", C.c_double)]


class E3State(C.Structure):
    _fields_ = [(n, DP) for n in ("rho", "mx", "my", "mz", "E")]


def m2_params(**kw):
    """Reference defaults of M2:44-86 (mhd2d_resistive_hlld_ct_gravity.cpp), overridable."""
    d = dict(gamma=5.0 / 3.0, cfl=0.45, x0=0.0, x1=1.0, y0=0.0, y1=1.0, Nx=192, Ny=192, floor_rho=1e-10,
             floor_P=1e-10, periodic_x=1, periodic_y=1, eta=5e-4, gravity_on=1, G=8.0, sor_iters=600,
             sor_omega=1.8, hlld_guard=0.0, ct_comp=0)
    d.update(kw)
    return M2Params(**d)


def m3_params(**kw):
    """Reference defaults of M3:43-89 (mhd3d_resistive_hlld_ct.cpp), overridable."""
    d = dict(gamma1=1.30, gamma2=2.50, K1=0.05, rho_nuc=80.0, gamma_th=1.50, rho_bounce=120.0, gamma_bounce=3.0,
             Kbounce=0.02, gamma_init=5.0 / 3.0, cfl=0.4, x0=0.0, x1=1.0, y0=0.0, y1=1.0, z0=0.0, z1=1.0,
             Nx=48, Ny=48, Nz=48, floor_rho=1e-10, floor_P=1e-10, periodic_x=1, periodic_y=1, periodic_z=1,
             eta=5e-4, ideal_eos=0, gamma_ideal=5.0 / 3.0, hlld_guard=0.0, ct_comp=0)
    d.update(kw)
    return M3Params(**d)


def eh_params(**kw):
    """Reference defaults of EH:29-56 plus the SOR constants of EH:329-331."""
    d = dict(gamma=1.2, cfl=0.45, x0=0.0, x1=1.0, y0=0.0, y1=1.0, Nx=50, Ny=50, floor_rho=1e-12, floor_P=1e-12,
             periodic_x=1, periodic_y=1, gravity_on=1, G=1.0, max_iter=50000, tol=1e-10, omega=1.7)
    d.update(kw)
    return EHParams(**d)


def ns_params(**kw):
    """Reference defaults of NS:54-86."""
    d = dict(gamma=1.4, cfl=0.35, x0=0.0, x1=1.0, y0=0.0, y1=1.0, Nx=128, Ny=128, floor_rho=1e-10,
             floor_P=1e-10, viscosity_on=1, mu=1e-6, cfl_visc=0.05, gravity_on=1, G=10.0)
    d.update(kw)
    return NSParams(**d)


def e3_params(**kw):
    """The uniform-sphere benchmark's parameters (periodic [0,1)^3, gamma 5/3, G 1, CFL 0.4)."""
    d = dict(gamma=5.0 / 3.0, cfl=0.4, x0=0.0, x1=1.0, y0=0.0, y1=1.0, z0=0.0, z1=1.0, Nx=32, Ny=32, Nz=32,
             floor_rho=1e-10, floor_P=1e-10, gravity_on=1, G=1.0)
    d.update(kw)
    return E3Params(**d)


def ptr(a):
    assert a.dtype == np.float64 and a.flags.c_contiguous
    return a.ctypes.data_as(DP)


def fptr(a):
    assert a.dtype == np.float32 and a.flags.c_contiguous
    return a.ctypes.data_as(FP)


class Lib:
    """Either the restatement (kind='oracle') or a reference harness (kind='ref_m2', 'ref_m3', ...)."""

    def __init__(self, kind="oracle"):
        if kind == "oracle":
            # CC_ORACLE_LIB: another build of the same sources (the sanitizer build of the CPU suite)
            path = os.environ.get("CC_ORACLE_LIB") or os.path.join(HERE, "liboracle.so")
        else:
            path = os.path.join(HERE, "_ref", "lib%s.so" % kind)
        if not os.path.exists(path):
            raise FileNotFoundError(path)
        self.path = path
        self.kind = kind
        self.lib = C.CDLL(path)
        for name in ("or_m2_mc_limiter",):
            if hasattr(self.lib, name):
                getattr(self.lib, name).restype = C.c_double
        for name in ("or_m2_fast_speed_x", "or_m2_fast_speed_y", "or_m2_compute_dt", "or_m2_run",
                     "or_m3_sound_speed2", "or_m3_total_energy", "or_m3_compute_dt", "or_m3_run",
                     "or_eh_compute_dt", "or_eh_run", "or_ns_compute_dt", "or_ns_run", "or_e3_compute_dt",
                     "or_e3_run"):
            try:
                getattr(self.lib, name).restype = C.c_double
            except AttributeError:
                pass

    def __getattr__(self, name):
        return getattr(self.lib, name)


class M2Arrays:
    """Reference-layout 2D MHD state held in numpy arrays (M2:88-95), plus the FP32 low words of the faces
    that the compensated face update carries (zeros unless ct_comp is on)."""

    FIELDS = ("rho", "mx", "my", "mz", "E", "Bz", "Bx_f", "By_f")

    def __init__(self, Nx, Ny, arrays=None):
        self.Nx, self.Ny = Nx, Ny
        shapes = {"Bx_f": (Nx + 1) * Ny, "By_f": Nx * (Ny + 1)}
        self.a = {}
        for f in self.FIELDS:
            n = shapes.get(f, Nx * Ny)
            if arrays is not None:
                self.a[f] = np.ascontiguousarray(arrays[f], dtype=np.float64).reshape(-1).copy()
                assert self.a[f].size == n, (f, self.a[f].size, n)
            else:
                self.a[f] = np.zeros(n)
        self.lo = {"Bx_lo": np.zeros((Nx + 1) * Ny, np.float32), "By_lo": np.zeros(Nx * (Ny + 1), np.float32)}
        if arrays is not None:
            for f in self.lo:
                if f in arrays:
                    self.lo[f] = np.ascontiguousarray(arrays[f], dtype=np.float32).reshape(-1).copy()

    def copy(self):
        c = M2Arrays(self.Nx, self.Ny, self.a)
        c.lo = {k: v.copy() for k, v in self.lo.items()}
        return c

    def cstate(self):
        return M2State(*([ptr(self.a[f]) for f in self.FIELDS] + [fptr(self.lo[f]) for f in ("Bx_lo", "By_lo")]))

    def __getitem__(self, k):
        return self.a[k]


class M3Arrays:
    FIELDS = ("rho", "mx", "my", "mz", "E", "Bx_f", "By_f", "Bz_f")

    def __init__(self, Nx, Ny, Nz, arrays=None):
        self.Nx, self.Ny, self.Nz = Nx, Ny, Nz
        shapes = {"Bx_f": (Nx + 1) * Ny * Nz, "By_f": Nx * (Ny + 1) * Nz, "Bz_f": Nx * Ny * (Nz + 1)}
        self.a = {}
        for f in self.FIELDS:
            n = shapes.get(f, Nx * Ny * Nz)
            if arrays is not None:
                self.a[f] = np.ascontiguousarray(arrays[f], dtype=np.float64).reshape(-1).copy()
                assert self.a[f].size == n, (f, self.a[f].size, n)
            else:
                self.a[f] = np.zeros(n)
        self.lo = {k.replace("_f", "_lo"): np.zeros(shapes[k], np.float32) for k in ("Bx_f", "By_f", "Bz_f")}

    def copy(self):
        c = M3Arrays(self.Nx, self.Ny, self.Nz, self.a)
        c.lo = {k: v.copy() for k, v in self.lo.items()}
        return c

    def cstate(self):
        return M3State(*([ptr(self.a[f]) for f in self.FIELDS] + [fptr(self.lo[f]) for f in ("Bx_lo", "By_lo", "Bz_lo")]))

    def __getitem__(self, k):
        return self.a[k]


class EHArrays:
    FIELDS = ("rho", "mx", "my", "E")

    def __init__(self, Nx, Ny, arrays=None):
        self.Nx, self.Ny = Nx, Ny
        self.a = {}
        for f in self.FIELDS:
            if arrays is not None:
                self.a[f] = np.ascontiguousarray(arrays[f], dtype=np.float64).reshape(-1).copy()
            else:
                self.a[f] = np.zeros(Nx * Ny)

    def copy(self):
        return EHArrays(self.Nx, self.Ny, self.a)

    def cstate(self):
        return EHState(*[ptr(self.a[f]) for f in self.FIELDS])

    def __getitem__(self, k):
        return self.a[k]


class E3Arrays:
    FIELDS = ("rho", "mx", "my", "mz", "E")

    def __init__(self, Nx, Ny, Nz, arrays=None):
        self.Nx, self.Ny, self.Nz = Nx, Ny, Nz
        self.a = {}
        for f in self.FIELDS:
            if arrays is not None:
                self.a[f] = np.ascontiguousarray(arrays[f], dtype=np.float64).reshape(-1).copy()
                assert self.a[f].size == Nx * Ny * Nz, f
            else:
                self.a[f] = np.zeros(Nx * Ny * Nz)

    def copy(self):
        return E3Arrays(self.Nx, self.Ny, self.Nz, self.a)

    def cstate(self):
        return E3State(*[ptr(self.a[f]) for f in self.FIELDS])

    def __getitem__(self, k):
        return self.a[k]
