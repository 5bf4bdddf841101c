"""ctypes binding of libccamd.so (the gfx950 solver-step library, C ABI in include/ccamd.h).

The library is built in-tree by __graft_entry__.build() (make -C core_collapse_amd/csrc).  There is
no CPU fallback: if the shared library is missing or no HIP device is visible, every entry point
raises, so a test or benchmark can never pass on a silent substitute.
"""
import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libccamd.so")

DP = C.POINTER(C.c_double)

CC_OK = 0
CC_FP64 = 0
CC_MIXED = 1
STATUS = {0: "CC_OK", 1: "CC_ERR_INVALID_ARG", 2: "CC_ERR_NON_POW2", 3: "CC_ERR_BAD_DT", 4: "CC_ERR_HIP",
          5: "CC_ERR_OOM", 6: "CC_ERR_NO_DEVICE"}


class CCError(RuntimeError):
    def __init__(self, status, msg):
        super().__init__("%s: %s" % (STATUS.get(status, status), msg))
        self.status = status


class Mhd2dParams(C.Structure):
    _fields_ = [("gamma", C.c_double), ("cfl", C.c_double), ("x0", C.c_double), ("x1", C.c_double),
                ("y0", C.c_double), ("y1", C.c_double), ("Nx", C.c_int), ("Ny", C.c_int),
                ("floor_rho", C.c_double), ("floor_P", C.c_double), ("periodic_x", C.c_int),
                ("periodic_y", C.c_int), ("eta", C.c_double), ("gravity_on", C.c_int), ("G", C.c_double),
                ("sor_iters", C.c_int), ("sor_omega", C.c_double), ("precision", C.c_int),
                ("hlld_guard", C.c_double), ("ct_compensated", C.c_int)]


class Mhd3dParams(C.Structure):
    _fields_ = [(n, C.c_double) for n in ("gamma1", "gamma2", "K1", "rho_nuc", "gamma_th", "rho_bounce",
                                           "gamma_bounce", "Kbounce", "cfl", "x0", "x1", "y0", "y1", "z0",
                                           "z1")] + \
               [("Nx", C.c_int), ("Ny", C.c_int), ("Nz", C.c_int), ("floor_rho", C.c_double),
                ("floor_P", C.c_double), ("periodic_x", C.c_int), ("periodic_y", C.c_int),
                ("periodic_z", C.c_int), ("eta", C.c_double), ("eos", C.c_int), ("gamma", C.c_double),
                ("precision", C.c_int), ("degen_tau", C.c_double), ("hlld_guard", C.c_double),
                ("ct_compensated", C.c_int)]


class Euler2dParams(C.Structure):
    _fields_ = [("gamma", C.c_double), ("cfl", C.c_double), ("x0", C.c_double), ("x1", C.c_double),
                ("y0", C.c_double), ("y1", C.c_double), ("Nx", C.c_int), ("Ny", C.c_int),
                ("floor_rho", C.c_double), ("floor_P", C.c_double), ("periodic_x", C.c_int),
                ("periodic_y", C.c_int), ("gravity_on", C.c_int), ("G", C.c_double), ("max_iter", C.c_int),
                ("tol", C.c_double), ("omega", C.c_double), ("precision", C.c_int)]


class Ns2dParams(C.Structure):
    _fields_ = [("gamma", C.c_double), ("cfl", C.c_double), ("x0", C.c_double), ("x1", C.c_double),
                ("y0", C.c_double), ("y1", C.c_double), ("Nx", C.c_int), ("Ny", C.c_int),
                ("floor_rho", C.c_double), ("floor_P", C.c_double), ("viscosity_on", C.c_int),
                ("mu", C.c_double), ("cfl_visc", C.c_double), ("gravity_on", C.c_int), ("G", C.c_double),
                ("precision", C.c_int)]


class Euler3dParams(C.Structure):
    _fields_ = [("Nx", C.c_int), ("Ny", C.c_int), ("Nz", C.c_int)] + \
               [(n, C.c_double) for n in ("x0", "x1", "y0", "y1", "z0", "z1", "gamma", "cfl", "floor_rho", "floor_P")] + \
               [("gravity_on", C.c_int), ("G", C.c_double), ("poisson_tol", C.c_double),
                ("poisson_max_iter", C.c_int), ("precision", C.c_int)]


class Sphere3dIc(C.Structure):
    _fields_ = [("seed", C.c_uint)] + [(n, C.c_double) for n in ("rho_in", "rho_out", "radius", "noise", "P0")]


class Mhd2dIc(C.Structure):
    _fields_ = [(n, C.c_double) for n in ("rho_bg", "rho_core", "r_core", "r_edge", "P0", "vin", "rot_omega",
                                           "B0x", "Bshear", "gamma_eff")]


class Collapse2dIc(C.Structure):
    _fields_ = [("seed", C.c_uint)] + [(n, C.c_double) for n in ("contrast", "width", "P0", "vamp", "B0")]


class Mhd3dIc(C.Structure):
    _fields_ = [(n, C.c_double) for n in ("rho_bg", "rho_core", "r_core", "r_edge", "P0", "vin", "rot_omega",
                                           "B0x", "B0z", "Bshear", "gamma_init")]


class Turbulence3dIc(C.Structure):
    _fields_ = [("seed", C.c_uint), ("kmin", C.c_int), ("kmax", C.c_int), ("mach", C.c_double),
                ("beta", C.c_double)]


class Euler2dIc(C.Structure):
    _fields_ = [("seed", C.c_uint)] + [(n, C.c_double) for n in ("rho0", "P0", "amp_rho", "amp_P", "vmax",
                                                                  "sigma", "bump_amp")]


class Ns2dIc(C.Structure):
    _fields_ = [("seed", C.c_uint)] + [(n, C.c_double) for n in ("rho0", "P0", "A", "sigma_frac", "Arho",
                                                                  "Mach0")]


FRAME_FIELDS = {
    "mhd2d": ("rho", "P", "Mach", "u", "v", "vabs", "Bx", "By", "Bz", "divB"),
    "mhd3d": ("rho", "P", "Mach", "u", "v", "w", "vabs", "Bx", "By", "Bz", "divB"),
    "euler2d": ("rho", "P", "Mach"),
    "ns2d": ("rho", "P", "Mach", "u", "v", "momx", "momy", "Phi"),
}

# name -> (restype, argtypes); every function returns cc_status unless stated
_PROTOS = {
    "cc_last_error": (C.c_char_p, []),
    "cc_version": (C.c_int, []),
    "cc_device_info": (C.c_int, [C.c_char_p, C.c_int, C.POINTER(C.c_int)]),
    "cc_mhd2d_create": (C.c_int, [C.POINTER(Mhd2dParams), C.POINTER(C.c_void_p)]),
    "cc_mhd2d_destroy": (C.c_int, [C.c_void_p]),
    "cc_mhd2d_upload": (C.c_int, [C.c_void_p] + [DP] * 8),
    "cc_mhd2d_download": (C.c_int, [C.c_void_p] + [DP] * 8),
    "cc_mhd2d_gravity": (C.c_int, [C.c_void_p]),
    "cc_mhd2d_get_gravity": (C.c_int, [C.c_void_p, DP, DP]),
    "cc_mhd2d_set_gravity": (C.c_int, [C.c_void_p, DP, DP]),
    "cc_mhd2d_compute_dt": (C.c_int, [C.c_void_p, DP]),
    "cc_mhd2d_stage": (C.c_int, [C.c_void_p, C.c_double]),
    "cc_mhd2d_rk2_step": (C.c_int, [C.c_void_p, C.c_double]),
    "cc_mhd2d_run": (C.c_int, [C.c_void_p, C.c_int, DP, C.c_double, DP]),
    "cc_mhd2d_divb_max": (C.c_int, [C.c_void_p, DP]),
    "cc_mhd2d_sync": (C.c_int, [C.c_void_p]),
    "cc_mhd2d_set_profiling": (C.c_int, [C.c_void_p, C.c_int]),
    "cc_mhd2d_kernel_time": (C.c_int, [C.c_void_p, C.c_char_p, C.POINTER(C.c_int), DP]),
    "cc_mhd2d_reset_profiling": (C.c_int, [C.c_void_p]),
    "cc_hlld2d_batch": (C.c_int, [C.c_int, C.c_int, DP, DP, DP, C.c_double, C.c_double, C.c_int, DP]),
    "cc_mhd3d_create": (C.c_int, [C.POINTER(Mhd3dParams), C.POINTER(C.c_void_p)]),
    "cc_mhd3d_destroy": (C.c_int, [C.c_void_p]),
    "cc_mhd3d_upload": (C.c_int, [C.c_void_p] + [DP] * 8),
    "cc_mhd3d_download": (C.c_int, [C.c_void_p] + [DP] * 8),
    "cc_mhd3d_compute_dt": (C.c_int, [C.c_void_p, DP]),
    "cc_mhd3d_stage": (C.c_int, [C.c_void_p, C.c_double]),
    "cc_mhd3d_rk2_step": (C.c_int, [C.c_void_p, C.c_double]),
    "cc_mhd3d_run": (C.c_int, [C.c_void_p, C.c_int, DP, C.c_double, DP]),
    "cc_mhd3d_divb_max": (C.c_int, [C.c_void_p, DP]),
    "cc_mhd3d_sync": (C.c_int, [C.c_void_p]),
    "cc_mhd3d_set_profiling": (C.c_int, [C.c_void_p, C.c_int]),
    "cc_mhd3d_kernel_time": (C.c_int, [C.c_void_p, C.c_char_p, C.POINTER(C.c_int), DP]),
    "cc_mhd3d_reset_profiling": (C.c_int, [C.c_void_p]),
    "cc_hlld3d_batch": (C.c_int, [C.POINTER(Mhd3dParams), C.c_int, DP, DP, DP, DP]),
    "cc_euler2d_create": (C.c_int, [C.POINTER(Euler2dParams), C.POINTER(C.c_void_p)]),
    "cc_euler2d_destroy": (C.c_int, [C.c_void_p]),
    "cc_euler2d_upload": (C.c_int, [C.c_void_p] + [DP] * 4),
    "cc_euler2d_download": (C.c_int, [C.c_void_p] + [DP] * 4),
    "cc_euler2d_gravity": (C.c_int, [C.c_void_p, DP, DP, DP, C.POINTER(C.c_int)]),
    "cc_euler2d_compute_dt": (C.c_int, [C.c_void_p, DP]),
    "cc_euler2d_rhs": (C.c_int, [C.c_void_p] + [DP] * 4),
    "cc_euler2d_rk2_step": (C.c_int, [C.c_void_p, C.c_double]),
    "cc_euler2d_run": (C.c_int, [C.c_void_p, C.c_int, DP, C.c_double, DP]),
    "cc_euler2d_sweep_stats": (C.c_int, [C.c_void_p, C.POINTER(C.c_int), C.POINTER(C.c_longlong),
                                         C.POINTER(C.c_int)]),
    "cc_euler2d_sync": (C.c_int, [C.c_void_p]),
    "cc_euler2d_set_profiling": (C.c_int, [C.c_void_p, C.c_int]),
    "cc_euler2d_kernel_time": (C.c_int, [C.c_void_p, C.c_char_p, C.POINTER(C.c_int), DP]),
    "cc_euler2d_reset_profiling": (C.c_int, [C.c_void_p]),
    "cc_hllc2d_batch": (C.c_int, [C.c_int, C.c_int, C.c_int, DP, DP, C.c_double, C.c_double, C.c_int, DP]),
    "cc_ns2d_create": (C.c_int, [C.POINTER(Ns2dParams), C.POINTER(C.c_void_p)]),
    "cc_ns2d_destroy": (C.c_int, [C.c_void_p]),
    "cc_ns2d_upload": (C.c_int, [C.c_void_p, DP]),
    "cc_ns2d_download": (C.c_int, [C.c_void_p, DP]),
    "cc_ns2d_poisson": (C.c_int, [C.c_void_p, DP, DP]),
    "cc_ns2d_rhs": (C.c_int, [C.c_void_p, DP]),
    "cc_ns2d_compute_dt": (C.c_int, [C.c_void_p, DP]),
    "cc_ns2d_heun_step": (C.c_int, [C.c_void_p, C.c_double]),
    "cc_ns2d_run": (C.c_int, [C.c_void_p, C.c_int, DP, C.c_double, DP]),
    "cc_ns2d_sync": (C.c_int, [C.c_void_p]),
    "cc_ns2d_set_profiling": (C.c_int, [C.c_void_p, C.c_int]),
    "cc_ns2d_kernel_time": (C.c_int, [C.c_void_p, C.c_char_p, C.POINTER(C.c_int), DP]),
    "cc_ns2d_reset_profiling": (C.c_int, [C.c_void_p]),
    "cc_fft2d": (C.c_int, [C.c_int, C.c_int, C.c_int, DP]),
    "cc_fft2d_noalloc": (C.c_int, [C.c_int, C.c_int, C.c_int, C.c_void_p, C.c_void_p]),
    "cc_copy_calibration": (C.c_int, [C.c_size_t, C.c_int, DP]),
    "cc_poisson2d_sor": (C.c_int, [C.c_int, C.c_int, C.c_int, C.c_int, C.c_double, C.c_double, C.c_double,
                                   C.c_int, C.c_double, C.c_int, DP, DP, DP, DP]),
    "cc_mhd2d_frame": (C.c_int, [C.c_void_p, DP]),
    "cc_mhd2d_totals": (C.c_int, [C.c_void_p, DP]),
    "cc_mhd2d_compare": (C.c_int, [C.c_void_p, C.c_void_p, DP]),
    "cc_mhd3d_frame": (C.c_int, [C.c_void_p, DP]),
    "cc_mhd3d_totals": (C.c_int, [C.c_void_p, DP]),
    "cc_mhd3d_compare": (C.c_int, [C.c_void_p, C.c_void_p, DP]),
    "cc_euler2d_frame": (C.c_int, [C.c_void_p, DP]),
    "cc_euler2d_totals": (C.c_int, [C.c_void_p, DP]),
    "cc_euler2d_compare": (C.c_int, [C.c_void_p, C.c_void_p, DP]),
    "cc_ns2d_frame": (C.c_int, [C.c_void_p, DP]),
    "cc_ns2d_totals": (C.c_int, [C.c_void_p, DP]),
    "cc_ns2d_compare": (C.c_int, [C.c_void_p, C.c_void_p, DP]),
    "cc_mhd2d_ic_defaults": (None, [C.POINTER(Mhd2dIc)]),
    "cc_mhd2d_ic_reference": (C.c_int, [C.POINTER(Mhd2dParams), C.POINTER(Mhd2dIc)] + [DP] * 8),
    "cc_collapse2d_ic_defaults": (None, [C.POINTER(Collapse2dIc)]),
    "cc_mhd2d_ic_collapse": (C.c_int, [C.POINTER(Mhd2dParams), C.POINTER(Collapse2dIc)] + [DP] * 8),
    "cc_mhd3d_ic_defaults": (None, [C.POINTER(Mhd3dIc)]),
    "cc_mhd3d_ic_reference": (C.c_int, [C.POINTER(Mhd3dParams), C.POINTER(Mhd3dIc)] + [DP] * 8),
    "cc_turbulence3d_ic_defaults": (None, [C.POINTER(Turbulence3dIc)]),
    "cc_mhd3d_ic_turbulence": (C.c_int, [C.POINTER(Mhd3dParams), C.POINTER(Turbulence3dIc)] + [DP] * 8),
    "cc_euler3d_create": (C.c_int, [C.POINTER(Euler3dParams), C.POINTER(C.c_void_p)]),
    "cc_euler3d_destroy": (C.c_int, [C.c_void_p]),
    "cc_euler3d_upload": (C.c_int, [C.c_void_p] + [DP] * 5),
    "cc_euler3d_download": (C.c_int, [C.c_void_p] + [DP] * 5),
    "cc_euler3d_gravity": (C.c_int, [C.c_void_p] + [DP] * 4),
    "cc_euler3d_poisson_stats": (C.c_int, [C.c_void_p, C.POINTER(C.c_int), C.POINTER(C.c_int), C.POINTER(C.c_int),
                                           C.POINTER(C.c_int), DP, DP, C.POINTER(C.c_int)]),
    "cc_euler3d_rhs": (C.c_int, [C.c_void_p] + [DP] * 5),
    "cc_euler3d_compute_dt": (C.c_int, [C.c_void_p, DP]),
    "cc_euler3d_rk2_step": (C.c_int, [C.c_void_p, C.c_double]),
    "cc_euler3d_run": (C.c_int, [C.c_void_p, C.c_int, DP, C.c_double, DP]),
    "cc_euler3d_totals": (C.c_int, [C.c_void_p, DP]),
    "cc_euler3d_sync": (C.c_int, [C.c_void_p]),
    "cc_euler3d_set_profiling": (C.c_int, [C.c_void_p, C.c_int]),
    "cc_euler3d_kernel_time": (C.c_int, [C.c_void_p, C.c_char_p, C.POINTER(C.c_int), DP]),
    "cc_euler3d_reset_profiling": (C.c_int, [C.c_void_p]),
    "cc_sphere3d_ic_defaults": (None, [C.POINTER(Sphere3dIc)]),
    "cc_euler3d_ic_sphere": (C.c_int, [C.POINTER(Euler3dParams), C.POINTER(Sphere3dIc)] + [DP] * 5),
    "cc_euler2d_ic_defaults": (None, [C.POINTER(Euler2dIc)]),
    "cc_euler2d_ic_reference": (C.c_int, [C.POINTER(Euler2dParams), C.POINTER(Euler2dIc)] + [DP] * 4),
    "cc_sor_autotune": (C.c_int, [C.c_int]),
    "cc_sor_tile_choice": (C.c_int, [C.c_int, C.c_int, C.c_int, C.c_int] + [C.POINTER(C.c_int)] * 4),
    "cc_sor_flow_schedule": (C.c_int, [C.c_int] * 7 + [C.POINTER(C.c_int)] * 2 + [C.c_void_p] * 3),
    "cc_mhd2d_batch_create": (C.c_int, [C.POINTER(C.c_void_p), C.c_int, C.POINTER(C.c_void_p)]),
    "cc_mhd2d_batch_destroy": (C.c_int, [C.c_void_p]),
    "cc_mhd2d_batch_info": (C.c_int, [C.c_void_p] + [C.POINTER(C.c_int)] * 5),
    "cc_mhd2d_batch_run": (C.c_int, [C.c_void_p, C.c_int, DP, DP]),
    "cc_collapse2d_batch_sample": (C.c_int, [C.c_uint, C.c_int, C.c_int, C.c_int, C.POINTER(C.c_int),
                                             C.POINTER(Collapse2dIc), DP]),
    "cc_ns2d_ic_defaults": (None, [C.POINTER(Ns2dIc)]),
    "cc_ns2d_ic_reference": (C.c_int, [C.POINTER(Ns2dParams), C.POINTER(Ns2dIc), DP]),
}

_lib = None


def lib():
    """Load libccamd.so once; raise if it is absent (no fallback)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError("libccamd.so not built at %s: run __graft_entry__.build() "
                               "(make -C core_collapse_amd/csrc)" % LIB_PATH)
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in _PROTOS.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def exported_symbols():
    return list(_PROTOS)


def check(status):
    if status != CC_OK:
        raise CCError(status, lib().cc_last_error().decode(errors="replace"))


def device_info():
    name = C.create_string_buffer(256)
    cu = C.c_int(0)
    check(lib().cc_device_info(name, 256, C.byref(cu)))
    return name.value.decode(), cu.value


# ---------------------------------------------------------------------------------------- diagnostics
TOTALS = {
    "mhd2d": ("mass", "momx", "momy", "momz", "energy", "divB_max"),
    "mhd3d": ("mass", "momx", "momy", "momz", "energy", "divB_max"),
    "euler2d": ("mass", "momx", "momy", "energy"),
    "ns2d": ("mass", "momx", "momy", "energy"),
    "euler3d": ("mass", "momx", "momy", "momz", "energy"),
}


def frame_of(kind, h, shape):
    """The snapshot frame of handle h (cc_<kind>_frame) as {field: array of `shape`}."""
    import numpy as np
    names = FRAME_FIELDS[kind]
    buf = np.empty((len(names),) + tuple(shape))
    check(getattr(lib(), "cc_%s_frame" % kind)(h, buf.ctypes.data_as(DP)))
    return {n: buf[i] for i, n in enumerate(names)}


def totals_of(kind, h):
    """In-situ totals of handle h (cc_<kind>_totals) as {name: value}."""
    import numpy as np
    names = TOTALS[kind]
    out = np.empty(len(names))
    check(getattr(lib(), "cc_%s_totals" % kind)(h, out.ctypes.data_as(DP)))
    return dict(zip(names, out.tolist()))


def compare_of(kind, ha, hb):
    """Frame comparison of two handles (cc_<kind>_compare): {field: (mean|a-b|, sum|a-b|/sum|b|, max|a-b|)}."""
    import numpy as np
    names = FRAME_FIELDS[kind]
    out = np.empty(3 * len(names))
    check(getattr(lib(), "cc_%s_compare" % kind)(ha, hb, out.ctypes.data_as(DP)))
    return {n: tuple(out[3 * i:3 * i + 3].tolist()) for i, n in enumerate(names)}
