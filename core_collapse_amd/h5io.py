"""Reading the programs' HDF5 snapshot files from Python without h5py: a ctypes binding of the HDF5
C library the programs link (conda's libhdf5, 1.10).

    info(path)                 -> {"/rho": {"shape": (Nt, Ny, Nx), "f64le": True, "layout": "CONTIGUOUS"}, ...}
    read(path, name)           -> numpy array of the whole dataset
    frames(path, name)         -> iterator over the dataset's leading-axis slices (one frame at a time)
"""
import ctypes as C
import os

import numpy as np

_HDF5_CANDIDATES = (os.environ.get("CC_HDF5_LIB", ""), "/opt/conda/lib/libhdf5.so", "libhdf5.so")
_LAYOUTS = {0: "COMPACT", 1: "CONTIGUOUS", 2: "CHUNKED", 3: "VIRTUAL"}
_h5 = None
hid_t = C.c_int64
hsize_t = C.c_ulonglong


class _GInfo(C.Structure):
    _fields_ = [("storage_type", C.c_int), ("nlinks", hsize_t), ("max_corder", C.c_int64), ("mounted", C.c_bool)]


def _lib():
    global _h5
    if _h5 is None:
        err = None
        for cand in _HDF5_CANDIDATES:
            if not cand:
                continue
            try:
                L = C.CDLL(cand)
                break
            except OSError as e:
                err = e
        else:
            raise OSError("libhdf5 not found (set CC_HDF5_LIB): %s" % err)
        L.H5open()
        for name, res, args in (
                ("H5Fopen", hid_t, [C.c_char_p, C.c_uint, hid_t]), ("H5Fclose", C.c_int, [hid_t]),
                ("H5Dopen2", hid_t, [hid_t, C.c_char_p, hid_t]), ("H5Dclose", C.c_int, [hid_t]),
                ("H5Dget_space", hid_t, [hid_t]), ("H5Dget_type", hid_t, [hid_t]),
                ("H5Dget_create_plist", hid_t, [hid_t]), ("H5Pget_layout", C.c_int, [hid_t]),
                ("H5Pclose", C.c_int, [hid_t]), ("H5Tequal", C.c_int, [hid_t, hid_t]), ("H5Tclose", C.c_int, [hid_t]),
                ("H5Sget_simple_extent_ndims", C.c_int, [hid_t]),
                ("H5Sget_simple_extent_dims", C.c_int, [hid_t, C.POINTER(hsize_t), C.POINTER(hsize_t)]),
                ("H5Screate_simple", hid_t, [C.c_int, C.POINTER(hsize_t), C.POINTER(hsize_t)]),
                ("H5Sselect_hyperslab", C.c_int, [hid_t, C.c_int, C.POINTER(hsize_t), C.POINTER(hsize_t),
                                                  C.POINTER(hsize_t), C.POINTER(hsize_t)]),
                ("H5Sclose", C.c_int, [hid_t]),
                ("H5Dread", C.c_int, [hid_t, hid_t, hid_t, hid_t, hid_t, C.c_void_p]),
                ("H5Gget_info", C.c_int, [hid_t, C.POINTER(_GInfo)]),
                ("H5Lget_name_by_idx", C.c_ssize_t, [hid_t, C.c_char_p, C.c_int, C.c_int, hsize_t, C.c_char_p,
                                                     C.c_size_t, hid_t])):
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        L.native_double = hid_t.in_dll(L, "H5T_NATIVE_DOUBLE_g").value
        L.ieee_f64le = hid_t.in_dll(L, "H5T_IEEE_F64LE_g").value
        _h5 = L
    return _h5


class _File:
    def __init__(self, path):
        self.L = _lib()
        self.f = self.L.H5Fopen(os.fsencode(path), 0, 0)  # H5F_ACC_RDONLY, H5P_DEFAULT
        if self.f < 0:
            raise OSError("cannot open %s" % path)

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.L.H5Fclose(self.f)

    def names(self):
        gi = _GInfo()
        self.L.H5Gget_info(self.f, C.byref(gi))
        out = []
        for i in range(gi.nlinks):
            buf = C.create_string_buffer(512)
            self.L.H5Lget_name_by_idx(self.f, b".", 0, 0, i, buf, 512, 0)
            out.append("/" + buf.value.decode())
        return out

    def shape(self, d):
        sp = self.L.H5Dget_space(d)
        n = self.L.H5Sget_simple_extent_ndims(sp)
        dims = (hsize_t * max(n, 1))()
        self.L.H5Sget_simple_extent_dims(sp, dims, None)
        self.L.H5Sclose(sp)
        return tuple(int(dims[i]) for i in range(n))


def info(path):
    """Dataset name -> shape, whether the stored type is IEEE F64LE, storage layout."""
    out = {}
    with _File(path) as f:
        L = f.L
        for name in f.names():
            d = L.H5Dopen2(f.f, name.encode(), 0)
            t = L.H5Dget_type(d)
            pl = L.H5Dget_create_plist(d)
            out[name] = {"shape": f.shape(d), "f64le": L.H5Tequal(t, L.ieee_f64le) > 0,
                         "layout": _LAYOUTS.get(L.H5Pget_layout(pl), "?")}
            L.H5Pclose(pl)
            L.H5Tclose(t)
            L.H5Dclose(d)
    return out


def read(path, name):
    """The whole dataset `name` as float64."""
    with _File(path) as f:
        d = f.L.H5Dopen2(f.f, name.encode(), 0)
        if d < 0:
            raise KeyError(name)
        a = np.empty(f.shape(d))
        st = f.L.H5Dread(d, f.L.native_double, 0, 0, 0, a.ctypes.data)  # H5S_ALL, H5S_ALL, H5P_DEFAULT
        f.L.H5Dclose(d)
        if st < 0:
            raise OSError("read of %s failed" % name)
        return a


def frames(path, name):
    """Yield the leading-axis slices of dataset `name` one at a time (bounded memory)."""
    with _File(path) as f:
        L = f.L
        d = L.H5Dopen2(f.f, name.encode(), 0)
        if d < 0:
            raise KeyError(name)
        shape = f.shape(d)
        rank = len(shape)
        count = (hsize_t * rank)(*((1,) + shape[1:]))
        mem = L.H5Screate_simple(rank, count, None)
        buf = np.empty(shape[1:])
        try:
            for k in range(shape[0]):
                start = (hsize_t * rank)(*((k,) + (0,) * (rank - 1)))
                sp = L.H5Dget_space(d)
                L.H5Sselect_hyperslab(sp, 0, start, None, count, None)  # H5S_SELECT_SET
                st = L.H5Dread(d, L.native_double, mem, sp, 0, buf.ctypes.data)
                L.H5Sclose(sp)
                if st < 0:
                    raise OSError("read of %s frame %d failed" % (name, k))
                yield buf.copy()
        finally:
            L.H5Sclose(mem)
            L.H5Dclose(d)
