"""Minimal read-only HDF5 access for the tests (h5py is not installed).

ctypes over the serial HDF5 1.10 library the driver links (libhdf5.so and
libhdf5_hl.so under NS2D_HDF5_PREFIX, default /opt/conda): list a group's
links, read a dataset (double, int or the {r, i} compound as complex128), read
a double attribute and report a compound type's member names.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

PREFIX = os.environ.get("NS2D_HDF5_PREFIX", "/opt/conda")
_h5 = None

hid_t = ctypes.c_int64
hsize_t = ctypes.c_ulonglong
H5P_DEFAULT = 0
H5F_ACC_RDONLY = 0
H5S_ALL = 0
H5T_COMPOUND = 6


def lib():
    global _h5
    if _h5 is None:
        _h5 = ctypes.CDLL(os.path.join(PREFIX, "lib", "libhdf5.so"))
        _h5.H5open()
        for name, res, args in [
            ("H5Fopen", hid_t, [ctypes.c_char_p, ctypes.c_uint, hid_t]),
            ("H5Fclose", ctypes.c_int, [hid_t]),
            ("H5Oopen", hid_t, [hid_t, ctypes.c_char_p, hid_t]),
            ("H5Oclose", ctypes.c_int, [hid_t]),
            ("H5Dopen2", hid_t, [hid_t, ctypes.c_char_p, hid_t]),
            ("H5Dclose", ctypes.c_int, [hid_t]),
            ("H5Dget_space", hid_t, [hid_t]),
            ("H5Dget_type", hid_t, [hid_t]),
            ("H5Dread", ctypes.c_int, [hid_t, hid_t, hid_t, hid_t, hid_t, ctypes.c_void_p]),
            ("H5Sget_simple_extent_ndims", ctypes.c_int, [hid_t]),
            ("H5Sget_simple_extent_dims", ctypes.c_int, [hid_t, ctypes.POINTER(hsize_t), ctypes.POINTER(hsize_t)]),
            ("H5Sclose", ctypes.c_int, [hid_t]),
            ("H5Tget_class", ctypes.c_int, [hid_t]),
            ("H5Tget_size", ctypes.c_size_t, [hid_t]),
            ("H5Tget_nmembers", ctypes.c_int, [hid_t]),
            ("H5Tget_member_name", ctypes.c_void_p, [hid_t, ctypes.c_uint]),
            ("H5Tget_member_offset", ctypes.c_size_t, [hid_t, ctypes.c_uint]),
            ("H5Tcreate", hid_t, [ctypes.c_int, ctypes.c_size_t]),
            ("H5Tinsert", ctypes.c_int, [hid_t, ctypes.c_char_p, ctypes.c_size_t, hid_t]),
            ("H5Tclose", ctypes.c_int, [hid_t]),
            ("H5free_memory", ctypes.c_int, [ctypes.c_void_p]),
            ("H5Lexists", ctypes.c_int, [hid_t, ctypes.c_char_p, hid_t]),
            ("H5Lget_name_by_idx", ctypes.c_ssize_t,
             [hid_t, ctypes.c_char_p, ctypes.c_int, ctypes.c_int, hsize_t, ctypes.c_char_p, ctypes.c_size_t, hid_t]),
            ("H5Aopen", hid_t, [hid_t, ctypes.c_char_p, hid_t]),
            ("H5Aread", ctypes.c_int, [hid_t, hid_t, ctypes.c_void_p]),
            ("H5Aclose", ctypes.c_int, [hid_t]),
        ]:
            f = getattr(_h5, name)
            f.restype = res
            f.argtypes = args
        _h5.H5Eset_auto2.argtypes = [hid_t, ctypes.c_void_p, ctypes.c_void_p]
        _h5.H5Eset_auto2(0, None, None)   # silence the error stack printer
    return _h5


def _native(name: str) -> int:
    return hid_t.in_dll(lib(), name).value


class File:
    def __init__(self, path: str):
        self.h = lib().H5Fopen(path.encode(), H5F_ACC_RDONLY, H5P_DEFAULT)
        if self.h < 0:
            raise OSError(f"cannot open {path}")

    def close(self):
        if self.h >= 0:
            lib().H5Fclose(self.h)
            self.h = -1

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def keys(self, group: str = "/") -> list[str]:
        out, i = [], 0
        buf = ctypes.create_string_buffer(512)
        while True:
            n = lib().H5Lget_name_by_idx(self.h, group.encode(), 0, 0, i, buf, 512, H5P_DEFAULT)
            if n < 0:
                return out
            out.append(buf.value.decode())
            i += 1

    def exists(self, path: str) -> bool:
        return lib().H5Lexists(self.h, path.encode(), H5P_DEFAULT) > 0

    def members(self, path: str) -> list[tuple[str, int]]:
        d = lib().H5Dopen2(self.h, path.encode(), H5P_DEFAULT)
        t = lib().H5Dget_type(d)
        out = []
        if lib().H5Tget_class(t) == H5T_COMPOUND:
            for m in range(lib().H5Tget_nmembers(t)):
                p = lib().H5Tget_member_name(t, m)
                out.append((ctypes.string_at(p).decode(), int(lib().H5Tget_member_offset(t, m))))
                lib().H5free_memory(p)
        lib().H5Tclose(t)
        lib().H5Dclose(d)
        return out

    def read(self, path: str) -> np.ndarray:
        L = lib()
        d = L.H5Dopen2(self.h, path.encode(), H5P_DEFAULT)
        if d < 0:
            raise KeyError(path)
        sp = L.H5Dget_space(d)
        nd = L.H5Sget_simple_extent_ndims(sp)
        dims = (hsize_t * max(nd, 1))()
        L.H5Sget_simple_extent_dims(sp, dims, None)
        L.H5Sclose(sp)
        shape = tuple(int(dims[k]) for k in range(nd))
        t = L.H5Dget_type(d)
        cls = L.H5Tget_class(t)
        size = L.H5Tget_size(t)
        L.H5Tclose(t)
        if cls == H5T_COMPOUND:
            mt = L.H5Tcreate(H5T_COMPOUND, 16)
            dbl = _native("H5T_NATIVE_DOUBLE_g")
            L.H5Tinsert(mt, b"r", 0, dbl)
            L.H5Tinsert(mt, b"i", 8, dbl)
            out = np.empty(shape, np.complex128)
            rc = L.H5Dread(d, mt, H5S_ALL, H5S_ALL, H5P_DEFAULT, out.ctypes.data)
            L.H5Tclose(mt)
        elif cls == 0:   # H5T_INTEGER
            out = np.empty(shape, np.int32)
            rc = L.H5Dread(d, _native("H5T_NATIVE_INT_g"), H5S_ALL, H5S_ALL, H5P_DEFAULT, out.ctypes.data)
        else:
            out = np.empty(shape, np.float64)
            rc = L.H5Dread(d, _native("H5T_NATIVE_DOUBLE_g"), H5S_ALL, H5S_ALL, H5P_DEFAULT, out.ctypes.data)
        L.H5Dclose(d)
        if rc < 0:
            raise OSError(f"read of {path} failed (type class {cls}, size {size})")
        return out

    def attr(self, obj: str, name: str) -> float:
        L = lib()
        o = L.H5Oopen(self.h, obj.encode(), H5P_DEFAULT)
        a = L.H5Aopen(o, name.encode(), H5P_DEFAULT)
        v = ctypes.c_double()
        rc = L.H5Aread(a, _native("H5T_NATIVE_DOUBLE_g"), ctypes.byref(v))
        L.H5Aclose(a)
        L.H5Oclose(o)
        if rc < 0:
            raise KeyError(f"{obj}@{name}")
        return v.value


def write_double_dataset(path: str, name: str, data: np.ndarray) -> None:
    """Create `path` holding one double dataset (test inputs such as the
    MAX_PALIN stream function file)."""
    L = lib()
    L.H5Fcreate.restype = hid_t
    L.H5Fcreate.argtypes = [ctypes.c_char_p, ctypes.c_uint, hid_t, hid_t]
    L.H5Screate_simple.restype = hid_t
    L.H5Screate_simple.argtypes = [ctypes.c_int, ctypes.POINTER(hsize_t), ctypes.POINTER(hsize_t)]
    L.H5Dcreate2.restype = hid_t
    L.H5Dcreate2.argtypes = [hid_t, ctypes.c_char_p, hid_t, hid_t, hid_t, hid_t, hid_t]
    L.H5Dwrite.restype = ctypes.c_int
    L.H5Dwrite.argtypes = [hid_t, hid_t, hid_t, hid_t, hid_t, ctypes.c_void_p]
    a = np.ascontiguousarray(data, dtype=np.float64)
    dims = (hsize_t * a.ndim)(*a.shape)
    f = L.H5Fcreate(path.encode(), 2, H5P_DEFAULT, H5P_DEFAULT)   # H5F_ACC_TRUNC
    sp = L.H5Screate_simple(a.ndim, dims, None)
    dbl = _native("H5T_NATIVE_DOUBLE_g")
    d = L.H5Dcreate2(f, name.encode(), dbl, sp, H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT)
    rc = L.H5Dwrite(d, dbl, H5S_ALL, H5S_ALL, H5P_DEFAULT, a.ctypes.data)
    L.H5Dclose(d)
    L.H5Sclose(sp)
    L.H5Fclose(f)
    if f < 0 or rc < 0:
        raise OSError(f"writing {path} failed")
