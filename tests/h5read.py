"""Minimal HDF5 reader over the HDF5 C library (ctypes), for checking dump
files in tests without the writer's own code (h5py is not installed)."""
import ctypes
from ctypes import POINTER, byref, c_char_p, c_int, c_int64, c_uint64, c_void_p

import numpy as np

HDF5_LIB = "/opt/conda/lib/libhdf5.so"
_L = None


def lib():
    global _L
    if _L is None:
        L = ctypes.CDLL(HDF5_LIB)
        L.H5open.restype = c_int
        L.H5open()
        for name in ("H5Fopen", "H5Dopen2", "H5Dget_space", "H5Aopen", "H5Aget_space", "H5Oopen"):
            getattr(L, name).restype = c_int64
        L.H5Fopen.argtypes = [c_char_p, ctypes.c_uint, c_int64]
        L.H5Dopen2.argtypes = [c_int64, c_char_p, c_int64]
        L.H5Dget_space.argtypes = [c_int64]
        L.H5Dread.argtypes = [c_int64, c_int64, c_int64, c_int64, c_int64, c_void_p]
        L.H5Oopen.argtypes = [c_int64, c_char_p, c_int64]
        L.H5Aopen.argtypes = [c_int64, c_char_p, c_int64]
        L.H5Aget_space.argtypes = [c_int64]
        L.H5Aread.argtypes = [c_int64, c_int64, c_void_p]
        L.H5Sget_simple_extent_ndims.argtypes = [c_int64]
        L.H5Sget_simple_extent_dims.argtypes = [c_int64, POINTER(c_uint64), POINTER(c_uint64)]
        L.H5Lexists.argtypes = [c_int64, c_char_p, c_int64]
        for name in ("H5Dclose", "H5Sclose", "H5Aclose", "H5Oclose", "H5Fclose"):
            getattr(L, name).argtypes = [c_int64]
        _L = L
    return _L


def _type(kind):
    L = lib()
    return c_int64.in_dll(L, "H5T_NATIVE_FLOAT_g" if kind == np.float32 else "H5T_NATIVE_DOUBLE_g").value


def _dims(space):
    L = lib()
    nd = L.H5Sget_simple_extent_ndims(space)
    d = (c_uint64 * max(nd, 1))()
    L.H5Sget_simple_extent_dims(space, d, None)
    return tuple(int(d[i]) for i in range(nd))


class File:
    def __init__(self, path):
        self.L = lib()
        self.f = self.L.H5Fopen(str(path).encode(), 0, 0)
        assert self.f >= 0, path

    def exists(self, name):
        # walk the path so missing intermediate groups read as absent
        parts = [p for p in name.split("/") if p]
        for k in range(1, len(parts) + 1):
            if self.L.H5Lexists(self.f, ("/" + "/".join(parts[:k])).encode(), 0) <= 0:
                return False
        return True

    def dataset(self, name, dtype=np.float32):
        L = self.L
        d = L.H5Dopen2(self.f, name.encode(), 0)
        assert d >= 0, name
        sp = L.H5Dget_space(d)
        shape = _dims(sp)
        out = np.empty(shape, dtype=dtype)
        assert L.H5Dread(d, _type(dtype), 0, 0, 0, out.ctypes.data_as(c_void_p)) >= 0
        L.H5Sclose(sp)
        L.H5Dclose(d)
        return out

    def attribute(self, obj, name, dtype=np.float64):
        L = self.L
        o = L.H5Oopen(self.f, obj.encode(), 0)
        assert o >= 0, obj
        a = L.H5Aopen(o, name.encode(), 0)
        assert a >= 0, (obj, name)
        sp = L.H5Aget_space(a)
        shape = _dims(sp)
        out = np.empty(shape, dtype=dtype)
        assert L.H5Aread(a, _type(dtype), out.ctypes.data_as(c_void_p)) >= 0
        L.H5Sclose(sp)
        L.H5Aclose(a)
        L.H5Oclose(o)
        return out

    def close(self):
        if self.f >= 0:
            self.L.H5Fclose(self.f)
            self.f = -1
