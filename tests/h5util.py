"""Minimal HDF5 reader for the tests (ctypes over libhdf5 >= 1.10), kept apart
from the library's own loader so the files are checked by independent code."""
from __future__ import annotations

import ctypes
import os

import numpy as np

_CANDIDATES = (os.environ.get("METALBM_HDF5_LIBRARY"), "libhdf5.so", "libhdf5.so.103",
               "/opt/conda/lib/libhdf5.so", "/opt/conda/lib/libhdf5.so.103")
_lib = None


def available() -> bool:
    try:
        lib()
        return True
    except OSError:
        return False


def lib():
    global _lib
    if _lib is None:
        err = None
        for c in _CANDIDATES:
            if not c:
                continue
            try:
                L = ctypes.CDLL(c)
                break
            except OSError as e:
                err = e
        else:
            raise OSError(f"no libhdf5: {err}")
        hid = ctypes.c_int64
        L.H5open.restype = ctypes.c_int
        L.H5open()
        L.H5Eset_auto2.argtypes = [hid, ctypes.c_void_p, ctypes.c_void_p]
        L.H5Eset_auto2(0, None, None)
        L.H5Fopen.argtypes = [ctypes.c_char_p, ctypes.c_uint, hid]
        L.H5Fopen.restype = hid
        L.H5Fclose.argtypes = [hid]
        L.H5Lexists.argtypes = [hid, ctypes.c_char_p, hid]
        L.H5Lexists.restype = ctypes.c_int
        L.H5Dopen2.argtypes = [hid, ctypes.c_char_p, hid]
        L.H5Dopen2.restype = hid
        L.H5Dget_space.argtypes = [hid]
        L.H5Dget_space.restype = hid
        L.H5Sget_simple_extent_ndims.argtypes = [hid]
        L.H5Sget_simple_extent_dims.argtypes = [hid, ctypes.POINTER(ctypes.c_ulonglong),
                                               ctypes.c_void_p]
        L.H5Dread.argtypes = [hid, hid, hid, hid, hid, ctypes.c_void_p]
        L.H5Dclose.argtypes = [hid]
        L.H5Sclose.argtypes = [hid]
        L.H5Dget_type.argtypes = [hid]
        L.H5Dget_type.restype = hid
        L.H5Tget_size.argtypes = [hid]
        L.H5Tget_size.restype = ctypes.c_size_t
        L.H5Tget_class.argtypes = [hid]
        L.H5Tclose.argtypes = [hid]
        _lib = L
    return _lib


class H5File:
    """Read-only file: `names` via exists(), `read(name)` -> float64 array."""

    def __init__(self, path: str):
        self.L = lib()
        self.id = self.L.H5Fopen(os.fsencode(path), 0, 0)
        if self.id < 0:
            raise OSError(f"cannot open {path}")

    def close(self):
        if self.id >= 0:
            self.L.H5Fclose(self.id)
            self.id = -1

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def exists(self, name: str) -> bool:
        return self.L.H5Lexists(self.id, name.encode(), 0) > 0

    def read(self, name: str) -> np.ndarray:
        L = self.L
        d = L.H5Dopen2(self.id, name.encode(), 0)
        assert d >= 0, name
        s = L.H5Dget_space(d)
        nd = L.H5Sget_simple_extent_ndims(s)
        dims = (ctypes.c_ulonglong * nd)()
        L.H5Sget_simple_extent_dims(s, dims, None)
        t = L.H5Dget_type(d)
        # H5T_FLOAT is class 1; the reference writes H5T_NATIVE_DOUBLE (8 bytes)
        assert L.H5Tget_class(t) == 1 and L.H5Tget_size(t) == 8, name
        L.H5Tclose(t)
        out = np.empty(tuple(int(x) for x in dims), dtype=np.float64)
        native_double = ctypes.c_int64.in_dll(L, "H5T_NATIVE_DOUBLE_g").value
        rc = L.H5Dread(d, native_double, 0, 0, 0, out.ctypes.data)  # H5S_ALL = 0
        L.H5Sclose(s)
        L.H5Dclose(d)
        assert rc >= 0, name
        return out
