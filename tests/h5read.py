"""Minimal HDF5 dataset reader over the serial HDF5 C library (ctypes), for
checking the snapshot files the library writes (h5py is not installed)."""
import ctypes
import os

import numpy as np

_CANDIDATES = (os.environ.get("OWT_HDF5_LIB"), "libhdf5.so", "libhdf5.so.103", "/opt/conda/lib/libhdf5.so")
_lib = None


def lib():
    """The HDF5 library, or None when it cannot be loaded."""
    global _lib
    if _lib is None:
        for name in _CANDIDATES:
            if not name:
                continue
            try:
                L = ctypes.CDLL(name)
            except OSError:
                continue
            hid, P = ctypes.c_int64, ctypes.c_void_p
            L.H5Fopen.argtypes = [ctypes.c_char_p, ctypes.c_uint, hid]
            L.H5Fopen.restype = hid
            L.H5Fclose.argtypes = [hid]
            L.H5Dopen2.argtypes = [hid, ctypes.c_char_p, hid]
            L.H5Dopen2.restype = hid
            L.H5Dclose.argtypes = [hid]
            L.H5Dget_space.argtypes = [hid]
            L.H5Dget_space.restype = hid
            L.H5Sget_simple_extent_ndims.argtypes = [hid]
            L.H5Sget_simple_extent_dims.argtypes = [hid, P, P]
            L.H5Sclose.argtypes = [hid]
            L.H5Dread.argtypes = [hid, hid, hid, hid, hid, P]
            L.H5open()
            _lib = L
            break
    return _lib


def read(path, name):
    """Dataset `name` of file `path` as a float64 array of its own shape."""
    L = lib()
    f = L.H5Fopen(os.fsencode(path), 0, 0)  # H5F_ACC_RDONLY, H5P_DEFAULT
    assert f >= 0, path
    try:
        d = L.H5Dopen2(f, name.encode(), 0)
        assert d >= 0, name
        s = L.H5Dget_space(d)
        nd = L.H5Sget_simple_extent_ndims(s)
        dims = (ctypes.c_uint64 * nd)()
        L.H5Sget_simple_extent_dims(s, dims, None)
        L.H5Sclose(s)
        out = np.empty(tuple(dims), dtype=np.float64)
        native_double = ctypes.c_int64.in_dll(L, "H5T_NATIVE_DOUBLE_g").value
        assert L.H5Dread(d, native_double, 0, 0, 0, out.ctypes.data_as(ctypes.c_void_p)) >= 0
        L.H5Dclose(d)
        return out
    finally:
        L.H5Fclose(f)
