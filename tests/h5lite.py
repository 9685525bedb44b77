"""Minimal HDF5 access for the tests through the HDF5 C library (ctypes):
reading datasets and attributes to check the kpolaris_image v1 contract, and
overwriting or replacing datasets to build the corrupted iHARM fixtures the
reference's contract tests build with h5py (tests/test_hdf5_contract.py:
1296-1424).  h5py is not part of this image."""
from __future__ import annotations

import ctypes as C
from pathlib import Path

import numpy as np

_L = None
hid_t = C.c_int64
H5F_ACC_RDONLY, H5F_ACC_RDWR = 0, 1
H5P_DEFAULT = hid_t(0)
H5S_ALL = hid_t(0)
H5T_FLOAT, H5T_INTEGER, H5T_STRING = 1, 0, 3


def lib():
    global _L
    if _L is not None:
        return _L
    L = C.CDLL(str(Path("/opt/conda/lib/libhdf5.so")))
    L.H5open()
    for fn, args, res in (
        ("H5Fopen", [C.c_char_p, C.c_uint, hid_t], hid_t), ("H5Fclose", [hid_t], C.c_int),
        ("H5Dopen2", [hid_t, C.c_char_p, hid_t], hid_t), ("H5Dclose", [hid_t], C.c_int),
        ("H5Dget_space", [hid_t], hid_t), ("H5Dget_type", [hid_t], hid_t), ("H5Sclose", [hid_t], C.c_int),
        ("H5Tclose", [hid_t], C.c_int), ("H5Tget_class", [hid_t], C.c_int), ("H5Tget_size", [hid_t], C.c_size_t),
        ("H5Sget_simple_extent_ndims", [hid_t], C.c_int),
        ("H5Sget_simple_extent_dims", [hid_t, C.POINTER(C.c_uint64), C.c_void_p], C.c_int),
        ("H5Dread", [hid_t, hid_t, hid_t, hid_t, hid_t, C.c_void_p], C.c_int),
        ("H5Dwrite", [hid_t, hid_t, hid_t, hid_t, hid_t, C.c_void_p], C.c_int),
        ("H5Ldelete", [hid_t, C.c_char_p, hid_t], C.c_int),
        ("H5Lexists", [hid_t, C.c_char_p, hid_t], C.c_int),
        ("H5Screate_simple", [C.c_int, C.POINTER(C.c_uint64), C.c_void_p], hid_t),
        ("H5Dcreate2", [hid_t, C.c_char_p, hid_t, hid_t, hid_t, hid_t, hid_t], hid_t),
        ("H5Oopen", [hid_t, C.c_char_p, hid_t], hid_t), ("H5Oclose", [hid_t], C.c_int),
        ("H5Aexists", [hid_t, C.c_char_p], C.c_int), ("H5Aopen", [hid_t, C.c_char_p, hid_t], hid_t),
        ("H5Aclose", [hid_t], C.c_int), ("H5Aget_type", [hid_t], hid_t),
        ("H5Aread", [hid_t, hid_t, C.c_void_p], C.c_int),
    ):
        f = getattr(L, fn)
        f.argtypes = args
        f.restype = res
    _L = L
    return L


def _native(name):
    return hid_t.in_dll(lib(), name).value


class File:
    def __init__(self, path, write=False):
        self.id = lib().H5Fopen(str(path).encode(), H5F_ACC_RDWR if write else H5F_ACC_RDONLY, H5P_DEFAULT)
        if self.id < 0:
            raise OSError(f"cannot open {path}")

    def close(self):
        if self.id >= 0:
            lib().H5Fclose(self.id)
            self.id = -1

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def exists(self, path: str) -> bool:
        parts = [p for p in path.strip("/").split("/") if p]
        cur = ""
        for p in parts:
            cur += "/" + p
            if lib().H5Lexists(self.id, cur.encode(), H5P_DEFAULT) <= 0:
                return False
        return True

    def dataset(self, path: str):
        """(numpy array, type class, element size) of a dataset."""
        L = lib()
        d = L.H5Dopen2(self.id, path.encode(), H5P_DEFAULT)
        if d < 0:
            raise KeyError(path)
        try:
            s = L.H5Dget_space(d)
            n = L.H5Sget_simple_extent_ndims(s)
            dims = (C.c_uint64 * max(n, 1))()
            L.H5Sget_simple_extent_dims(s, dims, None)
            L.H5Sclose(s)
            shape = tuple(int(dims[i]) for i in range(n))
            t = L.H5Dget_type(d)
            cls, size = L.H5Tget_class(t), L.H5Tget_size(t)
            L.H5Tclose(t)
            if cls == H5T_INTEGER:
                out = np.zeros(shape, dtype=np.int64)
                L.H5Dread(d, _native("H5T_NATIVE_LLONG_g"), H5S_ALL, H5S_ALL, H5P_DEFAULT, out.ctypes.data)
            else:
                out = np.zeros(shape, dtype=np.float64)
                L.H5Dread(d, _native("H5T_NATIVE_DOUBLE_g"), H5S_ALL, H5S_ALL, H5P_DEFAULT, out.ctypes.data)
            return out, cls, size
        finally:
            L.H5Dclose(d)

    def attr(self, obj: str, name: str):
        """A scalar numeric or fixed-length string attribute of object obj."""
        L = lib()
        o = L.H5Oopen(self.id, obj.encode(), H5P_DEFAULT)
        try:
            if L.H5Aexists(o, name.encode()) <= 0:
                raise KeyError(f"{obj}@{name}")
            a = L.H5Aopen(o, name.encode(), H5P_DEFAULT)
            t = L.H5Aget_type(a)
            cls, size = L.H5Tget_class(t), L.H5Tget_size(t)
            try:
                if cls == H5T_STRING:
                    buf = C.create_string_buffer(size + 1)
                    L.H5Aread(a, t, buf)
                    return buf.value.decode()
                if cls == H5T_INTEGER:
                    v = C.c_longlong()
                    L.H5Aread(a, _native("H5T_NATIVE_LLONG_g"), C.byref(v))
                    return int(v.value)
                v = C.c_double()
                L.H5Aread(a, _native("H5T_NATIVE_DOUBLE_g"), C.byref(v))
                return float(v.value)
            finally:
                L.H5Tclose(t)
                L.H5Aclose(a)
        finally:
            L.H5Oclose(o)

    def overwrite_scalar(self, path: str, value) -> None:
        """Writes a numeric scalar dataset in place (its own type class)."""
        L = lib()
        d = L.H5Dopen2(self.id, path.encode(), H5P_DEFAULT)
        t = L.H5Dget_type(d)
        cls = L.H5Tget_class(t)
        L.H5Tclose(t)
        if cls == H5T_INTEGER:
            v = C.c_longlong(int(value))
            L.H5Dwrite(d, _native("H5T_NATIVE_LLONG_g"), H5S_ALL, H5S_ALL, H5P_DEFAULT, C.byref(v))
        else:
            v = C.c_double(float(value))
            L.H5Dwrite(d, _native("H5T_NATIVE_DOUBLE_g"), H5S_ALL, H5S_ALL, H5P_DEFAULT, C.byref(v))
        L.H5Dclose(d)

    def replace_float32(self, path: str, data: np.ndarray) -> None:
        """Deletes a dataset and writes `data` (float32) in its place."""
        L = lib()
        L.H5Ldelete(self.id, path.encode(), H5P_DEFAULT)
        a = np.ascontiguousarray(data, dtype=np.float32)
        dims = (C.c_uint64 * a.ndim)(*a.shape)
        s = L.H5Screate_simple(a.ndim, dims, None)
        d = L.H5Dcreate2(self.id, path.encode(), _native("H5T_IEEE_F32LE_g"), s, H5P_DEFAULT, H5P_DEFAULT,
                         H5P_DEFAULT)
        L.H5Dwrite(d, _native("H5T_NATIVE_FLOAT_g"), H5S_ALL, H5S_ALL, H5P_DEFAULT, a.ctypes.data)
        L.H5Dclose(d)
        L.H5Sclose(s)
