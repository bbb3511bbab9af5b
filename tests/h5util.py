"""Minimal HDF5 reader/writer over ctypes (libhdf5 1.10 ABI) -- test infrastructure.

Writes and reads TAMM's tensor file format independently of the engine: one 1-D double
dataset "tensor" of a given extent, with `data` stored from offset 0 (write_to_disk's tammio
layout, tamm_io.hpp:223-460).
"""
import ctypes as C
import os

import numpy as np

hid_t = C.c_int64
hsize_t = C.c_ulonglong
_lib = None


def lib():
    global _lib
    if _lib is None:
        for n in ("libhdf5.so", "libhdf5.so.103", "/opt/conda/lib/libhdf5.so"):
            try:
                _lib = C.CDLL(n)
                break
            except OSError:
                continue
        if _lib is None:
            raise OSError("libhdf5 not found")
        L = _lib
        L.H5open.restype = C.c_int
        L.H5open()
        for f, r, a in [("H5Fcreate", hid_t, [C.c_char_p, C.c_uint, hid_t, hid_t]),
                        ("H5Fopen", hid_t, [C.c_char_p, C.c_uint, hid_t]),
                        ("H5Fclose", C.c_int, [hid_t]),
                        ("H5Screate_simple", hid_t, [C.c_int, C.POINTER(hsize_t), C.c_void_p]),
                        ("H5Sclose", C.c_int, [hid_t]),
                        ("H5Sget_simple_extent_dims", C.c_int, [hid_t, C.POINTER(hsize_t), C.c_void_p]),
                        ("H5Dcreate2", hid_t, [hid_t, C.c_char_p, hid_t, hid_t, hid_t, hid_t, hid_t]),
                        ("H5Dopen2", hid_t, [hid_t, C.c_char_p, hid_t]),
                        ("H5Dclose", C.c_int, [hid_t]),
                        ("H5Dget_space", hid_t, [hid_t]),
                        ("H5Dwrite", C.c_int, [hid_t, hid_t, hid_t, hid_t, hid_t, C.c_void_p]),
                        ("H5Dread", C.c_int, [hid_t, hid_t, hid_t, hid_t, hid_t, C.c_void_p])]:
            getattr(L, f).restype = r
            getattr(L, f).argtypes = a
    return _lib


def _double():
    return hid_t.in_dll(lib(), "H5T_NATIVE_DOUBLE_g").value


def write_tensor_file(path: str, extent: int, data: np.ndarray) -> None:
    """Dataset "tensor" of `extent` doubles; `data` at offset 0, the rest zero (fill value)."""
    L = lib()
    full = np.zeros(extent, dtype=np.float64)
    full[:data.size] = data
    f = L.H5Fcreate(path.encode(), 2, 0, 0)
    dims = hsize_t(extent)
    sp = L.H5Screate_simple(1, C.byref(dims), None)
    ds = L.H5Dcreate2(f, b"tensor", _double(), sp, 0, 0, 0)
    assert f >= 0 and sp >= 0 and ds >= 0
    assert L.H5Dwrite(ds, _double(), 0, 0, 0, full.ctypes.data) >= 0
    L.H5Dclose(ds); L.H5Sclose(sp); L.H5Fclose(f)


def read_tensor_file(path: str):
    """Returns (extent, all values of dataset "tensor")."""
    L = lib()
    f = L.H5Fopen(path.encode(), 0, 0)
    ds = L.H5Dopen2(f, b"tensor", 0)
    sp = L.H5Dget_space(ds)
    dims = hsize_t(0)
    assert L.H5Sget_simple_extent_dims(sp, C.byref(dims), None) == 1
    out = np.empty(dims.value, dtype=np.float64)
    assert L.H5Dread(ds, _double(), 0, 0, 0, out.ctypes.data) >= 0
    L.H5Sclose(sp); L.H5Dclose(ds); L.H5Fclose(f)
    return dims.value, out


def available() -> bool:
    try:
        lib()
        return True
    except OSError:
        return False
