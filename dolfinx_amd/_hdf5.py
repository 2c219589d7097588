"""Datasets in HDF5 files through the HDF5 C library (libhdf5 1.10+,
loaded with ctypes): what io::hdf5 (cpp/dolfinx/io/HDF5Interface.h) does
for the XDMF writer and reader on one process -- open or create a file,
write a whole 2D dataset at a path (intermediate groups created), read a
dataset back with its shape. Host-side I/O only.
"""

from __future__ import annotations

import ctypes as C
import ctypes.util
import os

import numpy as np

_hid = C.c_int64      # hid_t (int64_t since HDF5 1.10)
_herr = C.c_int       # herr_t
_hsize = C.c_uint64   # hsize_t

_ACC_RDONLY, _ACC_RDWR, _ACC_TRUNC = 0x0000, 0x0001, 0x0002
_P_DEFAULT = 0
_S_ALL = 0
_T_INTEGER, _T_FLOAT = 0, 1

_lib = None


def _candidates():
    env = os.environ.get("DFX_HDF5_LIB")
    if env:
        yield env
    found = ctypes.util.find_library("hdf5")
    if found:
        yield found
    for prefix in (os.environ.get("CONDA_PREFIX"), "/opt/conda", "/usr"):
        if prefix:
            yield os.path.join(prefix, "lib", "libhdf5.so")


def lib():
    """libhdf5, loaded once (raises RuntimeError when it is absent)."""
    global _lib
    if _lib is not None:
        return _lib
    err = None
    for name in _candidates():
        try:
            L = C.CDLL(name)
            break
        except OSError as e:
            err = e
    else:
        raise RuntimeError(f"the HDF5 library (libhdf5.so) was not found: {err}")
    sig = {
        "H5open": (_herr, []),
        "H5Eset_auto2": (_herr, [_hid, C.c_void_p, C.c_void_p]),
        "H5Fcreate": (_hid, [C.c_char_p, C.c_uint, _hid, _hid]),
        "H5Fopen": (_hid, [C.c_char_p, C.c_uint, _hid]),
        "H5Fclose": (_herr, [_hid]),
        "H5Fflush": (_herr, [_hid, C.c_int]),
        "H5Pcreate": (_hid, [_hid]),
        "H5Pset_create_intermediate_group": (_herr, [_hid, C.c_uint]),
        "H5Pclose": (_herr, [_hid]),
        "H5Screate_simple": (_hid, [C.c_int, C.POINTER(_hsize), C.POINTER(_hsize)]),
        "H5Sget_simple_extent_ndims": (C.c_int, [_hid]),
        "H5Sget_simple_extent_dims": (C.c_int, [_hid, C.POINTER(_hsize), C.POINTER(_hsize)]),
        "H5Sclose": (_herr, [_hid]),
        "H5Dcreate2": (_hid, [_hid, C.c_char_p, _hid, _hid, _hid, _hid, _hid]),
        "H5Dopen2": (_hid, [_hid, C.c_char_p, _hid]),
        "H5Dwrite": (_herr, [_hid, _hid, _hid, _hid, _hid, C.c_void_p]),
        "H5Dread": (_herr, [_hid, _hid, _hid, _hid, _hid, C.c_void_p]),
        "H5Dget_space": (_hid, [_hid]),
        "H5Dget_type": (_hid, [_hid]),
        "H5Dclose": (_herr, [_hid]),
        "H5Tget_class": (C.c_int, [_hid]),
        "H5Tclose": (_herr, [_hid]),
        "H5Lexists": (C.c_int, [_hid, C.c_char_p, _hid]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(L, name)
        fn.restype = res
        fn.argtypes = args
    if L.H5open() < 0:
        raise RuntimeError("H5open failed")
    # errors are reported by the return codes checked below
    L.H5Eset_auto2(0, None, None)
    _lib = L
    return L


def _global_id(name):
    return _hid.in_dll(lib(), name).value


def _check(v, what):
    if v < 0:
        raise RuntimeError(f"HDF5: {what} failed")
    return v


class File:
    """An HDF5 file: mode "w" creates (truncates), "a" opens for writing
    (creating it if missing), "r" opens read-only."""

    def __init__(self, path, mode):
        L = lib()
        self.path = str(path)
        p = self.path.encode()
        if mode == "w" or (mode == "a" and not os.path.exists(self.path)):
            self.id = _check(L.H5Fcreate(p, _ACC_TRUNC, _P_DEFAULT, _P_DEFAULT), f"creating {path}")
        elif mode == "a":
            self.id = _check(L.H5Fopen(p, _ACC_RDWR, _P_DEFAULT), f"opening {path}")
        elif mode == "r":
            self.id = _check(L.H5Fopen(p, _ACC_RDONLY, _P_DEFAULT), f"opening {path}")
        else:
            raise ValueError(f"file mode {mode!r}")

    def close(self):
        if getattr(self, "id", -1) >= 0:
            lib().H5Fclose(self.id)
            self.id = -1

    def __del__(self):
        self.close()

    def exists(self, path: str) -> bool:
        """Whether every link of path exists (H5Lexists along the path)."""
        L = lib()
        parts = [q for q in path.split("/") if q]
        cur = ""
        for q in parts:
            cur += "/" + q
            if L.H5Lexists(self.id, cur.encode(), _P_DEFAULT) <= 0:
                return False
        return True

    def write(self, path: str, a) -> None:
        """Dataset `path` (groups created as needed) holding the 2D array a:
        int64 as native 64-bit integers, floats as native doubles."""
        L = lib()
        a = np.ascontiguousarray(a)
        if np.issubdtype(a.dtype, np.integer):
            a = a.astype(np.int64, copy=False)
            t = _global_id("H5T_NATIVE_LLONG_g")
        else:
            a = a.astype(np.float64, copy=False)
            t = _global_id("H5T_NATIVE_DOUBLE_g")
        if self.exists(path):
            raise RuntimeError(f"HDF5 dataset {path} already exists")
        dims = (_hsize * a.ndim)(*a.shape)
        space = _check(L.H5Screate_simple(a.ndim, dims, None), "dataspace")
        lcpl = _check(L.H5Pcreate(_global_id("H5P_CLS_LINK_CREATE_ID_g")), "property list")
        try:
            _check(L.H5Pset_create_intermediate_group(lcpl, 1), "intermediate groups")
            d = _check(L.H5Dcreate2(self.id, path.encode(), t, space, lcpl, _P_DEFAULT,
                                    _P_DEFAULT), f"creating dataset {path}")
            try:
                _check(L.H5Dwrite(d, t, _S_ALL, _S_ALL, _P_DEFAULT,
                                  a.ctypes.data_as(C.c_void_p)), f"writing {path}")
            finally:
                L.H5Dclose(d)
        finally:
            L.H5Pclose(lcpl)
            L.H5Sclose(space)
        L.H5Fflush(self.id, 0)

    def read(self, path: str):
        """The dataset at path as int64 (integer data) or float64."""
        L = lib()
        d = _check(L.H5Dopen2(self.id, path.encode(), _P_DEFAULT), f"opening dataset {path}")
        try:
            space = _check(L.H5Dget_space(d), "dataspace")
            n = L.H5Sget_simple_extent_ndims(space)
            dims = (_hsize * max(n, 1))()
            L.H5Sget_simple_extent_dims(space, dims, None)
            L.H5Sclose(space)
            ft = L.H5Dget_type(d)
            cls = L.H5Tget_class(ft)
            L.H5Tclose(ft)
            if cls == _T_INTEGER:
                out, t = np.empty([dims[i] for i in range(n)], np.int64), "H5T_NATIVE_LLONG_g"
            elif cls == _T_FLOAT:
                out, t = np.empty([dims[i] for i in range(n)], np.float64), "H5T_NATIVE_DOUBLE_g"
            else:
                raise RuntimeError(f"HDF5 dataset {path}: unsupported type class {cls}")
            _check(L.H5Dread(d, _global_id(t), _S_ALL, _S_ALL, _P_DEFAULT,
                             out.ctypes.data_as(C.c_void_p)), f"reading {path}")
            return out
        finally:
            L.H5Dclose(d)
