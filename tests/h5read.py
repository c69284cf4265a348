"""A small HDF5 reader for the tests, written against the HDF5 C library with
ctypes (h5py is not installed).  It shares no code with the library's writer:
it lists what a file holds (datasets of the root group, the attributes of a
dataset with their type class, size and length) and reads values as stored."""
import ctypes as C
import os
from pathlib import Path

import numpy as np

HDF5_DIR = Path(os.environ.get("HDF5_DIR", "/opt/conda"))
hid_t = C.c_int64
hsize_t = C.c_uint64
H5P_DEFAULT = 0
H5S_ALL = 0
H5F_ACC_RDONLY = 0
H5T_INTEGER, H5T_FLOAT = 0, 1

_h5 = None


def lib():
    global _h5
    if _h5 is None:
        L = C.CDLL(str(HDF5_DIR / "lib" / "libhdf5.so"))
        L.H5open()
        L.H5Eset_auto2.argtypes = [hid_t, C.c_void_p, C.c_void_p]
        L.H5Eset_auto2(0, None, None)
        for fn, res, args in (
                ("H5Fopen", hid_t, [C.c_char_p, C.c_uint, hid_t]),
                ("H5Fclose", C.c_int, [hid_t]),
                ("H5Dopen2", hid_t, [hid_t, C.c_char_p, hid_t]),
                ("H5Dclose", C.c_int, [hid_t]),
                ("H5Dget_space", hid_t, [hid_t]),
                ("H5Dget_type", hid_t, [hid_t]),
                ("H5Dread", C.c_int, [hid_t, hid_t, hid_t, hid_t, hid_t, C.c_void_p]),
                ("H5Sget_simple_extent_ndims", C.c_int, [hid_t]),
                ("H5Sget_simple_extent_dims", C.c_int, [hid_t, C.POINTER(hsize_t), C.POINTER(hsize_t)]),
                ("H5Sclose", C.c_int, [hid_t]),
                ("H5Aopen", hid_t, [hid_t, C.c_char_p, hid_t]),
                ("H5Aclose", C.c_int, [hid_t]),
                ("H5Aget_type", hid_t, [hid_t]),
                ("H5Aget_space", hid_t, [hid_t]),
                ("H5Aread", C.c_int, [hid_t, hid_t, C.c_void_p]),
                ("H5Tget_class", C.c_int, [hid_t]),
                ("H5Tget_size", C.c_size_t, [hid_t]),
                ("H5Tclose", C.c_int, [hid_t])):
            f = getattr(L, fn)
            f.restype = res
            f.argtypes = args
        _h5 = L
    return _h5


def _native(name):
    return hid_t.in_dll(lib(), name).value


_ATTR_CB = C.CFUNCTYPE(C.c_int, hid_t, C.c_char_p, C.c_void_p, C.c_void_p)
_LINK_CB = C.CFUNCTYPE(C.c_int, hid_t, C.c_char_p, C.c_void_p, C.c_void_p)


def _dims(space):
    L = lib()
    n = L.H5Sget_simple_extent_ndims(space)
    d = (hsize_t * max(n, 1))()
    L.H5Sget_simple_extent_dims(space, d, None)
    return tuple(int(d[q]) for q in range(n))


def _read(obj, space_of, type_of, reader):
    """(values, type class, type size) of a dataset or an attribute."""
    L = lib()
    sp = space_of(obj)
    tp = type_of(obj)
    try:
        shape = _dims(sp)
        cls, size = L.H5Tget_class(tp), int(L.H5Tget_size(tp))
        if cls == H5T_FLOAT:
            out = np.zeros(shape, dtype=np.float64)
            mem = _native("H5T_NATIVE_DOUBLE_g")
        elif cls == H5T_INTEGER:
            out = np.zeros(shape, dtype=np.int32)
            mem = _native("H5T_NATIVE_INT_g")
        else:
            raise TypeError(f"type class {cls}")
        if reader(obj, mem, out.ctypes.data) < 0:
            raise IOError("read failed")
        return out, cls, size
    finally:
        L.H5Tclose(tp)
        L.H5Sclose(sp)


class File:
    def __init__(self, path):
        L = lib()
        self.f = L.H5Fopen(str(path).encode(), H5F_ACC_RDONLY, H5P_DEFAULT)
        if self.f < 0:
            raise IOError(f"cannot open {path}")

    def close(self):
        if self.f >= 0:
            lib().H5Fclose(self.f)
            self.f = -1

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def names(self):
        """Names of the links of the root group."""
        L = lib()
        out = []

        def cb(g, name, info, data):
            out.append(name.decode())
            return 0

        fn = _LINK_CB(cb)
        L.H5Literate.argtypes = [hid_t, C.c_int, C.c_int, C.POINTER(hsize_t), _LINK_CB, C.c_void_p]
        L.H5Literate(self.f, 0, 0, None, fn, None)
        return out

    def dataset(self, name):
        """(values, type class, type size); values keep the file's C order."""
        L = lib()
        d = L.H5Dopen2(self.f, name.encode(), H5P_DEFAULT)
        if d < 0:
            raise KeyError(name)
        try:
            return _read(d, L.H5Dget_space, L.H5Dget_type,
                         lambda o, mem, buf: L.H5Dread(o, mem, H5S_ALL, H5S_ALL, H5P_DEFAULT, buf))
        finally:
            L.H5Dclose(d)

    def attributes(self, dset):
        """{name: (values, type class, type size)} of a dataset's attributes."""
        L = lib()
        d = L.H5Dopen2(self.f, dset.encode(), H5P_DEFAULT)
        if d < 0:
            raise KeyError(dset)
        names = []

        def cb(loc, name, info, data):
            names.append(name)
            return 0

        fn = _ATTR_CB(cb)
        L.H5Aiterate2.argtypes = [hid_t, C.c_int, C.c_int, C.POINTER(hsize_t), _ATTR_CB, C.c_void_p]
        try:
            L.H5Aiterate2(d, 0, 0, None, fn, None)
            out = {}
            for nm in names:
                a = L.H5Aopen(d, nm, H5P_DEFAULT)
                try:
                    out[nm.decode()] = _read(a, L.H5Aget_space, L.H5Aget_type, L.H5Aread)
                finally:
                    L.H5Aclose(a)
            return out
        finally:
            L.H5Dclose(d)
