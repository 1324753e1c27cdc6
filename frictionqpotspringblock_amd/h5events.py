"""HDF5 event logs in the schema of the reference's examples.

The examples write and read their per-event records with h5py
(examples/Line1d_Cuspy_Laplace.py:63-67): one file with the datasets
"x_frame" (float64 [nstep]), "f_frame" (float64 [nstep]) and "S" (int64
[nstep]); the random-forcing example adds "t_insta"
(examples/Line1d_System_Cuspy_Laplace_RandomForcing.py:76-79). ``write``
stores the records of ``System.quasistatic`` / ``Batch.quasistatic`` (and any
other 1-d or 2-d float64 / int64 arrays) in such a file, ``read`` loads them
back, so that a reference user's h5py script reads the same names.

No h5py for this interpreter: the HDF5 C library is called through ctypes
(libhdf5 of the environment: FQSB_LIBHDF5, else the loader's search path,
else /opt/conda/lib). Host-side I/O only; nothing here touches the GPU.
"""

from __future__ import annotations

import ctypes as C
import ctypes.util
import os

import numpy as np

_H5F_ACC_RDONLY = 0x0000
_H5F_ACC_TRUNC = 0x0002
_H5P_DEFAULT = 0
_H5S_ALL = 0

_hid = C.c_int64
_herr = C.c_int
_hsize = C.c_ulonglong

_lib = None


def _candidates():
    env = os.environ.get("FQSB_LIBHDF5")
    if env:
        yield env
    found = ctypes.util.find_library("hdf5")
    if found:
        yield found
    for d in ("/opt/conda/lib", "/usr/lib/x86_64-linux-gnu", "/usr/lib64", "/usr/lib"):
        for name in ("libhdf5.so", "libhdf5_serial.so"):
            p = os.path.join(d, name)
            if os.path.exists(p):
                yield p


def library():
    """The HDF5 C library (loaded once); RuntimeError if none is found."""
    global _lib
    if _lib is not None:
        return _lib
    errors = []
    for path in _candidates():
        try:
            L = C.CDLL(path)
            break
        except OSError as e:
            errors.append(str(e))
    else:
        raise RuntimeError("HDF5 event logs need libhdf5 (set FQSB_LIBHDF5 to its path): " + "; ".join(errors))
    sig = {
        "H5open": (_herr, []),
        "H5Fcreate": (_hid, [C.c_char_p, C.c_uint, _hid, _hid]),
        "H5Fopen": (_hid, [C.c_char_p, C.c_uint, _hid]),
        "H5Fclose": (_herr, [_hid]),
        "H5Screate_simple": (_hid, [C.c_int, C.POINTER(_hsize), C.POINTER(_hsize)]),
        "H5Sclose": (_herr, [_hid]),
        "H5Sget_simple_extent_ndims": (C.c_int, [_hid]),
        "H5Sget_simple_extent_dims": (C.c_int, [_hid, C.POINTER(_hsize), C.POINTER(_hsize)]),
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
    _lib = L
    return L


def _native(L, dtype):
    """Memory (and file) type of a numpy dtype: the library's native double
    or 64-bit integer (H5T_NATIVE_DOUBLE_g / H5T_NATIVE_INT64_g, valid after
    H5open)."""
    if dtype == np.float64:
        return _hid.in_dll(L, "H5T_NATIVE_DOUBLE_g").value
    if dtype == np.int64:
        return _hid.in_dll(L, "H5T_NATIVE_INT64_g").value
    raise TypeError(f"HDF5 event logs hold float64 and int64 arrays, not {dtype}")


def _ok(code, what):
    if code < 0:
        raise RuntimeError(f"HDF5 error in {what}")
    return code


def write(path, datasets):
    """Write {name: array} as datasets of a new HDF5 file (overwritten if it
    exists). Arrays are stored as float64 or int64 with their shapes."""
    L = library()
    fid = _ok(L.H5Fcreate(os.fsencode(path), _H5F_ACC_TRUNC, _H5P_DEFAULT, _H5P_DEFAULT), "H5Fcreate")
    try:
        for name, value in datasets.items():
            a = np.asarray(value)
            a = np.ascontiguousarray(a, dtype=np.int64 if a.dtype.kind in "iub" else np.float64)
            dims = (_hsize * max(a.ndim, 1))(*(a.shape if a.ndim else (1,)))
            sid = _ok(L.H5Screate_simple(max(a.ndim, 1), dims, None), "H5Screate_simple")
            try:
                t = _native(L, a.dtype)
                did = _ok(L.H5Dcreate2(fid, name.encode(), t, sid, _H5P_DEFAULT, _H5P_DEFAULT, _H5P_DEFAULT),
                          f"H5Dcreate2({name})")
                try:
                    _ok(L.H5Dwrite(did, t, _H5S_ALL, _H5S_ALL, _H5P_DEFAULT, a.ctypes.data), f"H5Dwrite({name})")
                finally:
                    L.H5Dclose(did)
            finally:
                L.H5Sclose(sid)
    finally:
        L.H5Fclose(fid)


def read(path, names=None):
    """Datasets of an HDF5 file as numpy arrays ({name: array}); `names`
    defaults to the examples' records present in the file."""
    L = library()
    fid = _ok(L.H5Fopen(os.fsencode(path), _H5F_ACC_RDONLY, _H5P_DEFAULT), "H5Fopen")
    out = {}
    try:
        if names is None:
            names = [n for n in ("x_frame", "f_frame", "S", "inc", "t_insta")
                     if L.H5Lexists(fid, n.encode(), _H5P_DEFAULT) > 0]
        for name in names:
            did = _ok(L.H5Dopen2(fid, name.encode(), _H5P_DEFAULT), f"H5Dopen2({name})")
            try:
                tid = L.H5Dget_type(did)
                cls = L.H5Tget_class(tid)  # 0: integer, 1: float
                L.H5Tclose(tid)
                sid = L.H5Dget_space(did)
                nd = L.H5Sget_simple_extent_ndims(sid)
                dims = (_hsize * max(nd, 1))()
                L.H5Sget_simple_extent_dims(sid, dims, None)
                L.H5Sclose(sid)
                shape = tuple(int(dims[i]) for i in range(nd))
                a = np.empty(shape, dtype=np.int64 if cls == 0 else np.float64)
                _ok(L.H5Dread(did, _native(L, a.dtype), _H5S_ALL, _H5S_ALL, _H5P_DEFAULT, a.ctypes.data),
                    f"H5Dread({name})")
                out[name] = a
            finally:
                L.H5Dclose(did)
    finally:
        L.H5Fclose(fid)
    return out
