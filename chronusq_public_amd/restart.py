"""Restart files in the reference's layout, for the READ guess.

The reference keeps the SCF densities in an HDF5 file named after the input
(``<input basename>.bin``, src/fileio/fileio.cpp:58-69), group ``/SCF``, datasets
``DensityScalar`` and, for open shells, ``DensityMz``: N x N doubles holding the Eigen
(column-major) buffer (include/singleslater/singleslater_scfutils.h:584-660).  Its READ guess
reads them and forms the Fock matrix (include/singleslater/singleslater_guess.h:320-329), and
fails with "FATAL: READ Guess cannot find DensityScalar" when the dataset is missing
(singleslater_scfutils.h:633-634).

This module writes and reads the same datasets through the HDF5 C library (ctypes; the
image's libhdf5 1.10).  Matrices are written as the reference's column-major buffer: the
dataset's row-major element [i][j] is M(j, i).
"""
from __future__ import annotations

import ctypes
import ctypes.util
import os
from typing import Optional, Tuple

import numpy as np

hid_t = ctypes.c_int64
herr_t = ctypes.c_int
hsize_t = ctypes.c_ulonglong
H5P_DEFAULT = 0
H5S_ALL = 0
H5F_ACC_RDONLY, H5F_ACC_RDWR, H5F_ACC_TRUNC = 0x0000, 0x0001, 0x0002
GROUP = "/SCF"

_H5 = None


class RestartError(RuntimeError):
    pass


def _lib():
    global _H5
    if _H5 is not None:
        return _H5
    names = [ctypes.util.find_library("hdf5"), "/opt/conda/lib/libhdf5.so", "libhdf5.so", "libhdf5_serial.so"]
    lib = None
    for n in names:
        if not n:
            continue
        try:
            lib = ctypes.CDLL(n)
            break
        except OSError:
            continue
    if lib is None:
        raise RestartError("restart files need the HDF5 C library (libhdf5.so), which was not found")
    c = ctypes.c_char_p
    sig = {
        "H5open": ([], herr_t), "H5Eset_auto2": ([hid_t, ctypes.c_void_p, ctypes.c_void_p], herr_t),
        "H5Fcreate": ([c, ctypes.c_uint, hid_t, hid_t], hid_t), "H5Fopen": ([c, ctypes.c_uint, hid_t], hid_t),
        "H5Fclose": ([hid_t], herr_t),
        "H5Gcreate2": ([hid_t, c, hid_t, hid_t, hid_t], hid_t), "H5Gopen2": ([hid_t, c, hid_t], hid_t),
        "H5Gclose": ([hid_t], herr_t),
        "H5Screate_simple": ([ctypes.c_int, ctypes.POINTER(hsize_t), ctypes.POINTER(hsize_t)], hid_t),
        "H5Sclose": ([hid_t], herr_t),
        "H5Sget_simple_extent_ndims": ([hid_t], ctypes.c_int),
        "H5Sget_simple_extent_dims": ([hid_t, ctypes.POINTER(hsize_t), ctypes.POINTER(hsize_t)], ctypes.c_int),
        "H5Dcreate2": ([hid_t, c, hid_t, hid_t, hid_t, hid_t, hid_t], hid_t),
        "H5Dopen2": ([hid_t, c, hid_t], hid_t), "H5Dclose": ([hid_t], herr_t),
        "H5Dget_space": ([hid_t], hid_t),
        "H5Dwrite": ([hid_t, hid_t, hid_t, hid_t, hid_t, ctypes.c_void_p], herr_t),
        "H5Dread": ([hid_t, hid_t, hid_t, hid_t, hid_t, ctypes.c_void_p], herr_t),
        "H5Lexists": ([hid_t, c, hid_t], ctypes.c_int),
    }
    for name, (args, res) in sig.items():
        f = getattr(lib, name)
        f.argtypes = args
        f.restype = res
    if lib.H5open() < 0:
        raise RestartError("H5open failed")
    lib.H5Eset_auto2(0, None, None)      # errors come back as return codes, not printed
    lib.native_double = ctypes.c_int64.in_dll(lib, "H5T_NATIVE_DOUBLE_g").value
    _H5 = lib
    return lib


def _ok(v, what):
    if v < 0:
        raise RestartError(f"HDF5: {what} failed")
    return v


def _dataset_dims(L, d):
    sp = _ok(L.H5Dget_space(d), "H5Dget_space")
    try:
        nd = L.H5Sget_simple_extent_ndims(sp)
        dims = (hsize_t * max(nd, 1))()
        L.H5Sget_simple_extent_dims(sp, dims, None)
        return tuple(int(dims[i]) for i in range(nd))
    finally:
        L.H5Sclose(sp)


def write_densities(path: str, P: np.ndarray, Mz: Optional[np.ndarray] = None) -> None:
    """Store P (and Mz) as /SCF/DensityScalar (/SCF/DensityMz), creating the file or group as
    needed and overwriting datasets of the same shape."""
    L = _lib()
    f = L.H5Fopen(path.encode(), H5F_ACC_RDWR, H5P_DEFAULT) if os.path.exists(path) else \
        L.H5Fcreate(path.encode(), H5F_ACC_TRUNC, H5P_DEFAULT, H5P_DEFAULT)
    _ok(f, f"opening {path}")
    try:
        exists = L.H5Lexists(f, GROUP.encode(), H5P_DEFAULT) > 0
        g = _ok(L.H5Gopen2(f, GROUP.encode(), H5P_DEFAULT) if exists else
                L.H5Gcreate2(f, GROUP.encode(), H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT), "opening /SCF")
        try:
            for name, M in (("DensityScalar", P), ("DensityMz", Mz)):
                if M is None:
                    continue
                M = np.asarray(M, dtype=np.float64)
                buf = np.ascontiguousarray(M.T)           # the column-major buffer
                n0, n1 = M.shape
                if L.H5Lexists(g, name.encode(), H5P_DEFAULT) > 0:
                    d = _ok(L.H5Dopen2(g, name.encode(), H5P_DEFAULT), f"opening {name}")
                    if _dataset_dims(L, d) != (n0, n1):
                        L.H5Dclose(d)
                        raise RestartError(f"{path}:{GROUP}/{name} has another shape than {n0} x {n1}")
                else:
                    dims = (hsize_t * 2)(n0, n1)
                    sp = _ok(L.H5Screate_simple(2, dims, None), "H5Screate_simple")
                    d = L.H5Dcreate2(g, name.encode(), L.native_double, sp, H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT)
                    L.H5Sclose(sp)
                    _ok(d, f"creating {name}")
                try:
                    _ok(L.H5Dwrite(d, L.native_double, H5S_ALL, H5S_ALL, H5P_DEFAULT,
                                   buf.ctypes.data_as(ctypes.c_void_p)), f"writing {name}")
                finally:
                    L.H5Dclose(d)
        finally:
            L.H5Gclose(g)
    finally:
        L.H5Fclose(f)


def read_densities(path: str, n: int, open_shell: bool = False) -> Tuple[np.ndarray, Optional[np.ndarray]]:
    """P (and Mz when open_shell) from /SCF; the reference's READ-guess errors when missing."""
    L = _lib()
    if not os.path.exists(path):
        raise RestartError(f"FATAL: READ Guess cannot find DensityScalar (no restart file {path})")
    f = _ok(L.H5Fopen(path.encode(), H5F_ACC_RDONLY, H5P_DEFAULT), f"opening {path}")
    try:
        if L.H5Lexists(f, GROUP.encode(), H5P_DEFAULT) <= 0:
            raise RestartError("FATAL: READ Guess cannot find DensityScalar")
        g = _ok(L.H5Gopen2(f, GROUP.encode(), H5P_DEFAULT), "opening /SCF")
        try:
            out = []
            for name in ("DensityScalar",) + (("DensityMz",) if open_shell else ()):
                if L.H5Lexists(g, name.encode(), H5P_DEFAULT) <= 0:
                    raise RestartError(f"FATAL: READ Guess cannot find {name}")
                d = _ok(L.H5Dopen2(g, name.encode(), H5P_DEFAULT), f"opening {name}")
                try:
                    if _dataset_dims(L, d) != (n, n):
                        raise RestartError(f"{path}:{GROUP}/{name} is {_dataset_dims(L, d)}, the basis has {n}")
                    buf = np.empty((n, n), dtype=np.float64)
                    _ok(L.H5Dread(d, L.native_double, H5S_ALL, H5S_ALL, H5P_DEFAULT,
                                  buf.ctypes.data_as(ctypes.c_void_p)), f"reading {name}")
                    out.append(np.ascontiguousarray(buf.T))
                finally:
                    L.H5Dclose(d)
        finally:
            L.H5Gclose(g)
    finally:
        L.H5Fclose(f)
    return out[0], (out[1] if open_shell else None)
