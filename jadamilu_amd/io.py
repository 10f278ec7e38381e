"""MatrixMarket input and output through the native library (jdhip_mm_*).

read_mm(path) returns the reference's input format: 1-based upper-triangle
CSR (ia, ja, a) with every diagonal entry stored; write_mm(path, ia, ja, a)
writes a symmetric coordinate file from it.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import lib


def _p(a, t):
    return a.ctypes.data_as(C.POINTER(t))


def read_mm(path):
    L = lib()
    n, nnz = C.c_int(), C.c_long()
    rc = L.jdhip_mm_info(str(path).encode(), C.byref(n), C.byref(nnz))
    if rc != 0:
        raise ValueError(f"jdhip_mm_info({path}): {rc}")
    ia = np.zeros(n.value + 1, np.int32)
    ja = np.zeros(max(nnz.value, 1), np.int32)
    a = np.zeros(max(nnz.value, 1))
    rc = L.jdhip_mm_read(str(path).encode(), n.value, nnz.value, _p(ia, C.c_int),
                         _p(ja, C.c_int), _p(a, C.c_double))
    if rc != 0:
        raise ValueError(f"jdhip_mm_read({path}): {rc}")
    return ia, ja[: nnz.value], a[: nnz.value]


def write_mm(path, ia, ja, a):
    L = lib()
    ia = np.ascontiguousarray(ia, dtype=np.int32)
    ja = np.ascontiguousarray(ja, dtype=np.int32)
    a = np.ascontiguousarray(a, dtype=np.float64)
    rc = L.jdhip_mm_write(str(path).encode(), ia.size - 1, _p(ia, C.c_int), _p(ja, C.c_int),
                          _p(a, C.c_double))
    if rc != 0:
        raise OSError(f"jdhip_mm_write({path}): {rc}")
