"""jadamilu_amd -- MI355X-native Jacobi-Davidson eigensolver behind the
DPJD / DPJDREVCOM / DJDREVCOM calling sequence.

The native library (libjadamilu_amd.so, built by `python -m jadamilu_amd.build`)
holds the HIP kernels and the C-ABI of include/jadamilu_amd.h.  This package
only loads it and offers thin Python mirrors of the reference drivers; there
is no CPU fallback: every entry point raises if the library is missing.
"""
from __future__ import annotations

import ctypes as C
import os

__all__ = ["lib", "LIB_PATH", "NativeLibraryMissing"]

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libjadamilu_amd.so")
_lib = None


class NativeLibraryMissing(RuntimeError):
    pass


_i = C.POINTER(C.c_int)
_l = C.POINTER(C.c_long)
_d = C.POINTER(C.c_double)
_v = C.c_void_p


class JdhipStats(C.Structure):
    _fields_ = [("inner_iters", C.c_long), ("host_syncs", C.c_long),
                ("n_precond_setups", C.c_int), ("spmv_lanes", C.c_int),
                ("setup_shifts", C.c_double * 64), ("precond_seconds", C.c_double),
                ("fill", C.c_double), ("precond_levels", C.c_int),
                ("precond_negpiv", C.c_int), ("precond_sweeps", C.c_int),
                ("setup_droptols", C.c_double * 64), ("setup_fills", C.c_double * 64)]


class JdhipBatchItem(C.Structure):
    """Mirror of jdhip_batch_item (include/jadamilu_amd.h)."""
    _fields_ = [("n", C.c_int), ("ia", _i), ("ja", _i), ("a", _d),
                ("neig", C.c_int), ("isearch", C.c_int), ("madspace", C.c_int),
                ("maxiter", C.c_int), ("sigma", C.c_double), ("tol", C.c_double),
                ("shift", C.c_double), ("droptol", C.c_double), ("mem", C.c_double),
                ("icntl", C.c_int * 5), ("eigs", _d), ("res", _d), ("x", _d),
                ("info", C.c_int), ("iter", C.c_int), ("neig_out", C.c_int),
                ("gap", C.c_double)]


# name -> (restype, argtypes)
SIGNATURES = {
    "dpjd_": (None, [_i, _d, _i, _i, _d, _d, _d, _i, _i, _d, _i, _i, _i, _i, _d, _d, _d, _d,
                     _i, _i, _i, _d]),
    "dpjdrevcom_": (None, [_i, _d, _i, _i, _d, _d, _d, _i, _i, _d, _i, _i, _i, _i, _d, _d, _d,
                           _d, _i, _i, _i, _i, _i, _i, _d]),
    "djdrevcom_": (None, [_i, _d, _d, _d, _i, _i, _d, _i, _i, _i, _i, _d, _i, _i, _i, _i, _i,
                          _d]),
    "dpjdcleanup_": (None, []),
    "dpjd": (None, None), "dpjdrevcom": (None, None), "djdrevcom": (None, None),
    "dpjdcleanup": (None, []),
    "dpjd_gep_": (None, [_i, _d, _i, _i, _d, _i, _i, _d, _d, _d, _i, _i, _d, _i, _i, _i, _i,
                         _d, _d, _d, _d, _i, _i, _i, _d]),
    "dpjdrevcom_gep_": (None, [_i, _d, _i, _i, _d, _i, _i, _d, _d, _d, _i, _i, _d, _i, _i, _i,
                               _i, _d, _d, _d, _d, _i, _i, _i, _i, _i, _i, _i, _d]),
    "djdrevcom_gep_": (None, [_i, _d, _d, _d, _i, _i, _d, _i, _i, _i, _i, _d, _i, _i, _i, _i,
                              _i, _i, _d]),
    "dpjd_gep": (None, None), "dpjdrevcom_gep": (None, None), "djdrevcom_gep": (None, None),
    # complex Hermitian drivers: the same argument types, DOUBLE COMPLEX arrays
    # as interleaved doubles
    "zpjd_": None, "zpjdrevcom_": None, "zjdrevcom_": None, "zpjdcleanup_": (None, []),
    "zpjd_gep_": None, "zpjdrevcom_gep_": None, "zjdrevcom_gep_": None,
    "zpjd": (None, None), "zpjdrevcom": (None, None), "zjdrevcom": (None, None),
    "zpjdcleanup": (None, []),
    "zpjd_gep": (None, None), "zpjdrevcom_gep": (None, None), "zjdrevcom_gep": (None, None),
    "jdhip_create": (_v, [C.c_int]),
    "jdhip_destroy": (None, [_v]),
    "jdhip_pjd": (None, [_v, C.c_int, _d, _i, _i, _d, _d, _v, C.c_long, _i, C.c_double, C.c_int,
                         C.c_int, C.c_int, _i, C.c_double, _d, _d, _d, _i, C.c_int, _i, _d]),
    "jdhip_pjdrevcom": (None, [_v, C.c_int, _d, _i, _i, _d, _d, _v, C.c_long, _i, C.c_double,
                               C.c_int, C.c_int, C.c_int, _i, C.c_double, _d, _d, _d, _i, _i,
                               _l, _l, C.c_int, _i, _d]),
    "jdhip_jdrevcom": (None, [_v, C.c_int, _d, _d, _v, C.c_long, _i, C.c_double, C.c_int,
                              C.c_int, C.c_int, _i, C.c_double, _i, _l, _l, C.c_int, _i, _d]),
    "jdhip_pjd_gep": (None, [_v, C.c_int, _d, _i, _i, _d, _i, _i, _d, _d, _v, C.c_long, _i,
                             C.c_double, C.c_int, C.c_int, C.c_int, _i, C.c_double, _d, _d, _d,
                             _i, C.c_int, _i, _d]),
    "jdhip_pjdrevcom_gep": (None, [_v, C.c_int, _d, _i, _i, _d, _i, _i, _d, _d, _v, C.c_long,
                                   _i, C.c_double, C.c_int, C.c_int, C.c_int, _i, C.c_double,
                                   _d, _d, _d, _i, _i, _l, _l, _l, C.c_int, _i, _d]),
    "jdhip_jdrevcom_gep": (None, [_v, C.c_int, _d, _d, _v, C.c_long, _i, C.c_double, C.c_int,
                                  C.c_int, C.c_int, _i, C.c_double, _i, _l, _l, _l, C.c_int,
                                  _i, _d]),
    "jdhip_cleanup": (None, [_v]),
    "jdhip_dense_eig": (C.c_int, [C.c_int, _d, C.c_int, _d, _d, _d, C.c_int]),
    "jdhip_reset": (C.c_int, [_v]),
    "jdhip_order_amd": (C.c_int, [C.c_int, _i, _i, _i]),
    "jdhip_legacy_context": (_v, []),
    "jdhip_stream": (_v, [_v]),
    "jdhip_set_matrix": (C.c_int, [_v, C.c_int, _i, _i, _d, C.c_int]),
    "jdhip_spmv": (C.c_int, [_v, _v, _v]),
    "jdhip_spmm": (C.c_int, [_v, C.c_int, _v, C.c_long, _v, C.c_long]),
    "jdhip_gram": (C.c_int, [_v, C.c_long, C.c_int, C.c_int, _v, C.c_long, _v, C.c_long, _v,
                             C.c_int]),
    "jdhip_gram_noalloc": (C.c_int, [_v, C.c_long, C.c_int, C.c_int, _v, C.c_long, _v, C.c_long,
                                     _v, C.c_int]),
    "jdhip_tall_gemm": (C.c_int, [_v, C.c_long, C.c_int, C.c_int, _v, C.c_long, _v, C.c_int,
                                  _v, C.c_long]),
    "jdhip_get_stats": (None, [_v, C.POINTER(JdhipStats)]),
    "jdhip_set_option": (C.c_int, [_v, C.c_char_p, C.c_double]),
    "jdhip_precond_apply": (C.c_int, [_v, _v, _v]),
    "jdhip_ildl_factor": (_v, [C.c_int, _i, _i, _d, C.c_double, C.c_double, C.c_double, C.c_int,
                               C.c_int, C.c_int, C.c_int, C.c_int, _i]),
    "jdhip_ildl_coupling": (C.c_int, [_v]),
    "jdhip_ildl_save": (C.c_int, [_v, C.c_char_p]),
    "jdhip_ildl_load": (_v, [C.c_char_p, _i]),
    "jdhip_mm_info": (C.c_int, [C.c_char_p, _i, _l]),
    "jdhip_mm_read": (C.c_int, [C.c_char_p, C.c_int, C.c_long, _i, _i, _d]),
    "jdhip_mm_write": (C.c_int, [C.c_char_p, C.c_int, _i, _i, _d]),
    "jdhip_ildl_free": (None, [_v]),
    "jdhip_ildl_nlevels": (C.c_int, [_v, _i]),
    "jdhip_ildl_level_sizes": (C.c_int, [_v, C.c_int, _i, _i, _l, _l]),
    "jdhip_ildl_level_data": (C.c_int, [_v, C.c_int, _i, _d, _i, _i, _d, _i, _i, _d, _d]),
    "jdhip_ildl_dense": (C.c_int, [_v, _d]),
    "jdhip_ildl_level_pivots": (C.c_int, [_v, C.c_int, _d, _d, _i]),
    "jdhip_ildl_stats": (C.c_int, [_v, _d, _i, _d, _d]),
    "jdhip_ildl_attach": (C.c_int, [_v, _v, C.c_int]),
    "jdhip_ildl_device_info": (C.c_int, [_v, _i, _i, _l]),
    "jdhip_ildl_time_sweep": (C.c_int, [_v, C.c_int, _d, _l]),
    "jdhip_ildl_level_sweeps": (C.c_int, [_v, _i, C.c_int]),
    "jdhip_ildl_level_forms": (C.c_int, [_v, _i, C.c_int]),
    "jdhip_adjoint_csr": (C.c_int, [_v, C.c_int, _v, _v, _v, C.c_long, C.c_int, _d, _v]),
    "jdhip_adjoint_diag": (C.c_int, [_v, C.c_int, _v, C.c_long, C.c_int, _d, _v]),
    "jdhip_last_error": (C.c_char_p, [_v]),
    "jdhip_pjd_batch": (C.c_int, [C.c_int, C.c_int, C.POINTER(JdhipBatchItem), C.c_int]),
    "jdhip_pjd_batch_opts": (C.c_int, [C.c_int, C.c_int, C.POINTER(JdhipBatchItem), C.c_int,
                                       C.c_int, C.POINTER(C.c_char_p), _d]),
    "jdhip_complex_context": (_v, []),
    "jdhip_hermitian_real_form": (C.c_long, [C.c_int, _d, _i, _i, _i, _i, _d, C.c_long]),
}
for _z in ("zpjd_", "zpjdrevcom_", "zjdrevcom_", "zpjd_gep_", "zpjdrevcom_gep_",
           "zjdrevcom_gep_"):
    SIGNATURES[_z] = SIGNATURES["d" + _z[1:]]


def lib():
    """Load the native library (raises NativeLibraryMissing if not built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise NativeLibraryMissing(
                f"{LIB_PATH} not found: build it with `python -m jadamilu_amd.build`")
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            f = getattr(L, name)
            f.restype = res
            if args is not None:
                f.argtypes = args
        _lib = L
    return _lib
