"""ctypes front-end of the CPU oracle (test infrastructure only).

The oracle is a plain-C restatement of the reference's Jacobi-Davidson
drivers (src/jd/JDaux.F, src/jd/PJD.F, src/jd/dglprec*.c).  It is imported
only by tests/, by __graft_entry__.smoke() and by bench.py's cpu_baseline leg
as the checker -- never by the product path in jadamilu_amd/.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "_build", "libjdoracle.so")
_NATIVE_PATH = os.path.join(_HERE, "_build", "libjdoracle_native.so")
_lib = None

_i = C.POINTER(C.c_int)
_d = C.POINTER(C.c_double)


def build() -> str:
    subprocess.check_call(["make", "-s", "-C", _HERE])
    return _LIB_PATH


def use_native() -> bool:
    """Build and load the -O3 -march=native copy of the oracle (for timing on
    the machine it runs on).  Must be called before the first lib() call;
    returns False (and keeps the portable -O3 build) if it does not build."""
    global _LIB_PATH
    if _lib is not None:
        return _LIB_PATH == _NATIVE_PATH
    try:
        subprocess.check_call(["make", "-s", "-B", "-C", _HERE, "native"])
    except (OSError, subprocess.CalledProcessError):
        return False
    _LIB_PATH = _NATIVE_PATH
    return True


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = C.CDLL(_LIB_PATH)
        L.jdo_create.restype = C.c_void_p
        L.jdo_destroy.argtypes = [C.c_void_p]
        L.jdo_pjdrevcom.argtypes = [C.c_void_p, C.c_int, _d, _i, _i, _d, _d, _d,
                                    C.c_int, _i, C.c_double, C.c_int, C.c_int,
                                    C.c_int, _i, C.c_double, _d, _d, _d, _i, _i,
                                    _i, _i, C.c_int, _i, _d]
        L.jdo_jdrevcom.argtypes = [C.c_void_p, C.c_int, _d, _d, _d, C.c_int, _i,
                                   C.c_double, C.c_int, C.c_int, C.c_int, _i,
                                   C.c_double, _i, _i, _i, C.c_int, _i, _d]
        L.jdo_pjd.argtypes = [C.c_void_p, C.c_int, _d, _i, _i, _d, _d, _d, C.c_int,
                              _i, C.c_double, C.c_int, C.c_int, C.c_int, _i,
                              C.c_double, _d, _d, _d, _i, C.c_int, _i, _d]
        L.jdo_matveca.argtypes = [C.c_int, _i, _i, _d, _d, _d]
        L.jdo_matvec_any.argtypes = [C.c_int, _i, _i, _d, _d, _d]
        L.jdo_nshift.argtypes = [C.c_void_p]
        L.jdo_shift.argtypes = [C.c_void_p, C.c_int]
        L.jdo_shift.restype = C.c_double
        L.jdo_inner_iters.argtypes = [C.c_void_p]
        L.jdo_inner_iters.restype = C.c_long
        L.jdo_nprecond.argtypes = [C.c_void_p]
        L.jdo_syev.argtypes = [C.c_int, _d, C.c_int, _d]
        L.jdo_sygv.argtypes = [C.c_int, _d, C.c_int, _d, C.c_int, _d]
        L.jdo_sortrx.argtypes = [C.c_int, _d, _i]
        L.jdo_cleanup.argtypes = [C.c_void_p]
        L.jdo_ildl_apply.argtypes = [C.c_int, C.c_void_p, C.c_int, _d, _d, _d]
        L.jdo_ildl_set_sweeps.argtypes = [C.c_int]
        L.jdo_ildl_set_sweeps.restype = None
        L.jdo_ildl_set_coupling.argtypes = [C.c_int]
        L.jdo_ildl_set_coupling.restype = None
        L.jdo_set_psolve.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p]
        _lib = L
    return _lib


def _p(a, t):
    return a.ctypes.data_as(C.POINTER(t))


def lx_optimal(n, neig, madspace):
    """Optimal workspace length (PJD.F:99-103)."""
    return n * (3 * madspace + neig + 1) + 3 * madspace ** 2 + max(madspace ** 2, neig)


class Result:
    def __init__(self, **kw):
        self.__dict__.update(kw)


def pjdrevcom_diag(n, diag, matvec, *, neig, isearch=0, sigma=0.0, ninit=0,
                   madspace=20, maxiter=1000, tol=1e-10, icntl=(0, 0, 0, 0, 0),
                   shift=0.0, x0=None, eigs0=None, lx=None):
    """DPJDREVCOM with the diagonal preconditioner (JA(1)<0), user matvec.

    `matvec(xin, xout)` writes A@xin into xout (numpy float64 views)."""
    L = lib()
    s = L.jdo_create()
    try:
        lx = lx or lx_optimal(n, neig, madspace)
        x = np.zeros(lx)
        if x0 is not None:
            x[: x0.size] = x0.ravel(order="F")
        d = np.ascontiguousarray(diag, dtype=np.float64).copy()
        ja = np.array([-1], dtype=np.int32)
        ia = np.array([-1], dtype=np.int32)
        eigs = np.zeros(max(neig, 1)) if eigs0 is None else np.array(eigs0, dtype=np.float64)
        res = np.zeros(max(neig, 1))
        ic = np.array(icntl, dtype=np.int32)
        ineig, iit = C.c_int(neig), C.c_int(maxiter)
        ijob, n1, n2, info = C.c_int(0), C.c_int(0), C.c_int(0), C.c_int(0)
        sh, dt, mem, gap = C.c_double(shift), C.c_double(0.0), C.c_double(0.0), C.c_double(0.0)
        nmv = 0
        while True:
            L.jdo_pjdrevcom(s, n, _p(d, C.c_double), _p(ja, C.c_int), _p(ia, C.c_int),
                            _p(eigs, C.c_double), _p(res, C.c_double), _p(x, C.c_double),
                            lx, C.byref(ineig), sigma, isearch, ninit, madspace,
                            C.byref(iit), tol, C.byref(sh), C.byref(dt), C.byref(mem),
                            _p(ic, C.c_int), C.byref(ijob), C.byref(n1), C.byref(n2), 0,
                            C.byref(info), C.byref(gap))
            if ijob.value != 1:
                break
            a1, a2 = n1.value - 1, n2.value - 1
            matvec(x[a1:a1 + n], x[a2:a2 + n])
            nmv += 1
        k = ineig.value if info.value >= 0 else 0
        shifts = [L.jdo_shift(s, i) for i in range(L.jdo_nshift(s))]
        return Result(eigs=eigs[:k].copy(), res=res[:k].copy(),
                      x=x[: n * k].reshape(k, n).T.copy(), neig=k, iter=iit.value,
                      info=info.value, gap=gap.value, shift=sh.value, shifts=shifts,
                      inner=L.jdo_inner_iters(s), nprec=L.jdo_nprecond(s), nmv=nmv)
    finally:
        L.jdo_destroy(s)


def jdrevcom(n, matvec, psolve, *, neig, isearch=0, sigma=0.0, madspace=20,
             maxiter=1000, tol=1e-10, ninit=0, x0=None, eigs0=None):
    """DJDREVCOM: both matvec and preconditioner by reverse communication."""
    L = lib()
    s = L.jdo_create()
    try:
        lx = lx_optimal(n, neig, madspace)
        x = np.zeros(lx)
        if x0 is not None:
            x[: x0.size] = x0.ravel(order="F")
        eigs = np.zeros(max(neig, 1)) if eigs0 is None else np.array(eigs0, dtype=np.float64)
        res = np.zeros(max(neig, 1))
        ineig, iit = C.c_int(neig), C.c_int(maxiter)
        ijob, n1, n2, info, gap = C.c_int(0), C.c_int(0), C.c_int(0), C.c_int(0), C.c_double(0)
        nmv = nps = 0
        while True:
            L.jdo_jdrevcom(s, n, _p(eigs, C.c_double), _p(res, C.c_double), _p(x, C.c_double),
                           lx, C.byref(ineig), sigma, isearch, ninit, madspace, C.byref(iit),
                           tol, C.byref(ijob), C.byref(n1), C.byref(n2), 0, C.byref(info),
                           C.byref(gap))
            a1, a2 = n1.value - 1, n2.value - 1
            if ijob.value == 1:
                matvec(x[a1:a1 + n], x[a2:a2 + n])
                nmv += 1
            elif ijob.value == 2:
                psolve(x[a2:a2 + n], x[a1:a1 + n])
                nps += 1
            else:
                break
        k = ineig.value
        return Result(eigs=eigs[:k].copy(), res=res[:k].copy(),
                      x=x[: n * k].reshape(k, n).T.copy(), neig=k, iter=iit.value,
                      info=info.value, gap=gap.value, inner=L.jdo_inner_iters(s),
                      nmv=nmv, nps=nps)
    finally:
        L.jdo_destroy(s)


def matvec_upper(ia, ja, a):
    """Closure computing y = A x for a 1-based upper CSR (any diagonal slot)."""
    L = lib()
    ia = np.ascontiguousarray(ia, dtype=np.int32)
    ja = np.ascontiguousarray(ja, dtype=np.int32)
    a = np.ascontiguousarray(a, dtype=np.float64)
    n = ia.size - 1

    def mv(xin, xout):
        xi = np.ascontiguousarray(xin)
        xo = np.empty(n)
        L.jdo_matvec_any(n, _p(ia, C.c_int), _p(ja, C.c_int), _p(a, C.c_double),
                         _p(xi, C.c_double), _p(xo, C.c_double))
        xout[:] = xo
    return mv


class _Level(C.Structure):
    _fields_ = [("n", C.c_int), ("nB", C.c_int), ("perm", _i), ("scal", _d),
                ("lb_ptr", _i), ("lb_idx", _i), ("lb_val", _d),
                ("lc_ptr", _i), ("lc_idx", _i), ("lc_val", _d), ("d", _d),
                ("doff", _d), ("dpair", _i)]


class _IldlCtx(C.Structure):
    _fields_ = [("nlev", C.c_int), ("levels", C.c_void_p), ("m", C.c_int), ("minv", _d)]


def _levels_struct(factor):
    levs = factor["levels"]
    arr = (_Level * max(len(levs), 1))()
    keep = [arr]
    for k, lv in enumerate(levs):
        fields = {}
        for name in ("perm", "lb_ptr", "lb_idx", "lc_ptr", "lc_idx", "dpair"):
            a = np.ascontiguousarray(lv.get(name, np.full(max(lv["nB"], 1), -1)),
                                     dtype=np.int32)
            if a.size == 0:
                a = np.zeros(1, np.int32)
            keep.append(a)
            fields[name] = a.ctypes.data_as(_i)
        for name in ("scal", "lb_val", "lc_val", "d", "doff"):
            a = np.ascontiguousarray(lv[name], dtype=np.float64)
            if a.size == 0:
                a = np.zeros(1)
            keep.append(a)
            fields[name] = a.ctypes.data_as(_d)
        arr[k] = _Level(lv["n"], lv["nB"], **fields)
    minv = np.ascontiguousarray(factor["minv"], dtype=np.float64).ravel()
    if minv.size == 0:
        minv = np.zeros(1)
    keep.append(minv)
    return arr, minv, keep


def pjd_ildl(n, ia, ja, a, factor, *, neig, isearch=0, sigma=0.0, ninit=0, madspace=20,
             maxiter=1000, tol=1e-10, icntl=(0, 1, 0, 0, 0), x0=None, sweeps=0):
    """DPJD of the oracle (internal MATVECA) preconditioned by the oracle's
    multilevel apply of an exported factor (the factor is computed once,
    outside; the oracle has no adaptive re-factorisation).  sweeps > 0
    replaces the triangular solves by that many Jacobi sweeps."""
    L = lib()
    L.jdo_ildl_set_sweeps(int(sweeps))
    L.jdo_ildl_set_coupling(int(factor.get("coupling", 1)))
    arr, minv, keep = _levels_struct(factor)
    ctx = _IldlCtx(len(factor["levels"]), C.cast(arr, C.c_void_p), factor["m"],
                   minv.ctypes.data_as(_d))
    s = L.jdo_create()
    try:
        fn = C.cast(L.jdo_ildl_psolve, C.c_void_p)
        L.jdo_set_psolve(s, fn, C.byref(ctx))
        lx = lx_optimal(n, neig, madspace)
        x = np.zeros(lx)
        if x0 is not None:
            x[: x0.size] = x0.ravel(order="F")
        aa = np.array(a, dtype=np.float64)
        jj = np.array(ja, dtype=np.int32)
        ii = np.array(ia, dtype=np.int32)
        eigs = np.zeros(max(neig, 1))
        res = np.zeros(max(neig, 1))
        ic = np.array(icntl, dtype=np.int32)
        ineig, iit, info = C.c_int(neig), C.c_int(maxiter), C.c_int(0)
        sh, dt, mem, gap = C.c_double(0), C.c_double(1e-3), C.c_double(0), C.c_double(0)
        L.jdo_pjd(s, n, _p(aa, C.c_double), _p(jj, C.c_int), _p(ii, C.c_int), _p(eigs, C.c_double),
                  _p(res, C.c_double), _p(x, C.c_double), lx, C.byref(ineig), sigma, isearch,
                  ninit, madspace, C.byref(iit), tol, C.byref(sh), C.byref(dt), C.byref(mem),
                  _p(ic, C.c_int), 0, C.byref(info), C.byref(gap))
        k = ineig.value if info.value >= 0 else 0
        return Result(eigs=eigs[:k].copy(), res=res[:k].copy(),
                      x=x[: n * k].reshape(k, n).T.copy(), neig=k, iter=iit.value,
                      info=info.value, gap=gap.value, inner=L.jdo_inner_iters(s))
    finally:
        L.jdo_destroy(s)
        L.jdo_ildl_set_sweeps(0)


def ildl_apply(factor, r, sweeps=0):
    """z = M^-1 r for an exported multilevel factor (dict from
    jadamilu_amd.ildl.IldlFactor.export()), sequential C restatement of the
    multilevel substitution (src/ilupack/symAMGsol.c)."""
    L = lib()
    levs = factor["levels"]
    arr = (_Level * max(len(levs), 1))()
    keep = []
    for k, lv in enumerate(levs):
        fields = {}
        for name in ("perm", "lb_ptr", "lb_idx", "lc_ptr", "lc_idx", "dpair"):
            a = np.ascontiguousarray(lv.get(name, np.full(max(lv["nB"], 1), -1)),
                                     dtype=np.int32)
            if a.size == 0:
                a = np.zeros(1, np.int32)
            keep.append(a)
            fields[name] = a.ctypes.data_as(_i)
        for name in ("scal", "lb_val", "lc_val", "d", "doff"):
            a = np.ascontiguousarray(lv.get(name, np.zeros(max(lv["nB"], 1))), dtype=np.float64)
            if a.size == 0:
                a = np.zeros(1)
            keep.append(a)
            fields[name] = a.ctypes.data_as(_d)
        arr[k] = _Level(lv["n"], lv["nB"], **fields)
    minv = np.ascontiguousarray(factor["minv"], dtype=np.float64).ravel()
    if minv.size == 0:
        minv = np.zeros(1)
    r = np.ascontiguousarray(r, dtype=np.float64)
    z = np.zeros_like(r)
    L.jdo_ildl_set_sweeps(int(sweeps))
    L.jdo_ildl_set_coupling(int(factor.get("coupling", 1)))
    L.jdo_ildl_apply(len(levs), arr, factor["m"], minv.ctypes.data_as(_d), _p(r, C.c_double),
                     _p(z, C.c_double))
    L.jdo_ildl_set_sweeps(0)
    return z


def syev(m):
    L = lib()
    j = m.shape[0]
    a = np.asfortranarray(m, dtype=np.float64).copy(order="F")
    w = np.zeros(j)
    L.jdo_syev(j, a.ctypes.data_as(_d), j, _p(w, C.c_double))
    return w, a


def sygv(m, b):
    L = lib()
    j = m.shape[0]
    a = np.asfortranarray(m, dtype=np.float64).copy(order="F")
    bb = np.asfortranarray(b, dtype=np.float64).copy(order="F")
    w = np.zeros(j)
    info = L.jdo_sygv(j, a.ctypes.data_as(_d), j, bb.ctypes.data_as(_d), j, _p(w, C.c_double))
    return info, w, a


def sortrx(data):
    L = lib()
    d = np.ascontiguousarray(data, dtype=np.float64)
    idx = np.zeros(d.size, dtype=np.int32)
    L.jdo_sortrx(d.size, _p(d, C.c_double), _p(idx, C.c_int))
    return idx
