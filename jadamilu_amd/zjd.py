"""Host wrappers of the complex Hermitian drivers (ZPJD, ZPJDREVCOM, ZJDREVCOM
and their generalized forms, Doc/ZPJD_prototypes.f) through the legacy
symbols of the native library.

Matrices are 1-based CSR of one triangle of a Hermitian matrix with complex
values (numpy complex128, every diagonal entry present and real); X is a
complex128 workspace of LX entries, as DOUBLE COMPLEX in the reference.
Reverse-communication callbacks receive and fill complex128 vectors of
length n.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import lib
from .jd import JDResult, _icntl, lx_optimal, lx_optimal_gep

_i, _d = C.c_int, C.c_double


def _pd(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def _pi(a):
    return a.ctypes.data_as(C.POINTER(C.c_int))


def _cplx(a):
    return np.ascontiguousarray(a, dtype=np.complex128)


def _int(a):
    return np.ascontiguousarray(a, dtype=np.int32)


def _result(x, n, eigs, res, NE, IT, INFO, GAP, SH=None, DT=None):
    k = NE.value if INFO.value >= 0 else 0
    r = JDResult(eigs=eigs[:k].copy(), res=res[:k].copy(),
                 x=x[: n * k].reshape(k, n).T.copy(), neig=k, iter=IT.value,
                 info=INFO.value, gap=GAP.value,
                 shift=SH.value if SH is not None else 0.0,
                 droptol=DT.value if DT is not None else 0.0)
    return r


def zpjd_host(n, a, ja, ia, *, neig, isearch=0, sigma=0.0, ninit=0, madspace=20, maxiter=1000,
              tol=1e-10, shift=0.0, droptol=1e-3, mem=20.0, icntl=(0, 0, 0, 0, 0), x0=None,
              eigs0=None, lx=None, iprint=0):
    """ZPJD through the legacy symbol `zpjd_`."""
    L = lib()
    lx = lx or lx_optimal(n, neig, madspace)
    x = np.zeros(lx, dtype=np.complex128)
    if x0 is not None:
        x[: np.size(x0)] = np.asarray(x0).ravel(order="F")
    a, ja, ia = _cplx(a), _int(ja), _int(ia)
    eigs = np.zeros(max(neig, 1)) if eigs0 is None else np.array(eigs0, dtype=np.float64)
    res = np.zeros(max(neig, 1))
    ic = _icntl(icntl)
    N, LX, NE, IS, NI, MS, IT = _i(n), _i(lx), _i(neig), _i(isearch), _i(ninit), _i(madspace), _i(maxiter)
    SG, TL, SH, DT, ME = _d(sigma), _d(tol), _d(shift), _d(droptol), _d(mem)
    IP, INFO, GAP = _i(iprint), _i(0), _d(0)
    L.zpjd_(C.byref(N), _pd(a), _pi(ja), _pi(ia), _pd(eigs), _pd(res), _pd(x), C.byref(LX),
            C.byref(NE), C.byref(SG), C.byref(IS), C.byref(NI), C.byref(MS), C.byref(IT),
            C.byref(TL), C.byref(SH), C.byref(DT), C.byref(ME), _pi(ic), C.byref(IP),
            C.byref(INFO), C.byref(GAP))
    return _result(x, n, eigs, res, NE, IT, INFO, GAP, SH, DT)


def zpjdrevcom_host(n, a, ja, ia, matvec, *, neig, isearch=0, sigma=0.0, ninit=0, madspace=20,
                    maxiter=1000, tol=1e-10, shift=0.0, droptol=1e-3, mem=20.0,
                    icntl=(0, 0, 0, 0, 0), lx=None, iprint=0):
    """ZPJDREVCOM through `zpjdrevcom_`; matvec(x, y) writes H x into y.  The
    preconditioner is built from (a, ja, ia); ja = [-1]: diagonal."""
    L = lib()
    lx = lx or lx_optimal(n, neig, madspace)
    x = np.zeros(lx, dtype=np.complex128)
    a, ja, ia = _cplx(a), _int(ja), _int(ia)
    eigs, res = np.zeros(max(neig, 1)), np.zeros(max(neig, 1))
    ic = _icntl(icntl)
    N, LX, NE, IS, NI, MS, IT = _i(n), _i(lx), _i(neig), _i(isearch), _i(ninit), _i(madspace), _i(maxiter)
    SG, TL, SH, DT, ME = _d(sigma), _d(tol), _d(shift), _d(droptol), _d(mem)
    IJ, N1, N2, IP, INFO, GAP = _i(0), _i(0), _i(0), _i(iprint), _i(0), _d(0)
    calls = 0
    while True:
        L.zpjdrevcom_(C.byref(N), _pd(a), _pi(ja), _pi(ia), _pd(eigs), _pd(res), _pd(x),
                      C.byref(LX), C.byref(NE), C.byref(SG), C.byref(IS), C.byref(NI),
                      C.byref(MS), C.byref(IT), C.byref(TL), C.byref(SH), C.byref(DT),
                      C.byref(ME), _pi(ic), C.byref(IJ), C.byref(N1), C.byref(N2),
                      C.byref(IP), C.byref(INFO), C.byref(GAP))
        if IJ.value != 1:
            break
        calls += 1
        a1, a2 = N1.value - 1, N2.value - 1
        y = np.empty(n, dtype=np.complex128)
        matvec(x[a1:a1 + n].copy(), y)
        x[a2:a2 + n] = y
    r = _result(x, n, eigs, res, NE, IT, INFO, GAP, SH, DT)
    r.matvec_calls = calls
    return r


def zjdrevcom_host(n, matvec, psolve, *, neig, isearch=0, sigma=0.0, ninit=0, madspace=20,
                   maxiter=1000, tol=1e-10, x0=None, lx=None, iprint=0):
    """ZJDREVCOM through `zjdrevcom_`; psolve(src, dst) writes P^-1 src."""
    L = lib()
    lx = lx or lx_optimal(n, neig, madspace)
    x = np.zeros(lx, dtype=np.complex128)
    if x0 is not None:
        x[: np.size(x0)] = np.asarray(x0).ravel(order="F")
    eigs, res = np.zeros(max(neig, 1)), np.zeros(max(neig, 1))
    N, LX, NE, IS, NI, MS, IT = _i(n), _i(lx), _i(neig), _i(isearch), _i(ninit), _i(madspace), _i(maxiter)
    SG, TL = _d(sigma), _d(tol)
    IJ, N1, N2, IP, INFO, GAP = _i(0), _i(0), _i(0), _i(iprint), _i(0), _d(0)
    while True:
        L.zjdrevcom_(C.byref(N), _pd(eigs), _pd(res), _pd(x), C.byref(LX), C.byref(NE),
                     C.byref(SG), C.byref(IS), C.byref(NI), C.byref(MS), C.byref(IT),
                     C.byref(TL), C.byref(IJ), C.byref(N1), C.byref(N2), C.byref(IP),
                     C.byref(INFO), C.byref(GAP))
        a1, a2 = N1.value - 1, N2.value - 1
        y = np.empty(n, dtype=np.complex128)
        if IJ.value == 1:
            matvec(x[a1:a1 + n].copy(), y)
            x[a2:a2 + n] = y
        elif IJ.value == 2:
            psolve(x[a2:a2 + n].copy(), y)
            x[a1:a1 + n] = y
        else:
            break
    return _result(x, n, eigs, res, NE, IT, INFO, GAP)


def zpjd_gep_host(n, a, ja, ia, b, jb, ib, *, neig, isearch=0, sigma=0.0, ninit=0, madspace=20,
                  maxiter=1000, tol=1e-10, shift=0.0, droptol=1e-3, mem=20.0,
                  icntl=(0, 0, 0, 0, 0), lx=None, iprint=0):
    """ZPJD_GEP through `zpjd_gep_`: H x = lambda M x, M Hermitian positive
    definite."""
    L = lib()
    lx = lx or lx_optimal_gep(n, neig, madspace)
    x = np.zeros(lx, dtype=np.complex128)
    a, ja, ia = _cplx(a), _int(ja), _int(ia)
    b, jb, ib = _cplx(b), _int(jb), _int(ib)
    eigs, res = np.zeros(max(neig, 1)), np.zeros(max(neig, 1))
    ic = _icntl(icntl)
    N, LX, NE, IS, NI, MS, IT = _i(n), _i(lx), _i(neig), _i(isearch), _i(ninit), _i(madspace), _i(maxiter)
    SG, TL, SH, DT, ME = _d(sigma), _d(tol), _d(shift), _d(droptol), _d(mem)
    IP, INFO, GAP = _i(iprint), _i(0), _d(0)
    L.zpjd_gep_(C.byref(N), _pd(a), _pi(ja), _pi(ia), _pd(b), _pi(jb), _pi(ib), _pd(eigs),
                _pd(res), _pd(x), C.byref(LX), C.byref(NE), C.byref(SG), C.byref(IS),
                C.byref(NI), C.byref(MS), C.byref(IT), C.byref(TL), C.byref(SH), C.byref(DT),
                C.byref(ME), _pi(ic), C.byref(IP), C.byref(INFO), C.byref(GAP))
    return _result(x, n, eigs, res, NE, IT, INFO, GAP, SH, DT)


def zpjdrevcom_gep_host(n, a, ja, ia, b, jb, ib, matvec_a, matvec_b, *, neig, isearch=0,
                        sigma=0.0, ninit=0, madspace=20, maxiter=1000, tol=1e-10, shift=0.0,
                        droptol=1e-3, mem=20.0, icntl=(0, 0, 0, 0, 0), lx=None, iprint=0):
    """ZPJDREVCOM_GEP through `zpjdrevcom_gep_`: the preconditioner is built
    from (a, ja, ia) and (b, jb, ib); the caller's products serve IJOB 1 (A
    and B), 3 (A only) and 4 (B only)."""
    L = lib()
    lx = lx or lx_optimal_gep(n, neig, madspace)
    x = np.zeros(lx, dtype=np.complex128)
    a, ja, ia = _cplx(a), _int(ja), _int(ia)
    b, jb, ib = _cplx(b), _int(jb), _int(ib)
    eigs, res = np.zeros(max(neig, 1)), np.zeros(max(neig, 1))
    ic = _icntl(icntl)
    N, LX, NE, IS, NI, MS, IT = _i(n), _i(lx), _i(neig), _i(isearch), _i(ninit), _i(madspace), _i(maxiter)
    SG, TL, SH, DT, ME = _d(sigma), _d(tol), _d(shift), _d(droptol), _d(mem)
    IJ, N1, N2, N3, IP, INFO, GAP = _i(0), _i(0), _i(0), _i(0), _i(iprint), _i(0), _d(0)
    counts = {1: 0, 3: 0, 4: 0}
    while True:
        L.zpjdrevcom_gep_(C.byref(N), _pd(a), _pi(ja), _pi(ia), _pd(b), _pi(jb), _pi(ib),
                          _pd(eigs), _pd(res), _pd(x), C.byref(LX), C.byref(NE), C.byref(SG),
                          C.byref(IS), C.byref(NI), C.byref(MS), C.byref(IT), C.byref(TL),
                          C.byref(SH), C.byref(DT), C.byref(ME), _pi(ic), C.byref(IJ),
                          C.byref(N1), C.byref(N2), C.byref(N3), C.byref(IP), C.byref(INFO),
                          C.byref(GAP))
        j = IJ.value
        if j not in counts:
            break
        counts[j] += 1
        a1, a2, a3 = N1.value - 1, N2.value - 1, N3.value - 1
        xi = x[a1:a1 + n].copy()
        y = np.empty(n, dtype=np.complex128)
        if j in (1, 3):
            matvec_a(xi, y)
            x[a2:a2 + n] = y
        if j in (1, 4):
            matvec_b(xi, y)
            x[a3:a3 + n] = y
    r = _result(x, n, eigs, res, NE, IT, INFO, GAP, SH, DT)
    r.ijob_counts = counts
    return r


def zjdrevcom_gep_host(n, matvec_a, matvec_b, psolve, *, neig, isearch=0, sigma=0.0, ninit=0,
                       madspace=20, maxiter=1000, tol=1e-10, lx=None, iprint=0):
    """ZJDREVCOM_GEP through `zjdrevcom_gep_`: IJOB 1 (A and B products),
    2 (preconditioner), 3 (A only), 4 (B only)."""
    L = lib()
    lx = lx or lx_optimal_gep(n, neig, madspace)
    x = np.zeros(lx, dtype=np.complex128)
    eigs, res = np.zeros(max(neig, 1)), np.zeros(max(neig, 1))
    N, LX, NE, IS, NI, MS, IT = _i(n), _i(lx), _i(neig), _i(isearch), _i(ninit), _i(madspace), _i(maxiter)
    SG, TL = _d(sigma), _d(tol)
    IJ, N1, N2, N3, IP, INFO, GAP = _i(0), _i(0), _i(0), _i(0), _i(iprint), _i(0), _d(0)
    counts = {1: 0, 2: 0, 3: 0, 4: 0}
    while True:
        L.zjdrevcom_gep_(C.byref(N), _pd(eigs), _pd(res), _pd(x), C.byref(LX), C.byref(NE),
                         C.byref(SG), C.byref(IS), C.byref(NI), C.byref(MS), C.byref(IT),
                         C.byref(TL), C.byref(IJ), C.byref(N1), C.byref(N2), C.byref(N3),
                         C.byref(IP), C.byref(INFO), C.byref(GAP))
        j = IJ.value
        if j not in counts:
            break
        counts[j] += 1
        a1, a2, a3 = N1.value - 1, N2.value - 1, N3.value - 1
        y = np.empty(n, dtype=np.complex128)
        if j == 2:
            psolve(x[a2:a2 + n].copy(), y)
            x[a1:a1 + n] = y
            continue
        xi = x[a1:a1 + n].copy()
        if j in (1, 3):
            matvec_a(xi, y)
            x[a2:a2 + n] = y
        if j in (1, 4):
            matvec_b(xi, y)
            x[a3:a3 + n] = y
    r = _result(x, n, eigs, res, NE, IT, INFO, GAP)
    r.ijob_counts = counts
    return r


def cleanup():
    """ZPJDCLEANUP: release the complex drivers' preconditioner."""
    lib().zpjdcleanup_()


def hermitian_dense(n, a, ja, ia):
    """Dense Hermitian matrix of a 1-based one-triangle complex CSR."""
    H = np.zeros((n, n), dtype=np.complex128)
    for i in range(n):
        for k in range(ia[i] - 1, ia[i + 1] - 1):
            j = ja[k] - 1
            if i == j:
                H[i, i] = a[k].real
            else:
                H[i, j] = a[k]
                H[j, i] = np.conj(a[k])
    return H


def hermitian_real_form(n, a, ja, ia):
    """The real symmetric form of order 2n solved by the complex drivers
    (jdhip_hermitian_real_form): 1-based one-triangle CSR (ia2, ja2, a2)."""
    L = lib()
    a, ja = _cplx(a), _int(ja)
    ia = _int(ia) if ia is not None else np.array([-1], dtype=np.int32)
    cnt = L.jdhip_hermitian_real_form(n, _pd(a), _pi(ja), _pi(ia), None, None, None, 0)
    if cnt < 0:
        raise ValueError("jdhip_hermitian_real_form: bad input")
    ia2 = np.zeros(2 * n + 1, dtype=np.int32)
    ja2 = np.zeros(max(cnt, 1), dtype=np.int32)
    a2 = np.zeros(max(cnt, 1))
    L.jdhip_hermitian_real_form(n, _pd(a), _pi(ja), _pi(ia), _pi(ia2), _pi(ja2), _pd(a2), cnt)
    return ia2, ja2[:cnt], a2[:cnt]


def set_option(name, value):
    """jdhip_set_option on the complex drivers' context."""
    L = lib()
    rc = L.jdhip_set_option(L.jdhip_complex_context(), name.encode(), float(value))
    if rc != 0:
        raise ValueError(f"jdhip_set_option({name}={value}): {rc}")


def reset():
    """The complex drivers' context back to the state of a fresh process."""
    L = lib()
    L.zpjdcleanup_()
    L.jdhip_reset(L.jdhip_complex_context())
