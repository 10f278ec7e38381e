"""Python mirrors of the reference drivers over the native C-ABI.

Same argument names and meaning as src/jd/PJD.F (DPJD, DPJDREVCOM, DJDREVCOM,
DPJDCLEANUP); reverse communication is serviced by a Python callable for the
legacy host-array entry points, or on the device (jdhip_spmv) for the
context API.  All numerics run in libjadamilu_amd.so.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, field

import numpy as np

from . import JdhipStats, lib

_i = C.c_int
_d = C.c_double


def lx_optimal(n, neig, madspace):
    """Optimal workspace length N*(3*MADSPACE+NEIG+1)+3*MADSPACE**2+MAX(MADSPACE**2,NEIG)
    (PJD.F:99-103)."""
    return n * (3 * madspace + neig + 1) + 3 * madspace ** 2 + max(madspace ** 2, neig)


def lx_optimal_gep(n, neig, madspace):
    """Workspace of the generalized drivers, N*(4*MADSPACE+2*NEIG+1)+3*MADSPACE**2+
    MAX(MADSPACE**2,NEIG) (PJD.F:544-545)."""
    return n * (4 * madspace + 2 * neig + 1) + 3 * madspace ** 2 + max(madspace ** 2, neig)


@dataclass
class JDResult:
    eigs: np.ndarray
    res: np.ndarray
    x: object
    neig: int
    iter: int
    info: int
    gap: float
    shift: float = 0.0
    droptol: float = 0.0
    stats: dict = field(default_factory=dict)


def _p(a, t=C.c_double):
    return a.ctypes.data_as(C.POINTER(t))


def _icntl(icntl):
    return np.array(list(icntl) + [0] * (5 - len(icntl)), dtype=np.int32)


def pjdrevcom_host(n, a, ja, ia, matvec, *, neig, isearch=0, sigma=0.0, ninit=0, madspace=20,
                   maxiter=1000, tol=1e-10, shift=0.0, droptol=1e-3, mem=20.0,
                   icntl=(0, 0, 0, 0, 0), x0=None, eigs0=None, lx=None, iprint=0):
    """DPJDREVCOM through the legacy symbol `dpjdrevcom_` (host X).

    `matvec(xin, xout)` must write A@xin into xout.  `ja[0] < 0` selects the
    diagonal preconditioner with `a` the diagonal (PJD.F:309-317)."""
    L = lib()
    lx = lx or lx_optimal(n, neig, madspace)
    x = np.zeros(lx)
    if x0 is not None:
        x[: np.size(x0)] = np.asarray(x0).ravel(order="F")
    a = np.array(a, dtype=np.float64)
    ja = np.array(ja, dtype=np.int32)
    ia = np.array(ia, dtype=np.int32)
    eigs = np.zeros(max(neig, 1)) if eigs0 is None else np.array(eigs0, dtype=np.float64)
    res = np.zeros(max(neig, 1))
    ic = _icntl(icntl)
    N, LX, NE, IS, NI, MS, IT = _i(n), _i(lx), _i(neig), _i(isearch), _i(ninit), _i(madspace), _i(maxiter)
    SG, TL, SH, DT, ME = _d(sigma), _d(tol), _d(shift), _d(droptol), _d(mem)
    IJ, N1, N2, IP, INFO, GAP = _i(0), _i(0), _i(0), _i(iprint), _i(0), _d(0)
    nmv = 0
    while True:
        L.dpjdrevcom_(C.byref(N), _p(a), _p(ja, C.c_int), _p(ia, C.c_int), _p(eigs), _p(res),
                      _p(x), C.byref(LX), C.byref(NE), C.byref(SG), C.byref(IS), C.byref(NI),
                      C.byref(MS), C.byref(IT), C.byref(TL), C.byref(SH), C.byref(DT),
                      C.byref(ME), _p(ic, C.c_int), C.byref(IJ), C.byref(N1), C.byref(N2),
                      C.byref(IP), C.byref(INFO), C.byref(GAP))
        if IJ.value != 1:
            break
        a1, a2 = N1.value - 1, N2.value - 1
        matvec(x[a1:a1 + n], x[a2:a2 + n])
        nmv += 1
    k = NE.value if INFO.value >= 0 else 0
    return JDResult(eigs=eigs[:k].copy(), res=res[:k].copy(),
                    x=x[: n * k].reshape(k, n).T.copy(), neig=k, iter=IT.value,
                    info=INFO.value, gap=GAP.value, shift=SH.value, droptol=DT.value,
                    stats={"nmv": nmv})


def jdrevcom_host(n, matvec, psolve, *, neig, isearch=0, sigma=0.0, ninit=0, madspace=20,
                  maxiter=1000, tol=1e-10, x0=None, eigs0=None, lx=None, iprint=0):
    """DJDREVCOM through the legacy symbol `djdrevcom_` (host X).

    `psolve(src, dst)` must write P^-1 src into dst."""
    L = lib()
    lx = lx or lx_optimal(n, neig, madspace)
    x = np.zeros(lx)
    if x0 is not None:
        x[: np.size(x0)] = np.asarray(x0).ravel(order="F")
    eigs = np.zeros(max(neig, 1)) if eigs0 is None else np.array(eigs0, dtype=np.float64)
    res = np.zeros(max(neig, 1))
    N, LX, NE, IS, NI, MS, IT = _i(n), _i(lx), _i(neig), _i(isearch), _i(ninit), _i(madspace), _i(maxiter)
    SG, TL = _d(sigma), _d(tol)
    IJ, N1, N2, IP, INFO, GAP = _i(0), _i(0), _i(0), _i(iprint), _i(0), _d(0)
    while True:
        L.djdrevcom_(C.byref(N), _p(eigs), _p(res), _p(x), C.byref(LX), C.byref(NE),
                     C.byref(SG), C.byref(IS), C.byref(NI), C.byref(MS), C.byref(IT),
                     C.byref(TL), C.byref(IJ), C.byref(N1), C.byref(N2), C.byref(IP),
                     C.byref(INFO), C.byref(GAP))
        a1, a2 = N1.value - 1, N2.value - 1
        if IJ.value == 1:
            matvec(x[a1:a1 + n], x[a2:a2 + n])
        elif IJ.value == 2:
            psolve(x[a2:a2 + n], x[a1:a1 + n])
        else:
            break
    k = NE.value if INFO.value >= 0 else 0
    return JDResult(eigs=eigs[:k].copy(), res=res[:k].copy(),
                    x=x[: n * k].reshape(k, n).T.copy(), neig=k, iter=IT.value,
                    info=INFO.value, gap=GAP.value)


def pjd_host(n, a, ja, ia, *, neig, isearch=0, sigma=0.0, ninit=0, madspace=20, maxiter=1000,
             tol=1e-10, shift=0.0, droptol=1e-3, mem=20.0, icntl=(0, 0, 0, 0, 0), x0=None,
             eigs0=None, lx=None, iprint=0):
    """DPJD through the legacy symbol `dpjd_` (host arrays, device solve)."""
    L = lib()
    lx = lx or lx_optimal(n, neig, madspace)
    x = np.zeros(lx)
    if x0 is not None:
        x[: np.size(x0)] = np.asarray(x0).ravel(order="F")
    a = np.array(a, dtype=np.float64)
    ja = np.array(ja, dtype=np.int32)
    ia = np.array(ia, dtype=np.int32)
    eigs = np.zeros(max(neig, 1)) if eigs0 is None else np.array(eigs0, dtype=np.float64)
    res = np.zeros(max(neig, 1))
    ic = _icntl(icntl)
    N, LX, NE, IS, NI, MS, IT = _i(n), _i(lx), _i(neig), _i(isearch), _i(ninit), _i(madspace), _i(maxiter)
    SG, TL, SH, DT, ME = _d(sigma), _d(tol), _d(shift), _d(droptol), _d(mem)
    IP, INFO, GAP = _i(iprint), _i(0), _d(0)
    L.dpjd_(C.byref(N), _p(a), _p(ja, C.c_int), _p(ia, C.c_int), _p(eigs), _p(res), _p(x),
            C.byref(LX), C.byref(NE), C.byref(SG), C.byref(IS), C.byref(NI), C.byref(MS),
            C.byref(IT), C.byref(TL), C.byref(SH), C.byref(DT), C.byref(ME), _p(ic, C.c_int),
            C.byref(IP), C.byref(INFO), C.byref(GAP))
    k = NE.value if INFO.value >= 0 else 0
    return JDResult(eigs=eigs[:k].copy(), res=res[:k].copy(),
                    x=x[: n * k].reshape(k, n).T.copy(), neig=k, iter=IT.value,
                    info=INFO.value, gap=GAP.value, shift=SH.value, droptol=DT.value)


def pjd_gep_host(n, a, ja, ia, b, jb, ib, *, neig, isearch=0, sigma=0.0, ninit=0, madspace=20,
                 maxiter=1000, tol=1e-10, shift=0.0, droptol=1e-3, mem=20.0,
                 icntl=(0, 0, 0, 0, 0), x0=None, eigs0=None, lx=None, iprint=0):
    """DPJD_GEP through the legacy symbol `dpjd_gep_`: A x = lambda B x."""
    L = lib()
    lx = lx or lx_optimal_gep(n, neig, madspace)
    x = np.zeros(lx)
    if x0 is not None:
        x[: np.size(x0)] = np.asarray(x0).ravel(order="F")
    a, b = np.array(a, dtype=np.float64), np.array(b, dtype=np.float64)
    ja, jb = np.array(ja, dtype=np.int32), np.array(jb, dtype=np.int32)
    ia, ib = np.array(ia, dtype=np.int32), np.array(ib, dtype=np.int32)
    eigs = np.zeros(max(neig, 1)) if eigs0 is None else np.array(eigs0, dtype=np.float64)
    res = np.zeros(max(neig, 1))
    ic = _icntl(icntl)
    N, LX, NE, IS, NI, MS, IT = _i(n), _i(lx), _i(neig), _i(isearch), _i(ninit), _i(madspace), _i(maxiter)
    SG, TL, SH, DT, ME = _d(sigma), _d(tol), _d(shift), _d(droptol), _d(mem)
    IP, INFO, GAP = _i(iprint), _i(0), _d(0)
    L.dpjd_gep_(C.byref(N), _p(a), _p(ja, C.c_int), _p(ia, C.c_int), _p(b), _p(jb, C.c_int),
                _p(ib, C.c_int), _p(eigs), _p(res), _p(x), C.byref(LX), C.byref(NE),
                C.byref(SG), C.byref(IS), C.byref(NI), C.byref(MS), C.byref(IT), C.byref(TL),
                C.byref(SH), C.byref(DT), C.byref(ME), _p(ic, C.c_int), C.byref(IP),
                C.byref(INFO), C.byref(GAP))
    k = NE.value if INFO.value >= 0 else 0
    return JDResult(eigs=eigs[:k].copy(), res=res[:k].copy(),
                    x=x[: n * k].reshape(k, n).T.copy(), neig=k, iter=IT.value,
                    info=INFO.value, gap=GAP.value, shift=SH.value, droptol=DT.value)


def pjdrevcom_gep_host(n, a, ja, ia, b, jb, ib, matvec_a, matvec_b, *, neig, isearch=0,
                       sigma=0.0, ninit=0, madspace=20, maxiter=1000, tol=1e-10, shift=0.0,
                       droptol=1e-3, mem=20.0, icntl=(0, 0, 0, 0, 0), lx=None, iprint=0):
    """DPJDREVCOM_GEP through `dpjdrevcom_gep_`; matvec_a(x, y) and
    matvec_b(x, y) write A x and B x into y (host numpy arrays).  Diagonal
    preconditioning: a = diagonal of A with ja = ia = [-1] (and b, jb, ib
    likewise), as in samples/EXAMPLE9_GEP.f."""
    L = lib()
    lx = lx or lx_optimal_gep(n, neig, madspace)
    x = np.zeros(lx)
    a, b = np.array(a, dtype=np.float64), np.array(b, dtype=np.float64)
    ja, jb = np.array(ja, dtype=np.int32), np.array(jb, dtype=np.int32)
    ia, ib = np.array(ia, dtype=np.int32), np.array(ib, dtype=np.int32)
    eigs, res = np.zeros(max(neig, 1)), np.zeros(max(neig, 1))
    ic = _icntl(icntl)
    N, LX, NE, IS, NI, MS, IT = _i(n), _i(lx), _i(neig), _i(isearch), _i(ninit), _i(madspace), _i(maxiter)
    SG, TL, SH, DT, ME = _d(sigma), _d(tol), _d(shift), _d(droptol), _d(mem)
    IJ, N1, N2, N3, IP, INFO, GAP = _i(0), _i(0), _i(0), _i(0), _i(iprint), _i(0), _d(0)
    counts = {1: 0, 3: 0, 4: 0}
    while True:
        L.dpjdrevcom_gep_(C.byref(N), _p(a), _p(ja, C.c_int), _p(ia, C.c_int), _p(b),
                          _p(jb, C.c_int), _p(ib, C.c_int), _p(eigs), _p(res), _p(x),
                          C.byref(LX), C.byref(NE), C.byref(SG), C.byref(IS), C.byref(NI),
                          C.byref(MS), C.byref(IT), C.byref(TL), C.byref(SH), C.byref(DT),
                          C.byref(ME), _p(ic, C.c_int), C.byref(IJ), C.byref(N1), C.byref(N2),
                          C.byref(N3), C.byref(IP), C.byref(INFO), C.byref(GAP))
        j = IJ.value
        if j not in (1, 3, 4):
            break
        counts[j] += 1
        xi = x[N1.value - 1: N1.value - 1 + n].copy()
        if j in (1, 3):
            y = np.empty(n)
            matvec_a(xi, y)
            x[N2.value - 1: N2.value - 1 + n] = y
        if j in (1, 4):
            y = np.empty(n)
            matvec_b(xi, y)
            x[N3.value - 1: N3.value - 1 + n] = y
    k = NE.value if INFO.value >= 0 else 0
    r = JDResult(eigs=eigs[:k].copy(), res=res[:k].copy(),
                 x=x[: n * k].reshape(k, n).T.copy(), neig=k, iter=IT.value,
                 info=INFO.value, gap=GAP.value, shift=SH.value, droptol=DT.value)
    r.ijob_counts = counts
    return r


def pjd_batch(problems, *, device=0, workers=8, want_x=False, options=None):
    """jdhip_pjd_batch: solve independent problems on one device.

    problems: list of dicts with keys ia, ja, a (1-based upper CSR) and
    optionally neig (5), isearch (0), sigma (0.0), madspace (20), maxiter
    (1000), tol (1e-10), shift (0.0), droptol (1e-3), mem (20.0), icntl
    ((0, 1, 0, 0, 0)).  options: {name: value} for every worker context
    (jdhip_pjd_batch_opts).  Returns a list of JDResult (x only if want_x)."""
    from . import JdhipBatchItem
    L = lib()
    items = (JdhipBatchItem * len(problems))()
    keep = []
    for k, pr in enumerate(problems):
        ia = np.ascontiguousarray(pr["ia"], dtype=np.int32)
        ja = np.ascontiguousarray(pr["ja"], dtype=np.int32)
        a = np.ascontiguousarray(pr["a"], dtype=np.float64)
        n = ia.size - 1
        neig = int(pr.get("neig", 5))
        eigs, res = np.zeros(neig), np.zeros(neig)
        x = np.zeros(n * neig) if want_x else None
        keep += [ia, ja, a, eigs, res, x]
        it = items[k]
        it.n, it.ia, it.ja, it.a = n, _p(ia, C.c_int), _p(ja, C.c_int), _p(a)
        it.neig, it.isearch = neig, int(pr.get("isearch", 0))
        it.madspace, it.maxiter = int(pr.get("madspace", 20)), int(pr.get("maxiter", 1000))
        it.sigma, it.tol = float(pr.get("sigma", 0.0)), float(pr.get("tol", 1e-10))
        it.shift, it.droptol = float(pr.get("shift", 0.0)), float(pr.get("droptol", 1e-3))
        it.mem = float(pr.get("mem", 20.0))
        for c, v in enumerate(pr.get("icntl", (0, 1, 0, 0, 0))):
            it.icntl[c] = int(v)
        it.eigs, it.res = _p(eigs), _p(res)
        it.x = _p(x) if want_x else C.cast(None, C.POINTER(C.c_double))
    opts = dict(options or {})
    names = (C.c_char_p * max(len(opts), 1))(*[k.encode() for k in opts])
    vals = np.array([float(v) for v in opts.values()] or [0.0])
    rc = L.jdhip_pjd_batch_opts(device, len(problems), items, workers, len(opts), names,
                                _p(vals))
    if rc < 0:
        raise ValueError("jdhip_pjd_batch: invalid arguments")
    out = []
    for k in range(len(problems)):
        it = items[k]
        n, ne = it.n, it.neig_out
        eigs, res, x = keep[6 * k + 3], keep[6 * k + 4], keep[6 * k + 5]
        out.append(JDResult(eigs=eigs[:ne].copy(), res=res[:ne].copy(),
                            x=None if x is None else x[: n * ne].reshape(ne, n).T.copy(),
                            neig=ne, iter=it.iter, info=it.info, gap=it.gap, shift=0.0,
                            droptol=0.0))
    return out


def cleanup():
    lib().dpjdcleanup_()


def legacy_stats():
    """Statistics (Context.stats() fields) of the last legacy-symbol call."""
    from . import JdhipStats
    L = lib()
    s = JdhipStats()
    L.jdhip_get_stats(L.jdhip_legacy_context(), C.byref(s))
    k = s.n_precond_setups
    return {"n_precond_setups": k, "setup_shifts": list(s.setup_shifts[:k]),
            "setup_droptols": list(s.setup_droptols[:k]), "setup_fills": list(s.setup_fills[:k]),
            "inner_iters": s.inner_iters, "precond_sweeps": s.precond_sweeps}


def legacy_set_option(name, value):
    """jdhip_set_option on the context behind the legacy symbols."""
    L = lib()
    rc = L.jdhip_set_option(L.jdhip_legacy_context(), name.encode(), float(value))
    if rc != 0:
        raise ValueError(f"option {name}: {rc}")


def reset_legacy():
    """Return the process-global context behind the legacy symbols to the
    state of a fresh process (preconditioner released, saved COMMON state
    zeroed), so a run does not inherit shifts from earlier problems."""
    L = lib()
    L.jdhip_reset(L.jdhip_legacy_context())


class Context:
    """A solver context on one HIP device (jdhip_create); X lives in device
    memory (a torch CUDA tensor is used as the allocator)."""

    def __init__(self, device=0):
        self._L = lib()
        self.device = device
        self.h = self._L.jdhip_create(device)
        if not self.h:
            raise RuntimeError("jdhip_create failed")

    def close(self):
        if self.h:
            self._L.jdhip_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def error(self):
        return self._L.jdhip_last_error(self.h).decode()

    def stats(self):
        s = JdhipStats()
        self._L.jdhip_get_stats(self.h, C.byref(s))
        return {"inner_iters": s.inner_iters, "host_syncs": s.host_syncs,
                "n_precond_setups": s.n_precond_setups, "spmv_lanes": s.spmv_lanes,
                "setup_shifts": list(s.setup_shifts[: s.n_precond_setups]),
                "setup_droptols": list(s.setup_droptols[: s.n_precond_setups]),
                "setup_fills": list(s.setup_fills[: s.n_precond_setups]),
                "precond_seconds": s.precond_seconds, "fill": s.fill,
                "precond_levels": s.precond_levels, "precond_negpiv": s.precond_negpiv,
                "precond_sweeps": s.precond_sweeps}

    def stream_ptr(self):
        return self._L.jdhip_stream(self.h)

    def set_option(self, name, value):
        rc = self._L.jdhip_set_option(self.h, name.encode(), float(value))
        if rc != 0:
            raise ValueError(f"unknown option {name}")

    def attach_ildl(self, factor, sweeps=0):
        """Upload an IldlFactor (jadamilu_amd.ildl) as the context's preconditioner."""
        rc = self._L.jdhip_ildl_attach(self.h, factor.h, sweeps)
        if rc != 0:
            raise RuntimeError(self.error())

    def precond_apply(self, z_ptr, r_ptr):
        rc = self._L.jdhip_precond_apply(self.h, C.c_void_p(z_ptr), C.c_void_p(r_ptr))
        if rc != 0:
            raise RuntimeError(self.error())

    def ildl_time_sweep(self, reps=20):
        """(avg ms, algorithmic bytes) of one Jacobi sweep launch over the
        first level's L_B, timed with HIP events on the context stream."""
        ms, nb = C.c_double(), C.c_long()
        rc = self._L.jdhip_ildl_time_sweep(self.h, reps, C.byref(ms), C.byref(nb))
        if rc != 0:
            raise RuntimeError(self.error() or "no ILDL preconditioner")
        return ms.value, nb.value

    def ildl_level_sweeps(self):
        """Per sparse level: 0 exact triangular solves, k Jacobi sweeps."""
        buf = (C.c_int * 64)()
        k = self._L.jdhip_ildl_level_sweeps(self.h, buf, 64)
        return list(buf[: max(k, 0)])

    def ildl_level_forms(self):
        """Per sparse level: kernel forms (forward sweep, backward sweep, L_CB,
        L_CB^T); 0 SELL one lane per row, 1 CSR-stream, k >= 4 CSR with k lanes
        per row."""
        buf = (C.c_int * 256)()
        k = self._L.jdhip_ildl_level_forms(self.h, buf, 256)
        return [list(buf[4 * l: 4 * l + 4]) for l in range(max(k, 0))]

    def ildl_device_info(self):
        f, b, nb = C.c_int(), C.c_int(), C.c_long()
        if self._L.jdhip_ildl_device_info(self.h, C.byref(f), C.byref(b), C.byref(nb)) != 0:
            return None
        return {"fwd_levels": f.value, "bwd_levels": b.value, "apply_bytes": nb.value}

    def set_matrix(self, ia, ja, a, lanes_per_row=0):
        ia = np.ascontiguousarray(ia, dtype=np.int32)
        ja = np.ascontiguousarray(ja, dtype=np.int32)
        a = np.ascontiguousarray(a, dtype=np.float64)
        self._keep = (ia, ja, a)
        rc = self._L.jdhip_set_matrix(self.h, ia.size - 1, _p(ia, C.c_int), _p(ja, C.c_int),
                                      _p(a), lanes_per_row)
        if rc != 0:
            raise RuntimeError(self.error())

    def spmm(self, k, x_ptr, ldx, y_ptr, ldy):
        """Y = A X for k device vectors (column-major)."""
        rc = self._L.jdhip_spmm(self.h, k, C.c_void_p(x_ptr), ldx, C.c_void_p(y_ptr), ldy)
        if rc != 0:
            raise RuntimeError(self.error() or f"jdhip_spmm: {rc}")

    def gram(self, n, k, m, v_ptr, ldv, w_ptr, ldw, c_ptr, ldc):
        """C = V^T W on the FP64 matrix cores (device arrays, column-major)."""
        rc = self._L.jdhip_gram(self.h, n, k, m, C.c_void_p(v_ptr), ldv, C.c_void_p(w_ptr), ldw,
                                C.c_void_p(c_ptr), ldc)
        if rc != 0:
            raise RuntimeError(self.error() or f"jdhip_gram: {rc}")

    def tall_gemm(self, n, k, j, v_ptr, ldv, q_ptr, ldq, y_ptr, ldy):
        """Y = V Q on the FP64 matrix cores (device arrays, column-major)."""
        rc = self._L.jdhip_tall_gemm(self.h, n, k, j, C.c_void_p(v_ptr), ldv, C.c_void_p(q_ptr),
                                     ldq, C.c_void_p(y_ptr), ldy)
        if rc != 0:
            raise RuntimeError(self.error() or f"jdhip_tall_gemm: {rc}")

    def spmv(self, x_ptr, y_ptr):
        rc = self._L.jdhip_spmv(self.h, C.c_void_p(x_ptr), C.c_void_p(y_ptr))
        if rc != 0:
            raise RuntimeError(self.error())

    def pjdrevcom(self, n, a, ja, ia, xdev_ptr, lx, *, neig, matvec=None, isearch=0, sigma=0.0,
                  ninit=0, madspace=20, maxiter=1000, tol=1e-10, shift=0.0, droptol=1e-3,
                  mem=20.0, icntl=(0, 0, 0, 0, 0), eigs0=None, iprint=0):
        """jdhip_pjdrevcom with X on the device.  IJOB=1 is serviced by
        `matvec(xin_ptr, xout_ptr)` or, if None, by the context's SpMV."""
        L = self._L
        a = np.array(a, dtype=np.float64)
        ja = np.array(ja, dtype=np.int32)
        ia = np.array(ia, dtype=np.int32)
        eigs = np.zeros(max(neig, 1)) if eigs0 is None else np.array(eigs0, dtype=np.float64)
        res = np.zeros(max(neig, 1))
        ic = _icntl(icntl)
        NE, IT, SH, DT, ME = _i(neig), _i(maxiter), _d(shift), _d(droptol), _d(mem)
        IJ, N1, N2, INFO, GAP = _i(0), C.c_long(0), C.c_long(0), _i(0), _d(0)
        nmv = 0
        while True:
            L.jdhip_pjdrevcom(self.h, n, _p(a), _p(ja, C.c_int), _p(ia, C.c_int), _p(eigs),
                              _p(res), C.c_void_p(xdev_ptr), lx, C.byref(NE), sigma, isearch,
                              ninit, madspace, C.byref(IT), tol, C.byref(SH), C.byref(DT),
                              C.byref(ME), _p(ic, C.c_int), C.byref(IJ), C.byref(N1),
                              C.byref(N2), iprint, C.byref(INFO), C.byref(GAP))
            if IJ.value != 1:
                break
            xi = xdev_ptr + 8 * (N1.value - 1)
            xo = xdev_ptr + 8 * (N2.value - 1)
            if matvec is None:
                self.spmv(xi, xo)
            else:
                matvec(xi, xo)
            nmv += 1
        k = NE.value if INFO.value >= 0 else 0
        st = self.stats()
        st["nmv"] = nmv
        return JDResult(eigs=eigs[:k].copy(), res=res[:k].copy(), x=None, neig=k,
                        iter=IT.value, info=INFO.value, gap=GAP.value, shift=SH.value,
                        droptol=DT.value, stats=st)

    def pjd(self, n, a, ja, ia, xdev_ptr, lx, *, neig, isearch=0, sigma=0.0, ninit=0,
            madspace=20, maxiter=1000, tol=1e-10, shift=0.0, droptol=1e-3, mem=20.0,
            icntl=(0, 0, 0, 0, 0), eigs0=None, iprint=0):
        """jdhip_pjd: internal device SpMV and preconditioner."""
        L = self._L
        a = np.array(a, dtype=np.float64)
        ja = np.array(ja, dtype=np.int32)
        ia = np.array(ia, dtype=np.int32)
        eigs = np.zeros(max(neig, 1)) if eigs0 is None else np.array(eigs0, dtype=np.float64)
        res = np.zeros(max(neig, 1))
        ic = _icntl(icntl)
        NE, IT, SH, DT, ME = _i(neig), _i(maxiter), _d(shift), _d(droptol), _d(mem)
        INFO, GAP = _i(0), _d(0)
        L.jdhip_pjd(self.h, n, _p(a), _p(ja, C.c_int), _p(ia, C.c_int), _p(eigs), _p(res),
                    C.c_void_p(xdev_ptr), lx, C.byref(NE), sigma, isearch, ninit, madspace,
                    C.byref(IT), tol, C.byref(SH), C.byref(DT), C.byref(ME), _p(ic, C.c_int),
                    iprint, C.byref(INFO), C.byref(GAP))
        k = NE.value if INFO.value >= 0 else 0
        return JDResult(eigs=eigs[:k].copy(), res=res[:k].copy(), x=None, neig=k,
                        iter=IT.value, info=INFO.value, gap=GAP.value, shift=SH.value,
                        droptol=DT.value, stats=self.stats())

    def pjd_gep(self, n, a, ja, ia, b, jb, ib, xdev_ptr, lx, *, neig, isearch=0, sigma=0.0,
                ninit=0, madspace=20, maxiter=1000, tol=1e-10, shift=0.0, droptol=1e-3,
                mem=20.0, icntl=(0, 0, 0, 0, 0), eigs0=None, iprint=0):
        """jdhip_pjd_gep: A x = lambda B x with device products by A and B."""
        L = self._L
        a, b = np.array(a, dtype=np.float64), np.array(b, dtype=np.float64)
        ja, jb = np.array(ja, dtype=np.int32), np.array(jb, dtype=np.int32)
        ia, ib = np.array(ia, dtype=np.int32), np.array(ib, dtype=np.int32)
        eigs = np.zeros(max(neig, 1)) if eigs0 is None else np.array(eigs0, dtype=np.float64)
        res = np.zeros(max(neig, 1))
        ic = _icntl(icntl)
        NE, IT, SH, DT, ME = _i(neig), _i(maxiter), _d(shift), _d(droptol), _d(mem)
        INFO, GAP = _i(0), _d(0)
        L.jdhip_pjd_gep(self.h, n, _p(a), _p(ja, C.c_int), _p(ia, C.c_int), _p(b),
                        _p(jb, C.c_int), _p(ib, C.c_int), _p(eigs), _p(res),
                        C.c_void_p(xdev_ptr), lx, C.byref(NE), sigma, isearch, ninit, madspace,
                        C.byref(IT), tol, C.byref(SH), C.byref(DT), C.byref(ME),
                        _p(ic, C.c_int), iprint, C.byref(INFO), C.byref(GAP))
        k = NE.value if INFO.value >= 0 else 0
        return JDResult(eigs=eigs[:k].copy(), res=res[:k].copy(), x=None, neig=k,
                        iter=IT.value, info=INFO.value, gap=GAP.value, shift=SH.value,
                        droptol=DT.value, stats=self.stats())

    def jdrevcom(self, n, xdev_ptr, lx, *, neig, matvec, psolve, isearch=0, sigma=0.0, ninit=0,
                 madspace=20, maxiter=1000, tol=1e-10, eigs0=None, iprint=0):
        """jdhip_jdrevcom: matvec(xin_ptr, xout_ptr), psolve(src_ptr, dst_ptr)."""
        L = self._L
        eigs = np.zeros(max(neig, 1)) if eigs0 is None else np.array(eigs0, dtype=np.float64)
        res = np.zeros(max(neig, 1))
        NE, IT = _i(neig), _i(maxiter)
        IJ, N1, N2, INFO, GAP = _i(0), C.c_long(0), C.c_long(0), _i(0), _d(0)
        while True:
            L.jdhip_jdrevcom(self.h, n, _p(eigs), _p(res), C.c_void_p(xdev_ptr), lx,
                             C.byref(NE), sigma, isearch, ninit, madspace, C.byref(IT), tol,
                             C.byref(IJ), C.byref(N1), C.byref(N2), iprint, C.byref(INFO),
                             C.byref(GAP))
            p1 = xdev_ptr + 8 * (N1.value - 1)
            p2 = xdev_ptr + 8 * (N2.value - 1)
            if IJ.value == 1:
                matvec(p1, p2)
            elif IJ.value == 2:
                psolve(p2, p1)
            else:
                break
        k = NE.value if INFO.value >= 0 else 0
        return JDResult(eigs=eigs[:k].copy(), res=res[:k].copy(), x=None, neig=k,
                        iter=IT.value, info=INFO.value, gap=GAP.value, stats=self.stats())
