"""Python access to the multilevel ILDL factor of the native library.

`factor(...)` runs the host factorisation (no GPU needed) and `export()`
copies every level out as numpy arrays (used by the tests to check the
factor against the mathematics and the device apply against the oracle).
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import lib

_ip = C.POINTER(C.c_int)
_dp = C.POINTER(C.c_double)


def _p(a, t):
    return a.ctypes.data_as(C.POINTER(t))


class IldlFactor:
    def __init__(self, n, ia, ja, a, *, shift=0.0, droptol=1e-3, condest=5.0, spd=False,
                 single=False, ordering=2, dense_max=1500, coupling=0):
        L = lib()
        self._L = L
        ia = np.ascontiguousarray(ia, dtype=np.int32)
        ja = np.ascontiguousarray(ja, dtype=np.int32)
        a = np.ascontiguousarray(a, dtype=np.float64)
        err = C.c_int(0)
        self.h = L.jdhip_ildl_factor(n, _p(ia, C.c_int), _p(ja, C.c_int), _p(a, C.c_double),
                                     shift, droptol, condest, int(spd), int(single), ordering,
                                     dense_max, int(coupling), C.byref(err))
        self.ierr = err.value
        self.n = n
        self.single = single

    def save(self, path):
        """Write the factor to a file (jdhip_ildl_save)."""
        rc = self._L.jdhip_ildl_save(self.h, str(path).encode())
        if rc != 0:
            raise OSError(f"jdhip_ildl_save({path}): {rc}")

    @classmethod
    def load(cls, path):
        """Read a factor written by save() (jdhip_ildl_load)."""
        L = lib()
        err = C.c_int(0)
        h = L.jdhip_ildl_load(str(path).encode(), C.byref(err))
        if not h:
            raise OSError(f"jdhip_ildl_load({path}): {err.value}")
        f = cls.__new__(cls)
        f._L = L
        f.h = h
        f.ierr = 0
        m = C.c_int()
        nlev = L.jdhip_ildl_nlevels(h, C.byref(m))
        if nlev > 0:
            n, nB, nlb, nlc = C.c_int(), C.c_int(), C.c_long(), C.c_long()
            L.jdhip_ildl_level_sizes(h, 0, C.byref(n), C.byref(nB), C.byref(nlb), C.byref(nlc))
            f.n = n.value
        else:
            f.n = m.value
        f.single = None
        return f

    def close(self):
        if self.h:
            self._L.jdhip_ildl_free(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def stats(self):
        fill, cond, sec = C.c_double(), C.c_double(), C.c_double()
        nneg = C.c_int()
        ierr = self._L.jdhip_ildl_stats(self.h, C.byref(fill), C.byref(nneg), C.byref(cond),
                                        C.byref(sec))
        return {"ierr": ierr, "fill": fill.value, "nneg": nneg.value, "condest": cond.value,
                "seconds": sec.value}

    def export(self):
        L = self._L
        m = C.c_int()
        nlev = L.jdhip_ildl_nlevels(self.h, C.byref(m))
        levels = []
        for l in range(nlev):
            n, nB = C.c_int(), C.c_int()
            nlb, nlc = C.c_long(), C.c_long()
            L.jdhip_ildl_level_sizes(self.h, l, C.byref(n), C.byref(nB), C.byref(nlb), C.byref(nlc))
            n, nB, nlb, nlc = n.value, nB.value, nlb.value, nlc.value
            nC = n - nB
            lev = {
                "n": n, "nB": nB,
                "perm": np.zeros(n, np.int32), "scal": np.zeros(n),
                "lb_ptr": np.zeros(nB + 1, np.int32), "lb_idx": np.zeros(max(nlb, 1), np.int32),
                "lb_val": np.zeros(max(nlb, 1)),
                "lc_ptr": np.zeros(nC + 1, np.int32), "lc_idx": np.zeros(max(nlc, 1), np.int32),
                "lc_val": np.zeros(max(nlc, 1)), "d": np.zeros(max(nB, 1)),
            }
            L.jdhip_ildl_level_data(self.h, l, _p(lev["perm"], C.c_int), _p(lev["scal"], C.c_double),
                                    _p(lev["lb_ptr"], C.c_int), _p(lev["lb_idx"], C.c_int),
                                    _p(lev["lb_val"], C.c_double), _p(lev["lc_ptr"], C.c_int),
                                    _p(lev["lc_idx"], C.c_int), _p(lev["lc_val"], C.c_double),
                                    _p(lev["d"], C.c_double))
            doff = np.zeros(max(nB, 1))
            dpair = np.full(max(nB, 1), -1, np.int32)
            L.jdhip_ildl_level_pivots(self.h, l, _p(lev["d"], C.c_double), _p(doff, C.c_double),
                                      _p(dpair, C.c_int))
            lev["doff"], lev["dpair"] = doff[:nB], dpair[:nB]
            lev["lb_idx"], lev["lb_val"] = lev["lb_idx"][:nlb], lev["lb_val"][:nlb]
            lev["lc_idx"], lev["lc_val"] = lev["lc_idx"][:nlc], lev["lc_val"][:nlc]
            lev["d"] = lev["d"][:nB]
            levels.append(lev)
        minv = np.zeros(m.value * m.value)
        if m.value:
            L.jdhip_ildl_dense(self.h, _p(minv, C.c_double))
        return {"levels": levels, "m": m.value, "minv": minv.reshape(m.value, m.value),
                "coupling": L.jdhip_ildl_coupling(self.h)}
