"""Host factor against the user guide's printed preconditioner set-ups.

Every set-up printed in Doc/userguide.ps (fixture tests/golden/userguide_runs.json,
the "precond_setups" entries: used DROPTOL, used SHIFT, fill factor) is redone
with the host factorisation at that DROPTOL and SHIFT and compared with the
printed fill to the printed two decimals.  The fill factor is the reference's
(nnzU + 2n) / nnz(A_upper) with a dense last level counted as in
src/ilupack/symilupacknnz.c:80-125; JDaux.F:1226 prints it.

No GPU is needed: the factorisation runs on the host.
"""
import numpy as np
import pytest

from jadamilu_amd import matrices as M


def _factor(n, ia, ja, a, shift, droptol, spd):
    from jadamilu_amd.ildl import IldlFactor
    f = IldlFactor(n, ia, ja, a, shift=shift, droptol=droptol, condest=5.0, spd=spd)
    st = f.stats()
    assert st["ierr"] == 0
    return st["fill"]


def _ex_matrix(name):
    n = 1000
    if name in ("EXAMPLE1", "EXAMPLE2"):
        return (n,) + M.tridiag(n, np.arange(n) * 1.0, 5.0)  # tridiag(5, i-1, 5)
    if name == "EXAMPLE11":
        return (n,) + M.tridiag(n, np.arange(1, n + 1) * 1.0, 5.0)  # tridiag(5, i, 5)
    if name.startswith("LISTING5"):
        return (n,) + M.tridiag(n, np.zeros(n), 5.0)  # the ILU is of tridiag(5, 0, 5)
    raise KeyError(name)


# ISEARCH of each guide run: the SPD conversion of the pivots applies for
# ISEARCH <= 1 (JDaux.F:1092-1130).  Listing 5 calls with ISEARCH = 0.
_ISEARCH = {"EXAMPLE1": 2, "EXAMPLE2": 0, "EXAMPLE11": 0, "LISTING5_ADAPTIVE": 0,
            "LISTING5_STATIC": 0}

# Set-ups whose fill the host factor does not reproduce yet (strict=False:
# an xpass is reported, not failed)
_OPEN = {}


def _case(name, k):
    if name in _OPEN:
        return pytest.param(name, k, marks=pytest.mark.xfail(reason=_OPEN[name], strict=False))
    return (name, k)


_CASES = [_case(name, k) for name, cnt in [("EXAMPLE1", 1), ("EXAMPLE2", 3), ("EXAMPLE11", 1),
                                             ("LISTING5_ADAPTIVE", 3), ("LISTING5_STATIC", 1)]
          for k in range(cnt)]


@pytest.mark.parametrize("name,k", _CASES)
def test_guide_setup_fill(golden, name, k):
    s = golden[name]["calls"][0]["precond_setups"][k]
    n, ia, ja, a = _ex_matrix(name)
    fill = _factor(n, ia, ja, a, s["shift"], s["droptol"], _ISEARCH[name] <= 1)
    print(name, k, "fill", round(fill, 4), "guide", s["fill"])
    assert abs(fill - s["fill"]) <= 0.005 + 1e-12, (fill, s["fill"])
