"""CPU checks of the host multilevel ILDL factorisation.

What pins the factor (no reference build is available):
* exactness: with droptol = 0 nothing is dropped, so the multilevel factor
  (sparse levels + dense tail) is an exact factorisation and M^-1 A x = x to
  rounding, on tridiagonal, 2D-Laplace and indefinite Anderson matrices, with
  several sparse levels forced (small dense switch);
* inertia (Sylvester): with droptol = 0 the number of negative pivots equals
  the number of eigenvalues of A - shift below zero (dense LAPACK count);
* FP32 factor (the mixed path): exact factor to FP32 rounding;
* the oracle apply (oracle/jdo_ildl.c) is the checker: a planted sign or
  index error in it breaks M^-1 A x = x.
"""
import numpy as np
import pytest
import scipy.sparse as sp

from jadamilu_amd import matrices as M
from oracle import oracle as O

ildl = pytest.importorskip("jadamilu_amd.ildl")


def _relerr(ia, ja, a, shift, **kw):
    n = ia.size - 1
    f = ildl.IldlFactor(n, ia, ja, a, shift=shift, **kw)
    assert f.ierr == 0
    ex = f.export()
    S = M.to_scipy(ia, ja, a) - shift * sp.eye(n)
    x = np.random.default_rng(1).standard_normal(n)
    z = O.ildl_apply(ex, S @ x)
    return np.linalg.norm(z - x) / np.linalg.norm(x), f, ex


CASES = [
    ("tridiag", lambda: M.tridiag(300, np.arange(300) * 1.0, 5.0), 0.0),
    ("laplace2d", lambda: M.laplace((14, 13)), 0.0),
    ("anderson6", lambda: M.anderson(6, 16.5, 1), 0.0),
    ("anderson7_shift", lambda: M.anderson(7, 12.0, 2), 0.37),
]


@pytest.mark.parametrize("name,gen,shift", CASES)
@pytest.mark.parametrize("dense_max", [20, 1500])
def test_exact_factor(name, gen, shift, dense_max):
    ia, ja, a = gen()
    err, f, ex = _relerr(ia, ja, a, shift, droptol=0.0, dense_max=dense_max)
    assert err < 1e-11, err
    if dense_max == 20 and name != "tridiag":
        assert len(ex["levels"]) >= 1


@pytest.mark.parametrize("name,gen,shift", CASES)
def test_inertia(name, gen, shift):
    ia, ja, a = gen()
    n = ia.size - 1
    w = np.linalg.eigvalsh(M.to_dense(ia, ja, a) - shift * np.eye(n))
    f = ildl.IldlFactor(n, ia, ja, a, shift=shift, droptol=0.0, dense_max=20)
    assert f.stats()["nneg"] == int((w < 0).sum())


def test_single_precision_factor():
    ia, ja, a = M.anderson(6, 16.5, 1)
    err, f, ex = _relerr(ia, ja, a, 0.0, droptol=0.0, dense_max=20, single=True)
    assert err < 5e-3, err
    errd, _, _ = _relerr(ia, ja, a, 0.0, droptol=0.0, dense_max=20, single=False)
    assert errd < err  # FP32 arithmetic visibly coarser


def test_spd_conversion_positive():
    """spd=True replaces pivots by |d|: the preconditioner is SPD."""
    ia, ja, a = M.anderson(6, 16.5, 1)
    n = ia.size - 1
    f = ildl.IldlFactor(n, ia, ja, a, shift=0.0, droptol=1e-3, spd=True, dense_max=20)
    ex = f.export()
    P = np.column_stack([O.ildl_apply(ex, e) for e in np.eye(n)])
    np.testing.assert_allclose(P, P.T, atol=1e-10 * np.abs(P).max())
    assert np.linalg.eigvalsh(0.5 * (P + P.T)).min() > 0


def test_droptol_reduces_fill_and_stays_useful():
    ia, ja, a = M.laplace((24, 24, 24))
    err, f, ex = _relerr(ia, ja, a, 0.0, droptol=1e-3, spd=True)
    st = f.stats()
    assert st["fill"] < 20 and err < 0.2
    err0, f0, _ = _relerr(ia, ja, a, 0.0, droptol=1e-6, spd=True)
    assert f0.stats()["fill"] > st["fill"] and err0 < err


def _amd(ia, ja):
    import ctypes as C
    from jadamilu_amd import lib
    n = ia.size - 1
    perm = np.zeros(n, dtype=np.int32)
    ia = np.ascontiguousarray(ia, dtype=np.int32)
    ja = np.ascontiguousarray(ja, dtype=np.int32)
    P = C.POINTER(C.c_int)
    assert lib().jdhip_order_amd(n, ia.ctypes.data_as(P), ja.ctypes.data_as(P),
                                 perm.ctypes.data_as(P)) == 0
    return perm - 1


def _chol_nnz(S, p):
    Lc = np.linalg.cholesky(S[p][:, p].toarray())
    return int((np.abs(Lc) > 1e-14).sum())


def test_amd_ordering_is_a_fill_reducing_permutation():
    """The AMD ordering (jdhip_order_amd) is a permutation; on a tree it
    produces no fill at all (minimum degree eliminates leaves first), and on
    2D/3D grids the exact Cholesky factor in its order is far sparser than in
    the natural or reverse Cuthill-McKee order."""
    from scipy.sparse.csgraph import reverse_cuthill_mckee
    # random tree (each node i>0 hangs off a random earlier node)
    n = 400
    rng = np.random.default_rng(7)
    par = np.array([rng.integers(0, i) for i in range(1, n)])
    r = np.r_[np.arange(n), np.minimum(np.arange(1, n), par)]
    c = np.r_[np.arange(n), np.maximum(np.arange(1, n), par)]
    T = sp.csr_matrix((np.r_[np.full(n, 4.0), -np.ones(n - 1)], (r, c)), shape=(n, n))
    T.sort_indices()
    p = _amd((T.indptr + 1).astype(np.int32), (T.indices + 1).astype(np.int32))
    assert sorted(p.tolist()) == list(range(n))
    F = (T + T.T - sp.diags(T.diagonal())).tocsr()
    assert _chol_nnz(F, p) == 2 * n - 1
    for shape in [(30, 30), (10, 10, 10)]:
        ia, ja, a = M.laplace(shape)
        n = ia.size - 1
        S = M.to_scipy(ia, ja, a).tocsr()
        p = _amd(ia, ja)
        assert sorted(p.tolist()) == list(range(n))
        f_amd = _chol_nnz(S, p)
        f_rcm = _chol_nnz(S, reverse_cuthill_mckee(S, symmetric_mode=True))
        f_nat = _chol_nnz(S, np.arange(n))
        assert f_amd < 0.7 * f_rcm and f_amd < 0.5 * f_nat, (shape, f_amd, f_rcm, f_nat)


def _depth(lv):
    """Longest dependency chain of a level's unit lower triangle L_B."""
    nB, p, idx = lv["nB"], lv["lb_ptr"], lv["lb_idx"]
    d = np.zeros(max(nB, 1), np.int64)
    for i in range(nB):
        cols = idx[p[i]:p[i + 1]]
        d[i] = (d[cols].max() + 1) if cols.size else 0
    return int(d.max()) if nB else 0


@pytest.mark.parametrize("gen,spd", [(lambda: M.laplace((9, 8, 7)), True),
                                     (lambda: M.anderson(6, 16.5, 3), False)])
def test_oracle_sweep_mode_is_exact_beyond_the_depth(gen, spd):
    """The oracle's Jacobi sweep mode (x <- b - N x from x = b, the checker
    of the device sweeps): N is strictly triangular with longest chain
    `depth` edges, so N^(depth+1) = 0 and k sweeps from x = b give the
    Neumann sum up to N^k: `depth` sweeps reproduce the exact substitution
    on every level, one sweep does not (starting from 0 instead of b, or an
    extra or missing sweep, breaks one of the two on these factors)."""
    ia, ja, a = gen()
    n = ia.size - 1
    f = ildl.IldlFactor(n, ia, ja, a, droptol=1e-3, spd=spd, dense_max=50)
    assert f.ierr == 0
    ex = f.export()
    depth = max(_depth(lv) for lv in ex["levels"])
    assert depth >= 2
    r = np.random.default_rng(4).standard_normal(n)
    z0 = O.ildl_apply(ex, r)
    zk = O.ildl_apply(ex, r, sweeps=depth)
    np.testing.assert_allclose(zk, z0, rtol=0, atol=1e-12 * np.abs(z0).max())
    z1 = O.ildl_apply(ex, r, sweeps=1)
    assert np.linalg.norm(z1 - z0) > 1e-6 * np.linalg.norm(z0)


def test_two_by_two_pivots_from_the_matching():
    """On indefinite matrices the symmetric weighted matching pairs rows into
    2x2 pivots (symwm.c, sympiluc.F:1042-1200): the blocks are adjacent
    (dpair[b] = b +- 1, mutual), nonsingular, counted into the inertia, and
    |D| of every 2x2 block is positive definite under the SPD conversion.
    The exact (droptol 0) factor with them reproduces A (M^-1 A x = x)."""
    ia, ja, a = M.anderson(8, 16.5, 6)
    n = ia.size - 1
    err, f, ex = _relerr(ia, ja, a, 0.0, droptol=0.0, dense_max=20)
    assert err < 1e-11, err
    npairs = 0
    for lv in ex["levels"]:
        dp, d, o = lv["dpair"], lv["d"], lv["doff"]
        for b in range(lv["nB"]):
            q = dp[b]
            if q < 0:
                assert o[b] == 0.0
                continue
            assert abs(q - b) == 1 and dp[q] == b and o[q] == o[b]
            if q == b + 1:
                npairs += 1
                assert d[b] * d[q] - o[b] ** 2 != 0.0
    assert npairs > 0
    w = np.linalg.eigvalsh(M.to_dense(ia, ja, a))
    assert f.stats()["nneg"] == int((w < 0).sum())
    g = ildl.IldlFactor(n, ia, ja, a, shift=0.0, droptol=1e-3, spd=True, dense_max=20)
    for lv in g.export()["levels"]:
        dp, d, o = lv["dpair"], lv["d"], lv["doff"]
        for b in range(lv["nB"]):
            if dp[b] < 0:
                assert d[b] > 0
            elif dp[b] == b + 1:
                q = b + 1
                assert d[b] > 0 and d[q] > 0 and d[b] * d[q] - o[b] ** 2 > 0
