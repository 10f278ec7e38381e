"""Pins the CPU oracle (oracle/) before it is trusted as the checker.

* Known answers: the user guide's printed runs of samples/EXAMPLE9.f and
  EXAMPLE10.f (tests/golden/userguide_runs.json, extracted from
  Doc/userguide.ps by tests/golden/extract_userguide.py): eigenvalues,
  printed residual norms, MATVEC count, estimated gap, suggested shift and
  the sequence of adaptive diagonal-preconditioner shifts.
* Mathematics: eigenvalues vs dense LAPACK (scipy dsyevd), Laplacian closed
  forms, residual/orthonormality invariants.
* The dense kernels (Jacobi syev/sygv, SORTRX, Householder QR) vs numpy/scipy.
A planted mistake in the oracle (dropped term, wrong sign or index) moves the
MATVEC count or the eigenvalues and fails these checks.
"""
import numpy as np
import pytest
import scipy.linalg as sla

from jadamilu_amd import matrices as M
from oracle import oracle as O


def ex9_matrix(n=1000):
    return M.tridiag(n, np.arange(1, n + 1, dtype=float), 5.0)


def fmt(x):
    return "%.1E" % x


@pytest.mark.parametrize("name,icntl", [("EXAMPLE9", (0, 0, 0, 0, 0)),
                                        ("EXAMPLE10", (2, 1, 0, 0, 0))])
def test_guide_diag_precond(golden, name, icntl):
    g = golden[name]["calls"][0]
    ia, ja, a = ex9_matrix()
    r = O.pjdrevcom_diag(1000, np.arange(1, 1001, dtype=float), O.matvec_upper(ia, ja, a),
                         neig=5, isearch=0, madspace=20, maxiter=1000, tol=1e-10,
                         icntl=icntl)
    assert r.info == g["info"] == 0
    assert r.iter == g["matvec"]
    np.testing.assert_allclose(r.eigs, [t["eig"] for t in g["table"]], rtol=0, atol=3e-15)
    assert "%.2E" % r.gap == g["gap"]
    assert "%.2E" % r.shift == g["suggested_shift"]
    np.testing.assert_allclose(r.shifts, [s["shift"] for s in g["precond_setups"]], rtol=1e-4)
    assert (r.res <= 1e-10).all()
    if name == "EXAMPLE9":
        # every printed residual norm reproduces
        assert [fmt(x) for x in r.res] == [t["res"] for t in g["table"]]
    else:
        assert [fmt(x) for x in r.res[:2]] == [t["res"] for t in g["table"][:2]]


def test_guide_eigs_vs_dense(golden):
    ia, ja, a = ex9_matrix()
    w = sla.eigh(M.to_dense(ia, ja, a), eigvals_only=True, driver="evd")
    np.testing.assert_allclose(w[:5], [t["eig"] for t in golden["EXAMPLE9"]["calls"][0]["table"]],
                               atol=5e-14)
    # EXAMPLE2 prints the smallest eigenvalues of tridiag(5, i-1, 5)
    np.testing.assert_allclose(w[:5] - 1, [t["eig"] for t in golden["EXAMPLE2"]["calls"][0]["table"]],
                               atol=5e-14)


def test_eigenvectors_invariants():
    ia, ja, a = ex9_matrix(400)
    A = M.to_dense(ia, ja, a)
    r = O.pjdrevcom_diag(400, np.arange(1, 401, dtype=float), O.matvec_upper(ia, ja, a),
                         neig=4, madspace=12, maxiter=2000, tol=1e-10)
    X = r.x
    np.testing.assert_allclose(X.T @ X, np.eye(4), atol=1e-12)
    for k in range(4):
        assert np.linalg.norm(A @ X[:, k] - r.eigs[k] * X[:, k]) <= 1.01e-10
    w = np.linalg.eigvalsh(A)
    np.testing.assert_allclose(r.eigs, w[:4], atol=1e-12)


def test_laplace2d_closed_form_small():
    shape = (30, 30)
    ia, ja, a = M.laplace(shape)
    lam, nxt = M.laplace_eigs(shape, 6)
    x0 = np.random.default_rng(0).random(900)
    r = O.pjdrevcom_diag(900, np.full(900, 4.0), O.matvec_upper(ia, ja, a), neig=6,
                         madspace=20, maxiter=20000, tol=1e-10, ninit=1, x0=x0)
    assert r.info == 0
    np.testing.assert_allclose(r.eigs, lam, rtol=1e-10)


def test_harmonic_interior_jdrevcom():
    """ISEARCH=2 (harmonic Ritz, sQMR inner solver) via DJDREVCOM with an
    exact shifted solve as the user preconditioner; interior eigenvalues
    nearest sigma vs dense LAPACK."""
    ia, ja, a = ex9_matrix(300)
    A = M.to_dense(ia, ja, a)
    sigma = 150.3
    P = np.linalg.inv(A - (sigma - 0.7) * np.eye(300))
    r = O.jdrevcom(300, O.matvec_upper(ia, ja, a), lambda src, dst: dst.__setitem__(slice(None), P @ src),
                   neig=5, isearch=2, sigma=sigma, madspace=15, maxiter=1000, tol=1e-10)
    assert r.info == 0 and r.neig == 5
    w = np.linalg.eigvalsh(A)
    near = np.sort(w[np.argsort(np.abs(w - sigma))[:5]])
    np.testing.assert_allclose(np.sort(r.eigs), near, atol=1e-11)
    assert (r.res <= 1e-10).all()


def test_input_validation():
    mv = lambda x, y: y.__setitem__(slice(None), x)
    for kw, code in [(dict(neig=0), -5), (dict(neig=2, madspace=2), -9),
                     (dict(neig=2, maxiter=0), -11), (dict(neig=2, tol=0.0), -12),
                     (dict(neig=2, lx=10), -10)]:
        r = O.pjdrevcom_diag(10, np.ones(10), mv, **kw)
        assert r.info == code, (kw, r.info)


def test_syev_sygv_vs_lapack():
    rng = np.random.default_rng(3)
    for j in (1, 2, 5, 13, 25):
        B = rng.standard_normal((j, j))
        S = B + B.T
        w, V = O.syev(S)
        np.testing.assert_allclose(w, np.linalg.eigvalsh(S), atol=1e-13 * max(1, abs(w).max()))
        np.testing.assert_allclose(V.T @ V, np.eye(j), atol=1e-13)
        np.testing.assert_allclose(S @ V, V * w, atol=1e-12)
        N = B @ B.T + j * np.eye(j)
        info, w2, Y = O.sygv(S, N)
        assert info == 0
        np.testing.assert_allclose(w2, sla.eigh(S, N, eigvals_only=True), atol=1e-12)
        np.testing.assert_allclose(Y.T @ N @ Y, np.eye(j), atol=1e-11)
    info, _, _ = O.sygv(np.eye(3), -np.eye(3))
    assert info == 4


def test_sortrx():
    rng = np.random.default_rng(0)
    for n in (1, 5, 9, 10, 37, 200):
        d = rng.standard_normal(n)
        idx = O.sortrx(d) - 1
        assert np.all(np.diff(d[idx]) >= 0)
        assert sorted(idx) == list(range(n))


@pytest.mark.parametrize("name,isearch,shift,droptol,tol", [
    ("EXAMPLE1", 2, 0.0, 1e-3, 1e-10),
    ("EXAMPLE11", 0, -8.0, 0.0, 1e-3),
])
def test_guide_ildl_runs(golden, name, isearch, shift, droptol, tol):
    """jdo_pjd (the oracle's DPJD) preconditioned by the host multilevel
    factor at the guide's printed set-up (one set-up in these runs, so the
    oracle's missing re-factorisation logic does not enter): eigenvalues to
    1e-10 of the printed ones; the MATVEC count within 3 of the guide's for
    the exact factor (EXAMPLE11, DROPTOL 0) and within 10 % where dropping
    makes the count depend on the factor's values, which follow the
    reference's fill but not its matching and ordering bit for bit (the
    maximum-product matching's dual scaling is not unique, and the
    reference's MC64 and AMD are not available to compare with)."""
    from jadamilu_amd.ildl import IldlFactor
    g = golden[name]["calls"][0]
    n = 1000
    diag = np.arange(n) * 1.0 if name == "EXAMPLE1" else np.arange(1, n + 1) * 1.0
    ia, ja, a = M.tridiag(n, diag, 5.0)
    f = IldlFactor(n, ia, ja, a, shift=shift, droptol=droptol, condest=5.0, spd=isearch <= 1)
    assert f.stats()["ierr"] == 0
    r = O.pjd_ildl(n, ia, ja, a, f.export(), neig=5, isearch=isearch, sigma=0.0, madspace=20,
                   maxiter=1000, tol=tol, icntl=(0, 1, 0, 0, 0))
    print(name, "oracle matvecs", r.iter, "guide", g["matvec"])
    assert r.info == 0 and r.neig == 5
    rt = 1e-10 if tol <= 1e-10 else 1e-6
    np.testing.assert_allclose(np.sort(r.eigs), np.sort([t["eig"] for t in g["table"]]),
                               rtol=rt, atol=rt)
    slack = 3 if droptol == 0.0 else max(3, int(0.1 * g["matvec"]))
    assert abs(r.iter - g["matvec"]) <= slack, (r.iter, g["matvec"])
