"""CPU checks of the small real nonsymmetric eigensolver that replaces DGEEV
in SOLPROJEP's harmonic branch of the generalized problem
(src/jd/JDaux.F:3293-3300), through jdhip_dense_eig: eigenvalues against
numpy.linalg.eigvals (LAPACK), and A v = lambda v with unit-norm vectors for
real eigenvalues and complex pairs."""
import ctypes as C

import numpy as np
import pytest


def _eig(A):
    from jadamilu_amd import lib
    L = lib()
    n = A.shape[0]
    a = np.asfortranarray(A.copy())
    wr, wi = np.zeros(max(n, 1)), np.zeros(max(n, 1))
    v = np.zeros((max(n, 1), max(n, 1)), order="F")
    p = lambda x: x.ctypes.data_as(C.POINTER(C.c_double))
    rc = L.jdhip_dense_eig(n, p(a), max(n, 1), p(wr), p(wi), p(v), max(n, 1))
    return rc, wr[:n], wi[:n], v[:n, :n]


def _check(A, tol=1e-10):
    n = A.shape[0]
    rc, wr, wi, v = _eig(A)
    assert rc == 0
    lam = wr + 1j * wi
    ref = np.linalg.eigvals(A)
    scale = max(np.abs(ref).max(), 1.0)
    # every computed eigenvalue matches a LAPACK one and vice versa
    for l in lam:
        assert np.min(np.abs(ref - l)) <= tol * scale, (l, ref)
    for l in ref:
        assert np.min(np.abs(lam - l)) <= tol * scale, (l, lam)
    k = 0
    while k < n:
        if wi[k] == 0.0:
            x = v[:, k]
            assert abs(np.linalg.norm(x) - 1.0) < 1e-12
            assert np.linalg.norm(A @ x - wr[k] * x) <= 1e-9 * scale
            k += 1
        else:
            assert wi[k] > 0 and wi[k + 1] == -wi[k] and wr[k + 1] == wr[k]
            x = v[:, k] + 1j * v[:, k + 1]
            assert abs(np.linalg.norm(x) - 1.0) < 1e-12
            assert abs(x[np.argmax(np.abs(x))].imag) < 1e-14
            assert np.linalg.norm(A @ x - lam[k] * x) <= 1e-9 * scale
            k += 2


@pytest.mark.parametrize("n", [1, 2, 3, 5, 8, 13, 20, 25])
@pytest.mark.parametrize("seed", [0, 1, 2])
def test_random_matrices(n, seed):
    rng = np.random.default_rng(100 * n + seed)
    _check(rng.standard_normal((n, n)))


def test_structured_matrices():
    rng = np.random.default_rng(7)
    S = rng.standard_normal((12, 12))
    _check(S + S.T)                                   # symmetric: real spectrum
    _check(np.triu(rng.standard_normal((9, 9))))       # already triangular
    R = np.zeros((6, 6))
    for k in range(0, 6, 2):                           # rotations: complex pairs
        c, s = np.cos(0.3 * (k + 1)), np.sin(0.3 * (k + 1))
        R[k:k + 2, k:k + 2] = [[c, -s], [s, c]]
    Q, _ = np.linalg.qr(rng.standard_normal((6, 6)))
    _check(Q @ R @ Q.T)
    _check(np.diag([3.0, -1.0, 2.0, 2.0 + 1e-3]))      # near-equal eigenvalues
    _check(np.zeros((4, 4)))
    # the harmonic projected matrix of the generalized problem: NN^-1 MM with
    # MM = W^T B V, NN = W^T W, W = (A - sigma B) V
    n, j = 60, 12
    A = rng.standard_normal((n, n)); A = A + A.T
    B = rng.standard_normal((n, n)); B = B @ B.T + n * np.eye(n)
    V = rng.standard_normal((n, j))
    W = (A - 0.3 * B) @ V
    _check(np.linalg.solve(W.T @ W, W.T @ B @ V))


def test_rejects_bad_input():
    from jadamilu_amd import lib
    assert lib().jdhip_dense_eig(-1, None, 1, None, None, None, 1) == -1
