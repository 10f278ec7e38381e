"""The complex Hermitian drivers on the GPU (ZPJD, ZPJDREVCOM, ZJDREVCOM,
ZPJD_GEP, ZJDREVCOM_GEP; Doc/ZPJD_prototypes.f).

The reference holds no printed run of its complex samples, so the checks are
what the mathematics fixes:
* eigenvalues against the dense Hermitian spectrum (numpy.linalg.eigh, or
  scipy.linalg.eigh(H, M) for a pencil) to 1e-10 relative to the norm;
* residuals ||H z - lambda z|| <= TOL recomputed on the host from the dense
  matrix, RES within a few TOL of them, and complex-orthonormal eigenvectors
  (Z^H Z = I), also for an eigenvalue of complex multiplicity two;
* the closed form of samples/EXAMPLE4_double_complex.f (diagonal 0, upper
  neighbour i: eigenvalues 2 cos(k pi / (n + 1)));
* the matrices and settings of samples/EXAMPLE1-4, 9, 1_GEP
  _double_complex.f, a Hofstadter torus (phases that do not gauge away), and
  reverse communication with the caller's product and preconditioner.
"""
import numpy as np
import pytest
import scipy.linalg as sla

from tests.test_complex_form import hofstadter, tri_example

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")

N = 1000
TOL = 1e-10


def _dev():
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")


@pytest.fixture(autouse=True)
def _fresh_complex():
    _dev()
    from jadamilu_amd import zjd
    zjd.reset()
    zjd.set_option("ildl_sweeps", 0)  # exact triangular solves
    zjd.set_option("cost_clock", 1)
    yield
    zjd.cleanup()


def _check(H, r, want_eigs, *, tol=TOL, slack=5.0):
    assert r.info == 0, r.info
    k = len(want_eigs)
    assert r.neig == k
    scale = max(1.0, np.abs(np.linalg.eigvalsh(H)).max())
    assert np.allclose(np.sort(r.eigs), np.sort(want_eigs), rtol=0, atol=1e-10 * scale)
    Z = r.x
    assert Z.shape == (H.shape[0], k)
    assert np.allclose(Z.conj().T @ Z, np.eye(k), atol=1e-8)
    true_res = np.linalg.norm(H @ Z - Z * r.eigs, axis=0)
    assert np.all(true_res <= slack * tol), true_res
    assert np.all(np.abs(r.res - true_res) <= slack * tol)


def _nearest(lam, sigma, k):
    return lam[np.argsort(np.abs(lam - sigma), kind="stable")[:k]]


def test_example1_interior_eigenvalues():
    """samples/EXAMPLE1_double_complex.f: the 5 eigenvalues nearest 0 with
    ILU preconditioning (ISEARCH = 2, MEM 20, DROPTOL 1e-3)."""
    from jadamilu_amd import zjd
    ia, ja, a = tri_example(N)
    H = zjd.hermitian_dense(N, a, ja, ia)
    lam = np.linalg.eigvalsh(H)
    r = zjd.zpjd_host(N, a, ja, ia, neig=5, isearch=2, sigma=0.0, madspace=20, maxiter=1000,
                      tol=TOL, droptol=1e-3, mem=20.0)
    _check(H, r, _nearest(lam, 0.0, 5))
    assert r.iter <= 1000


def test_example2_smallest_eigenvalues():
    """samples/EXAMPLE2_double_complex.f: the 5 smallest (ISEARCH = 0)."""
    from jadamilu_amd import zjd
    ia, ja, a = tri_example(N)
    H = zjd.hermitian_dense(N, a, ja, ia)
    r = zjd.zpjd_host(N, a, ja, ia, neig=5, isearch=0, madspace=20, maxiter=1000, tol=TOL,
                      droptol=1e-3, mem=20.0)
    _check(H, r, np.linalg.eigvalsh(H)[:5])


def test_example3_sigma_and_shift():
    """samples/EXAMPLE3_double_complex.f: ISEARCH = 1 with SIGMA = -7 and
    SHIFT = -7.1."""
    from jadamilu_amd import zjd
    ia, ja, a = tri_example(N)
    H = zjd.hermitian_dense(N, a, ja, ia)
    r = zjd.zpjd_host(N, a, ja, ia, neig=5, isearch=1, sigma=-7.0, shift=-7.1, madspace=20,
                      maxiter=1000, tol=TOL, droptol=1e-3, mem=20.0)
    _check(H, r, np.linalg.eigvalsh(H)[:5])


def test_example4_closed_form():
    """samples/EXAMPLE4_double_complex.f: diagonal 0, upper neighbour i; the
    eigenvalues are 2 cos(k pi / (n + 1)), the 5 smallest for ISEARCH = 1,
    SIGMA = -1.9, SHIFT = -1."""
    from jadamilu_amd import zjd
    ia, ja, a = tri_example(N, diag=lambda i: 0.0, off=1j)
    H = zjd.hermitian_dense(N, a, ja, ia)
    closed = np.sort(2 * np.cos(np.arange(1, N + 1) * np.pi / (N + 1)))[:5]
    r = zjd.zpjd_host(N, a, ja, ia, neig=5, isearch=1, sigma=-1.9, shift=-1.0, madspace=20,
                      maxiter=1000, tol=TOL, droptol=1e-3, mem=20.0)
    assert r.info == 0 and r.neig == 5
    assert np.allclose(np.sort(r.eigs), closed, rtol=0, atol=1e-10)
    _check(H, r, closed)


def test_example9_revcom_diagonal_preconditioner():
    """samples/EXAMPLE9_double_complex.f: ZPJDREVCOM with JA(1) = -1
    (diagonal preconditioning from D) and the caller's product."""
    from jadamilu_amd import zjd
    ia, ja, a = tri_example(N, diag=float)  # diagonal i
    H = zjd.hermitian_dense(N, a, ja, ia)
    d = np.diag(H).astype(np.complex128)
    r = zjd.zpjdrevcom_host(N, d, np.array([-1]), np.array([-1]),
                            lambda x, y: np.copyto(y, H @ x), neig=5, isearch=0, madspace=20,
                            maxiter=1000, tol=TOL)
    _check(H, r, np.linalg.eigvalsh(H)[:5])
    assert r.matvec_calls == r.iter


def test_hofstadter_torus_smallest_and_lower_triangle():
    """A Hofstadter torus (flux 1/8, random diagonal): complex phases that do
    not gauge away; the same eigenpairs from the upper and from the lower
    triangle."""
    from jadamilu_amd import zjd
    n, ia, ja, a = hofstadter(32, 0.125, 3)
    H = zjd.hermitian_dense(n, a, ja, ia)
    lam = np.linalg.eigvalsh(H)
    r = zjd.zpjd_host(n, a, ja, ia, neig=6, isearch=0, madspace=20, maxiter=2000, tol=TOL,
                      droptol=1e-3, mem=20.0)
    _check(H, r, lam[:6])
    zjd.reset()
    zjd.set_option("ildl_sweeps", 0)
    zjd.set_option("cost_clock", 1)
    n, ib, jb, b = hofstadter(32, 0.125, 3, lower=True)
    assert np.allclose(zjd.hermitian_dense(n, b, jb, ib), H)
    r2 = zjd.zpjd_host(n, b, jb, ib, neig=6, isearch=0, madspace=20, maxiter=2000, tol=TOL,
                       droptol=1e-3, mem=20.0)
    _check(H, r2, lam[:6])


def test_multiple_eigenvalue_gives_orthonormal_vectors():
    """Two uncoupled copies of a Hofstadter torus: every eigenvalue has
    complex multiplicity two (four real copies); the drivers return two
    complex-orthonormal eigenvectors for it."""
    from jadamilu_amd import zjd
    n1, ia, ja, a = hofstadter(12, 0.25, 5)
    n = 2 * n1
    ia2 = np.concatenate([ia, ia[1:] + ia[-1] - 1])
    ja2 = np.concatenate([ja, ja + n1])
    a2 = np.concatenate([a, a])
    H = zjd.hermitian_dense(n, a2, ja2, ia2)
    lam = np.linalg.eigvalsh(H)
    r = zjd.zpjd_host(n, a2, ja2, ia2, neig=4, isearch=0, madspace=20, maxiter=2000, tol=TOL,
                      droptol=1e-3, mem=20.0)
    _check(H, r, lam[:4])
    assert abs(r.eigs[0] - r.eigs[1]) < 1e-9 * max(1.0, abs(r.eigs[0]))


def test_zjdrevcom_with_callers_preconditioner():
    """ZJDREVCOM: the caller's product and an exact shifted solve as the
    preconditioner, interior eigenvalues of the Hofstadter torus."""
    from jadamilu_amd import zjd
    n, ia, ja, a = hofstadter(24, 0.2, 7)
    H = zjd.hermitian_dense(n, a, ja, ia)
    lam = np.linalg.eigvalsh(H)
    lu = sla.lu_factor(H - 0.1 * np.eye(n))
    r = zjd.zjdrevcom_host(n, lambda x, y: np.copyto(y, H @ x),
                           lambda s, d: np.copyto(d, sla.lu_solve(lu, s)), neig=4, isearch=2,
                           sigma=0.1, madspace=20, maxiter=2000, tol=TOL)
    _check(H, r, _nearest(lam, 0.1, 4))


def test_example1_gep_pencil():
    """samples/EXAMPLE1_GEP_double_complex.f: H as in EXAMPLE1, M tridiagonal
    with diagonal 1 and upper neighbour 0.25 or 0.25i alternately; the 5
    eigenvalues nearest 0 against scipy.linalg.eigh(H, M)."""
    from jadamilu_amd import zjd
    ia, ja, a = tri_example(N)
    b = np.array([1.0 if ja[k] == i + 1 else (0.25 if (i + 1) % 2 == 0 else 0.25j)
                  for i in range(N) for k in range(ia[i] - 1, ia[i + 1] - 1)],
                 dtype=np.complex128)
    H = zjd.hermitian_dense(N, a, ja, ia)
    Mm = zjd.hermitian_dense(N, b, ja, ia)
    lam = sla.eigh(H, Mm, eigvals_only=True)
    r = zjd.zpjd_gep_host(N, a, ja, ia, b, ja, ia, neig=5, isearch=2, sigma=0.0, madspace=20,
                          maxiter=1000, tol=TOL, droptol=1e-3, mem=20.0)
    assert r.info == 0 and r.neig == 5
    # every value an eigenvalue of the pencil among the 7 nearest SIGMA (as
    # for the real harmonic drivers, the set is not guaranteed to be the 5
    # nearest where |5.47| and |-5.52| nearly tie)
    near7 = _nearest(lam, 0.0, 7)
    for e in r.eigs:
        assert np.min(np.abs(near7 - e)) <= 1e-9 * max(1.0, abs(e)), e
    assert len(set(np.round(r.eigs, 6))) == 5
    # RES is the true residual ||H z - lambda M z|| / ||M z||.  The harmonic
    # generalized branch declares convergence on its updated residual; on
    # the exactly double eigenvalues of the real form the recomputed true
    # residual of the returned copy reaches ~12 TOL here (up to ~50 TOL for
    # the other copy), so the bound is 20 TOL
    Z = r.x
    mz = np.linalg.norm(Mm @ Z, axis=0)
    res = np.linalg.norm(H @ Z - (Mm @ Z) * r.eigs, axis=0) / mz
    assert np.allclose(r.res, res, rtol=1e-3, atol=1e-13)
    assert np.all(res <= 20 * TOL), res
    # B-normalised: z^H M z = 1
    assert np.allclose(np.einsum("ij,ij->j", Z.conj(), Mm @ Z).real, 1.0, atol=1e-8)


def test_zjdrevcom_gep_smallest():
    """ZJDREVCOM_GEP on the Hofstadter torus against a Hermitian positive
    definite mass matrix (I + 0.1 times the torus' hopping part), diagonal
    preconditioner from the caller."""
    from jadamilu_amd import zjd
    n, ia, ja, a = hofstadter(16, 0.25, 9)
    H = zjd.hermitian_dense(n, a, ja, ia)
    Mm = np.eye(n) + 0.1 * (H - np.diag(np.diag(H)))
    assert np.linalg.eigvalsh(Mm).min() > 0.5
    lam = sla.eigh(H, Mm, eigvals_only=True)
    dinv = 1.0 / (np.diag(H).real - lam[0] + 1.0)
    r = zjd.zjdrevcom_gep_host(n, lambda x, y: np.copyto(y, H @ x),
                               lambda x, y: np.copyto(y, Mm @ x),
                               lambda s, d: np.copyto(d, dinv * s), neig=4, isearch=0,
                               madspace=20, maxiter=4000, tol=TOL)
    assert r.info == 0 and r.neig == 4
    assert np.allclose(np.sort(r.eigs), lam[:4], rtol=0, atol=1e-9)
    assert r.ijob_counts[2] > 0


def test_argument_errors():
    """INFO codes of the argument checks (JDaux.F:1502-1524), in the caller's
    complex units."""
    from jadamilu_amd import zjd
    ia, ja, a = tri_example(10)
    r = zjd.zpjd_host(1, a, ja, ia, neig=1, madspace=20)
    assert r.info == -1
    r = zjd.zpjd_host(10, a, ja, ia, neig=2, madspace=20, lx=10)
    assert r.info == -10
    r = zjd.zpjd_host(10, a, ja, ia, neig=2, madspace=2)
    assert r.info == -9


def test_magnetic_anderson_interior():
    """The magnetic Anderson model of the complex bench workload at L = 8
    (n = 512, W = 16.5, flux 1/8): the 6 eigenvalues nearest 0 through ZPJD
    with the default (tuned) triangular solves."""
    from jadamilu_amd import matrices as M
    from jadamilu_amd import zjd
    zjd.set_option("ildl_sweeps", -1)
    ia, ja, a = M.magnetic_anderson(8, 16.5, 0.125, 0)
    n = 512
    H = zjd.hermitian_dense(n, a, ja, ia)
    lam = np.linalg.eigvalsh(H)
    r = zjd.zpjd_host(n, a, ja, ia, neig=6, isearch=2, sigma=0.0, madspace=25, maxiter=4000,
                      tol=TOL, droptol=1e-3, mem=15.0)
    assert r.info == 0 and r.neig == 6
    near8 = _nearest(lam, 0.0, 8)
    for e in r.eigs:
        assert np.min(np.abs(near8 - e)) <= 1e-10 * max(1.0, abs(e)), e
    Z = r.x
    assert np.allclose(Z.conj().T @ Z, np.eye(6), atol=1e-8)
    true_res = np.linalg.norm(H @ Z - Z * r.eigs, axis=0)
    assert np.all(true_res <= 5 * TOL), true_res


def test_zpjdrevcom_gep_smallest_with_callers_products():
    """ZPJDREVCOM_GEP on the pencil of samples/EXAMPLE1_GEP_double_complex.f:
    the preconditioner from the matrices, the products from the caller (IJOB
    1, 3 and 4), the 5 smallest eigenvalues against scipy.linalg.eigh(H, M)."""
    from jadamilu_amd import zjd
    n = 400
    ia, ja, a = tri_example(n)
    b = np.array([1.0 if ja[k] == i + 1 else (0.25 if (i + 1) % 2 == 0 else 0.25j)
                  for i in range(n) for k in range(ia[i] - 1, ia[i + 1] - 1)],
                 dtype=np.complex128)
    H = zjd.hermitian_dense(n, a, ja, ia)
    Mm = zjd.hermitian_dense(n, b, ja, ia)
    lam = sla.eigh(H, Mm, eigvals_only=True)
    r = zjd.zpjdrevcom_gep_host(n, a, ja, ia, b, ja, ia, lambda x, y: np.copyto(y, H @ x),
                                lambda x, y: np.copyto(y, Mm @ x), neig=5, isearch=0,
                                madspace=20, maxiter=2000, tol=TOL, droptol=1e-3, mem=20.0)
    assert r.info == 0 and r.neig == 5
    assert np.allclose(np.sort(r.eigs), lam[:5], rtol=0, atol=1e-9)
    assert r.ijob_counts[1] > 0
    Z = r.x
    res = np.linalg.norm(H @ Z - (Mm @ Z) * r.eigs, axis=0) / np.linalg.norm(Mm @ Z, axis=0)
    assert np.allclose(r.res, res, rtol=1e-3, atol=1e-13)
    assert np.all(res <= 5 * TOL), res
