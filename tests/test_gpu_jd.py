"""GPU parity of the JD engine (HIP) against the CPU oracle and the user guide.

Every test here calls the native library through its C-ABI (legacy
dpjdrevcom_/djdrevcom_ symbols with host arrays, or the jdhip_* context API
with device X).  Tolerances: eigenvalues to 1e-10 relative of the oracle /
closed form / guide (the north-star bound for the FP64 path); residual norms
<= TOL as the reference defines them (PJD.F:141-145).
"""
import numpy as np
import pytest

from jadamilu_amd import matrices as M
from oracle import oracle as O

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def _dev():
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")
    return torch.device("cuda:0")


def ex9(n=1000):
    ia, ja, a = M.tridiag(n, np.arange(1, n + 1, dtype=float), 5.0)
    return ia, ja, a


@pytest.mark.parametrize("name,icntl", [("EXAMPLE9", (0, 0, 0, 0, 0)), ("EXAMPLE10", (2, 1, 0, 0, 0))])
def test_guide_diag_legacy(golden, name, icntl):
    from jadamilu_amd import jd
    _dev()
    g = golden[name]["calls"][0]
    ia, ja, a = ex9()
    mv = O.matvec_upper(ia, ja, a)
    d = np.arange(1, 1001, dtype=float)
    # the guide's runs start from a fresh process: clear the saved state
    # (diagmin, shifts) that earlier tests left in the global context
    jd.reset_legacy()
    r = jd.pjdrevcom_host(1000, d, [-1], [-1], mv, neig=5, isearch=0, madspace=20,
                          maxiter=1000, tol=1e-10, icntl=icntl)
    ro = O.pjdrevcom_diag(1000, d, mv, neig=5, isearch=0, madspace=20, maxiter=1000,
                          tol=1e-10, icntl=icntl)
    print(name, "gpu iter", r.iter, "oracle iter", ro.iter, "guide", g["matvec"])
    assert r.info == 0
    golden_eigs = np.array([t["eig"] for t in g["table"]])
    np.testing.assert_allclose(r.eigs, golden_eigs, rtol=1e-10, atol=0)
    np.testing.assert_allclose(r.eigs, ro.eigs, rtol=1e-10, atol=0)
    assert (r.res <= 1e-10).all()
    # diagonal preconditioner: the guide's count and the oracle's, exactly
    assert r.iter == g["matvec"], (r.iter, g["matvec"])
    assert r.iter == ro.iter, (r.iter, ro.iter)
    assert "%.2E" % r.shift == g["suggested_shift"]
    A = M.to_dense(ia, ja, a)
    X = r.x
    for k in range(5):
        assert np.linalg.norm(A @ X[:, k] - r.eigs[k] * X[:, k]) <= 1.01e-10


def test_jdrevcom_harmonic_legacy():
    """ISEARCH=2 (harmonic Ritz + sQMR) through djdrevcom_, user precond."""
    from jadamilu_amd import jd
    _dev()
    ia, ja, a = ex9(300)
    A = M.to_dense(ia, ja, a)
    sigma = 150.3
    P = np.linalg.inv(A - (sigma - 0.7) * np.eye(300))
    mv = O.matvec_upper(ia, ja, a)
    ps = lambda src, dst: dst.__setitem__(slice(None), P @ src)
    r = jd.jdrevcom_host(300, mv, ps, neig=5, isearch=2, sigma=sigma, madspace=15,
                         maxiter=1000, tol=1e-10)
    ro = O.jdrevcom(300, mv, ps, neig=5, isearch=2, sigma=sigma, madspace=15, maxiter=1000,
                    tol=1e-10)
    assert r.info == 0 and r.neig == 5
    np.testing.assert_allclose(np.sort(r.eigs), np.sort(ro.eigs), rtol=1e-10)
    assert abs(r.iter - ro.iter) <= 3
    assert (r.res <= 1e-10).all()


def _far_coupled(n, seed=5):
    """Ragged upper CSR with couplings farther than 2^15 rows (exercises the
    32-bit column slices of the sliced-ELL format) and rows of 1..9 entries."""
    rng = np.random.default_rng(seed)
    rows, cols = [np.arange(n)], [np.arange(n)]
    r = np.arange(n - 1)
    rows.append(r)
    cols.append(r + 1)
    for _ in range(3):
        i = rng.integers(0, n - 40000, n // 7)
        rows.append(i)
        cols.append(i + rng.integers(33000, 40000, i.size))
    R, Cc = np.concatenate(rows), np.concatenate(cols)
    key = np.unique(R.astype(np.int64) * n + Cc)
    R, Cc = key // n, key % n
    vals = np.where(R == Cc, 4.0, -0.5) + 0.25 * rng.standard_normal(R.size)
    ia = np.zeros(n + 1, np.int32)
    np.add.at(ia, R + 1, 1)
    ia = (np.cumsum(ia) + 1).astype(np.int32)
    return ia, (Cc + 1).astype(np.int32), vals


def test_spmv_vs_matveca():
    """Device SpMV (full CSR built from the upper triangle) vs MATVECA."""
    dev = _dev()
    from jadamilu_amd.jd import Context
    ctx = Context(0)
    rng = np.random.default_rng(7)
    for ia, ja, a in [M.laplace((20, 17, 9)), M.anderson(12, 16.5, 3), M.tridiag(777, 2.0, -1.0),
                      _far_coupled(70001)]:
        n = ia.size - 1
        # 0 autotune, -1 sliced ELL, k > 0 CSR with k lanes per row
        for lpr in (0, -1, 1, 4, 16):
            ctx.set_matrix(ia, ja, a, lanes_per_row=lpr)
            x = rng.standard_normal(n)
            y_ref = np.empty(n)
            O.matvec_upper(ia, ja, a)(x, y_ref)
            xd = torch.tensor(x, device=dev)
            yd = torch.empty(n, dtype=torch.float64, device=dev)
            torch.cuda.synchronize()
            ctx.spmv(xd.data_ptr(), yd.data_ptr())
            torch.cuda.synchronize()
            np.testing.assert_allclose(yd.cpu().numpy(), y_ref, rtol=0, atol=1e-13 * np.abs(y_ref).max())
            # integer-valued inputs are exact in any summation order
            xi = rng.integers(-8, 8, n).astype(float)
            O.matvec_upper(ia, ja, np.round(a))(xi, y_ref)
            ctx.set_matrix(ia, ja, np.round(a), lanes_per_row=lpr)
            xd = torch.tensor(xi, device=dev)
            torch.cuda.synchronize()
            ctx.spmv(xd.data_ptr(), yd.data_ptr())
            torch.cuda.synchronize()
            assert np.array_equal(yd.cpu().numpy(), y_ref)
    ctx.close()


def _device_solve(ia, ja, a, diag, *, neig, madspace, isearch=0, sigma=0.0, maxiter=20000,
                  tol=1e-10, icntl=(0, 0, 0, 0, 0), x0=None):
    from jadamilu_amd.jd import Context, lx_optimal
    dev = _dev()
    n = ia.size - 1
    ctx = Context(0)
    ctx.set_matrix(ia, ja, a)
    lx = lx_optimal(n, neig, madspace)
    X = torch.zeros(lx, dtype=torch.float64, device=dev)
    ninit = 0
    if x0 is not None:
        X[:n] = torch.tensor(x0, device=dev)
        ninit = 1
    torch.cuda.synchronize()
    r = ctx.pjdrevcom(n, diag, [-1], [-1], X.data_ptr(), lx, neig=neig, isearch=isearch,
                      sigma=sigma, madspace=madspace, maxiter=maxiter, tol=tol, icntl=icntl,
                      ninit=ninit)
    torch.cuda.synchronize()
    Xh = X[: n * r.neig].cpu().numpy().reshape(r.neig, n).T
    ctx.close()
    return r, Xh


def test_laplace2d_device_vs_closed_form_and_oracle():
    shape = (40, 40)
    ia, ja, a = M.laplace(shape)
    n = ia.size - 1
    lam, nxt = M.laplace_eigs(shape, 8)
    # a seeded random start (NINIT=1) on both sides: the all-ones default is
    # symmetric on the grid and only reaches antisymmetric eigenvectors
    # through rounding noise
    x0 = np.random.default_rng(0).random(n)
    r, X = _device_solve(ia, ja, a, np.full(n, 4.0), neig=8, madspace=20, x0=x0)
    ro = O.pjdrevcom_diag(n, np.full(n, 4.0), O.matvec_upper(ia, ja, a), neig=8, madspace=20,
                          maxiter=20000, tol=1e-10, ninit=1, x0=x0)
    print("laplace40 gpu iter", r.iter, "oracle", ro.iter, r.stats)
    assert r.info == 0
    np.testing.assert_allclose(r.eigs, lam, rtol=1e-10)
    np.testing.assert_allclose(r.eigs, ro.eigs, rtol=1e-10)
    # over ~1000 MATVECs with adaptive shifts and degenerate clusters the
    # rounding paths of the two implementations separate; the count is a
    # cost figure, bounded here to 10 %
    assert abs(r.iter - ro.iter) <= max(3, 0.10 * ro.iter)
    S = M.to_scipy(ia, ja, a)
    for k in range(8):
        # the final true residual may exceed TOL slightly (JDaux.F:2317-2372)
        assert np.linalg.norm(S @ X[:, k] - r.eigs[k] * X[:, k]) <= 2e-10
    np.testing.assert_allclose(X.T @ X, np.eye(8), atol=1e-12)


def test_interior_diag_harmonic_device_vs_oracle():
    """ISEARCH=2 (harmonic extraction + sQMR) with the diagonal
    preconditioner: the 4 eigenvalues nearest sigma=500.3 of the EXAMPLE9
    matrix (tridiag, diag i, off-diagonal 5), n=1000."""
    n = 1000
    ia, ja, a = ex9(n)
    d = np.arange(1, n + 1, dtype=float)
    r, X = _device_solve(ia, ja, a, d, neig=4, madspace=20, isearch=2, sigma=500.3,
                         maxiter=5000)
    ro = O.pjdrevcom_diag(n, d, O.matvec_upper(ia, ja, a), neig=4, madspace=20, isearch=2,
                          sigma=500.3, maxiter=5000, tol=1e-10)
    w = np.linalg.eigvalsh(M.to_dense(ia, ja, a))
    near = np.sort(w[np.argsort(np.abs(w - 500.3))[:4]])
    print("interior gpu iter", r.iter, "oracle", ro.iter)
    assert r.info == 0 and ro.info == 0
    np.testing.assert_allclose(np.sort(r.eigs), near, rtol=1e-10)
    np.testing.assert_allclose(np.sort(r.eigs), np.sort(ro.eigs), rtol=1e-10)
    assert (r.res <= 1e-10).all()
    # The MATVEC count of this interior run is rounding-chaotic: the oracle
    # itself needs 649..1328 MATVECs when its all-ones start vector is
    # perturbed by 1e-13 (689 unperturbed), so only that envelope is checked.
    assert 0.5 * ro.iter <= r.iter <= 2.5 * ro.iter, (r.iter, ro.iter)


def test_input_errors_device():
    from jadamilu_amd import jd
    _dev()
    mv = lambda x, y: y.__setitem__(slice(None), x)
    for kw, code in [(dict(neig=0), -5), (dict(neig=2, madspace=2), -9),
                     (dict(neig=2, maxiter=0), -11), (dict(neig=2, tol=0.0), -12),
                     (dict(neig=2, lx=10), -10)]:
        r = jd.pjdrevcom_host(10, np.ones(10), [-1], [-1], mv, **kw)
        assert r.info == code, (kw, r.info)


def test_blocking_sync_same_result():
    """sync_blocking changes only how the host waits for the stream: the
    same kernels run in the same order, so eigenvalues, residuals and the
    MATVEC count are identical."""
    from jadamilu_amd.jd import Context, lx_optimal
    dev = _dev()
    ia, ja, a = ex9()
    n = ia.size - 1
    diag = np.arange(1, n + 1, dtype=float)
    out = []
    for blocking in (0, 1):
        ctx = Context(0)
        ctx.set_option("sync_blocking", blocking)
        ctx.set_matrix(ia, ja, a)
        lx = lx_optimal(n, 5, 20)
        X = torch.zeros(lx, dtype=torch.float64, device=dev)
        torch.cuda.synchronize()
        r = ctx.pjdrevcom(n, diag, [-1], [-1], X.data_ptr(), lx, neig=5, isearch=0,
                          sigma=0.0, madspace=20, maxiter=20000, tol=1e-10,
                          icntl=(0, 0, 0, 0, 0), ninit=0)
        torch.cuda.synchronize()
        out.append((r.info, r.iter, np.array(r.eigs), np.array(r.res),
                    X[: n * r.neig].cpu().numpy()))
        ctx.close()
    (i0, it0, e0, r0, x0), (i1, it1, e1, r1, x1) = out
    assert i0 == 0 and i1 == 0
    assert it0 == it1
    assert np.array_equal(e0, e1) and np.array_equal(r0, r1) and np.array_equal(x0, x1)
