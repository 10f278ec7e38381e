"""GPU checks of the mixed-precision path (ICNTL(5)=1: FP32 factor, FP64
outer iteration), with the inner solver in FP64 (default) and in FP32
(option inner_fp32).  Both must return eigenpairs whose true FP64
residuals meet TOL and whose eigenvalues match the closed form / dense
spectrum to 1e-10 relative, like the FP64 path."""
import numpy as np
import pytest

from jadamilu_amd import matrices as M

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")


def _dev():
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")


def _solve(ia, ja, a, *, neig, isearch, inner_fp32, spmv64=1, madspace=20, mem=20.0):
    from jadamilu_amd.jd import Context, lx_optimal
    n = ia.size - 1
    ctx = Context(0)
    try:
        ctx.set_option("inner_fp32", inner_fp32)
        ctx.set_option("inner_spmv64", spmv64)
        lx = lx_optimal(n, neig, madspace)
        X = torch.zeros(lx, dtype=torch.float64, device="cuda:0")
        X[:n] = torch.from_numpy(np.random.default_rng(0).random(n)).to("cuda:0")
        torch.cuda.synchronize()
        r = ctx.pjd(n, a, ja, ia, X.data_ptr(), lx, neig=neig, isearch=isearch, sigma=0.0,
                    ninit=1, madspace=madspace, maxiter=6000, tol=1e-10, droptol=1e-3,
                    mem=mem, icntl=(0, 1, 0, 0, 1))
        torch.cuda.synchronize()
        x = X[: n * r.neig].reshape(r.neig, n).T.cpu().numpy().copy()
    finally:
        ctx.close()
    return r, x


def _check_residuals(A, r, x, tol=1e-10):
    # convergence is declared on the updated residual; RES is the true one,
    # recomputed at termination, and may differ from it by rounding
    # (JDaux.F:2317-2372): a factor 2 covers that, not an unverified
    # convergence claim (those ended 10-50x above TOL)
    assert (r.res <= 2 * tol).all(), r.res
    for k in range(r.neig):
        rk = np.linalg.norm(A @ x[:, k] - r.eigs[k] * x[:, k]) / np.linalg.norm(x[:, k])
        assert rk <= 2 * tol, (k, rk)
        assert abs(rk - r.res[k]) <= 1e-3 * tol + 1e-6 * rk, (k, rk, r.res[k])


@pytest.mark.parametrize("inner_fp32", [0, 1])
def test_mixed_laplace3d_residuals_meet_tol(inner_fp32):
    _dev()
    shape = (20, 21, 22)
    ia, ja, a = M.laplace(shape)
    r, x = _solve(ia, ja, a, neig=8, isearch=1, inner_fp32=inner_fp32)
    print("laplace 20x21x22 mixed inner_fp32", inner_fp32, "matvecs", r.iter, "res", r.res.max())
    assert r.info == 0 and r.neig == 8
    lam, _ = M.laplace_eigs(shape, 8)
    np.testing.assert_allclose(np.sort(r.eigs), lam, rtol=1e-10)
    _check_residuals(M.to_scipy(ia, ja, a), r, x)


@pytest.mark.parametrize("inner_fp32,spmv64", [(0, 1), (1, 1), (1, 0)])
def test_mixed_anderson_interior_residuals_meet_tol(inner_fp32, spmv64):
    _dev()
    L = 12
    ia, ja, a = M.anderson(L, 16.5, 3)
    A = M.to_dense(ia, ja, a)
    w = np.linalg.eigvalsh(A)
    r, x = _solve(ia, ja, a, neig=5, isearch=2, inner_fp32=inner_fp32, spmv64=spmv64,
                  madspace=25, mem=25.0)
    print("anderson 12 mixed inner_fp32", inner_fp32, "spmv64", spmv64, "matvecs", r.iter)
    assert r.info == 0 and r.neig == 5
    near = np.sort(w[np.argsort(np.abs(w))[:5]])
    np.testing.assert_allclose(np.sort(r.eigs), near, rtol=1e-10, atol=1e-14)
    _check_residuals(A, r, x)
