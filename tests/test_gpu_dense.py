"""GPU checks of the FP64 matrix-core contractions (jdhip_gram, jdhip_tall_gemm)
and of the symmetric SpMM (jdhip_spmm), through the C-ABI.

References are computed on the host in float64 (numpy / scipy.sparse).  The
bound for the dense products is the FP64 rounding bound of a length-n inner
product, 2 n eps |V|^T |W| entrywise; the MFMA lane layout (which element
of A, B and D each lane holds) is exercised by every shape, since a wrong
layout permutes or mixes entries.  The SpMM must equal k single SpMVs to
the last bit on integer-valued data (every partial sum exact)."""
import numpy as np
import pytest

from jadamilu_amd import matrices as M

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")
EPS = np.finfo(float).eps


def _dev():
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")
    return torch.device("cuda:0")


def _colmajor(A, dev):
    """device copy of A (n x k) stored column-major (leading dimension n)"""
    return torch.tensor(np.asfortranarray(A).ravel(order="F"), device=dev)


@pytest.mark.parametrize("n,k,m", [(1, 1, 1), (100, 3, 5), (4097, 16, 16), (70001, 25, 25),
                                   (123457, 35, 2), (50000, 64, 64), (64, 17, 33)])
def test_gram_vs_host(n, k, m):
    from jadamilu_amd.jd import Context
    dev = _dev()
    rng = np.random.default_rng(n + 7 * k + 13 * m)
    V = rng.standard_normal((n, k))
    W = rng.standard_normal((n, m))
    dV, dW = _colmajor(V, dev), _colmajor(W, dev)
    ldc = k + 3  # a leading dimension larger than k
    dC = torch.full((ldc * m,), np.nan, dtype=torch.float64, device=dev)
    ctx = Context(0)
    ctx.gram(n, k, m, dV.data_ptr(), n, dW.data_ptr(), n, dC.data_ptr(), ldc)
    torch.cuda.synchronize()
    C = dC.cpu().numpy().reshape(m, ldc).T[:k]
    ref = V.T @ W
    bound = 2 * n * EPS * (np.abs(V).T @ np.abs(W))
    assert np.all(np.abs(C - ref) <= bound + 1e-300), np.max(np.abs(C - ref) / bound)
    assert np.all(np.isnan(dC.cpu().numpy().reshape(m, ldc).T[k:]))  # padding untouched
    # deterministic: a second call gives the same bits
    dC2 = torch.zeros_like(dC)
    ctx.gram(n, k, m, dV.data_ptr(), n, dW.data_ptr(), n, dC2.data_ptr(), ldc)
    torch.cuda.synchronize()
    c1 = dC.cpu().numpy().reshape(m, ldc).T[:k]
    c2 = dC2.cpu().numpy().reshape(m, ldc).T[:k]
    assert np.array_equal(c1, c2)
    ctx.close()


def test_gram_integer_exact():
    """Integer-valued data: every product and partial sum is exact, so the
    result equals the host's bit for bit (any mis-placed lane would show)."""
    from jadamilu_amd.jd import Context
    dev = _dev()
    rng = np.random.default_rng(3)
    n, k, m = 33333, 21, 19
    V = rng.integers(-8, 9, (n, k)).astype(float)
    W = rng.integers(-8, 9, (n, m)).astype(float)
    dV, dW = _colmajor(V, dev), _colmajor(W, dev)
    dC = torch.zeros(k * m, dtype=torch.float64, device=dev)
    ctx = Context(0)
    ctx.gram(n, k, m, dV.data_ptr(), n, dW.data_ptr(), n, dC.data_ptr(), k)
    torch.cuda.synchronize()
    assert np.array_equal(dC.cpu().numpy().reshape(m, k).T, V.T @ W)
    ctx.close()


@pytest.mark.parametrize("n,k,j", [(1, 1, 1), (17, 3, 2), (4096, 16, 16), (100003, 25, 13),
                                   (65537, 64, 64), (2049, 35, 18)])
def test_tall_gemm_vs_host(n, k, j):
    from jadamilu_amd.jd import Context
    dev = _dev()
    rng = np.random.default_rng(n + k + j)
    V = rng.integers(-8, 9, (n, k)).astype(float)
    Q = rng.integers(-8, 9, (k, j)).astype(float)
    ldy = n + 5
    dV, dQ = _colmajor(V, dev), _colmajor(Q, dev)
    dY = torch.full((ldy * j,), np.nan, dtype=torch.float64, device=dev)
    ctx = Context(0)
    ctx.tall_gemm(n, k, j, dV.data_ptr(), n, dQ.data_ptr(), k, dY.data_ptr(), ldy)
    torch.cuda.synchronize()
    Yall = dY.cpu().numpy().reshape(j, ldy).T
    assert np.array_equal(Yall[:n], V @ Q)  # integer data: exact
    assert np.all(np.isnan(Yall[n:]))
    # real data: rounding bound of a length-k inner product
    Vr = rng.standard_normal((n, k))
    Qr = rng.standard_normal((k, j))
    dV, dQ = _colmajor(Vr, dev), _colmajor(Qr, dev)
    ctx.tall_gemm(n, k, j, dV.data_ptr(), n, dQ.data_ptr(), k, dY.data_ptr(), ldy)
    torch.cuda.synchronize()
    Y = dY.cpu().numpy().reshape(j, ldy).T[:n]
    assert np.all(np.abs(Y - Vr @ Qr) <= 2 * k * EPS * (np.abs(Vr) @ np.abs(Qr)))
    ctx.close()


@pytest.mark.parametrize("lanes", [0, 4])
@pytest.mark.parametrize("k", [1, 3, 8, 11])
def test_spmm_equals_spmv(lanes, k):
    from jadamilu_amd.jd import Context
    dev = _dev()
    ia, ja, a = M.anderson(24, 7.0, 2)
    a = np.round(a * 4) / 4  # quarter-integers: exact sums
    n = ia.size - 1
    rng = np.random.default_rng(k)
    X = rng.integers(-16, 17, (n, k)).astype(float)
    ctx = Context(0)
    ctx.set_matrix(ia, ja, a, lanes_per_row=(-1 if lanes == 0 else lanes))
    dX = _colmajor(X, dev)
    ldy = n + 3
    dY = torch.zeros(ldy * k, dtype=torch.float64, device=dev)
    ctx.spmm(k, dX.data_ptr(), n, dY.data_ptr(), ldy)
    y1 = torch.zeros(n, dtype=torch.float64, device=dev)
    torch.cuda.synchronize()
    Y = dY.cpu().numpy().reshape(k, ldy).T[:n]
    S = M.to_scipy(ia, ja, a)
    assert np.array_equal(Y, S @ X)
    for c in range(k):
        ctx.spmv(dX[c * n:(c + 1) * n].data_ptr(), y1.data_ptr())
        torch.cuda.synchronize()
        assert np.array_equal(Y[:, c], y1.cpu().numpy())
    ctx.close()
