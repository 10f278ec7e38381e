"""jdhip_gram_noalloc (the entry point that allocates nothing) against
jdhip_gram, through the C-ABI: bit-identical results on two calls in a row,
the second with a smaller input (the static scratch keeps the first call's
partial sums, which the second must not read)."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")


def _dev():
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")
    return torch.device("cuda:0")


def _colmajor(A, dev):
    return torch.tensor(np.asfortranarray(A).ravel(order="F"), device=dev)


def test_gram_noalloc_matches_gram():
    from jadamilu_amd import lib
    from jadamilu_amd.jd import Context
    dev = _dev()
    L = lib()
    ctx = Context(0)
    stream = C.c_void_p(ctx.stream_ptr())
    rng = np.random.default_rng(12)
    for n, k, m in [(200003, 35, 30), (5001, 7, 3)]:
        V = rng.standard_normal((n, k))
        W = rng.standard_normal((n, m))
        dV, dW = _colmajor(V, dev), _colmajor(W, dev)
        c_ref = torch.full((k * m,), np.nan, dtype=torch.float64, device=dev)
        c_na = torch.full((k * m,), np.nan, dtype=torch.float64, device=dev)
        torch.cuda.synchronize()
        ctx.gram(n, k, m, dV.data_ptr(), n, dW.data_ptr(), n, c_ref.data_ptr(), k)
        rc = L.jdhip_gram_noalloc(stream, n, k, m, C.c_void_p(dV.data_ptr()), n,
                                  C.c_void_p(dW.data_ptr()), n, C.c_void_p(c_na.data_ptr()), k)
        assert rc == 0
        torch.cuda.synchronize()
        a, b = c_ref.cpu().numpy(), c_na.cpu().numpy()
        assert np.array_equal(a, b), (n, k, m, np.max(np.abs(a - b)))
        ref = (V.T @ W).ravel(order="F")
        assert np.max(np.abs(b - ref)) <= 2 * n * np.finfo(float).eps * np.max(
            (np.abs(V).T @ np.abs(W)))
    # sizes beyond the scratch's limit are rejected without a launch
    x = torch.zeros(65 * 10, dtype=torch.float64, device=dev)
    rc = L.jdhip_gram_noalloc(stream, 10, 65, 1, C.c_void_p(x.data_ptr()), 10,
                              C.c_void_p(x.data_ptr()), 10, C.c_void_p(x.data_ptr()), 65)
    assert rc == -1
    ctx.close()
