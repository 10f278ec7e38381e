"""Differentiable eigenvalues: torch.autograd.Function over the C-ABI.

`eigvals(ia, ja, values, ...)` solves with DPJD on the device and returns the
computed eigenvalues as a torch tensor; its backward calls the HIP adjoint
kernel (jdhip_adjoint_csr) that evaluates, on the stored upper-triangle
pattern,  dL/dA_e = (i==j ? 1 : 2) * sum_k g_k x_k(i) x_k(j)  with g = dL/dlambda.
`eigvals_potential(ia, ja, h, V, ...)` differentiates with respect to a
diagonal potential V (A = H + diag(V)) through jdhip_adjoint_diag:
dL/dV_i = sum_k g_k x_k(i)^2.  Both assume simple eigenvalues separated from
the rest of the spectrum (first-order perturbation theory).  The forward
solve computes `guard` extra eigenpairs and keeps the `neig` wanted ones.
"""
from __future__ import annotations

import ctypes as C

import numpy as np
import torch

from . import lib
from .jd import Context, lx_optimal


def _solve(ia, ja, a, *, neig, isearch, sigma, tol, madspace, maxiter, droptol, single,
           guard, device):
    """Solve for neig + guard eigenpairs and keep the neig wanted ones (the
    nearest to sigma for ISEARCH=2, else the smallest), ascending.  The guard
    pairs make the selection robust against the iteration skipping one of
    the wanted eigenvalues.  Returns (ctx, eigenvalues, X[n, neig] view)."""
    n = ia.size - 1
    want = neig + guard
    ctx = Context(device.index or 0)
    lx = lx_optimal(n, want, madspace)
    X = torch.zeros(lx, dtype=torch.float64, device=device)
    torch.cuda.synchronize()
    r = ctx.pjd(n, a, ja, ia, X.data_ptr(), lx, neig=want, isearch=isearch, sigma=sigma,
                madspace=madspace, maxiter=maxiter, tol=tol, droptol=droptol, mem=20.0,
                icntl=(0, 1, 0, 0, 1 if single else 0))
    torch.cuda.synchronize()
    ctx.close()
    if r.info != 0 or r.neig < neig:
        raise RuntimeError(f"eigensolver failed: info={r.info} neig={r.neig}")
    key = np.abs(r.eigs - sigma) if isearch >= 2 else r.eigs
    sel = np.sort(np.argsort(key, kind="stable")[:neig])
    sel = sel[np.argsort(r.eigs[sel], kind="stable")]
    Xk = X[: n * r.neig].view(r.neig, n)[torch.as_tensor(sel, device=device)].contiguous()
    return r.eigs[sel].copy(), Xk


def _torch_stream(device):
    return C.c_void_p(torch.cuda.current_stream(device).cuda_stream)


class _EigvalsCSR(torch.autograd.Function):
    @staticmethod
    def forward(fctx, values, ia, ja, opts):
        a = values.detach().cpu().numpy().astype(np.float64)
        eigs, X = _solve(ia, ja, a, device=values.device, **opts)
        fctx.X, fctx.k, fctx.n = X, eigs.size, ia.size - 1
        fctx.ia = torch.tensor(ia, dtype=torch.int32, device=values.device)
        fctx.ja = torch.tensor(ja, dtype=torch.int32, device=values.device)
        return torch.tensor(eigs, dtype=torch.float64, device=values.device)

    @staticmethod
    def backward(fctx, g):
        w = np.ascontiguousarray(g.detach().cpu().numpy(), dtype=np.float64)
        grad = torch.empty(fctx.ja.numel(), dtype=torch.float64, device=g.device)
        torch.cuda.synchronize()
        rc = lib().jdhip_adjoint_csr(_torch_stream(g.device), fctx.n, fctx.ia.data_ptr(),
                                     fctx.ja.data_ptr(), fctx.X.data_ptr(), fctx.n, fctx.k,
                                     w.ctypes.data_as(C.POINTER(C.c_double)), grad.data_ptr())
        torch.cuda.synchronize()
        if rc != 0:
            raise RuntimeError("jdhip_adjoint_csr failed")
        return grad, None, None, None


class _EigvalsPotential(torch.autograd.Function):
    @staticmethod
    def forward(fctx, V, ia, ja, h, opts):
        a = np.array(h, dtype=np.float64)
        dpos = ia[:-1] - 1
        assert np.all(ja[dpos] == np.arange(1, ia.size)), "diagonal must be first in each row"
        a[dpos] += V.detach().cpu().numpy()
        eigs, X = _solve(ia, ja, a, device=V.device, **opts)
        fctx.X, fctx.k, fctx.n = X, eigs.size, ia.size - 1
        return torch.tensor(eigs, dtype=torch.float64, device=V.device)

    @staticmethod
    def backward(fctx, g):
        w = np.ascontiguousarray(g.detach().cpu().numpy(), dtype=np.float64)
        grad = torch.empty(fctx.n, dtype=torch.float64, device=g.device)
        torch.cuda.synchronize()
        rc = lib().jdhip_adjoint_diag(_torch_stream(g.device), fctx.n, fctx.X.data_ptr(), fctx.n,
                                      fctx.k, w.ctypes.data_as(C.POINTER(C.c_double)),
                                      grad.data_ptr())
        torch.cuda.synchronize()
        if rc != 0:
            raise RuntimeError("jdhip_adjoint_diag failed")
        return grad, None, None, None, None


def _opts(neig, isearch, sigma, tol, madspace, maxiter, droptol, single, guard):
    return dict(neig=neig, isearch=isearch, sigma=sigma, tol=tol, madspace=madspace,
                maxiter=maxiter, droptol=droptol, single=single, guard=guard)


def eigvals(ia, ja, values, *, neig, isearch=2, sigma=0.0, tol=1e-12, madspace=25,
            maxiter=20000, droptol=1e-3, single=False, guard=2):
    """Eigenvalues (ascending) of the symmetric matrix given by the 1-based
    upper CSR (ia, ja) and the differentiable stored values."""
    ia = np.ascontiguousarray(ia, dtype=np.int32)
    ja = np.ascontiguousarray(ja, dtype=np.int32)
    return _EigvalsCSR.apply(values, ia, ja,
                             _opts(neig, isearch, sigma, tol, madspace, maxiter, droptol, single,
                                   guard))


def eigvals_potential(ia, ja, h, V, *, neig, isearch=2, sigma=0.0, tol=1e-12, madspace=25,
                      maxiter=20000, droptol=1e-3, single=False, guard=2):
    """Eigenvalues of H + diag(V), differentiable in V (device tensor, n)."""
    ia = np.ascontiguousarray(ia, dtype=np.int32)
    ja = np.ascontiguousarray(ja, dtype=np.int32)
    return _EigvalsPotential.apply(V, ia, ja, np.asarray(h, dtype=np.float64),
                                   _opts(neig, isearch, sigma, tol, madspace, maxiter, droptol,
                                         single, guard))
