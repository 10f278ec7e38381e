"""Matrix generators for the benchmark and test problems.

All return the reference input format: symmetric matrix as 1-based
upper-triangle CSR (ia[n+1], ja[nnz], a[nnz]) with the diagonal stored
first in every row (the layout PJDINIT produces, src/jd/JDaux.F:215-229).
Definitions follow samples/EXAMPLE1.f and EXAMPLE9.f (tridiagonal),
samples/laplace.f:121-193 (Laplacians) and samples/anderson.f:113-165
(periodic 3D Anderson model with a random diagonal).
"""
from __future__ import annotations

import numpy as np


def from_pairs(n, rows, cols, vals, diag):
    """Assemble upper CSR from strictly-upper (row<col) 0-based pairs."""
    rows = np.concatenate([np.arange(n), rows]).astype(np.int64)
    cols = np.concatenate([np.arange(n), cols]).astype(np.int64)
    vals = np.concatenate([np.asarray(diag, np.float64) * np.ones(n), vals])
    order = np.lexsort((cols, rows))
    rows, cols, vals = rows[order], cols[order], vals[order]
    ia = np.zeros(n + 1, np.int64)
    np.add.at(ia, rows + 1, 1)
    ia = np.cumsum(ia)
    return (ia + 1).astype(np.int32), (cols + 1).astype(np.int32), vals.astype(np.float64)


def tridiag(n, diag, off):
    """diag: array(n) or scalar, off: scalar; EXAMPLE1 uses diag=i-1, off=5."""
    r = np.arange(n - 1)
    return from_pairs(n, r, r + 1, np.full(n - 1, float(off)), diag)


def stencil(shape, diag, off, periodic=False):
    """Nearest-neighbour stencil on a tensor grid (first index fastest)."""
    shape = tuple(int(s) for s in shape)
    n = int(np.prod(shape))
    strides = np.cumprod((1,) + shape[:-1])
    idx = np.arange(n, dtype=np.int64)
    rows, cols = [], []
    for k, m in enumerate(shape):
        c = (idx // strides[k]) % m
        ok = c + 1 < m
        rows.append(idx[ok])
        cols.append(idx[ok] + strides[k])
        if periodic and m > 2:
            w = ~ok
            rows.append(idx[w] - (m - 1) * strides[k])
            cols.append(idx[w])
    rows = np.concatenate(rows)
    cols = np.concatenate(cols)
    return from_pairs(n, rows, cols, np.full(rows.size, float(off)), diag)


def laplace(shape, periodic=False):
    """-Laplacian, diag 2d, off-diagonal -1 (Dirichlet unless periodic)."""
    return stencil(shape, 2.0 * len(shape), -1.0, periodic)


def anderson_potential(L, W, seed):
    """Diagonal V_i = W (u_i - 1/2), u_i ~ U[0,1) from numpy PCG64(seed)."""
    rng = np.random.default_rng(seed)
    return W * (rng.random(L ** 3) - 0.5)


def anderson(L, W, seed, hopping=-1.0):
    """Periodic 3D Anderson model L^3 (n = L^3, 6 neighbours)."""
    return stencil((L, L, L), anderson_potential(L, W, seed), hopping, periodic=True)


def laplace_eigs(shape, k):
    """Smallest k eigenvalues of the Dirichlet -Laplacian (closed form), and
    the (k+1)-th."""
    import itertools
    lam1 = [np.sort(2 - 2 * np.cos(np.arange(1, m + 1) * np.pi / (m + 1)))[:k + 1] for m in shape]
    vals = sorted(float(sum(t)) for t in itertools.product(*lam1))
    return np.array(vals[:k]), vals[k]


def to_scipy(ia, ja, a):
    import scipy.sparse as sp
    n = ia.size - 1
    rows = np.repeat(np.arange(n), np.diff(ia))
    U = sp.csr_matrix((a, (rows, ja - 1)), shape=(n, n))
    D = sp.diags(U.diagonal())
    return (U + U.T - D).tocsr()


def to_dense(ia, ja, a):
    return to_scipy(ia, ja, a).toarray()


def magnetic_anderson(L, W, flux, seed):
    """Periodic 3D Anderson model L^3 in a magnetic field along z (Landau
    gauge): hopping -1 along x and z, -exp(2 pi i flux x) along y, diagonal
    V_i = W (u_i - 1/2) as in anderson().  A complex Hermitian matrix in the
    input format of the complex drivers: 1-based upper CSR, complex values,
    diagonal first in each row (flux * L should be an integer so that the
    phases close around the torus)."""
    n = L ** 3
    idx = np.arange(n, dtype=np.int64)
    x = idx % L
    rows, cols, vals = [], [], []
    for k, stride in enumerate((1, L, L * L)):
        c = (idx // stride) % L
        nb = np.where(c + 1 < L, idx + stride, idx - (L - 1) * stride)
        h = -np.exp(2j * np.pi * flux * x) if k == 1 else -np.ones(n, dtype=np.complex128)
        up = nb > idx  # store the (row < col) entry: H_ij, or conj(H_ji)
        rows += [idx[up], nb[~up]]
        cols += [nb[up], idx[~up]]
        vals += [h[up], np.conj(h[~up])]
    rows = np.concatenate([idx] + rows)
    cols = np.concatenate([idx] + cols)
    vals = np.concatenate([anderson_potential(L, W, seed).astype(np.complex128)] + vals)
    order = np.lexsort((cols, rows))
    rows, cols, vals = rows[order], cols[order], vals[order]
    ia = np.zeros(n + 1, np.int64)
    np.add.at(ia, rows + 1, 1)
    ia = np.cumsum(ia)
    return (ia + 1).astype(np.int32), (cols + 1).astype(np.int32), vals
