"""The real symmetric form that the complex Hermitian drivers (ZPJD family,
Doc/ZPJD_prototypes.f) solve, checked on the CPU: the matrix that
jdhip_hermitian_real_form builds from a one-triangle complex CSR must be the
interleaved real representation of the Hermitian matrix (entry a + ib as the
block [[a, -b], [b, a]]), whose spectrum is the Hermitian spectrum with every
eigenvalue twice, and whose action on an interleaved vector is the complex
product."""
import numpy as np
import pytest

from jadamilu_amd import zjd


def tri_example(n, diag=lambda i: i - 1.0, off=5j):
    """samples/EXAMPLE1_double_complex.f: diagonal i - 1 (i = 1..n), upper
    neighbour 5i (other samples: another diagonal or neighbour)."""
    ia, ja, a = [1], [], []
    for i in range(1, n + 1):
        ja.append(i)
        a.append(diag(i))
        if i < n:
            ja.append(i + 1)
            a.append(off)
        ia.append(len(ja) + 1)
    return np.array(ia), np.array(ja), np.array(a, dtype=np.complex128)


def hofstadter(L, flux, seed, lower=False):
    """L x L torus, hopping -1 along x and -exp(2 pi i flux x) along y, random
    diagonal in [-1, 1]: a Hermitian matrix whose phases do not gauge away."""
    rng = np.random.default_rng(seed)
    n = L * L
    rows = [[] for _ in range(n)]
    for y in range(L):
        for x in range(L):
            i = x + L * y
            rows[i].append((i, rng.uniform(-1, 1)))
            for j, h in (((x + 1) % L + L * y, -1.0),
                         (x + L * ((y + 1) % L), -np.exp(2j * np.pi * flux * x))):
                if lower:  # store H_ji = conj(h) in row j instead
                    r, c, v = (j, i, np.conj(h)) if j > i else (i, j, h)
                else:
                    r, c, v = (i, j, h) if j > i else (j, i, np.conj(h))
                rows[r].append((c, v))
    ia, ja, a = [1], [], []
    for r in rows:
        for c, v in r:
            ja.append(c + 1)
            a.append(v)
        ia.append(len(ja) + 1)
    return n, np.array(ia), np.array(ja), np.array(a, dtype=np.complex128)


def dense_real(n2, ia2, ja2, a2):
    R = np.zeros((n2, n2))
    for i in range(n2):
        for k in range(ia2[i] - 1, ia2[i + 1] - 1):
            j = ja2[k] - 1
            R[i, j] += a2[k]
            if i != j:
                R[j, i] += a2[k]
    return R


@pytest.mark.parametrize("lower", [False, True])
def test_real_form_is_the_interleaved_representation(lower):
    n, ia, ja, a = hofstadter(6, 0.25, 1, lower=lower)
    H = zjd.hermitian_dense(n, a, ja, ia)
    assert np.allclose(H, H.conj().T)
    ia2, ja2, a2 = zjd.hermitian_real_form(n, a, ja, ia)
    R = dense_real(2 * n, ia2, ja2, a2)
    E = np.zeros((2 * n, 2 * n))
    E[0::2, 0::2], E[0::2, 1::2] = H.real, -H.imag
    E[1::2, 0::2], E[1::2, 1::2] = H.imag, H.real
    assert np.array_equal(R, E)
    # stored in the triangle of the input
    rows = np.repeat(np.arange(2 * n), np.diff(ia2))
    cols = ja2 - 1
    assert np.all(cols <= rows) if lower else np.all(cols >= rows)
    # the complex product on an interleaved vector
    rng = np.random.default_rng(2)
    z = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    assert np.allclose((R @ z.view(np.float64)).view(np.complex128), H @ z, atol=1e-13)
    # doubled spectrum
    lh = np.linalg.eigvalsh(H)
    lr = np.linalg.eigvalsh(R)
    assert np.allclose(lr, np.repeat(lh, 2), atol=1e-12)


def test_real_form_of_the_sample_tridiagonal_drops_zero_parts():
    n = 50
    ia, ja, a = tri_example(n)
    ia2, ja2, a2 = zjd.hermitian_real_form(n, a, ja, ia)
    # every diagonal entry, and one real entry per zero part of 5i
    assert len(a2) == 2 * n + 2 * (n - 1)
    R = dense_real(2 * n, ia2, ja2, a2)
    H = zjd.hermitian_dense(n, a, ja, ia)
    assert np.allclose(np.linalg.eigvalsh(R), np.repeat(np.linalg.eigvalsh(H), 2), atol=1e-11)


def test_real_form_of_a_diagonal_matrix():
    n = 7
    d = np.arange(1, n + 1).astype(np.complex128)
    ia2, ja2, a2 = zjd.hermitian_real_form(n, d, np.array([-1]), None)
    assert ja2[0] == -1
    assert np.array_equal(a2, np.repeat(np.arange(1, n + 1), 2).astype(float))


def test_magnetic_anderson_generator():
    """matrices.magnetic_anderson against a site-by-site dense construction
    (Landau gauge: -exp(2 pi i flux x) on the y bonds, -1 on x and z)."""
    from jadamilu_amd import matrices as M
    L, W, flux = 6, 3.0, 1.0 / 3.0
    ia, ja, a = M.magnetic_anderson(L, W, flux, 4)
    n = L ** 3
    H = zjd.hermitian_dense(n, a, ja, ia)
    V = M.anderson_potential(L, W, 4)
    E = np.diag(V).astype(np.complex128)
    for s in range(n):
        x, y, z = s % L, (s // L) % L, s // (L * L)
        for k, (dx, dy, dz) in enumerate(((1, 0, 0), (0, 1, 0), (0, 0, 1))):
            t = (x + dx) % L + L * ((y + dy) % L) + L * L * ((z + dz) % L)
            h = -np.exp(2j * np.pi * flux * x) if k == 1 else -1.0
            E[s, t] += h
            E[t, s] += np.conj(h)
    assert np.allclose(H, E)
    # upper triangle, diagonal first in every row
    rows = np.repeat(np.arange(n), np.diff(ia))
    assert np.all(ja - 1 >= rows)
    assert np.all(ja[ia[:-1] - 1] - 1 == np.arange(n))
