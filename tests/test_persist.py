"""CPU checks of the on-disk formats (jadamilu_amd/csrc/persist.cpp).

* Factor files: save -> load reproduces every exported array bit for bit
  (FP64 and FP32 factors, both coupling forms), and the oracle's apply of the
  loaded factor equals that of the original one; a truncated file, a flipped
  byte (checksum) and a foreign file are rejected.
* MatrixMarket: files written by scipy.io.mmwrite read back as the reference's
  input (1-based upper CSR with every diagonal entry) and equal the matrix;
  files written by jdhip_mm_write read back in scipy identically; a missing
  diagonal entry becomes an explicit zero; general (unsymmetric storage)
  files of a symmetric matrix give its upper triangle.
"""
import os

import numpy as np
import pytest
import scipy.io as sio
import scipy.sparse as sp

from jadamilu_amd import io as jio
from jadamilu_amd import matrices as M
from jadamilu_amd.ildl import IldlFactor
from oracle import oracle as O


def _same_export(e1, e2):
    assert e1["m"] == e2["m"] and e1["coupling"] == e2["coupling"]
    assert len(e1["levels"]) == len(e2["levels"])
    for a, b in zip(e1["levels"], e2["levels"]):
        for k in a:
            if isinstance(a[k], np.ndarray):
                assert np.array_equal(a[k], b[k]), k
            else:
                assert a[k] == b[k], k
    assert np.array_equal(e1["minv"], e2["minv"])


@pytest.mark.parametrize("single", [False, True])
@pytest.mark.parametrize("coupling", [0, 1])
def test_factor_roundtrip(tmp_path, single, coupling):
    ia, ja, a = M.anderson(9, 16.5, 3)
    n = ia.size - 1
    f = IldlFactor(n, ia, ja, a, droptol=1e-3, single=single, coupling=coupling, dense_max=100)
    path = tmp_path / "f.jdf"
    f.save(path)
    g = IldlFactor.load(path)
    e1, e2 = f.export(), g.export()
    _same_export(e1, e2)
    assert g.stats()["fill"] == f.stats()["fill"]
    r = np.random.default_rng(2).standard_normal(n)
    assert np.array_equal(O.ildl_apply(e1, r), O.ildl_apply(e2, r))


def test_factor_file_rejected_when_damaged(tmp_path):
    ia, ja, a = M.laplace((9, 8, 7))
    n = ia.size - 1
    f = IldlFactor(n, ia, ja, a, droptol=1e-3, spd=True)
    path = tmp_path / "f.jdf"
    f.save(path)
    raw = path.read_bytes()
    bad = tmp_path / "bad.jdf"
    bad.write_bytes(raw[: len(raw) // 2])  # truncated
    with pytest.raises(OSError):
        IldlFactor.load(bad)
    flipped = bytearray(raw)
    flipped[len(raw) // 3] ^= 0x40  # one bit in the payload: checksum mismatch
    bad.write_bytes(bytes(flipped))
    with pytest.raises(OSError):
        IldlFactor.load(bad)
    bad.write_bytes(b"%%MatrixMarket matrix coordinate real symmetric\n")
    with pytest.raises(OSError):
        IldlFactor.load(bad)
    with pytest.raises(OSError):
        IldlFactor.load(tmp_path / "missing.jdf")


def test_mm_read_scipy_files(tmp_path):
    ia, ja, a = M.anderson(7, 12.0, 5)
    S = M.to_scipy(ia, ja, a)
    p = tmp_path / "a.mtx"
    sio.mmwrite(str(p), S, symmetry="symmetric")
    ia2, ja2, a2 = jio.read_mm(p)
    assert np.array_equal(M.to_scipy(ia2, ja2, a2).toarray(), S.toarray())
    # upper triangle with every diagonal entry, columns ascending per row
    n = ia2.size - 1
    for i in range(n):
        cols = ja2[ia2[i] - 1: ia2[i + 1] - 1]
        assert cols[0] == i + 1 and np.all(np.diff(cols) > 0)
    # a general file holding the same symmetric matrix
    p2 = tmp_path / "g.mtx"
    sio.mmwrite(str(p2), S, symmetry="general")
    ia3, ja3, a3 = jio.read_mm(p2)
    assert np.array_equal(ia3, ia2) and np.array_equal(ja3, ja2) and np.array_equal(a3, a2)


def test_mm_write_read_back(tmp_path):
    rng = np.random.default_rng(4)
    ia, ja, a = M.laplace((6, 5))
    a = a * (1.0 + 1e-7 * rng.standard_normal(a.size))  # full-precision values
    p = tmp_path / "w.mtx"
    jio.write_mm(p, ia, ja, a)
    S = sio.mmread(str(p)).tocsr()
    assert np.array_equal(S.toarray(), M.to_scipy(ia, ja, a).toarray())
    ia2, ja2, a2 = jio.read_mm(p)
    assert np.array_equal(ia2, ia) and np.array_equal(ja2, ja) and np.array_equal(a2, a)


def test_mm_missing_diagonal_and_errors(tmp_path):
    p = tmp_path / "d.mtx"
    p.write_text("%%MatrixMarket matrix coordinate real symmetric\n3 3 2\n1 1 2.0\n3 2 -1.5\n")
    ia, ja, a = jio.read_mm(p)
    assert list(ia) == [1, 2, 4, 5] and list(ja) == [1, 2, 3, 3]
    assert list(a) == [2.0, 0.0, -1.5, 0.0]
    p.write_text("%%MatrixMarket matrix coordinate complex hermitian\n2 2 1\n1 1 1 0\n")
    with pytest.raises(ValueError):
        jio.read_mm(p)
    p.write_text("%%MatrixMarket matrix coordinate real symmetric\n2 2 3\n1 1 1\n")
    with pytest.raises(ValueError):
        jio.read_mm(p)
