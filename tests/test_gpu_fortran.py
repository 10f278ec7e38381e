"""Fortran drop-in: a program written against the reference's DPJDREVCOM /
DPJDCLEANUP calling sequence (tests/fortran/dropin_revcom.f90), compiled with
ROCm flang and linked to libjadamilu_amd.so, reproduces the user guide's
EXAMPLE9 run (tridiag(5, i, 5), N=1000, 5 smallest eigenpairs, diagonal
preconditioner).  The binary is built by `make -C tests/fortran` (part of the
package build when flang is present)."""
import os
import subprocess

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
BIN = os.path.join(HERE, "fortran", "_build", "dropin_revcom")


def test_fortran_dropin_example9(golden):
    if not os.path.exists(BIN):
        pytest.skip("Fortran drop-in caller not built (no flang)")
    out = subprocess.run([BIN], capture_output=True, text=True, timeout=100)
    print(out.stdout, out.stderr)
    assert out.returncode == 0
    lines = out.stdout.strip().splitlines()
    head = lines[0].split()
    assert head[0] == "MATVEC" and head[2] == "INFO" and int(head[3]) == 0, lines[0]
    g = golden["EXAMPLE9"]["calls"][0]
    eigs = np.array([float(l.split()[0]) for l in lines[1:6]])
    res = np.array([float(l.split()[1]) for l in lines[1:6]])
    np.testing.assert_allclose(eigs, [t["eig"] for t in g["table"]], rtol=1e-10)
    assert (res <= 1e-10).all()
    assert abs(int(head[1]) - g["matvec"]) <= 3, (head[1], g["matvec"])
