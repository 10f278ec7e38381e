"""Build recipe of the native library (gfx950).

    python -m jadamilu_amd.build        # -> jadamilu_amd/libjadamilu_amd.so

Every translation unit is compiled with hipcc for gfx950 (the .cpp host units
only use the HIP runtime API) and linked into one shared library whose
exported symbols are the C-ABI of include/jadamilu_amd.h.
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
OBJ = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libjadamilu_amd.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("JDHIP_ARCH", "gfx950")

SOURCES = ["kernels_blas.hip", "kernels_inner.hip", "kernels_mfma.hip", "kernels_spmm.hip", "noalloc.hip", "ildl_kernels.hip", "engine.cpp",
           "dense.cpp", "driver.cpp", "ildl.cpp", "ildl_factor.cpp", "ordering.cpp", "matching.cpp", "batch.cpp", "persist.cpp", "complex.cpp",
           "adjoint.hip"]

COMMON = ["-O3", "-std=c++17", "-fPIC", "-Wall", "-Wno-unused-function", "-Wno-unused-value", "-Wno-unused-result",
          "-I", os.path.join(ROOT, "include")]
# Packed FP32 instructions are kept out of the generated code.
HIPFLAGS = [f"--offload-arch={ARCH}", "-Xclang", "-target-feature", "-Xclang",
            "-packed-fp32-ops", "-munsafe-fp-atomics"]


def _compile(src: str) -> str:
    path = os.path.join(CSRC, src)
    out = os.path.join(OBJ, os.path.splitext(src)[0] + ".o")
    cmd = [HIPCC, *HIPFLAGS, *COMMON, "-c", path, "-o", out]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"compile failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    return out


def build(verbose: bool = False) -> str:
    os.makedirs(OBJ, exist_ok=True)
    srcs = [s for s in SOURCES if os.path.exists(os.path.join(CSRC, s))]
    with cf.ThreadPoolExecutor(max_workers=min(8, len(srcs))) as ex:
        objs = list(ex.map(_compile, srcs))
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", LIB + ".tmp", *objs,
           "-Wl,--no-undefined"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    os.replace(LIB + ".tmp", LIB)
    if verbose:
        print("built", LIB)
    return LIB


if __name__ == "__main__":
    build(verbose=True)
    sys.exit(0)
