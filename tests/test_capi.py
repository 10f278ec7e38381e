"""CPU checks of the C-ABI: the native library loads without a GPU and
exports every function declared in include/*.h."""
import os
import re
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    names = set()
    inc = os.path.join(ROOT, "include")
    for f in os.listdir(inc):
        if not f.endswith(".h"):
            continue
        text = open(os.path.join(inc, f)).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        for m in re.finditer(r"^\s*[A-Za-z_][\w\s\*]*?\b(\w+)\s*\(", text, flags=re.M):
            name = m.group(1)
            if name not in ("if", "while", "for", "sizeof", "defined"):
                names.add(name)
    return names


def test_library_exports_all_declared():
    from jadamilu_amd import LIB_PATH, lib
    if not os.path.exists(LIB_PATH):
        pytest.skip("native library not built")
    lib()  # loads (no device call)
    out = subprocess.run(["nm", "-D", "--defined-only", LIB_PATH], capture_output=True,
                         text=True, check=True).stdout
    exported = {l.split()[-1] for l in out.splitlines() if " T " in l}
    missing = declared_symbols() - exported
    assert not missing, missing
    for s in ("dpjd_", "dpjdrevcom_", "djdrevcom_", "dpjdcleanup_", "jdhip_create"):
        assert s in exported


def test_python_signatures_cover_exports():
    from jadamilu_amd import SIGNATURES
    assert declared_symbols() <= set(SIGNATURES) | {"jdhip_ctx"}
