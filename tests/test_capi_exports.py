"""The C-ABI library loads and exports every entry point include/pra.h declares
(no compute calls: these run without a GPU)."""
import os
import re
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB = os.path.join(ROOT, "parallel-ra_amd", "libpra_amd.so")


def declared():
    names = []
    for h in os.listdir(os.path.join(ROOT, "include")):
        if not h.endswith(".h"):
            continue
        txt = open(os.path.join(ROOT, "include", h)).read()
        txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
        names += re.findall(r"^\s*(?:const\s+)?[a-z_0-9]+\s*\**\s*(pra_[a-z_0-9]+)\s*\(", txt, flags=re.M)
    return sorted(set(names))


def test_header_declares_entry_points():
    d = declared()
    assert "pra_ctx_create" in d and "pra_execute" in d and "pra_scc_add_join" in d
    assert len(d) >= 25


@pytest.mark.skipif(not os.path.exists(LIB), reason="libpra_amd.so not built")
def test_library_exports_every_declared_symbol():
    out = subprocess.run(["nm", "-D", "--defined-only", LIB], capture_output=True, text=True, check=True).stdout
    exported = set(re.findall(r"\bT (pra_[a-z_0-9]+)", out))
    missing = [n for n in declared() if n not in exported]
    assert not missing, missing


@pytest.mark.skipif(not os.path.exists(LIB), reason="libpra_amd.so not built")
def test_python_binding_covers_header():
    import pra_amd
    assert sorted(pra_amd.EXPORTS) == declared()
    L = pra_amd.lib()
    for n in declared():
        assert hasattr(L, n)
