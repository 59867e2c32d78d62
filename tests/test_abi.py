"""CPU checks of the C ABI: the library loads and exports every symbol that
include/fmx.h declares, and argument validation needs no GPU."""
import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    out = set()
    inc = os.path.join(ROOT, "include")
    for f in os.listdir(inc):
        if f.endswith(".h"):
            text = open(os.path.join(inc, f)).read()
            text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
            out |= set(re.findall(r"\b(fmx_[a-z0-9_]+)\s*\(", text))
    return out


def test_library_exports_every_declared_symbol():
    import flashx_amd.fmx as F
    L = F.lib()
    syms = declared_symbols()
    assert len(syms) >= 14
    missing = [s for s in sorted(syms) if not hasattr(L, s)]
    assert not missing, missing
    # and the Python binding covers the whole ABI
    assert syms <= set(F.SIGNATURES), sorted(syms - set(F.SIGNATURES))


def test_strerror_and_null_checks():
    import flashx_amd.fmx as F
    L = F.lib()
    assert L.fmx_strerror(0) == b"success"
    assert L.fmx_strerror(4) == b"malformed tile image"
    assert L.fmx_ctx_create(0, None) == F.EINVAL
    assert L.fmx_ctx_destroy(None) == F.EINVAL
    assert L.fmx_spm_destroy(None) == F.EINVAL
    assert L.fmx_spmm(None, None, 2, 16, None, 16, None, 16, None) == F.EINVAL
    info = F.fmx_spm_info()
    assert L.fmx_spm_info_get(None, ctypes.byref(info)) == F.EINVAL


def test_product_never_imports_oracle():
    # the product package and library must not route through the oracle
    pkg = os.path.join(ROOT, "flashx_amd")
    for dirpath, _, files in os.walk(pkg):
        for f in files:
            if f.endswith((".py", ".cpp", ".hip", ".h")):
                text = open(os.path.join(dirpath, f), errors="replace").read()
                assert "oracle" not in text.lower() or f == "__init__.py" and "oracle" not in text, f
    so = os.path.join(pkg, "lib", "libfmx.so")
    if os.path.exists(so):
        assert b"orc_" not in open(so, "rb").read()
