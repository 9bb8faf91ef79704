"""The C ABI: librbg_hip.so loads and exports every symbol include/rbg_hip.h
declares, and the Python binding declares the same set.  CPU only (no
compute calls)."""
import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def header_symbols():
    syms = set()
    for name in sorted(os.listdir(os.path.join(ROOT, "include"))):
        if not name.endswith(".h"):
            continue
        text = open(os.path.join(ROOT, "include", name)).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        syms |= set(re.findall(r"\b(rbg_\w+)\s*\(", text))
    return syms


def test_header_declares_functions():
    assert len(header_symbols()) >= 15


def test_library_exports_every_declared_symbol():
    from redblackgraph_amd import _lib
    if not os.path.exists(_lib.LIB_PATH):
        pytest.fail("librbg_hip.so is not built (python -m redblackgraph_amd.build)")
    lib = ctypes.CDLL(_lib.LIB_PATH)
    missing = [s for s in sorted(header_symbols()) if not hasattr(lib, s)]
    assert not missing, missing


def test_binding_matches_header():
    from redblackgraph_amd import _lib
    assert set(_lib.SIGNATURES) == header_symbols()


def test_host_scalar_entry_points():
    import redblackgraph_amd as rb
    assert rb.avos_product(9223372036854775807, 2) == 18446744073709551614
    with pytest.raises(OverflowError):
        rb.avos_product(4294967296, 4294967296)
    with pytest.raises(OverflowError):
        rb.avos_product(9223372036854775807, 3)
    assert rb.avos_sum(-1, 1) == -1
