"""The C-ABI library loads and exports every symbol the public headers declare
(no compute calls: this runs without a GPU)."""
import ctypes
import os
import re

import pytest

from conftest import PKG_DIR, REPO


def declared(header):
    text = open(os.path.join(REPO, "include", header)).read()
    return sorted(set(re.findall(r"^\s*(?:int|void|double)\s+\**\s*(lbmib_\w+)\s*\(", text, re.M)))


def test_header_declarations_are_complete(lbmib):
    names = declared("lbmib.h")
    assert set(lbmib.ALL_SYMBOLS) == set(names)


def test_library_exports_every_declared_symbol():
    path = os.path.join(PKG_DIR, "liblbmib.so")
    if not os.path.exists(path):
        pytest.fail("liblbmib.so not built")
    lib = ctypes.CDLL(path)
    missing = [n for n in declared("lbmib.h") if not hasattr(lib, n)]
    assert not missing, missing


def test_host_layer_exports_reference_names():
    path = os.path.join(PKG_DIR, "liblbmib_host.so")
    if not os.path.exists(path):
        pytest.skip("host layer not built")
    lib = ctypes.CDLL(path)
    text = open(os.path.join(REPO, "include", "lbmib_host.h")).read()
    names = re.findall(r"^\s*(?:void|int|Fibershape\s*\*|Fluidgrid\s*\*|double)\s*\**\s*(\w+)\s*\(", text, re.M)
    assert len(names) > 20
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_surrogate_header_symbols_are_exported():
    names = declared("lbmib_surrogate.h")
    assert len(names) >= 14
    import surrogate
    assert set(surrogate.SG_SYMBOLS) == set(names)
    path = os.path.join(PKG_DIR, "liblbmib.so")
    if not os.path.exists(path):
        pytest.fail("liblbmib.so not built")
    lib = ctypes.CDLL(path)
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
