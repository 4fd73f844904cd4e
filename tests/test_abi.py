"""The C ABI: libmhd3d.so loads on a host without a GPU and exports every
function include/mhd3d.h declares; the ctypes table covers the same set.
No compute call is made here."""
import ctypes
import os
import re

import pytest

from mhd3d import _lib as L

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "mhd3d.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(mhd3d_[a-z0-9_]+)\s*\(", text)))


def test_header_declares_entry_points():
    names = declared_functions()
    assert "mhd3d_step" in names and "mhd3d_ensemble_step" in names
    assert len(names) >= 20


def test_library_exports_every_declared_symbol():
    if not os.path.exists(L.LIB_PATH):
        pytest.skip("libmhd3d.so not built")
    lib = ctypes.CDLL(L.LIB_PATH)
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, missing


def test_ctypes_table_matches_header():
    assert sorted(L.SIGNATURES) == declared_functions()
