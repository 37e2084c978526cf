"""The C ABI library: it loads and exports every function include/hashgraph.h
declares (no device calls: this runs without a GPU)."""
import os
import re

from hashgraph_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    src = open(os.path.join(ROOT, "include", "hashgraph.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(hg_[a-z_0-9]+)\s*\(", src)))


def test_header_declarations_are_known():
    assert set(_declared()) == set(_lib.EXPORTS)


def test_library_exports_every_declared_symbol():
    assert os.path.exists(_lib.LIB_PATH), "build the library first (__graft_entry__.build())"
    assert _lib.exported_symbols() == _lib.EXPORTS


def test_version_string():
    L = _lib.load()
    assert b"gfx950" in L.hg_version()
