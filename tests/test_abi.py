"""The C-ABI library loads and exports every entry point include/p3hip.h declares
(CPU only: no device calls)."""

import ctypes
import os
import re

import p3hip

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared():
    text = open(os.path.join(ROOT, "include", "p3hip.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(p3h_[a-z0-9_]+)\s*\(", text)))


def test_header_declares_entry_points():
    names = declared()
    assert "p3h_coset_lde" in names and "p3h_mmcs_commit" in names and "p3h_fri_commit_phase" in names
    assert len(names) >= 30


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(p3hip.LIB_PATH)
    missing = [n for n in declared() if not hasattr(lib, n)]
    assert not missing, missing


def test_python_binding_covers_the_header():
    assert sorted(p3hip.exported_symbols()) == declared()
