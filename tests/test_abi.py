"""The C-ABI library loads and exports every symbol include/iss_hip.h declares
(no compute calls: these run without a GPU)."""
import ctypes
import os
import re

import iss_amd

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    src = open(os.path.join(ROOT, "include", "iss_hip.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"^\s*(?:const\s+)?[\w]+\s*\*?\s*(iss_\w+)\s*\(", src, flags=re.M)))


def test_header_declares_entry_points():
    syms = declared_symbols()
    assert "iss_ctx_create" in syms and "iss_sample" in syms
    assert len(syms) >= 14


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(iss_amd.LIB_PATH)
    missing = [s for s in declared_symbols() if not hasattr(lib, s)]
    assert not missing, missing


def test_binding_covers_every_declared_symbol():
    assert set(declared_symbols()) <= set(iss_amd._SIGNATURES)


def test_hadron_layout_matches_iss_hadron():
    assert iss_amd.HADRON_DTYPE.itemsize == 40
    assert list(iss_amd.HADRON_DTYPE.names) == ["pid", "mass", "E", "px", "py", "pz", "t", "x", "y", "z"]
