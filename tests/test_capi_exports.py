"""The C-ABI library loads on a host without a GPU and exports every symbol
declared in include/tk_hip.h (no compute calls)."""
import ctypes
import os
import re

import tk_hip

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    text = open(os.path.join(ROOT, "include", "tk_hip.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(tk_[a-z0-9_]+)\s*\(", text)))


def test_header_declares_entry_points():
    syms = declared_symbols()
    assert "tk_heff_apply" in syms and "tk_block_table" in syms and len(syms) >= 20


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(tk_hip.LIB_PATH)
    missing = [s for s in declared_symbols() if not hasattr(lib, s)]
    assert not missing, missing


def test_binding_covers_every_declared_symbol():
    assert set(declared_symbols()) <= set(tk_hip.EXPORTED) | {"tk_last_error"}


def test_version_string():
    assert tk_hip.lib().tk_version().decode().startswith("tk-hip")
