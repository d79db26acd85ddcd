"""The C-ABI library loads and exports every entry point declared in include/xtpgpu.h
(no compute calls: these run on the CPU-only machine too)."""
import ctypes
import os
import re

from votca_amd import pyxtp

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    txt = open(os.path.join(ROOT, "include", "xtpgpu.h")).read()
    return sorted(set(re.findall(r"\b(xtpgpu_[a-z0-9_]+)\s*\(", txt)))


def test_header_declares_entry_points():
    syms = declared_symbols()
    assert "xtpgpu_bse_matmul" in syms and "xtpgpu_rpa_epsilon" in syms and "xtpgpu_davidson_solve" in syms


def test_library_exports_every_declared_symbol():
    L = ctypes.CDLL(pyxtp.LIB_PATH)
    missing = [s for s in declared_symbols() if not hasattr(L, s)]
    assert not missing, missing


def test_binding_covers_header():
    assert set(declared_symbols()) == set(pyxtp.SIGNATURES), \
        set(declared_symbols()) ^ set(pyxtp.SIGNATURES)


def test_binding_loads():
    L = pyxtp.lib()
    assert L.xtpgpu_last_error() is not None
