"""The C-ABI library loads and exports every entry point include/motion_hip.h declares (no compute
call: this runs without a GPU)."""
import ctypes
import importlib
import re
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]


def test_header_symbols_exported():
    lib_mod = importlib.import_module("3fold-motion-engine_amd._lib")
    header = (ROOT / "include" / "motion_hip.h").read_text()
    declared = set(re.findall(r"^\s*(?:int|const char \*)\s*(me_\w+)\s*\(", header, re.M))
    assert declared == set(lib_mod.EXPORTED)
    L = ctypes.CDLL(str(lib_mod.LIB_PATH))
    for name in declared:
        assert hasattr(L, name), name


def test_workspace_size_and_argument_errors():
    lib_mod = importlib.import_module("3fold-motion-engine_amd._lib")
    L = lib_mod.lib()
    cfg = lib_mod.Config(2, 10000, 0, 16, 10, 0.2, 1.0, 1, 0.05, 0.0, 0, 0, 16, 256, 0)
    n = ctypes.c_size_t()
    assert L.me_workspace_size(ctypes.byref(cfg), ctypes.byref(n)) == 0
    assert 1 << 20 < n.value < 1 << 32
    bad = lib_mod.Config(7, 10, 0, 16, 10, 0.2, 1.0, 1, 0.05, 0.0, 0, 0, 0, 256, 0)
    assert L.me_workspace_size(ctypes.byref(bad), ctypes.byref(n)) == lib_mod.ME_EARG
    assert b"solver" in L.me_last_error()
    # bead geometry under a coloured solver: pair chunks with the manifold reduction, or per contact
    bead_cc = lib_mod.Config(2, 10, 18, 16, 10, 0.2, 1.0, 1, 0.05, 0.0, 0, 0, 0, 256, 0)
    assert L.me_workspace_size(ctypes.byref(bead_cc), ctypes.byref(n)) == 0
    bead_pc = lib_mod.Config(1, 10, 18, 16, 10, 0.2, 1.0, 1, 0.05, 0.0, 0, 0, 0, 256, lib_mod.FLAG_PER_CONTACT)
    assert L.me_workspace_size(ctypes.byref(bead_pc), ctypes.byref(n)) == 0
    # ColorCache always solves pair chunks
    cc_pc = lib_mod.Config(2, 10, 18, 16, 10, 0.2, 1.0, 1, 0.05, 0.0, 0, 0, 0, 256, lib_mod.FLAG_PER_CONTACT)
    assert L.me_workspace_size(ctypes.byref(cc_pc), ctypes.byref(n)) == lib_mod.ME_EARG
    assert b"per-contact" in L.me_last_error()
    # whole-step graphs need the tiled coloured solve
    g_plain = lib_mod.Config(2, 100, 0, 16, 10, 0.2, 1.0, 1, 0.05, 0.0, 0, 0, 16, 64, lib_mod.FLAG_GRAPH)
    assert L.me_workspace_size(ctypes.byref(g_plain), ctypes.byref(n)) == lib_mod.ME_EARG
    assert b"graph" in L.me_last_error()
    g_jac = lib_mod.Config(0, 100, 0, 16, 10, 0.2, 1.0, 1, 0.05, 0.0, 0, 0, 0, 64,
                           lib_mod.FLAG_GRAPH | lib_mod.FLAG_TILED)
    assert L.me_workspace_size(ctypes.byref(g_jac), ctypes.byref(n)) == lib_mod.ME_EARG
    g_ok = lib_mod.Config(2, 100, 0, 16, 10, 0.2, 1.0, 1, 0.05, 0.0, 0, 0, 16, 64,
                          lib_mod.FLAG_GRAPH | lib_mod.FLAG_TILED)
    assert L.me_workspace_size(ctypes.byref(g_ok), ctypes.byref(n)) == 0
    g_bead = lib_mod.Config(2, 100, 18, 16, 10, 0.2, 1.0, 1, 0.05, 0.0, 0, 0, 16, 64,
                            lib_mod.FLAG_GRAPH | lib_mod.FLAG_TILED)
    assert L.me_workspace_size(ctypes.byref(g_bead), ctypes.byref(n)) == 0


def test_fast_math_flag_needs_the_tiled_color_cache_solve():
    lib_mod = importlib.import_module("3fold-motion-engine_amd._lib")
    L = lib_mod.lib()
    n, m = ctypes.c_size_t(), ctypes.c_size_t()
    F, T = lib_mod.FLAG_FAST_MATH, lib_mod.FLAG_TILED
    for solver, flags in ((2, F), (1, F | T), (0, F | T)):
        c = lib_mod.Config(solver, 100, 0, 16, 10, 0.2, 1.0, 1, 0.05, 0.0, 0, 0, 16, 64, flags)
        assert L.me_workspace_size(ctypes.byref(c), ctypes.byref(n)) == lib_mod.ME_EARG
        assert b"fast arithmetic" in L.me_last_error()
    exact = lib_mod.Config(2, 1000, 0, 16, 10, 0.2, 1.0, 1, 0.05, 0.0, 0, 0, 16, 64, T)
    fast = lib_mod.Config(2, 1000, 0, 16, 10, 0.2, 1.0, 1, 0.05, 0.0, 0, 0, 16, 64, F | T)
    assert L.me_workspace_size(ctypes.byref(exact), ctypes.byref(n)) == 0
    assert L.me_workspace_size(ctypes.byref(fast), ctypes.byref(m)) == 0
    assert m.value > n.value  # the fast constants and records
    pkg = importlib.import_module("3fold-motion-engine_amd")
    with pytest.raises(ValueError, match="tiled"):
        pkg.ColorCacheContactEngine(fast_math=True)
