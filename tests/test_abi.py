"""The C-ABI libraries load and export every entry point their headers declare
(no device calls: these tests run without a GPU)."""
import ctypes
import os
import re

import iwpp

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared(header):
    text = open(os.path.join(ROOT, "include", header)).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    names = set(re.findall(r"\b(\w+)\s*\([^;{]*\)\s*;", text))
    return {n for n in names if n not in ("sizeof",)}


def test_iwpp_exports_every_declared_symbol():
    names = declared("iwpp.h")
    assert "iwpp_reconstruct" in names and "MorphologicalReconstruction" in names
    L = ctypes.CDLL(iwpp.LIB_PATH)
    for n in names:
        assert hasattr(L, n), n
    assert names == set(iwpp.SIGNATURES), names ^ set(iwpp.SIGNATURES)


def test_synth_exports_every_declared_symbol():
    names = declared("iwpp_synth.h")
    L = ctypes.CDLL(iwpp.SYNTH_PATH)
    for n in names:
        assert hasattr(L, n), n
    assert names == set(iwpp.SYNTH_SIGNATURES)


def test_io_exports_every_declared_symbol():
    names = declared("iwpp_io.h")
    L = ctypes.CDLL(iwpp.SYNTH_PATH)
    for n in names:
        assert hasattr(L, n), n
    assert names == set(iwpp.IO_SIGNATURES)


def test_library_carries_gfx950_code():
    data = open(iwpp.LIB_PATH, "rb").read()
    assert b"gfx950" in data


def test_default_options():
    o = iwpp.default_options()
    assert o.sweep_rounds == -1 and o.col_block_rows == 0 and o.max_local_iters == 8
    assert o.timeout_s == 60.0
