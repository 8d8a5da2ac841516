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


class _FakeImage:  # what the batch array reads from a DeviceImage (no device memory)
    def __init__(self, h, w, dtype, pitch, ptr):
        self.h, self.w, self.dtype, self.pitch, self.ptr = h, w, dtype, pitch, ptr


def test_batch_array_prebuilt_matches_list():
    pairs = [(_FakeImage(10 + k, 30 + 3 * k, 2, 256 * (k + 1), 0x10000 * (k + 1)),
              _FakeImage(10 + k, 30 + 3 * k, 2, 512, 0x900000 + 0x1000 * k)) for k in range(7)]
    arr, dtype = iwpp._batch_array(pairs)
    b = iwpp.Batch(pairs)
    arr_b, dtype_b = iwpp._batch_array(b)
    assert dtype == dtype_b == 2 and len(b) == 7 and b[3] is pairs[3] and list(b) == pairs
    assert bytes(arr) == bytes(arr_b)
    assert [(d.marker, d.mask, d.pitch_marker, d.pitch_mask, d.width, d.height) for d in arr_b] == \
        [(m.ptr, s.ptr, m.pitch, s.pitch, m.w, m.h) for m, s in pairs]
