"""Synthetic workload generators: determinism (committed SHA-256), ranges and
marker construction rules."""
import hashlib

import numpy as np

import iwpp
from iwpp import workloads as W


def sha(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def test_c1_bytes_are_pinned():
    c = W.c1()
    assert c.mask.shape == (1024, 1024) and c.mask.dtype == np.uint8
    assert sha(c.mask) == "7db479023179cdd2657db132ab6729e6eea4889e3e8d8ec27ec995445001b786"
    assert sha(c.marker) == "17f427ecb92f501b605ea45011f27ab466792ea5c3d86aa85964133743997e57"
    assert c.mask.min() == 0 and c.mask.max() == 255  # min-max rescale


def test_small_twins_are_pinned():
    m = iwpp.synth_mask(200, 300, iwpp.U16, 4095, 16.0, 5, 10, 40, 1315)
    assert sha(m) == "40c79f1453ac82883226516ab0d4a93b13310a400dfd2931f13866422441435c"
    assert m.max() <= 4095
    m = iwpp.synth_mask(256, 256, iwpp.U8, 255, 32.0, 10, 5, 30, 1317)
    s = iwpp.synth_marker_seeds(m, 6, 1317)
    assert sha(m) == "bb7d2ecbf45a1c14a8b8ee45ae4f6692649145d8f43629f856989cbd8485244a"
    assert sha(s) == "88267062789b6beb7fe8f74d15ac00c389c2021bab1fa17a0b09a98cb095c622"


def test_thread_count_does_not_change_bytes():
    a = iwpp.synth_mask(333, 517, iwpp.U8, 255, 8.0, 3, 5, 20, 99, nthreads=1)
    b = iwpp.synth_mask(333, 517, iwpp.U8, 255, 8.0, 3, 5, 20, 99, nthreads=7)
    assert np.array_equal(a, b)


def test_markers():
    m = iwpp.synth_mask(128, 96, iwpp.U8, 255, 4.0, 0, 0, 0, 5)
    hd = iwpp.synth_marker_hdome(m, 30)
    assert np.array_equal(hd, np.maximum(m.astype(int) - 30, 0).astype(np.uint8))
    s = iwpp.synth_marker_seeds(m, 100, 5)
    nz = s != 0
    assert np.array_equal(s[nz], m[nz])
    # exactly 100 distinct seed positions (zero-valued mask seeds are invisible)
    assert nz.sum() == 100 - ((s == 0) & (m == 0)).sum() or nz.sum() <= 100
    assert W.seeds_count(4096 * 4096, 1000) == 16777
    assert W.seeds_count(32768 * 32768, 100) == 107374


def test_batch_shapes():
    w, h, k, hh = iwpp.synth_batch_shapes(256, W.S0 + 4)
    s = np.maximum(w, h)
    assert s.min() >= 256 and s.max() <= 4096
    assert (np.minimum(w, h) * 4 >= s - 4).all()  # aspect within 1:4
    assert (k == np.arange(256) % 2).all()
    assert hh.min() >= 10 and hh.max() <= 60


def test_random_case_rule():
    j, i = iwpp.synth_random_case(64, 80, 1312)
    assert (i >= j).all()
