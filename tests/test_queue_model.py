"""The CPU model of the queue phase (oracle/queue_model.c) against the
oracle, and the reference's queue self-checks on it.  CPU only.

What it pins that pixel comparisons of the GPU path cannot:
  * the reference's double-queue known-answer test (IWPP MP DQ/src/
    globalQueue.cu:218-253): fill x4, swap, fill, swap, fill x2 leaves every
    coordinate exactly twice in the read queue;
  * the unique queue's flag table (IWPP MP UQ/src/
    MorphologicalReconstruction.cu:1152-1293, 1312-1322): unique entries
    extracted, every flag clear afterwards;
  * each discipline, under many interleavings of concurrent workers (stale
    snapshots, queue overflows and restarts, capped local sweeps), reaches
    the oracle's fixed point; the mutex queue never runs one megapixel twice
    at once and ends with every state clear; the unique queue visits a
    megapixel at most once per round and ends with every flag clear;
  * planted protocol mistakes are caught (lost re-notification, a flag never
    cleared).
"""
import zlib

import numpy as np
import pytest

import oracle


@pytest.mark.parametrize("g", [1, 7, 128])
def test_dq_fill_swap_kat(g):
    assert oracle.queue_fill_swap_kat(g) == 0


@pytest.mark.parametrize("g,dup", [(1, 1), (16, 3), (32, 5)])
def test_uq_unique_kat(g, dup):
    assert oracle.queue_unique_kat(g, dup, seed=g * 31 + dup) == 0


def _case(rng, h, w, dtype, kind):
    top = 255 if dtype == np.uint8 else 4095
    mask = rng.integers(0, top + 1, (h, w)).astype(dtype)
    if kind == "seeds":
        marker = np.zeros_like(mask)
        idx = rng.choice(h * w, max(1, h * w // 300), replace=False)
        marker.flat[idx] = mask.flat[idx]
    elif kind == "hdome":
        marker = np.maximum(mask.astype(np.int64) - top // 5, 0).astype(dtype)
    else:  # the reference's random test case: marker uniform, mask = marker + residue
        marker = rng.integers(0, top + 1, (h, w)).astype(dtype)
        mask = (marker.astype(np.int64) + rng.integers(0, top + 1, (h, w)) %
                (top + 1 - marker.astype(np.int64))).astype(dtype)
    return marker, mask


@pytest.mark.parametrize("variant", ["mutex", "dq", "uq"])
@pytest.mark.parametrize("conn", [4, 8])
def test_model_reaches_the_fixed_point(variant, conn):
    rng = np.random.default_rng(zlib.crc32(f"{variant}/{conn}".encode()))
    for (h, w), dtype, kind in (((70, 90), np.uint8, "seeds"), ((33, 120), np.uint16, "hdome"),
                                ((64, 64), np.uint8, "random"), ((1, 77), np.uint8, "seeds"),
                                ((50, 1), np.uint16, "seeds")):
        J, I = _case(rng, h, w, dtype, kind)
        want = oracle.reconstruct(J, I, conn)
        for mp, workers, seed, cap, sweeps in ((8, 1, 1, 0, 0), (8, 6, 2, 0, 0), (4, 16, 3, 0, 4),
                                                (16, 5, 4, 0, 6), (8, 6, 5, 7, 4)):
            got, st = oracle.queue_model(J, I, conn, variant, mp=mp, workers=workers, seed=seed,
                                         cap=cap if variant != "mutex" else 0, max_sweeps=sweeps)
            what = (variant, conn, (h, w), kind, mp, workers, seed, cap, sweeps, st)
            assert np.array_equal(got, want), what
            assert st["state_left"] == 0, what
            if variant == "mutex":
                assert st["overlap"] == 0, what
            if variant == "uq":
                assert st["repeat_in_round"] == 0, what
            if cap and variant != "mutex" and st["seeded"] > cap:
                assert st["restarts"] > 0, what


def test_dq_duplicates_are_allowed_and_harmless():
    rng = np.random.default_rng(5)
    J, I = _case(rng, 70, 90, np.uint8, "seeds")
    want = oracle.reconstruct(J, I, 8)
    got, st = oracle.queue_model(J, I, 8, "dq", mp=8, workers=8, seed=9)
    assert np.array_equal(got, want)
    assert st["repeat_in_round"] > 0  # the double queue's duplicates happen ...
    got, st2 = oracle.queue_model(J, I, 8, "uq", mp=8, workers=8, seed=9)
    assert np.array_equal(got, want) and st2["repeat_in_round"] == 0  # ... the unique queue's not
    assert st2["visits"] < st["visits"]


def test_planted_protocol_mistakes_are_caught():
    rng = np.random.default_rng(6)
    caught = {"mutex_drop_renotify": False, "uq_keep_flag": False}
    for trial in range(20):
        J, I = _case(rng, 60, 80, np.uint8, "seeds")
        want = oracle.reconstruct(J, I, 8)
        got, st = oracle.queue_model(J, I, 8, "mutex", mp=8, workers=8, seed=trial,
                                     fault="mutex_drop_renotify")
        caught["mutex_drop_renotify"] |= not np.array_equal(got, want)
        got, st = oracle.queue_model(J, I, 8, "uq", mp=8, workers=8, seed=trial,
                                     fault="uq_keep_flag")
        caught["uq_keep_flag"] |= (not np.array_equal(got, want)) or st["state_left"] > 0
    assert all(caught.values()), caught


def test_bad_arguments_are_refused():
    J = np.zeros((8, 8), np.uint8)
    with pytest.raises(ValueError):
        oracle.queue_model(J, J, 8, "mutex", max_sweeps=3)  # cannot certify convergence
    with pytest.raises(ValueError):
        oracle.queue_model(J, J, 6, "mutex")
