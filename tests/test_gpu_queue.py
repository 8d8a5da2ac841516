"""The queue variants' disciplines, checked on the device while they run
(iwpp_options.check_queue), and the per-phase clock split (trace_phases),
through the C ABI.  The pixels are checked against the oracle as everywhere
else; these tests pin what the pixels cannot show.

Reference self-checks restated as counters: the double queue's fill/swap
known-answer test (IWPP MP DQ/src/globalQueue.cu:218-253) and the unique
queue's "flags all zero after clear" (IWPP MP UQ/src/
MorphologicalReconstruction.cu:1312-1322):
  * mutex: no megapixel is visited by two waves at once, and every state
    word (PENDING / RUNNING bits) is clear when the phase ends;
  * UQ: no megapixel is visited twice in one round, and every queued flag is
    clear when the phase ends;
  * DQ: duplicates within a round are allowed (reported).
"""
import sys

import numpy as np
import pytest

import iwpp
import oracle
from iwpp import workloads as W

pytestmark = pytest.mark.gpu


def say(*a):
    print("[case]", *a, flush=True)
    sys.stdout.flush()


def up(a):
    return iwpp.DeviceImage.from_numpy(a)


def _cases():
    rng = np.random.default_rng(31)
    out = []
    for h, w in ((300, 517), (1000, 777), (64, 2048)):
        mask = rng.integers(0, 256, (h, w)).astype(np.uint8)
        marker = np.zeros_like(mask)
        idx = rng.choice(h * w, max(1, h * w // 500), replace=False)
        marker.flat[idx] = mask.flat[idx]
        out.append((f"seeds {h}x{w}", marker, mask, 8))
    mask16 = rng.integers(0, 4096, (700, 500)).astype(np.uint16)
    out.append(("hdome u16 700x500", np.maximum(mask16.astype(np.int64) - 600, 0).astype(np.uint16),
                mask16, 4))
    c = W.c2()
    out.append((c.name, c.marker, c.mask, c.connectivity))
    return out


@pytest.mark.parametrize("queue", ["mutex", "dq", "uq"])
def test_queue_discipline(torch_cuda, queue):
    for name, J, I, conn in _cases():
        want = oracle.reconstruct(J, I, conn)
        for sweeps in (0, 1):
            say("discipline", queue, name, sweeps)
            dj = up(J)
            st = iwpp.reconstruct(dj, up(I), conn, 0, queue, "plain",
                                  iwpp.default_options(sweep_rounds=sweeps, check_queue=1))
            assert np.array_equal(dj.numpy(), want), (name, st)
            say("  visits", st["tile_visits"], "rounds", st["queue_rounds"],
                "repeat", st["repeat_visits"], "state left", st["state_left"])
            if queue in ("mutex", "uq"):
                assert st["repeat_visits"] == 0, st
                assert st["state_left"] == 0, st
            else:
                assert st["repeat_visits"] >= 0


@pytest.mark.parametrize("queue", ["mutex", "uq"])
def test_queue_discipline_batch(torch_cuda, queue):
    cases = W.c4(n=24, max_side=1024)
    pairs = [(up(c.marker), up(c.mask)) for c in cases]
    say("discipline batch", queue)
    st = iwpp.reconstruct_batch(pairs, 8, 0, queue, "plain",
                                iwpp.default_options(sweep_rounds=0, check_queue=1))
    for (dj, _), c in zip(pairs, cases):
        assert np.array_equal(dj.numpy(), oracle.reconstruct(c.marker, c.mask, 8)), c.name
    assert st["repeat_visits"] == 0 and st["state_left"] == 0, st


@pytest.mark.parametrize("queue", ["mutex", "dq", "uq"])
def test_phase_clock_split(torch_cuda, queue):
    # the reference's clock64 split of a visit (IWPP MP/src/
    # MorphologicalReconstruction.cu:199-231,534-540): wave-microseconds per
    # phase, summed over the waves; their total is the waves' time in the
    # kernel, so it lies between one wave's and all waves' share of it
    c = W.c2()
    dj = up(c.marker)
    st = iwpp.reconstruct(dj, up(c.mask), 8, 0, queue, "plain",
                          iwpp.default_options(sweep_rounds=0, trace_phases=1))
    assert np.array_equal(dj.numpy(), oracle.reconstruct(c.marker, c.mask, 8))
    parts = [st[k] for k in ("us_wait", "us_fetch", "us_converge", "us_store", "us_notify")]
    say("phases", queue, [round(p) for p in parts], "queue ms", st["ms_queue"])
    assert all(p >= 0 for p in parts) and st["us_converge"] > 0 and st["us_fetch"] > 0
    total, kernel_us = sum(parts), st["ms_queue"] * 1e3
    assert kernel_us * 0.5 <= total <= kernel_us * 256 * 64, (total, kernel_us)
    # without tracing the counters stay zero
    dj = up(c.marker)
    st = iwpp.reconstruct(dj, up(c.mask), 8, 0, queue, "plain",
                          iwpp.default_options(sweep_rounds=0))
    assert st["us_converge"] == 0.0 and st["us_wait"] == 0.0


@pytest.mark.parametrize("queue", ["dq", "uq"])
def test_round_kernels_not_co_resident(torch_cuda, queue, monkeypatch):
    # The round-based kernels need every block running at once.  Forced here
    # to launch 4x more blocks than fit, without the cooperative launch: the
    # start gate sends the blocks back within milliseconds and the host
    # relaunches smaller grids until they fit, instead of timing out in a
    # grid barrier.  Every wait is bounded; the output is still exact.
    import time
    monkeypatch.setenv("IWPP_TEST_OVERSUBSCRIBE", "4")
    rng = np.random.default_rng(41)
    mask = rng.integers(0, 256, (500, 700)).astype(np.uint8)
    marker = np.zeros_like(mask)
    marker.flat[rng.choice(mask.size, 300, replace=False)] = 255
    marker = np.minimum(marker, mask)
    dj = up(marker)
    say("oversubscribed", queue)
    t0 = time.perf_counter()
    st = iwpp.reconstruct(dj, up(mask), 8, 0, queue, "plain",
                          iwpp.default_options(sweep_rounds=0, timeout_s=20.0))
    dt = time.perf_counter() - t0
    say("  seconds", round(dt, 3), st)
    assert np.array_equal(dj.numpy(), oracle.reconstruct(marker, mask, 8))
    assert dt < 5.0
