"""The allocation-free entry point (iwpp_reconstruct_noalloc, scratch in a
static device array) against the existing entry point (iwpp_reconstruct_ex,
workspace scratch) and the CPU oracle, through the C ABI: two calls in a row,
the second on a smaller image, bit-exact."""
import sys

import numpy as np
import pytest

import iwpp
import oracle

pytestmark = pytest.mark.gpu


def say(*a):
    print("[case]", *a, flush=True)
    sys.stdout.flush()


def _pair(rng, h, w, dtype, kind):
    top = 255 if dtype == np.uint8 else 4095
    mask = rng.integers(0, top + 1, (h, w)).astype(dtype)
    if kind == "seeds":
        marker = np.zeros_like(mask)
        idx = rng.choice(h * w, max(1, h * w // 400), replace=False)
        marker.flat[idx] = mask.flat[idx]
    else:
        marker = np.maximum(mask.astype(np.int64) - top // 6, 0).astype(dtype)
    return marker, mask


def _existing(J, I, conn, sweeps):
    dj, di = iwpp.DeviceImage.from_numpy(J), iwpp.DeviceImage.from_numpy(I)
    iwpp.reconstruct(dj, di, conn, 0, "mutex", "plain", iwpp.default_options(sweep_rounds=sweeps))
    return dj.numpy()


def test_noalloc_matches_existing_entry_point(torch_cuda):
    rng = np.random.default_rng(77)
    # (shape, dtype, connectivity, kind, sweep rounds): the second call of each
    # pair is smaller than the first, so it runs over what the first left in
    # the shared scratch
    calls = [((1500, 2100), np.uint8, 8, "seeds", 1), ((300, 200), np.uint16, 4, "hdome", 0),
             ((4096, 4096), np.uint8, 8, "seeds", -1), ((65, 129), np.uint8, 8, "hdome", 2)]
    free = []
    for (h, w), dtype, conn, kind, sweeps in calls:
        say("noalloc", h, w, dtype.__name__, conn, kind, sweeps)
        J, I = _pair(rng, h, w, dtype, kind)
        ref = _existing(J, I, conn, sweeps)
        dj, di = iwpp.DeviceImage.from_numpy(J), iwpp.DeviceImage.from_numpy(I)
        torch_cuda.cuda.synchronize()
        before = torch_cuda.cuda.mem_get_info()[0]
        iwpp.reconstruct_noalloc(dj, di, conn, iwpp.default_options(sweep_rounds=sweeps))
        free.append((before, torch_cuda.cuda.mem_get_info()[0]))
        got = dj.numpy()
        assert np.array_equal(got, ref), ((h, w), int((got != ref).sum()))
        assert np.array_equal(got, oracle.reconstruct(J, I, conn))
    # after the first call (which loads the library's module holding the
    # scratch) no call changes the device's free memory
    for before, after in free[1:]:
        assert after == before, free


def test_noalloc_rejects_what_it_cannot_hold(torch_cuda):
    rng = np.random.default_rng(78)
    J, I = _pair(rng, 64, 64, np.uint8, "seeds")
    dj, di = iwpp.DeviceImage.from_numpy(J), iwpp.DeviceImage.from_numpy(I)
    with pytest.raises(iwpp.IWPPError) as e:
        iwpp.reconstruct_noalloc(dj, di, 8, iwpp.default_options(queue_capacity=16))
    assert e.value.code == -1
    assert np.array_equal(dj.numpy(), J)  # nothing ran
    # too many megapixels: rejected from the size alone, before any device work
    # (the pointers are valid; the claimed size is never touched)
    big_w = 64 * 1024
    big_h = 64 * (iwpp.NOALLOC_MAX_MP // 1024 + 1)
    rc = iwpp.lib().iwpp_reconstruct_noalloc(dj.ptr, di.ptr, big_w, big_w, big_w, big_h, 1, 8,
                                             None, None)
    assert rc == -1, rc
    assert b"at most" in iwpp.lib().iwpp_last_error()
