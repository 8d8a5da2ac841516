"""GPU parity over random parameter combinations: image shape, pixel type,
pitch, connectivity, queue variant, sweep rounds, column block height and
fix-up rows drawn from seeded generators, single images and mixed-shape
batches, each output compared bit for bit with the CPU oracle.  The fixed
grids of test_gpu_parity.py cover each parameter; this covers how they
combine."""
import sys

import numpy as np
import pytest

import iwpp
import oracle

pytestmark = pytest.mark.gpu


def say(*a):
    print("[case]", *a, flush=True)
    sys.stdout.flush()


def rand_pair(rng, h, w, dtype, kind):
    top = 255 if dtype == np.uint8 else 4095
    mask = rng.integers(0, top + 1, (h, w)).astype(dtype)
    if kind == "seeds":
        marker = np.zeros_like(mask)
        idx = rng.choice(h * w, max(1, h * w // 400), replace=False)
        marker.flat[idx] = mask.flat[idx]
    elif kind == "hdome":
        marker = np.maximum(mask.astype(np.int64) - top // 6, 0).astype(dtype)
    else:
        marker = np.minimum(rng.integers(0, top + 1, (h, w)).astype(dtype), mask)
    return marker, mask


def draw_options(rng):
    return dict(sweep_rounds=int(rng.integers(0, 4)),
                col_block_rows=int(rng.choice([0, 32, 64, 96, 128])),
                seam_rows=int(rng.integers(-1, 48)))


@pytest.mark.parametrize("block", range(4))
def test_random_single(torch_cuda, block):
    for t in range(25):
        rng = np.random.default_rng(7000 + 25 * block + t)
        dtype = [np.uint8, np.uint16][rng.integers(2)]
        h, w = int(rng.integers(1, 600)), int(rng.integers(1, 2400))
        kind = ["random", "seeds", "hdome"][rng.integers(3)]
        conn = [4, 8][rng.integers(2)]
        queue = ["mutex", "dq", "uq"][rng.integers(3)]
        opt = draw_options(rng)
        pitch = None
        if rng.random() < 0.3:  # 4-byte aligned rows that are not 16-byte aligned
            pitch = (w * np.dtype(dtype).itemsize + 3) // 4 * 4
            pitch = pitch + 4 if pitch % 16 == 0 else pitch
        say("random single", block, t, dtype.__name__, h, w, kind, conn, queue, opt, pitch)
        J, I = rand_pair(rng, h, w, dtype, kind)
        dj, di = iwpp.DeviceImage.from_numpy(J, pitch), iwpp.DeviceImage.from_numpy(I, pitch)
        st = iwpp.reconstruct(dj, di, conn, 0, queue, "plain", iwpp.default_options(**opt))
        assert np.array_equal(dj.numpy(), oracle.reconstruct(J, I, conn)), (t, st)


@pytest.mark.parametrize("block", range(2))
def test_random_batch(torch_cuda, block):
    for t in range(20):
        rng = np.random.default_rng(8000 + 20 * block + t)
        dtype = [np.uint8, np.uint16][rng.integers(2)]
        conn = [4, 8][rng.integers(2)]
        queue = ["mutex", "dq", "uq"][rng.integers(3)]
        opt = draw_options(rng)
        cases = []
        for _ in range(int(rng.integers(1, 7))):
            h, w = int(rng.integers(1, 500)), int(rng.integers(1, 1800))
            cases.append(rand_pair(rng, h, w, dtype, ["random", "seeds", "hdome"][rng.integers(3)]))
        say("random batch", block, t, dtype.__name__, [J.shape for J, _ in cases], conn, queue, opt)
        pairs = [(iwpp.DeviceImage.from_numpy(J), iwpp.DeviceImage.from_numpy(I)) for J, I in cases]
        batch = iwpp.Batch(pairs) if t % 2 else pairs  # prebuilt descriptor array, or the list
        st = iwpp.reconstruct_batch(batch, conn, 0, queue, "plain", iwpp.default_options(**opt))
        for (J, I), (dj, _) in zip(cases, pairs):
            assert np.array_equal(dj.numpy(), oracle.reconstruct(J, I, conn)), (t, J.shape, st)
