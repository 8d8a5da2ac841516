"""GPU parity: every kernel and the full reconstruction, through the C ABI,
against the CPU oracle on the same inputs.  Integer work: bit-exact."""
import sys
import zlib

import numpy as np
import pytest

import iwpp
import oracle
from iwpp import workloads as W

pytestmark = pytest.mark.gpu


def say(*a):
    print("[case]", *a, flush=True)
    sys.stdout.flush()


def up(a, pitch=None):
    return iwpp.DeviceImage.from_numpy(a, pitch=pitch)


def rand_pair(rng, h, w, dtype, kind="random"):
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


def odd_pitch(w, dtype):
    # a multiple of 4 that is not a multiple of 16: selects the 4-byte kernels
    p = (w * np.dtype(dtype).itemsize + 3) // 4 * 4
    return p + 4 if p % 16 == 0 else p


ROW_SHAPES = [(3, 1), (2, 17), (5, 1023), (4, 4099), (3, 8200), (2, 16400), (2, 33000)]


@pytest.mark.parametrize("dtype", [np.uint8, np.uint16])
@pytest.mark.parametrize("pitch_kind", ["aligned", "odd"])
def test_row_pass(torch_cuda, dtype, pitch_kind):
    rng = np.random.default_rng(1)
    for h, w in ROW_SHAPES:
        say("row_pass", dtype.__name__, pitch_kind, h, w)
        J, I = rand_pair(rng, h, w, dtype, "seeds")
        pitch = odd_pitch(w, dtype) if pitch_kind == "odd" else None
        dj, di = up(J, pitch), up(I, pitch)
        raw_before = dj.raw_numpy()
        ch = torch_cuda.zeros(1, dtype=torch_cuda.int32, device="cuda")
        iwpp.sweep_rows(dj, di, ch)
        torch_cuda.cuda.synchronize()
        want = oracle.pass_x(J, I)
        assert np.array_equal(dj.numpy(), want), (h, w)
        # padding bytes beyond the row are never written
        rb = w * np.dtype(dtype).itemsize
        assert np.array_equal(dj.raw_numpy()[:, rb:], raw_before[:, rb:])
        assert int(ch.item()) == int(not np.array_equal(want, J))


COL_SHAPES = [(1, 5), (17, 3), (130, 1100), (300, 513), (64, 2048), (70, 16500), (33, 8200)]


@pytest.mark.parametrize("dtype", [np.uint8, np.uint16])
@pytest.mark.parametrize("conn", [4, 8])
@pytest.mark.parametrize("pitch_kind", ["aligned", "odd"])
def test_col_pass(torch_cuda, dtype, conn, pitch_kind):
    rng = np.random.default_rng(2)
    for h, w in COL_SHAPES:
        for rows in (1, 7, 64, 1000):
            for d in (1, -1):
                say("col_pass", dtype.__name__, conn, pitch_kind, h, w, rows, d)
                J, I = rand_pair(rng, h, w, dtype, "seeds")
                pitch = odd_pitch(w, dtype) if pitch_kind == "odd" else None
                src, msk = up(J, pitch), up(I, pitch)
                dst = iwpp.DeviceImage(h, w, src.dtype, pitch)
                bw = iwpp.sweep_cols_block_width(src, dst, msk)
                iwpp.sweep_cols(src, dst, msk, conn, d, rows)
                torch_cuda.cuda.synchronize()
                want = oracle.pass_y(J, I, conn, d, bw, rows)
                assert np.array_equal(dst.numpy(), want), (h, w, rows, d, bw)


@pytest.mark.parametrize("dtype", [np.uint8, np.uint16])
@pytest.mark.parametrize("conn", [4, 8])
@pytest.mark.parametrize("pitch_kind", ["aligned", "odd"])
def test_raster_pass(torch_cuda, dtype, conn, pitch_kind):
    # one out-of-place raster pass (each row's column step, then its scans
    # across the strip) against the oracle's pass, blocks of 1..1000 rows
    rng = np.random.default_rng(zlib.crc32(f"raster/{dtype.__name__}/{conn}/{pitch_kind}".encode()))
    for h, w in COL_SHAPES:
        for rows in (1, 7, 64, 1000):
            for d in (1, -1):
                for kind in ("seeds", "hdome"):
                    say("raster_pass", dtype.__name__, conn, pitch_kind, h, w, rows, d, kind)
                    J, I = rand_pair(rng, h, w, dtype, kind)
                    pitch = odd_pitch(w, dtype) if pitch_kind == "odd" else None
                    src, msk = up(J, pitch), up(I, pitch)
                    dst = iwpp.DeviceImage(h, w, src.dtype, pitch)
                    bw = iwpp.sweep_cols_block_width(src, dst, msk)
                    iwpp.sweep_raster(src, dst, msk, conn, d, rows)
                    torch_cuda.cuda.synchronize()
                    want = oracle.pass_raster(J, I, conn, d, bw, rows)
                    assert np.array_equal(dst.numpy(), want), (h, w, rows, d, bw, kind)


@pytest.mark.parametrize("dtype", [np.uint8, np.uint16])
@pytest.mark.parametrize("conn", [4, 8])
def test_unstable_megapixels(torch_cuda, dtype, conn):
    rng = np.random.default_rng(3)
    for h, w in [(1, 1), (64, 64), (65, 129), (200, 333), (512, 300), (130, 16500), (70, 8200)]:
        say("unstable", dtype.__name__, conn, h, w)
        J, I = rand_pair(rng, h, w, dtype, "seeds")
        for stage in range(3):
            if stage == 1:
                J = oracle.reconstruct(J, I, conn)
            if stage == 2:  # a few planted dips
                for _ in range(3):
                    J[rng.integers(0, h), rng.integers(0, w)] = 0
            flags, count = iwpp.unstable_megapixels(up(J), up(I), conn)
            want = oracle.unstable_blocks(J, I, conn, 64)
            assert np.array_equal(flags, want), (h, w, stage)
            assert count == int(want.sum())


VARIANTS = ["mutex", "dq", "uq"]


@pytest.mark.parametrize("queue", VARIANTS)
@pytest.mark.parametrize("conn", [4, 8])
@pytest.mark.parametrize("dtype", [np.uint8, np.uint16])
def test_reconstruct_random(torch_cuda, queue, conn, dtype):
    rng = np.random.default_rng(4)
    shapes = [(1, 1), (1, 300), (300, 1), (64, 64), (65, 129), (200, 333), (640, 257)]
    for h, w in shapes:
        for kind in ("random", "seeds", "hdome"):
            for sweeps in (0, 1, 2):
                say("reconstruct", queue, conn, dtype.__name__, h, w, kind, sweeps)
                J, I = rand_pair(rng, h, w, dtype, kind)
                dj, di = up(J), up(I)
                st = iwpp.reconstruct(dj, di, conn, 0, queue, "plain",
                                      iwpp.default_options(sweep_rounds=sweeps))
                want = oracle.reconstruct(J, I, conn)
                got = dj.numpy()
                assert np.array_equal(got, want), (h, w, kind, sweeps, st)


@pytest.mark.parametrize("queue", VARIANTS)
@pytest.mark.parametrize("conn", [4, 8])
@pytest.mark.parametrize("dtype", [np.uint8, np.uint16])
def test_reconstruct_wide_rows(torch_cuda, queue, conn, dtype):
    # rows of at least 16 KiB take the column pass with 16-byte lanes, whose
    # last pass also sets the seeding flags and the strips' border columns:
    # a ragged last strip, a full-width row and heights of several blocks
    rng = np.random.default_rng(zlib.crc32(f"wide/{queue}/{conn}/{dtype.__name__}".encode()))
    size = np.dtype(dtype).itemsize
    for h, w in ((150, 16397 // size), (97, 16384 // size)):
        for kind in ("seeds", "hdome"):
            for sweeps in (1, 2):
                say("wide rows", queue, conn, dtype.__name__, h, w, kind, sweeps)
                J, I = rand_pair(rng, h, w, dtype, kind)
                dj, di = up(J), up(I)
                st = iwpp.reconstruct(dj, di, conn, 0, queue, "plain",
                                      iwpp.default_options(sweep_rounds=sweeps))
                got = dj.numpy()
                assert np.array_equal(got, oracle.reconstruct(J, I, conn)), (h, w, kind, sweeps, st)


@pytest.mark.parametrize("queue", VARIANTS)
def test_reconstruct_unclipped_marker_and_odd_pitch(torch_cuda, queue):
    # The first pass clips the marker; a raster pass also clips the pixels it
    # reads from outside its block and strip (other waves' pixels, not yet
    # clipped or already raised): several blocks and strips, both lane widths.
    rng = np.random.default_rng(5)
    for dtype in (np.uint8, np.uint16):
        for conn in (4, 8):
            for h, w in ((150, 211), (150, 700), (97, 16500 // np.dtype(dtype).itemsize)):
                for sweeps in (0, 1, 2):
                    for split in (0, 1):
                        say("unclipped", queue, dtype.__name__, conn, h, w, sweeps, split)
                        I = rng.integers(0, 200, (h, w)).astype(dtype)
                        J = rng.integers(0, 255, (h, w)).astype(dtype)  # not clipped
                        pitch = odd_pitch(w, dtype) if w < 1000 else None
                        dj, di = up(J, pitch), up(I, pitch)
                        st = iwpp.reconstruct(dj, di, conn, 0, queue, "plain",
                                              iwpp.default_options(sweep_rounds=sweeps,
                                                                   col_block_rows=32,
                                                                   split_sweeps=split))
                        assert np.array_equal(dj.numpy(), oracle.reconstruct(J, I, conn)), st


@pytest.mark.parametrize("queue", VARIANTS)
@pytest.mark.parametrize("conn", [4, 8])
@pytest.mark.parametrize("dtype", [np.uint8, np.uint16])
def test_reconstruct_split_sweeps(torch_cuda, queue, conn, dtype):
    # the sweep rounds of a row pass and two column passes (split_sweeps)
    rng = np.random.default_rng(zlib.crc32(f"split/{queue}/{conn}/{dtype.__name__}".encode()))
    size = np.dtype(dtype).itemsize
    for h, w in ((1, 300), (300, 1), (200, 333), (640, 257), (150, 16397 // size)):
        for kind in ("random", "seeds", "hdome"):
            for sweeps in (1, 2):
                say("split sweeps", queue, conn, dtype.__name__, h, w, kind, sweeps)
                J, I = rand_pair(rng, h, w, dtype, kind)
                dj, di = up(J), up(I)
                st = iwpp.reconstruct(dj, di, conn, 0, queue, "plain",
                                      iwpp.default_options(sweep_rounds=sweeps, split_sweeps=1))
                assert np.array_equal(dj.numpy(), oracle.reconstruct(J, I, conn)), (h, w, kind, st)


@pytest.mark.parametrize("conn", [4, 8])
@pytest.mark.parametrize("dtype", [np.uint8, np.uint16])
def test_reconstruct_seam_rows(torch_cuda, conn, dtype):
    # raster fix-up windows of every size up to (and past) half a block,
    # one and two sweep rounds, and the sizes where a window meets the image
    # bottom (a last block shorter than the window) or there is one block
    rng = np.random.default_rng(zlib.crc32(f"seams/{conn}/{dtype.__name__}".encode()))
    size = np.dtype(dtype).itemsize
    for h, w in ((70, 300), (200, 333), (97, 4100 // size), (330, 2100 // size)):
        for kind in ("random", "seeds", "hdome"):
            J, I = rand_pair(rng, h, w, dtype, kind)
            want = oracle.reconstruct(J, I, conn)
            for seam in (0, 1, 7, 16, 40):
                for sweeps in (1, 2):
                    say("seam rows", conn, dtype.__name__, h, w, kind, seam, sweeps)
                    dj, di = up(J), up(I)
                    st = iwpp.reconstruct(dj, di, conn, 0, "mutex", "plain",
                                          iwpp.default_options(sweep_rounds=sweeps, seam_rows=seam,
                                                               col_block_rows=32))
                    assert np.array_equal(dj.numpy(), want), (h, w, kind, seam, sweeps, st)


@pytest.mark.parametrize("queue", VARIANTS)
def test_golden_jpeg_crops(torch_cuda, queue):
    import os
    d = np.load(os.path.join(os.path.dirname(__file__), "golden", "jpeg_crops.npz"))
    for k in range(5):
        for conn in (4, 8):
            say("golden", queue, k, conn)
            dj, di = up(d[f"crop{k}_marker"]), up(d[f"crop{k}_mask"])
            iwpp.reconstruct(dj, di, conn, 0, queue)
            assert np.array_equal(dj.numpy(), d[f"crop{k}_out{conn}"]), (k, conn)


def test_legacy_entry_point(torch_cuda):
    import os
    d = np.load(os.path.join(os.path.dirname(__file__), "golden", "jpeg_crops.npz"))
    for k in range(3):  # the square crops
        say("legacy", k)
        dj, di = up(d[f"crop{k}_marker"]), up(d[f"crop{k}_mask"])
        iwpp.MorphologicalReconstruction(dj.ptr, di.ptr, dj.pitch, di.pitch, dj.w)
        assert np.array_equal(dj.numpy(), d[f"crop{k}_out4"]), k


def test_overflow_restart(torch_cuda):
    rng = np.random.default_rng(6)
    J, I = rand_pair(rng, 700, 500, np.uint8, "seeds")
    want = oracle.reconstruct(J, I, 8)
    for queue in ("dq", "uq"):
        say("overflow", queue)
        dj, di = up(J), up(I)
        st = iwpp.reconstruct(dj, di, 8, 0, queue, "plain",
                              iwpp.default_options(sweep_rounds=0, queue_capacity=16))
        assert st["restarts"] > 0, st
        assert np.array_equal(dj.numpy(), want)


@pytest.mark.parametrize("queue", VARIANTS)
def test_config1_full(torch_cuda, queue):
    c = W.c1()
    want = oracle.reconstruct(c.marker, c.mask, c.connectivity)
    for sweeps in (0, 1):
        say("c1", queue, sweeps)
        dj, di = up(c.marker), up(c.mask)
        st = iwpp.reconstruct(dj, di, c.connectivity, c.tile, queue, "plain",
                              iwpp.default_options(sweep_rounds=sweeps))
        assert np.array_equal(dj.numpy(), want), st


@pytest.mark.parametrize("queue", VARIANTS)
def test_config2_full(torch_cuda, queue):
    c = W.c2()
    want = oracle.reconstruct(c.marker, c.mask, c.connectivity)
    say("c2", queue)
    dj, di = up(c.marker), up(c.mask)
    st = iwpp.reconstruct(dj, di, c.connectivity, c.tile, queue)
    say("c2 stats", st)
    assert np.array_equal(dj.numpy(), want), st


def _full_case_check(case, queues):
    want = oracle.reconstruct(case.marker, case.mask, case.connectivity)
    mask = up(case.mask)
    for queue in queues:
        say(case.name, queue)
        dj = up(case.marker)
        st = iwpp.reconstruct(dj, mask, case.connectivity, case.tile, queue)
        got = dj.numpy()
        assert np.array_equal(got, want), (queue, st, int((got != want).sum()))
        # size-independent properties on the device result
        assert (got <= case.mask).all() and (got >= np.minimum(case.marker, case.mask)).all()
        again = up(got)
        st2 = iwpp.reconstruct(again, mask, case.connectivity, case.tile, queue)
        assert st2["tile_visits"] == 0 or np.array_equal(again.numpy(), got)
        assert np.array_equal(again.numpy(), got)  # idempotent


def test_config3_full(torch_cuda):
    case = W.c3()
    _full_case_check(case, ("mutex", "dq", "uq"))
    from iwpp import learned
    learned.install()
    try:
        say(case.name, "learned")
        dj = up(case.marker)
        st = iwpp.reconstruct(dj, up(case.mask), case.connectivity, case.tile, "uq", "learned")
        assert st["schedule"] == iwpp.SCHEDULES["learned"], st
        assert np.array_equal(dj.numpy(), oracle.reconstruct(case.marker, case.mask,
                                                             case.connectivity)), st
    finally:
        iwpp.set_model(None)


@pytest.mark.parametrize("dtype", [np.uint8, np.uint16])
@pytest.mark.parametrize("pitch_kind", ["aligned", "odd"])
def test_clip(torch_cuda, dtype, pitch_kind):
    # the stand-alone clip entry point (IWPP MP/src/MorphologicalReconstruction.cu:827-836)
    rng = np.random.default_rng(21)
    top = 255 if dtype == np.uint8 else 4095
    for h, w in [(1, 1), (3, 17), (64, 1023), (130, 4099), (17, 16400)]:
        say("clip", dtype.__name__, pitch_kind, h, w)
        J = rng.integers(0, top + 1, (h, w)).astype(dtype)
        I = rng.integers(0, top + 1, (h, w)).astype(dtype)
        pitch = odd_pitch(w, dtype) if pitch_kind == "odd" else None
        dj, di = up(J, pitch), up(I, pitch)
        raw_before = dj.raw_numpy()
        iwpp.clip(dj, di)
        torch_cuda.cuda.synchronize()
        assert np.array_equal(dj.numpy(), oracle.clip(J, I)), (h, w)
        rb = w * np.dtype(dtype).itemsize
        assert np.array_equal(dj.raw_numpy()[:, rb:], raw_before[:, rb:])


# Batches for the fused seeding test: single images with 16-byte lanes (rows
# of at least 16 KiB, a ragged last strip), with 4-byte lanes (odd pitch, or
# rows under 16 KiB), and mixed-size batches whose column strips are busy
# enough for 16-byte lanes or not.
def _seed_batches(dtype):
    size = np.dtype(dtype).itemsize
    return [
        ("wide ragged", [(150, 16397 // size)], None),
        ("wide full", [(300, 16384 // size)], None),
        ("odd pitch", [(333, 1000)], "odd"),
        ("narrow", [(200, 3000)], None),
        ("batch busy", [(100, 4096 // size), (64, 2048 // size), (130, 3072 // size)], None),
        ("batch sparse", [(300, 200), (64, 700), (513, 65), (1, 50), (129, 129)], None),
    ]


def _fixup_bands(h, block_rows, seam_rows):
    """Megapixel rows that an upward fix-up pass can leave flagged though
    stable: those holding a row of a window [yb - k, yb) before a block
    boundary yb, or the row just above it (k = min(seam_rows, R / 2))."""
    k = min(seam_rows, block_rows // 2)
    bands = set()
    for yb in range(block_rows, h, block_rows):
        for y in range(yb - k - 1, yb):
            bands.add(y // 64)
    return sorted(bands)


# seam: rows of the raster fix-up passes (0 = none; the fused flags are then
# exactly the stability test's), or None for the library's choice.
@pytest.mark.parametrize("dtype", [np.uint8, np.uint16])
@pytest.mark.parametrize("conn", [4, 8])
@pytest.mark.parametrize("split,seam", [(False, 0), (False, None), (False, 24), (True, 0)])
def test_fused_seeding_flags(torch_cuda, dtype, conn, split, seam):
    # The last column pass of a sweep phase sets the seeding flags itself
    # (tests inside its blocks and strips, then across strip edges and block
    # boundaries).  Without fix-up passes its flags must be exactly those of
    # the stability test on the marker it leaves: a missing flag loses work,
    # an extra one wastes a visit.  The upward fix-up pass re-tests the rows
    # it changes and adds flags, so a flag set before it for a megapixel
    # that it then made stable stays: every unstable megapixel must still be
    # flagged, and extras may lie only beside a fix-up window.
    rng = np.random.default_rng(zlib.crc32(f"seedflags/{dtype.__name__}/{conn}/{split}/{seam}".encode()))
    seam_rows = -1 if seam is None else seam
    for name, shapes, pitch_kind in _seed_batches(dtype):
        for block_rows in (0, 32, 96):
            for kind in ("seeds", "hdome", "random"):
                say("fused seeding", dtype.__name__, conn, split, seam, name, block_rows, kind)
                cases = [rand_pair(rng, h, w, dtype, kind) for h, w in shapes]
                pairs = []
                for (J, I), (h, w) in zip(cases, shapes):
                    pitch = odd_pitch(w, dtype) if pitch_kind == "odd" else None
                    pairs.append((up(J, pitch), up(I, pitch)))
                flags = iwpp.sweep_round_seed(pairs, conn, block_rows, split, seam_rows)
                for (J, I), (dj, _), f in zip(cases, pairs, flags):
                    after = dj.numpy()
                    want = oracle.unstable_blocks(after, I, conn)
                    info = (name, J.shape, block_rows, kind, int(f.sum()), int(want.sum()))
                    if split or seam == 0:
                        assert np.array_equal(f, want), info
                    else:
                        assert set(np.unique(f)) <= {0, 1}, info
                        assert (f >= want).all(), info  # no unstable megapixel left out
                        if block_rows and seam is not None:
                            extra = np.nonzero((f > want).any(axis=1))[0]
                            allowed = _fixup_bands(J.shape[0], block_rows, seam)
                            assert set(extra.tolist()) <= set(allowed), (info, extra, allowed)
                    # the round only raised pixels towards the reconstruction
                    assert (after >= oracle.clip(J, I)).all()
                    assert (after <= oracle.reconstruct(J, I, conn)).all()


def _batch_inputs(rng, shapes, dtype):
    cases = []
    for i, (h, w) in enumerate(shapes):
        kind = ("random", "seeds", "hdome")[i % 3]
        cases.append(rand_pair(rng, h, w, dtype, kind))
    return cases


@pytest.mark.parametrize("queue", VARIANTS)
@pytest.mark.parametrize("conn", [4, 8])
@pytest.mark.parametrize("dtype", [np.uint8, np.uint16])
def test_batch_random(torch_cuda, queue, conn, dtype):
    rng = np.random.default_rng(11)
    shapes = [(1, 1), (1, 300), (300, 1), (64, 64), (65, 129), (200, 333), (640, 257),
              (77, 1500), (1029, 70), (512, 512)]
    cases = _batch_inputs(rng, shapes, dtype)
    for sweeps in (0, 2):
        say("batch", queue, conn, dtype.__name__, sweeps)
        pairs = [(up(J), up(I)) for J, I in cases]
        st = iwpp.reconstruct_batch(pairs, conn, 0, queue, "plain",
                                    iwpp.default_options(sweep_rounds=sweeps))
        for (J, I), (dj, _) in zip(cases, pairs):
            assert np.array_equal(dj.numpy(), oracle.reconstruct(J, I, conn)), (J.shape, st)


def test_batch_graph_replays(torch_cuda):
    rng = np.random.default_rng(12)
    shapes = [(300, 200), (64, 700), (513, 65), (128, 128), (1, 50)]
    cases = _batch_inputs(rng, shapes, np.uint8)
    wants = [oracle.reconstruct(J, I, 8) for J, I in cases]
    pairs = [(up(J), up(I)) for J, I in cases]
    inputs = [up(J) for J, _ in cases]
    ws = iwpp.Workspace()
    g = iwpp.BatchGraph(pairs, ws, 8, options=iwpp.default_options(sweep_rounds=1))
    for rep in range(4):
        say("graph replay", rep)
        for (dj, _), src in zip(pairs, inputs):
            dj.buf.copy_(src.buf)
        # unrelated eager work between replays
        scratch = torch_cuda.ones(1 << 20, device="cuda").sum()
        torch_cuda.cuda.synchronize()
        g.launch()
        st = g.stats()
        for (dj, _), want in zip(pairs, wants):
            assert np.array_equal(dj.numpy(), want), (rep, st)
        assert float(scratch.item()) == float(1 << 20)
    g.close()


def test_config4_batch(torch_cuda):
    cases = W.c4(n=48, max_side=1024)
    wants = [oracle.reconstruct(c.marker, c.mask, 8) for c in cases]
    for queue in ("mutex", "uq"):
        say("c4", queue)
        pairs = [(up(c.marker), up(c.mask)) for c in cases]
        st = iwpp.reconstruct_batch(pairs, 8, 0, queue)
        for (dj, _), want, c in zip(pairs, wants, cases):
            assert np.array_equal(dj.numpy(), want), (c.name, st)


def test_autotune_restores_markers_and_tunes(torch_cuda):
    rng = np.random.default_rng(13)
    cases = _batch_inputs(rng, [(200, 300), (64, 64), (130, 517)], np.uint8)
    pairs = [(up(J), up(I)) for J, I in cases]
    before = [dj.numpy() for dj, _ in pairs]
    say("autotune")
    best = iwpp.autotune(pairs, 8, 1)
    assert best["queue"] in ("mutex", "dq", "uq") and 0 <= best["sweep_rounds"] <= 4
    for (dj, _), b in zip(pairs, before):
        assert np.array_equal(dj.numpy(), b)
    assert best["seed_class"] in (0, 1, 2) and best["col_block_rows"] in (0, 128, 256, 512, 1024)
    st = iwpp.reconstruct_batch(pairs, 8, 0, "auto")
    assert st["queue"] == {"mutex": 0, "dq": 1, "uq": 2}[best["queue"]]
    assert st["sweep_rounds"] == best["sweep_rounds"]
    assert st["seed_class"] == best["seed_class"]
    if best["col_block_rows"]:
        assert st["col_block_rows"] == best["col_block_rows"], (st, best)
    for (J, I), (dj, _) in zip(cases, pairs):
        assert np.array_equal(dj.numpy(), oracle.reconstruct(J, I, 8))
    iwpp.lib().iwpp_autotune_clear()


def _class_case(rng, h, w, dtype, frac):
    # mask non-zero everywhere; on every 64th row (what the class samples) a
    # marker non-zero on exactly round(frac * sampled) pixels, elsewhere
    # non-zero on every pixel (rows the class must not count)
    top = 255 if dtype == np.uint8 else 4095
    mask = rng.integers(1, top + 1, (h, w)).astype(dtype)
    marker = mask.copy()
    rows = np.arange(0, h, 64)
    sampled = np.zeros((len(rows), w), dtype)
    k = int(round(frac * sampled.size))
    sampled.flat[rng.choice(sampled.size, k, replace=False)] = 1
    marker[rows] = np.where(sampled > 0, mask[rows], 0)
    return marker, mask


def _numpy_class(cases):
    nz = sum(int((np.minimum(J[::64], I[::64]) > 0).sum()) for J, I in cases)
    n = sum(J[::64].size for J, _ in cases)
    return 0 if nz * 100 < n else (1 if nz * 2 < n else 2)


def test_seed_class_counts(torch_cuda):
    # the autotuner's seed class (sparse < 1% of sampled pixels with
    # min(marker, mask) > 0, partial < 50%, dense) on fractions either side of
    # both thresholds, rows whose width ends inside a 32-bit word, both pixel
    # types, an odd pitch and a batch; an 'auto' call measures it once some
    # tuning is cached
    rng = np.random.default_rng(15)
    iwpp.autotune([(up(J), up(I)) for J, I in [rand_pair(rng, 70, 90, np.uint8, "seeds")]], 8, 1)
    try:
        for dtype, (h, w), pitch_kind in ((np.uint8, (130, 257), None), (np.uint16, (129, 1003), None),
                                          (np.uint8, (200, 333), "odd")):
            for frac in (0.009, 0.011, 0.49, 0.51):
                J, I = _class_case(rng, h, w, dtype, frac)
                pitch = odd_pitch(w, dtype) if pitch_kind == "odd" else None
                say("seed class", dtype.__name__, h, w, pitch_kind, frac)
                st = iwpp.reconstruct(up(J, pitch), up(I, pitch), 8, 0, "auto")
                assert st["seed_class"] == _numpy_class([(J, I)]), (dtype, h, w, frac, st)
        for frac in (0.009, 0.011, 0.49, 0.51):
            cases = [_class_case(rng, h, w, np.uint8, frac) for h, w in ((130, 257), (65, 70), (300, 41))]
            say("seed class batch", frac)
            st = iwpp.reconstruct_batch([(up(J), up(I)) for J, I in cases], 8, 0, "auto")
            assert st["seed_class"] == _numpy_class(cases), (frac, st)
    finally:
        iwpp.lib().iwpp_autotune_clear()


def test_autotune_keys_on_seed_class(torch_cuda):
    # two workloads of the same sizes, one of sparse seeds and one of h-dome
    # markers, get separate cache entries; each later 'auto' call finds its own
    rng = np.random.default_rng(14)
    shapes = [(300, 400), (256, 512)]
    seeds = [rand_pair(rng, h, w, np.uint8, "seeds") for h, w in shapes]
    hdome = [rand_pair(rng, h, w, np.uint8, "hdome") for h, w in shapes]
    got = {}
    for name, cases in (("seeds", seeds), ("hdome", hdome)):
        pairs = [(up(J), up(I)) for J, I in cases]
        say("autotune seed class", name)
        got[name] = iwpp.autotune(pairs, 8, 1)
    assert got["seeds"]["seed_class"] == 0 and got["hdome"]["seed_class"] == 2, got
    try:
        for name, cases in (("seeds", seeds), ("hdome", hdome)):
            pairs = [(up(J), up(I)) for J, I in cases]
            st = iwpp.reconstruct_batch(pairs, 8, 0, "auto")
            assert st["seed_class"] == got[name]["seed_class"], (name, st)
            assert st["sweep_rounds"] == got[name]["sweep_rounds"], (name, st, got)
            for (J, I), (dj, _) in zip(cases, pairs):
                assert np.array_equal(dj.numpy(), oracle.reconstruct(J, I, 8))
    finally:
        iwpp.lib().iwpp_autotune_clear()
