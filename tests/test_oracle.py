"""Pins the CPU oracle (oracle/) before anything is checked against it.

Each check is independent of the oracle's own code:
  - golden fixtures: crops of the reference's test pair, reconstructed by a
    brute-force fixed-point iteration (tests/golden/make_golden.py);
  - brute force here: J <- min(grey_dilation(J), I) until stable (scipy);
  - closed forms: 1-D rows as "max over sources of min(source, mask on the path)";
    a column pass as an explicit enumeration of monotone paths;
  - invariants: bounds, idempotence, marker = mask, marker = 0, monotonicity.
A planted mistake test shows the brute force catches a dropped diagonal.
"""
import itertools
import os

import numpy as np
import pytest
from scipy import ndimage

import oracle

GOLDEN = os.path.join(os.path.dirname(__file__), "golden", "jpeg_crops.npz")


def brute_force(marker, mask, conn):
    fp = np.ones((3, 3), bool) if conn == 8 else ndimage.generate_binary_structure(2, 1)
    J = np.minimum(marker, mask)
    while True:
        N = np.minimum(ndimage.grey_dilation(J, footprint=fp, mode="constant", cval=0), mask)
        if np.array_equal(N, J):
            return J
        J = N


def random_pair(rng, h, w, dtype, kind):
    top = 255 if dtype == np.uint8 else 4095
    mask = rng.integers(0, top + 1, (h, w)).astype(dtype)
    if kind == "smooth":
        mask = ndimage.uniform_filter(mask.astype(np.float64), 3).astype(dtype)
    if kind == "seeds":
        marker = np.zeros_like(mask)
        n = max(1, h * w // 50)
        idx = rng.choice(h * w, n, replace=False)
        marker.flat[idx] = mask.flat[idx]
    elif kind == "hdome":
        marker = np.maximum(mask.astype(np.int64) - top // 8, 0).astype(dtype)
    else:  # unclipped random marker: the clip is part of the call
        marker = rng.integers(0, top + 1, (h, w)).astype(dtype)
    return marker, mask


SHAPES = [(1, 1), (1, 17), (17, 1), (2, 3), (5, 7), (13, 29), (31, 64), (64, 65), (37, 100)]


@pytest.mark.parametrize("conn", [4, 8])
@pytest.mark.parametrize("dtype", [np.uint8, np.uint16])
@pytest.mark.parametrize("kind", ["random", "seeds", "hdome", "smooth"])
def test_oracle_matches_brute_force(conn, dtype, kind):
    rng = np.random.default_rng(hash((conn, kind, dtype == np.uint8)) % 2**32)
    for h, w in SHAPES:
        marker, mask = random_pair(rng, h, w, dtype, kind)
        got = oracle.reconstruct(marker, mask, conn)
        want = brute_force(marker, mask, conn)
        assert np.array_equal(got, want), (h, w)


@pytest.mark.parametrize("conn", [4, 8])
def test_oracle_golden_jpeg_crops(conn):
    d = np.load(GOLDEN)
    for k in range(5):
        got = oracle.reconstruct(d[f"crop{k}_marker"], d[f"crop{k}_mask"], conn)
        assert np.array_equal(got, d[f"crop{k}_out{conn}"]), k


def test_oracle_row_closed_form():
    # A single row: out[j] = max over k of min(J[k], min of I between k and j).
    rng = np.random.default_rng(7)
    for w in (1, 2, 9, 50):
        I = rng.integers(0, 256, w).astype(np.uint8)
        J = np.minimum(rng.integers(0, 256, w).astype(np.uint8), I)
        want = np.zeros(w, np.int64)
        for j in range(w):
            for k in range(w):
                a, b = min(j, k), max(j, k)
                want[j] = max(want[j], min(int(J[k]), int(I[a:b + 1].min())))
        for conn in (4, 8):
            got = oracle.reconstruct(J[None, :], I[None, :], conn)[0]
            assert np.array_equal(got, want.astype(np.uint8))
        # the row pass alone (forward then backward) already reaches it
        assert np.array_equal(oracle.pass_x(J[None, :], I[None, :])[0], want.astype(np.uint8))


def test_oracle_invariants():
    rng = np.random.default_rng(11)
    for conn in (4, 8):
        for h, w in [(23, 41), (64, 64)]:
            marker, mask = random_pair(rng, h, w, np.uint8, "random")
            J0 = np.minimum(marker, mask)
            R = oracle.reconstruct(marker, mask, conn)
            assert (R >= J0).all() and (R <= mask).all()
            assert np.array_equal(oracle.reconstruct(R, mask, conn), R)          # idempotent
            assert np.array_equal(oracle.reconstruct(mask, mask, conn), mask)    # marker = mask
            assert not oracle.reconstruct(np.zeros_like(mask), mask, conn).any()  # marker = 0
            lower = (J0 // 2).astype(np.uint8)                                  # monotone
            assert (oracle.reconstruct(lower, mask, conn) <= R).all()
            if conn == 8:  # 8-connected result dominates the 4-connected one
                assert (R >= oracle.reconstruct(marker, mask, 4)).all()


def test_oracle_constant_mask_component():
    # One seed at the mask's maximum floods exactly its connected component.
    mask = np.zeros((20, 20), np.uint8)
    mask[2:8, 2:8] = 200
    mask[8, 8] = 200          # touches the square only diagonally
    mask[12:18, 12:18] = 200  # separate square
    marker = np.zeros_like(mask)
    marker[4, 4] = 200
    r4 = oracle.reconstruct(marker, mask, 4)
    r8 = oracle.reconstruct(marker, mask, 8)
    assert (r4[2:8, 2:8] == 200).all() and r4[8, 8] == 0 and not r4[12:18, 12:18].any()
    assert (r8[2:8, 2:8] == 200).all() and r8[8, 8] == 200 and not r8[12:18, 12:18].any()


def _down_paths_max(S, I, conn, y, x):
    """max over monotone downward paths ending at (y, x) of min(S[start], I along the
    path after the start) -- explicit enumeration of paths."""
    h, w = S.shape
    best = int(S[y, x])
    steps = (-1, 0, 1) if conn == 8 else (0,)
    for length in range(1, y + 1):
        for moves in itertools.product(steps, repeat=length):
            xs = [x]
            for dx in moves:
                xs.append(xs[-1] - dx)
            if min(xs) < 0 or max(xs) >= w:
                continue
            y0 = y - length
            v = int(S[y0, xs[-1]])
            for k in range(length):
                v = min(v, int(I[y - k, xs[k]]))
            best = max(best, v)
    return best


@pytest.mark.parametrize("conn", [4, 8])
def test_oracle_pass_y_path_enumeration(conn):
    rng = np.random.default_rng(3 + conn)
    for h, w in [(5, 6), (6, 4), (1, 5)]:
        I = rng.integers(0, 256, (h, w)).astype(np.uint8)
        S = np.minimum(rng.integers(0, 256, (h, w)).astype(np.uint8), I)
        got = oracle.pass_y(S, I, conn, +1, 1 << 20, 1 << 20)
        want = np.array([[_down_paths_max(S, I, conn, y, x) for x in range(w)] for y in range(h)])
        assert np.array_equal(got, want.astype(np.uint8)), (h, w)
        up = oracle.pass_y(S[::-1], I[::-1], conn, +1, 1 << 20, 1 << 20)[::-1]
        assert np.array_equal(oracle.pass_y(S, I, conn, -1, 1 << 20, 1 << 20), up)


@pytest.mark.parametrize("conn", [4, 8])
def test_oracle_pass_y_blocks_are_independent(conn):
    # A blocked pass equals the full-length pass run on each block on its own,
    # with the block's outside neighbours (from src) as a fixed frame.
    rng = np.random.default_rng(21)
    h, w, bw, bh = 37, 45, 16, 10
    I = rng.integers(0, 256, (h, w)).astype(np.uint8)
    S = np.minimum(rng.integers(0, 256, (h, w)).astype(np.uint8), I)
    for d in (+1, -1):
        got = oracle.pass_y(S, I, conn, d, bw, bh)
        for by in range(0, h, bh):
            for bx in range(0, w, bw):
                y0, y1, x0, x1 = by, min(h, by + bh), bx, min(w, bx + bw)
                # frame: one row before (in sweep order) and one column each side
                fy0, fy1 = max(0, y0 - 1), min(h, y1 + 1)
                fx0, fx1 = max(0, x0 - 1), min(w, x1 + 1)
                Sf = S[fy0:fy1, fx0:fx1].copy()
                If = I[fy0:fy1, fx0:fx1].copy()
                # frame pixels keep their src value: a zero mask stops them rising
                inner = np.zeros_like(Sf, bool)
                inner[y0 - fy0:y1 - fy0, x0 - fx0:x1 - fx0] = True
                If[~inner] = 0
                # a frame column must not feed a later row of the block, only the
                # row right after it: isolate by keeping src values fixed (mask 0)
                # and sweeping the block rows only
                out = oracle.pass_y(Sf, If, conn, d, 1 << 20, 1 << 20)
                if conn == 8:
                    # frame columns stay at src, which is what the blocked pass reads
                    pass
                got_block = got[y0:y1, x0:x1]
                assert np.array_equal(got_block, out[y0 - fy0:y1 - fy0, x0 - fx0:x1 - fx0]), (
                    d, by, bx)


@pytest.mark.parametrize("conn", [4, 8])
def test_oracle_unstable_blocks_brute_force(conn):
    rng = np.random.default_rng(5)
    fp = np.ones((3, 3), bool) if conn == 8 else ndimage.generate_binary_structure(2, 1)
    fp = fp.copy()
    fp[1, 1] = False  # neighbours only
    for h, w, mp in [(40, 70, 16), (64, 64, 64), (33, 9, 8)]:
        I = rng.integers(0, 256, (h, w)).astype(np.uint8)
        J = np.minimum(rng.integers(0, 256, (h, w)).astype(np.uint8), I)
        J = oracle.reconstruct(J, I, conn)
        J[rng.integers(0, h), rng.integers(0, w)] = 0  # plant instability
        nbmax = ndimage.grey_dilation(J, footprint=fp, mode="constant", cval=0)
        px = np.minimum(nbmax, I) > J
        want = np.zeros(((h + mp - 1) // mp, (w + mp - 1) // mp), np.uint8)
        for y, x in zip(*np.nonzero(px)):
            want[y // mp, x // mp] = 1
        got = oracle.unstable_blocks(J, I, conn, mp)
        assert np.array_equal(got, want)
        assert not oracle.unstable_blocks(oracle.reconstruct(J, I, conn), I, conn, mp).any()


def test_planted_mistake_is_caught():
    # A 4-connected result presented as 8-connected must fail the brute force on
    # an input that needs a diagonal step.
    mask = np.array([[9, 0], [0, 9]], np.uint8)
    marker = np.array([[9, 0], [0, 0]], np.uint8)
    wrong = oracle.reconstruct(marker, mask, 4)
    assert not np.array_equal(wrong, brute_force(marker, mask, 8))
    assert np.array_equal(oracle.reconstruct(marker, mask, 8), brute_force(marker, mask, 8))


def test_pair50_window_hashes():
    # the committed hashes of the regression window are the oracle's output
    import hashlib
    d = np.load(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "pair50.npz"))
    for conn in (4, 8):
        out = oracle.reconstruct(d["marker"], d["mask"], conn)
        assert hashlib.sha256(out.tobytes()).hexdigest() == str(d[f"out{conn}_sha256"])


def _raster_by_thresholds(S, I, conn, d, bw, bh):
    """The raster pass by threshold decomposition: out(p) >= t iff p is
    reached at level t, i.e. by a walk of allowed moves through pixels whose
    mask is >= t, from a pixel whose src value is >= t.  Within a block and in
    sweep order, a row is entered from the row before (straight, or also
    diagonally with 8-connectivity; from src when that row lies outside the
    block or image edge), from its own src pixels, and from src's pixel left
    of the block (moving right) or right of it (moving left); in the row a
    walk moves right any number of steps, then left.  An independent
    formulation of the pass: reachability per level, no sequential update."""
    h, w = S.shape
    out = S.copy()
    levels = sorted(set(int(v) for v in np.unique(S)) | set(int(v) for v in np.unique(I)))
    for t in levels:
        if t == 0:
            continue
        for by0 in range(0, h, bh):
            by1 = min(h, by0 + bh)
            for bx0 in range(0, w, bw):
                bx1 = min(w, bx0 + bw)
                rows = range(by0, by1) if d > 0 else range(by1 - 1, by0 - 1, -1)
                final = {}  # row -> set of x reached at the end of the row (left moves done)
                for y in rows:
                    yp = y - d
                    start = set()
                    for x in range(bx0, bx1):
                        if I[y, x] < t:
                            continue
                        if S[y, x] >= t:
                            start.add(x)
                            continue
                        if 0 <= yp < h:
                            nbr = [x - 1, x, x + 1] if conn == 8 else [x]
                            for xq in nbr:
                                if not 0 <= xq < w:
                                    continue
                                inside = by0 <= yp < by1 and bx0 <= xq < bx1
                                if (xq in final[yp]) if inside else (S[yp, xq] >= t):
                                    start.add(x)
                                    break
                    right = set()
                    x = bx0
                    carry = bx0 > 0 and S[y, bx0 - 1] >= t
                    for x in range(bx0, bx1):
                        carry = I[y, x] >= t and (carry or x in start)
                        if carry:
                            right.add(x)
                    left = set()
                    carry = bx1 < w and S[y, bx1] >= t
                    for x in range(bx1 - 1, bx0 - 1, -1):
                        carry = I[y, x] >= t and (carry or x in right)
                        if carry:
                            left.add(x)
                    final[y] = left
                    for x in left:
                        out[y, x] = max(out[y, x], t)
    return out


@pytest.mark.parametrize("conn", [4, 8])
def test_oracle_pass_raster_thresholds(conn):
    rng = np.random.default_rng(40 + conn)
    for h, w, bw, bh in [(6, 7, 1 << 20, 1 << 20), (9, 11, 4, 3), (1, 9, 1 << 20, 1 << 20),
                         (8, 1, 1 << 20, 4), (12, 10, 5, 5)]:
        I = rng.integers(0, 12, (h, w)).astype(np.uint8)
        S = np.minimum(rng.integers(0, 12, (h, w)).astype(np.uint8), I)
        S[rng.random((h, w)) < 0.6] = 0
        for d in (+1, -1):
            got = oracle.pass_raster(S, I, conn, d, bw, bh)
            want = _raster_by_thresholds(S, I, conn, d, bw, bh)
            assert np.array_equal(got, want), (h, w, bw, bh, d)


def test_oracle_pass_raster_closed_forms():
    rng = np.random.default_rng(44)
    I = rng.integers(0, 256, (1, 40)).astype(np.uint8)
    S = np.minimum(rng.integers(0, 256, (1, 40)).astype(np.uint8), I)
    # one row, one block: the row pass
    for conn in (4, 8):
        assert np.array_equal(oracle.pass_raster(S, I, conn, 1, 1 << 20, 1 << 20),
                              oracle.pass_x(S, I))
    # one column: the column pass
    I = rng.integers(0, 256, (40, 1)).astype(np.uint8)
    S = np.minimum(rng.integers(0, 256, (40, 1)).astype(np.uint8), I)
    for d in (1, -1):
        assert np.array_equal(oracle.pass_raster(S, I, 8, d, 1 << 20, 7),
                              oracle.pass_y(S, I, 8, d, 1 << 20, 7))


@pytest.mark.parametrize("conn", [4, 8])
@pytest.mark.parametrize("dtype", [np.uint8, np.uint16])
def test_oracle_raster_passes_reach_the_reconstruction(conn, dtype):
    # Raster passes down and up, repeated until nothing changes, stop at the
    # reconstruction (checked by brute force), blocked or not; each pass stays
    # between its input and the reconstruction.
    rng = np.random.default_rng(50 + conn)
    for h, w, bw, bh in [(33, 47, 1 << 20, 1 << 20), (40, 29, 8, 6), (17, 64, 16, 32)]:
        marker, mask = random_pair(rng, h, w, dtype, "seeds")
        want = brute_force(marker, mask, conn)
        J = np.minimum(marker, mask)
        for it in range(10000):
            before = J
            J = oracle.pass_raster(J, mask, conn, +1, bw, bh)
            J = oracle.pass_raster(J, mask, conn, -1, bw, bh)
            assert (J >= before).all() and (J <= want).all()
            if np.array_equal(J, before):
                break
        assert np.array_equal(J, want), (h, w, bw, bh)


def test_oracle_pass_raster_planted_mistake_is_caught():
    # A raster pass without its row scans (the column pass alone) differs
    # from the threshold formulation on an input that needs a sideways step.
    I = np.array([[5, 5, 5], [0, 0, 5]], np.uint8)
    S = np.array([[5, 0, 0], [0, 0, 0]], np.uint8)
    want = _raster_by_thresholds(S, I, 4, 1, 1 << 20, 1 << 20)
    assert want[1, 2] == 5
    assert not np.array_equal(oracle.pass_y(S, I, 4, 1, 1 << 20, 1 << 20), want)
    assert np.array_equal(oracle.pass_raster(S, I, 4, 1, 1 << 20, 1 << 20), want)
