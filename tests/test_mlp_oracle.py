"""Pins the learned schedule's CPU oracle (oracle/mlp.py) by independent
routes: torch's own bfloat16 conversion, torch float64 layers and autograd,
central finite differences, torch.optim.Adam, and brute-force pixel loops for
the tile features.  CPU only."""
import numpy as np
import pytest
import torch

from oracle import mlp as M

import iwpp
from iwpp import learned


def test_layout_size():
    assert M.N_PARAMS == 5378 == iwpp.MLP_PARAMS
    # 16*64 + 64*64 + 64*2 multiply-adds per sample, 2 FLOP each
    assert 2 * (16 * 64 + 64 * 64 + 64 * 2) == 10496


SPECIAL = np.array([0.0, -0.0, 1.0, -1.0, np.inf, -np.inf, 3.4028235e38, -3.4028235e38,
                    1e-45, -1e-45, 1.1754944e-38, 65504.0, 1.0 + 2 ** -8, 1.0 + 3 * 2 ** -8,
                    1.0 + 2 ** -8 + 2 ** -20, 1.0 - 2 ** -9, 0.1, 1 / 3, -2.5e-30],
                   dtype=np.float32)


def _torch_bf16_bits(x):
    return torch.from_numpy(x).to(torch.bfloat16).view(torch.int16).numpy().view(np.uint16)


def test_bf16_rounding_matches_torch():
    rng = np.random.default_rng(0)
    rnd = (rng.standard_normal(200000) * np.exp(rng.uniform(-40, 40, 200000))).astype(np.float32)
    # exact halfway cases: low 16 bits 0x8000 with even and odd kept parts
    base = rng.integers(0, 2 ** 31, 5000, dtype=np.uint64).astype(np.uint32) & 0x7F7FFFFF
    half = ((base & 0xFFFF0000) | 0x8000).view(np.float32)
    for x in (SPECIAL, rnd, half, -half):
        x = x[np.isfinite(x) | np.isinf(x)]
        assert np.array_equal(M.bf16_bits(x), _torch_bf16_bits(x))


def test_bf16_nan_stays_nan():
    nans = np.array([0x7FC00000, 0x7F800001, 0xFFFFFFFF, 0x7FBFFFFF], dtype=np.uint32).view(np.float32)
    out = M.bf16_to_f32(M.bf16_bits(nans))
    assert np.isnan(out).all()


def test_product_bf16_matches_oracle():
    rng = np.random.default_rng(1)
    x = (rng.standard_normal(50000) * 10).astype(np.float32)
    x = np.concatenate([x, SPECIAL])
    assert np.array_equal(learned.bf16_bits(x), M.bf16_bits(x))


def _rand(seed, n=37):
    rng = np.random.default_rng(seed)
    p = learned.init_params(seed).astype(np.float64)
    p[M.O_B1:M.O_W2] = rng.normal(0, 0.1, 64)
    p[M.O_B2:M.O_W3] = rng.normal(0, 0.1, 64)
    p[M.O_B3:] = rng.normal(0, 0.1, 2)
    x = rng.uniform(0, 1, (n, 16))
    t = rng.uniform(0, 4, (n, 2))
    return p, x, t


def _torch_net(p):
    W1, b1, W2, b2, W3, b3 = (torch.tensor(a, dtype=torch.float64, requires_grad=True)
                              for a in M.unpack(p))
    return [W1, b1, W2, b2, W3, b3]


def _torch_forward(ts, x):
    W1, b1, W2, b2, W3, b3 = ts
    h1 = torch.relu(torch.nn.functional.linear(x, W1, b1))
    h2 = torch.relu(torch.nn.functional.linear(h1, W2, b2))
    return torch.nn.functional.linear(h2, W3, b3)


def test_forward_matches_torch():
    p, x, _ = _rand(2)
    y, _, _ = M.forward(p, x)
    want = _torch_forward(_torch_net(p), torch.tensor(x)).detach().numpy()
    np.testing.assert_allclose(y, want, rtol=1e-12, atol=1e-12)


@pytest.mark.parametrize("lw", [(1.0, 1.0), (1.0, 4.0), (0.0, 2.0)])
def test_backward_matches_autograd(lw):
    p, x, t = _rand(3)
    g = M.backward(p, x, t, lw)
    ts = _torch_net(p)
    y = _torch_forward(ts, torch.tensor(x))
    L = ((y - torch.tensor(t)) ** 2 * torch.tensor(lw, dtype=torch.float64)).sum() / (2 * x.shape[0])
    L.backward()
    want = np.concatenate([a.grad.numpy().ravel() for a in ts])
    np.testing.assert_allclose(g, want, rtol=1e-10, atol=1e-12)
    assert abs(M.loss(p, x, t, lw) - L.item()) < 1e-12


def test_backward_matches_finite_differences():
    p, x, t = _rand(4, n=11)
    g = M.backward(p, x, t, (1.0, 2.0))
    rng = np.random.default_rng(5)
    idx = np.concatenate([rng.choice(M.N_PARAMS, 300, replace=False),
                          np.arange(M.O_B3 - 3, M.N_PARAMS)])
    eps = 1e-6
    for i in idx:
        d = np.zeros(M.N_PARAMS)
        d[i] = eps
        fd = (M.loss(p + d, x, t, (1.0, 2.0)) - M.loss(p - d, x, t, (1.0, 2.0))) / (2 * eps)
        assert abs(fd - g[i]) <= 1e-6 + 1e-5 * abs(g[i]), (i, fd, g[i])


def test_adam_matches_torch():
    rng = np.random.default_rng(6)
    w = rng.standard_normal(M.N_PARAMS)
    tw = torch.tensor(w.copy(), requires_grad=True)
    opt = torch.optim.Adam([tw], lr=3e-3, betas=(0.9, 0.999), eps=1e-8)
    m = np.zeros_like(w)
    v = np.zeros_like(w)
    for step in range(1, 6):
        g = rng.standard_normal(M.N_PARAMS)
        w, m, v = M.adam(w, m, v, g, 3e-3, step)
        opt.zero_grad()
        tw.grad = torch.tensor(g)
        opt.step()
        np.testing.assert_allclose(w, tw.detach().numpy(), rtol=1e-12, atol=1e-14)


def _features_brute(J, I, tile, sub):
    H, W = J.shape
    ntx, nty = -(-W // tile), -(-H // tile)
    out = np.zeros((nty * ntx, 16))
    for t in range(nty * ntx):
        tx, ty = t % ntx, t // ntx
        x0, y0 = tx * tile, ty * tile
        x1, y1 = min(W, x0 + tile), min(H, y0 + tile)
        r = out[t]
        for y in range(y0, y1, sub):
            for x in range(x0, x1):
                j, i = float(J[y, x]), float(I[y, x])
                r[0] += i
                r[1] += j
                r[2] += i - j
                r[3] += (i - j) ** 2
                r[4] = max(r[4], i - j)
                r[5] += j == i
                r[6] += j == 0
                r[7] += j == i and j > 0
                r[8] = max(r[8], j)
                if (x - x0) % 2 == 0 and x + 1 < x1:
                    r[9] += abs(float(I[y, x + 1]) - i)
                r[14] += 1
                r[15] = max(r[15], i)
        for x in range(x0, x1):
            if y0 > 0:
                r[10] += max(0.0, min(float(J[y0 - 1, x]), float(I[y0, x])) - float(J[y0, x]))
            if y1 < H:
                r[11] += max(0.0, min(float(J[y1, x]), float(I[y1 - 1, x])) - float(J[y1 - 1, x]))
        for y in range(y0, y1, sub):
            if x0 > 0:
                r[12] += max(0.0, min(float(J[y, x0 - 1]), float(I[y, x0])) - float(J[y, x0]))
            if x1 < W:
                r[13] += max(0.0, min(float(J[y, x1]), float(I[y, x1 - 1])) - float(J[y, x1 - 1]))
    return out


@pytest.mark.parametrize("shape,tile,sub", [((70, 45), 16, 3), ((33, 64), 32, 1), ((20, 20), 64, 4)])
def test_features_match_brute_force(shape, tile, sub):
    rng = np.random.default_rng(7)
    I = rng.integers(0, 256, shape).astype(np.uint8)
    J = rng.integers(0, 256, shape).astype(np.uint8)
    J[rng.random(shape) < 0.2] = 0
    eq = rng.random(shape) < 0.2
    J[eq] = I[eq]
    np.testing.assert_array_equal(M.tile_features_raw(J, I, tile, sub), _features_brute(J, I, tile, sub))


def test_normalise_constant_image():
    # marker == mask == c everywhere, one full tile: gap 0, all pixels equal,
    # all are seeds, no border inflow, full coverage
    c, tile, sub = 200, 64, 4
    I = np.full((64, 64), c, np.uint8)
    f = M.normalise(M.tile_features_raw(I, I, tile, sub), tile, sub, np.uint8, 8, I.size)[0]
    np.testing.assert_allclose(f[[0, 8]], 1.0)
    np.testing.assert_allclose(f[[2, 3, 4, 6, 9, 10, 11, 12, 13]], 0.0)
    assert f[5] == 1.0 and f[7] == 100.0
    assert f[15] == 12.0 / 32.0  # 64 x 64 pixels = 2^12
    assert f[1] == 1.0 and f[14] == 0.0  # 8-connectivity, uint8
    g = M.normalise(M.tile_features_raw(I, I, tile, sub), tile, sub, np.uint16, 4, I.size)[0]
    assert g[1] == 0.0 and g[14] == 1.0


def test_planted_mistake_is_caught():
    # a sign error in the backward pass must fail the autograd comparison
    p, x, t = _rand(8)
    g = M.backward(p, x, t)
    g[M.O_W2:M.O_B2] *= -1
    ts = _torch_net(p)
    y = _torch_forward(ts, torch.tensor(x))
    (((y - torch.tensor(t)) ** 2).sum() / (2 * x.shape[0])).backward()
    want = np.concatenate([a.grad.numpy().ravel() for a in ts])
    assert not np.allclose(g, want, rtol=1e-6, atol=1e-9)


def test_labels_pick_fastest_round():
    tr = {"final_mean": np.array([10.0, 20.0, 30.0]), "maxval": 255.0}
    y = learned.labels(tr, np.array([5.0, 3.0, 2.5, 2.7, 9.0]))
    assert y[:, 0].tolist() == [2, 2, 2]
    # within 2% of the best, the fewest rounds win
    assert learned.labels(tr, np.array([5.0, 2.53, 2.5, 2.7, 9.0]))[0, 0] == 1  # 1.2% slower
    assert learned.labels(tr, np.array([5.0, 2.56, 2.5, 2.7, 9.0]))[0, 0] == 2  # 2.4% slower
    assert learned.labels(tr, np.array([2.54, 2.6, 2.5, 2.7, 9.0]))[0, 0] == 0
    np.testing.assert_allclose(y[:, 1], [10 / 255, 20 / 255, 30 / 255], rtol=1e-6)


def test_labels_per_tile_cost_model():
    # timings made exactly from t(k) = a + b k + c D(k): the fit recovers
    # (a, b, c), and each tile's label is the argmin of its own share of the
    # rounds against the queue work its dirty megapixels leave (closed form)
    a, b, c = 0.2, 1.5, 0.05
    dirty = np.array([[100, 90, 10, 5, 0],      # a tile a second round clears
                      [100, 2, 1, 1, 0],        # one round is enough
                      [0, 0, 0, 0, 0]],         # nothing to do
                     dtype=np.int64)
    px = np.array([1000, 1000, 1000])
    D = dirty.sum(0)
    t = a + b * np.arange(5) + c * D
    fit = learned.cost_model(t, D)
    np.testing.assert_allclose(fit, (a, b, c), rtol=1e-6)
    tr = {"final_mean": np.zeros(3), "maxval": 1.0, "dirty": dirty, "pixels": px}
    assert learned.labels(tr, t)[:, 0].tolist() == [2, 2, 2]  # the image rule: its k* for all
    y = learned.labels(tr, t, per_tile=True)
    share = px / px.sum()
    want = [int(np.argmin(np.arange(5) * b * share[i] + c * dirty[i])) for i in range(3)]
    assert y[:, 0].tolist() == want
    # per tile, a round costs b/3 = 0.5 ms; a dirty megapixel c = 0.05 ms
    assert want == [2, 1, 0]
    # an unphysical fit (time falling with more queue work) keeps the image rule
    assert learned.cost_model(a + b * np.arange(5) - c * D, D) is None


def test_tile_sums():
    flags = np.arange(6 * 5).reshape(6, 5)
    s = learned._tile_sums(flags, 128)  # 2 x 2 megapixels per tile
    want = [flags[0:2, 0:2].sum(), flags[0:2, 2:4].sum(), flags[0:2, 4:].sum(),
            flags[2:4, 0:2].sum(), flags[2:4, 2:4].sum(), flags[2:4, 4:].sum(),
            flags[4:6, 0:2].sum(), flags[4:6, 2:4].sum(), flags[4:6, 4:].sum()]
    assert s.tolist() == want
    assert learned._tile_pixels(300, 130, 128).tolist() == [128 * 128, 128 * 128, 128 * 44,
                                                            2 * 128, 2 * 128, 2 * 44]
