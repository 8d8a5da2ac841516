"""CPU oracle for the learned schedule's arithmetic: BF16 rounding, the tile
features, the MLP 16-64-64-2 forward and backward passes, and one Adam step.

TEST INFRASTRUCTURE ONLY.  Only tests/, __graft_entry__.smoke() and the CPU
baseline leg of bench.py may import this module; the product path
(iwpp-mp_amd/) never does.

No reference counterpart: the reference has no learned part (its schedule is
fixed at build time).  Each function is pinned in tests/test_mlp_oracle.py
by an independent route: BF16 rounding against torch's float32->bfloat16
conversion (round to nearest even) on special and random values; the forward
pass against torch.nn.functional in float64; the backward pass against
central finite differences and torch autograd in float64; Adam against
torch.optim.Adam; the features against brute-force per-pixel loops on small
images.
"""
import numpy as np

N_IN, N_HID, N_OUT = 16, 64, 2
O_W1 = 0
O_B1 = O_W1 + N_HID * N_IN
O_W2 = O_B1 + N_HID
O_B2 = O_W2 + N_HID * N_HID
O_W3 = O_B2 + N_HID
O_B3 = O_W3 + N_OUT * N_HID
N_PARAMS = O_B3 + N_OUT  # 5378


def bf16_bits(x):
    """IEEE-754 binary32 -> bfloat16 bit patterns, round to nearest even on
    the 16 dropped bits; NaNs stay NaN (quiet bit set)."""
    u = np.ascontiguousarray(x, dtype=np.float32).view(np.uint32).astype(np.uint64)
    nan = ((u & 0x7F800000) == 0x7F800000) & ((u & 0x007FFFFF) != 0)
    lsb = (u >> 16) & 1
    r = ((u + 0x7FFF + lsb) >> 16) & 0xFFFF
    r = np.where(nan, ((u >> 16) | 0x40) & 0xFFFF, r)
    return r.astype(np.uint16)


def bf16_to_f32(bits):
    return (np.asarray(bits, dtype=np.uint16).astype(np.uint32) << 16).view(np.float32)


def bf16_round(x):
    return bf16_to_f32(bf16_bits(x))


def unpack(p):
    p = np.asarray(p)
    return (p[O_W1:O_B1].reshape(N_HID, N_IN), p[O_B1:O_W2], p[O_W2:O_B2].reshape(N_HID, N_HID),
            p[O_B2:O_W3], p[O_W3:O_B3].reshape(N_OUT, N_HID), p[O_B3:])


def forward(p, x):
    """y (n, 2), h1 (n, 64), h2 (n, 64); float64 arithmetic on the given
    parameter values (round them with bf16_round first for the BF16 path)."""
    W1, b1, W2, b2, W3, b3 = (a.astype(np.float64) for a in unpack(p))
    x = np.asarray(x, dtype=np.float64)
    h1 = np.maximum(x @ W1.T + b1, 0.0)
    h2 = np.maximum(h1 @ W2.T + b2, 0.0)
    return h2 @ W3.T + b3, h1, h2


def loss(p, x, t, lw=(1.0, 1.0)):
    y, _, _ = forward(p, x)
    n = x.shape[0]
    e = y - np.asarray(t, dtype=np.float64)
    return float((np.asarray(lw) * e * e).sum() / (2 * n))


def backward(p, x, t, lw=(1.0, 1.0)):
    """Gradient of loss() with respect to the parameters (float64)."""
    W1, b1, W2, b2, W3, b3 = (a.astype(np.float64) for a in unpack(p))
    x = np.asarray(x, dtype=np.float64)
    n = x.shape[0]
    y, h1, h2 = forward(p, x)
    d3 = (y - np.asarray(t, dtype=np.float64)) * np.asarray(lw) / n
    d2 = (d3 @ W3) * (h2 > 0)
    d1 = (d2 @ W2) * (h1 > 0)
    g = np.zeros(N_PARAMS)
    g[O_W1:O_B1] = (d1.T @ x).ravel()
    g[O_B1:O_W2] = d1.sum(0)
    g[O_W2:O_B2] = (d2.T @ h1).ravel()
    g[O_B2:O_W3] = d2.sum(0)
    g[O_W3:O_B3] = (d3.T @ h2).ravel()
    g[O_B3:] = d3.sum(0)
    return g


def adam(w, m, v, g, lr, step, b1=0.9, b2=0.999, eps=1e-8):
    """One bias-corrected Adam step (Kingma & Ba 2015, Algorithm 1); float64."""
    m = b1 * m + (1 - b1) * g
    v = b2 * v + (1 - b2) * g * g
    mh = m / (1 - b1 ** step)
    vh = v / (1 - b2 ** step)
    return w - lr * mh / (np.sqrt(vh) + eps), m, v


def tile_features_raw(J, I, tile, sub):
    """Raw per-tile statistics (tiles in row-major order, 16 each):
    0 sum I, 1 sum J, 2 sum (I-J), 3 sum (I-J)^2, 4 max(0, max (I-J)), 5 #(J==I),
    6 #(J==0), 7 #(J==I>0), 8 max J, 9 sum |I(x+1)-I(x)| over pairs with x even
    and x+1 in the tile,
    10..13 top/bottom/left/right border inflow sum max(0, min(J(out), I(in)) - J(in)),
    14 pixels sampled, 15 max I.  Statistics 0-9, 12-15 use every `sub`-th
    row from the tile's first; the top and bottom border terms use every
    pixel of the row."""
    J = np.asarray(J, dtype=np.float64)
    I = np.asarray(I, dtype=np.float64)
    H, W = J.shape
    ntx, nty = -(-W // tile), -(-H // tile)
    out = np.zeros((nty * ntx, 16))
    for ty in range(nty):
        for tx in range(ntx):
            x0, y0 = tx * tile, ty * tile
            x1, y1 = min(W, x0 + tile), min(H, y0 + tile)
            j = J[y0:y1:sub, x0:x1]
            i = I[y0:y1:sub, x0:x1]
            g = i - j
            r = out[ty * ntx + tx]
            r[0], r[1], r[2], r[3] = i.sum(), j.sum(), g.sum(), (g * g).sum()
            r[4] = max(0.0, g.max())
            r[5] = (j == i).sum()
            r[6] = (j == 0).sum()
            r[7] = ((j == i) & (j > 0)).sum()
            r[8] = j.max()
            nx = (x1 - x0) // 2 * 2
            r[9] = np.abs(i[:, 1:nx:2] - i[:, 0:nx:2]).sum()
            if y0 > 0:
                r[10] = np.maximum(0, np.minimum(J[y0 - 1, x0:x1], I[y0, x0:x1]) - J[y0, x0:x1]).sum()
            if y1 < H:
                r[11] = np.maximum(0, np.minimum(J[y1, x0:x1], I[y1 - 1, x0:x1])
                                   - J[y1 - 1, x0:x1]).sum()
            if x0 > 0:
                r[12] = np.maximum(0, np.minimum(J[y0:y1:sub, x0 - 1], I[y0:y1:sub, x0])
                                   - J[y0:y1:sub, x0]).sum()
            if x1 < W:
                r[13] = np.maximum(0, np.minimum(J[y0:y1:sub, x1], I[y0:y1:sub, x1 - 1])
                                   - J[y0:y1:sub, x1 - 1]).sum()
            r[14] = j.size
            r[15] = i.max()
    return out


def normalise(raw, tile, sub, dtype, conn, pixels):
    """The model's inputs from raw statistics; M = the largest max I over all
    tiles; features 1 and 14 carry the connectivity (8 -> 1) and the pixel
    type (uint16 -> 1), feature 15 the image's size, min(log2(pixels), 28) / 32.
    The left/right border sums, taken on every `sub`-th row, are scaled by
    `sub` to the size of the top/bottom ones."""
    raw = np.asarray(raw, dtype=np.float64)
    M = max(raw[:, 15].max(), 1.0)
    n = np.maximum(raw[:, 14], 1.0)
    f = np.zeros_like(raw)
    f[:, 0] = raw[:, 0] / n / M
    f[:, 1] = 1.0 if conn == 8 else 0.0
    f[:, 2] = raw[:, 2] / n / M
    f[:, 3] = np.sqrt(raw[:, 3] / n) / M
    f[:, 4] = raw[:, 4] / M
    f[:, 5] = raw[:, 5] / n
    f[:, 6] = raw[:, 6] / n
    f[:, 7] = raw[:, 7] / n * 100.0
    f[:, 8] = raw[:, 8] / M
    f[:, 9] = raw[:, 9] / n / M * 10.0
    for k in range(10, 14):
        f[:, k] = raw[:, k] * (sub if k >= 12 else 1) / tile / M
    f[:, 14] = 1.0 if np.dtype(dtype).itemsize == 2 else 0.0
    f[:, 15] = min(np.log2(float(pixels)), 28.0) / 32.0
    return f
