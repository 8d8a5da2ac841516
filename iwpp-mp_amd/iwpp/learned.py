"""Learned schedule: traces generated by the exact method, labels, training on
the GPU with the library's own kernels, and the model file.

The model (MLP 16-64-64-2, ReLU) maps the 16 features of a tile to
  y0 = sweep rounds the tile takes before the queue phase (rounded, 0..4),
  y1 = the tile's seeding priority (higher enters the queue first).
The library applies it (iwpp_reconstruct_ex with IWPP_SCHED_LEARNED); the
schedule never changes the result, only the work done to reach it.

Traces.  For one training image, with the schedule's tile size:
  * the features of every tile on the input marker (what the library sees),
  * d_t(k): megapixels of tile t whose marker still differs from the
    reconstruction after k full sweep rounds, k = 0..K, using the library's
    own raster passes (the same kernels and column blocks as a
    reconstruction, out of place); the queue phase visits each of them at
    least once,
  * the tile's mean reconstructed value.
Labels.  For every training image the whole reconstruction is timed with
k = 0..4 uniform sweep rounds; every tile of the image gets the fastest k,
(the fewest rounds among those within 2% of it; labels(..., per_tile=True)
gives each tile its own k from a cost model fitted to the timings and the
dirty counts, for split sweeps), and priority_t = mean
reconstructed value of t / max mask value, so tiles
whose flood carries the highest values are seeded first (their propagation
overwrites lower floods, which would otherwise be propagated and then
overtaken).  The reference has no counterpart; its schedule is fixed
(IWPP MP/src/MorphologicalReconstruction.cu:785-806 seeds every megapixel).

Model file: iwpp-mp_amd/model/sched_mlp.f32 (5378 little-endian float32,
layout W1[64][16] b1[64] W2[64][64] b2[64] W3[2][64] b3[2]) and
sched_mlp.json (training metadata).  C callers read the .f32 file and pass
it to iwpp_sched_set_model.
"""
import json
import os
import time

import numpy as np

from . import (MLP_FEATURES, MLP_PARAMS, MP_SIZE, DeviceImage, Workspace, clip, default_options,
               mlp_adam, mlp_backward, mlp_forward, reconstruct, set_model, sweep_raster,
               tile_features)

_PKG = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
MODEL_DIR = os.path.join(_PKG, "model")
MODEL_PATH = os.path.join(MODEL_DIR, "sched_mlp.f32")
META_PATH = os.path.join(MODEL_DIR, "sched_mlp.json")

MAX_ROUNDS = 4
FEATURE_SUB = 64  # row subsampling of the features (the library uses the same)
FEATURE_SET = 4   # feature definitions (2: feature 15 = log2(image pixels) / 32;
                  # 4: left/right border inflow on the sampled rows only)


def schedule_tile(width, height):
    """Tile size of the schedule for an image (the library's plan: 2048 when
    the short side is at least 8192, else 1024)."""
    return 2048 if min(width, height) >= 8192 else 1024


def _torch():
    import torch
    return torch


# ---------------------------------------------------------------- model -----
def init_params(seed=0):
    """He-normal weights, zero biases (float32, MLP_PARAMS values)."""
    rng = np.random.default_rng(seed)
    p = np.zeros(MLP_PARAMS, dtype=np.float32)
    o = 0
    for fan_out, fan_in in ((64, 16), (64, 64), (2, 64)):
        p[o:o + fan_out * fan_in] = rng.normal(0, np.sqrt(2.0 / fan_in), fan_out * fan_in)
        o += fan_out * fan_in + fan_out
    return p


def bf16_bits(w):
    """float32 -> bfloat16 bit patterns, round to nearest even (NaN kept quiet);
    the same rule as the library's iwpp_sched_set_model."""
    u = np.ascontiguousarray(w, dtype=np.float32).view(np.uint32).astype(np.uint64)
    nan = ((u & 0x7F800000) == 0x7F800000) & ((u & 0x007FFFFF) != 0)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) & 0xFFFF
    return np.where(nan, ((u >> 16) | 0x40) & 0xFFFF, r).astype(np.uint16)


def load_model(path=None):
    """Weights of a model file.  Its .json beside it, when present, must name
    the feature subsampling the library computes (a model trained on other
    features is refused)."""
    path = path or MODEL_PATH
    w = np.fromfile(path, dtype="<f4")
    if w.size != MLP_PARAMS:
        raise ValueError(f"{path}: {w.size} values, expected {MLP_PARAMS}")
    meta_path = os.path.splitext(path)[0] + ".json"
    if os.path.exists(meta_path):
        with open(meta_path) as f:
            meta = json.load(f)
        sub = meta.get("feature_sub")
        if sub is not None and sub != FEATURE_SUB:
            raise ValueError(f"{path}: trained on features of every {sub}th row, "
                             f"the library samples every {FEATURE_SUB}th")
        if meta.get("feature_set", 1) != FEATURE_SET:
            raise ValueError(f"{path}: trained on feature set {meta.get('feature_set', 1)}, "
                             f"the library computes set {FEATURE_SET}")
    return w.astype(np.float32)


def install(path=None):
    """Load the model file and install it in the library; returns the weights."""
    w = load_model(path)
    set_model(w)
    return w


def save_model(w, meta, path=None, meta_path=None):
    path = path or MODEL_PATH
    meta_path = meta_path or (os.path.splitext(path)[0] + ".json")
    os.makedirs(os.path.dirname(path), exist_ok=True)
    np.asarray(w, dtype="<f4").tofile(path)
    with open(meta_path, "w") as f:
        json.dump(meta, f, indent=1, sort_keys=True)


# --------------------------------------------------------------- traces -----
def _tile_sums(flags, tile):
    """Per-tile sums of a (nmy, nmx) megapixel array, tiles in row-major order."""
    m = tile // MP_SIZE
    nmy, nmx = flags.shape
    nty, ntx = -(-nmy // m), -(-nmx // m)
    pad = np.zeros((nty * m, ntx * m), dtype=np.int64)
    pad[:nmy, :nmx] = flags
    return pad.reshape(nty, m, ntx, m).sum(axis=(1, 3)).ravel()


def _tile_pixels(w, h, tile):
    xs = np.minimum(tile, w - np.arange(0, w, tile))
    ys = np.minimum(tile, h - np.arange(0, h, tile))
    return np.outer(ys, xs).ravel()


def _dirty_megapixels(J, F):
    """(nmy, nmx) int64: megapixels where device images J and F differ."""
    torch = _torch()
    rb = J.w * J.dtype
    nmy, nmx = -(-J.h // MP_SIZE), -(-J.w // MP_SIZE)
    d = torch.zeros((nmy * MP_SIZE, nmx * MP_SIZE * J.dtype), dtype=torch.bool,
                    device=J.buf.device)
    d[:J.h, :rb] = J.buf[:, :rb] != F.buf[:, :rb]
    d = d.view(nmy, MP_SIZE, nmx, MP_SIZE * J.dtype).any(dim=3).any(dim=1)
    return d.cpu().numpy().astype(np.int64)


def trace(mask, marker, connectivity, rounds=MAX_ROUNDS, tile=None):
    """Trace of one image (numpy mask/marker): dict with
    features (T, 16) of the input marker; dirty (T, rounds+1): megapixels of
    tile t whose marker still differs from the reconstruction after k sweep
    rounds (each one the queue phase has to visit at least once); pixels (T,);
    final_mean (T,); maxval; tile."""
    torch = _torch()
    h, w = mask.shape
    tile = tile or schedule_tile(w, h)
    I = DeviceImage.from_numpy(mask)
    J = DeviceImage.from_numpy(marker)
    F = DeviceImage.from_numpy(marker)
    st = reconstruct(F, I, connectivity=connectivity, queue="mutex")
    _, feat = tile_features(J, I, tile, connectivity, FEATURE_SUB)
    feats = feat.cpu().numpy().astype(np.float32)
    P = DeviceImage(h, w, I.dtype)
    clip(J, I)
    dirty = []
    # the reconstruction's sweep rounds: raster passes down and up over its
    # column blocks (out of place here, so that the trace is deterministic)
    rows = st["col_block_rows"]
    for k in range(rounds + 1):
        dirty.append(_tile_sums(_dirty_megapixels(J, F), tile))
        if k < rounds:
            sweep_raster(J, P, I, connectivity, +1, rows)
            sweep_raster(P, J, I, connectivity, -1, rows)
    torch.cuda.synchronize()
    final = F.numpy().astype(np.float64)
    nty, ntx = -(-h // tile), -(-w // tile)
    pad = np.zeros((nty * tile, ntx * tile))
    pad[:h, :w] = final
    sums = pad.reshape(nty, tile, ntx, tile).sum(axis=(1, 3)).ravel()
    px = _tile_pixels(w, h, tile)
    return {"features": feats, "dirty": np.stack(dirty, 1), "pixels": px,
            "final_mean": sums / px, "maxval": float(mask.max()), "tile": tile}


def image_timings(mask, marker, conn, rounds=MAX_ROUNDS, trials=5, workspace=None):
    """Wall-clock ms of a whole reconstruction (launches, host work and
    synchronisation included) with k uniform sweep rounds, k = 0..rounds,
    best of `trials`, on the GPU."""
    torch = _torch()
    I = DeviceImage.from_numpy(mask)
    J0 = DeviceImage.from_numpy(marker)
    J = DeviceImage(mask.shape[0], mask.shape[1], I.dtype)
    out = np.full(rounds + 1, np.inf)
    for k in range(rounds + 1):
        opt = default_options(sweep_rounds=k)
        for _ in range(trials):
            J.buf.copy_(J0.buf)
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            reconstruct(J, I, connectivity=conn, queue="mutex", options=opt, workspace=workspace)
            torch.cuda.synchronize()
            out[k] = min(out[k], (time.perf_counter() - t0) * 1e3)
    return out


TIE = 1.02  # timings within 2% count as equal (the smaller k wins)


def cost_model(timings, dirty_total):
    """Fit of an image's whole-reconstruction time with k uniform sweep rounds,
    t(k) = a + b k + c D(k), D(k) = megapixels still dirty after k rounds:
    b = ms per sweep round over the whole image, c = ms per dirty megapixel
    left to the queue phase.  None when the fit is not physical (b or c <= 0)."""
    t = np.asarray(timings, dtype=np.float64)
    D = np.asarray(dirty_total, dtype=np.float64)[: t.size]
    A = np.stack([np.ones_like(t), np.arange(t.size, dtype=np.float64), D], 1)
    (a, b, c), *_ = np.linalg.lstsq(A, t, rcond=None)
    return (a, b, c) if b > 0 and c > 0 else None


def labels(tr, timings, per_tile=False):
    """(k, priority) per tile.
    k: the image's k* = the fewest uniform rounds within TIE of the fastest
    measured, for every tile: the library runs an image's rounds in every
    tile under raster sweeps (a masked raster round costs about a whole one).
    per_tile (for split sweeps, whose later rounds can be masked per tile):
    with a physical cost model of the image (cost_model), each tile's own
    argmin over k of  k * b * (tile pixels / image pixels) + c * d_t(k)
    (its share of the sweep rounds against the queue work its dirty
    megapixels leave), ties to fewer rounds; k* when the fit is not physical.
    priority: the tile's mean reconstructed value over the image's largest
    mask value."""
    t = np.asarray(timings, dtype=np.float64)
    k_img = float(np.flatnonzero(t <= TIE * t.min())[0])  # fewest rounds within TIE of the best
    prio = (tr["final_mean"] / max(tr["maxval"], 1.0)).astype(np.float32)
    d = np.asarray(tr["dirty"], dtype=np.float64)[:, : t.size] if "dirty" in tr else None
    fit = cost_model(t, d.sum(0)) if d is not None and per_tile else None
    if fit is None:
        k = np.full(prio.shape, k_img, np.float32)
    else:
        _, b, c = fit
        share = np.asarray(tr["pixels"], dtype=np.float64) / float(np.sum(tr["pixels"]))
        ks = np.arange(t.size, dtype=np.float64)
        cost = ks[None, :] * b * share[:, None] + c * d
        k = np.argmin(cost + 1e-12 * ks[None, :], axis=1).astype(np.float32)
    return np.stack([k, prio], 1)


# ------------------------------------------------------------- training -----
def train(X, Y, steps=3000, lr=2e-3, lw=(1.0, 4.0), seed=0, log_every=100, device="cuda"):
    """Full-batch Adam on the GPU with the library's kernels: BF16 forward,
    backward through the BF16 weights into FP32 master weights.
    Returns (weights float32, history [(step, loss)])."""
    torch = _torch()
    n = X.shape[0]
    x = torch.from_numpy(np.ascontiguousarray(X, dtype=np.float32)).to(device)
    t = torch.from_numpy(np.ascontiguousarray(Y, dtype=np.float32)).to(device)
    w0 = init_params(seed)
    w = torch.from_numpy(w0).to(device)
    m = torch.zeros_like(w)
    v = torch.zeros_like(w)
    wbf = torch.from_numpy(bf16_bits(w0).view(np.int16)).to(device)
    lwt = torch.tensor(lw, dtype=torch.float32, device=device)
    hist = []
    for step in range(1, steps + 1):
        y, acts = mlp_forward(wbf, x, bf16=True, acts=True)
        if step == 1 or step % log_every == 0 or step == steps:
            L = float(((y - t) ** 2 * lwt).sum() / (2 * n))
            hist.append((step - 1, L))
        g = mlp_backward(wbf, x, acts, y, t, bf16=True, lw=lw)
        mlp_adam(w, m, v, g, wbf, lr, step)
    torch.cuda.synchronize()
    return w.cpu().numpy().astype(np.float32), hist


def training_cases(counts=None, seed_base=None):
    """Training images: the generators of the first four workloads (h-dome
    1024^2; 0.1% seeds 4096^2; uint16 h-dome with disks 16384^2, 4-connected;
    the batch of mixed sizes) with seeds those workloads do not use, plus the
    same seeds / h-dome generators at 8192^2 with other seed densities, blur
    widths and disk counts.  Returns (name, mask, marker, connectivity) tuples."""
    from . import U8, U16, synth_mask, synth_marker_hdome, synth_marker_seeds
    from . import workloads as wl
    base = seed_base if seed_base is not None else wl.S0 + 1000
    n = {"c1": 4, "c2": 6, "c3": 2, "c4": 96, "seeds8k": 6, "hdome8k": 3, "sizes": 1}
    n.update(counts or {})
    out = []
    # The h-dome (first workload) and seeds (second) generators over a range
    # of sides, and the uint16 disk (third) and seeds-with-disks (fifth)
    # generators at smaller sides: the best round count depends on the size.
    for rep in range(n["sizes"]):
        for side in (2048, 3072, 6144, 8192, 12288):
            for fam, gen in (("c1", wl.c1), ("c2", wl.c2)):
                c = gen(side=side, seed=base + 100 + 31 * rep + side // 512 + (7 if fam == "c2" else 0))
                out.append((f"{c.name}_{side}", c.mask, c.marker, c.connectivity))
        for side in (8192, 12288):
            c = wl.c3(side=side, seed=base + 200 + 31 * rep + side // 1024, ndisks=200 * side * side // 16384 ** 2)
            out.append((f"{c.name}_{side}", c.mask, c.marker, c.connectivity))
        for side in (8192, 16384):
            c = wl.c5(side=side, seed=base + 300 + 31 * rep + side // 1024, ndisks=1000 * side * side // 32768 ** 2)
            out.append((f"{c.name}_{side}", c.mask, c.marker, c.connectivity))
    for i in range(n["c1"]):
        c = wl.c1(seed=base + 10 + i)
        out.append((c.name, c.mask, c.marker, c.connectivity))
    for i in range(n["c2"]):
        c = wl.c2(seed=base + 20 + i)
        out.append((c.name, c.mask, c.marker, c.connectivity))
    for i in range(n["c3"]):
        c = wl.c3(seed=base + 30 + i)
        out.append((c.name, c.mask, c.marker, c.connectivity))
    if n["c4"]:
        for c in wl.c4(n=n["c4"], seed=base + 40):
            out.append((c.name, c.mask, c.marker, c.connectivity))
    # (ppm, sigma, disks) of the 8192^2 seeds images, cycled
    seeds_params = [(1000, 8.0, 0), (300, 16.0, 100), (100, 32.0, 250), (30, 32.0, 250),
                    (100, 16.0, 0), (300, 32.0, 250)]
    for i in range(n["seeds8k"]):
        ppm, sigma, disks = seeds_params[i % len(seeds_params)]
        sd = base + 60 + i
        mask = synth_mask(8192, 8192, U8, 255, sigma, disks, 50, 400, sd)
        marker = synth_marker_seeds(mask, wl.seeds_count(mask.size, ppm), sd)
        out.append((f"seeds_8192_u8_ppm{ppm}_s{int(sigma)}", mask, marker, 8))
    hdome_params = [(50, 16.0, 50, U16, 4095, 4), (30, 8.0, 0, U8, 255, 8), (20, 32.0, 100, U8, 255, 8)]
    for i in range(n["hdome8k"]):
        h, sigma, disks, dt, top, conn = hdome_params[i % len(hdome_params)]
        sd = base + 80 + i
        mask = synth_mask(8192, 8192, dt, top, sigma, disks, 50, 400, sd)
        out.append((f"hdome_8192_{'u16' if dt == U16 else 'u8'}_h{h}", mask,
                    synth_marker_hdome(mask, h), conn))
    return out


def build_dataset(cases, rounds=MAX_ROUNDS, log=None):
    """Features and labels of every tile of `cases` (traces and timings on
    the GPU).  Returns X, Y, per-image information, and the traces' dirty
    counts (tiles x rounds + 1) and tile pixels (for relabelling)."""
    from . import Workspace as _W
    ws = _W()
    X, Y, info, Ds, Ps, Rs = [], [], [], [], [], []
    for name, mask, marker, conn in cases:
        t0 = time.time()
        tr = trace(mask, marker, conn, rounds)
        tm = image_timings(mask, marker, conn, rounds, workspace=ws)
        y = labels(tr, tm)
        X.append(tr["features"])
        Y.append(y)
        Ds.append(tr["dirty"])
        Ps.append(tr["pixels"])
        info.append({"image": name, "tiles": int(y.shape[0]), "ms_by_rounds": tm.round(4).tolist(),
                     "dirty_by_rounds": tr["dirty"].sum(0).tolist(),
                     "k_hist": np.bincount(y[:, 0].astype(int), minlength=rounds + 1).tolist()})
        if log:
            log(f"trace {name}: {tr['features'].shape[0]} tiles, dirty "
                f"{tr['dirty'].sum(0).tolist()}, ms by rounds {np.round(tm, 3).tolist()}, "
                f"{time.time() - t0:.1f} s")
    ws.close()
    return (np.concatenate(X), np.concatenate(Y), info, np.concatenate(Ds).astype(np.int64),
            np.concatenate(Ps).astype(np.int64))


def validation_cases(seed_base=None):
    """Held-out images for choosing between trained models: the h-dome and
    seeds generators at 6144^2 and 12288^2, the uint16 disk generator at
    12288^2 and the seeds-with-disks generator at 16384^2, with seeds that
    neither the training set nor the tests use."""
    from . import workloads as wl
    base = seed_base if seed_base is not None else wl.S0 + 3000
    out = []
    for side in (6144, 12288):
        out.append(wl.c1(side=side, seed=base + side // 1024))
        out.append(wl.c2(side=side, seed=base + 50 + side // 1024))
    out.append(wl.c3(side=12288, seed=base + 100, ndisks=112))
    out.append(wl.c5(side=16384, seed=base + 101, ndisks=250))
    return out


def validation_ms(w, cases, reps=3):
    """Total wall-clock ms of reconstructing `cases` under the learned
    schedule with weights w (median of reps per image)."""
    torch = _torch()
    import statistics
    from . import Workspace as _W
    set_model(w)
    ws = _W()
    total = 0.0
    for c in cases:
        I = DeviceImage.from_numpy(c.mask)
        J0 = DeviceImage.from_numpy(c.marker)
        J = DeviceImage(J0.h, J0.w, J0.dtype, J0.pitch)
        ts = []
        for _ in range(reps + 1):
            J.buf.copy_(J0.buf)
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            reconstruct(J, I, connectivity=c.connectivity, queue="auto", schedule="learned",
                        workspace=ws)
            torch.cuda.synchronize()
            ts.append((time.perf_counter() - t0) * 1e3)
        total += statistics.median(ts[1:])
    set_model(None)
    return total


def main(argv=None):
    """Train the schedule model on the GPU and write the model files:
        python -m iwpp.learned [--out DIR] [--steps N] [--lr LR] [--inits K]
    With K > 1, K networks from different initial weights are trained on the
    same data and the one with the lowest held-out reconstruction time
    (validation_cases) is kept."""
    import argparse
    ap = argparse.ArgumentParser(description=main.__doc__)
    ap.add_argument("--out", default=MODEL_DIR)
    ap.add_argument("--steps", type=int, default=4000)
    ap.add_argument("--lr", type=float, default=2e-3)
    ap.add_argument("--c4", type=int, default=96)
    ap.add_argument("--inits", type=int, default=1)
    ap.add_argument("--relabel", default=None,
                    help="reuse DIR/sched_dataset.npz and the per-image timings of "
                         "DIR/sched_mlp.json (labels recomputed with the current rule)")
    args = ap.parse_args(argv)

    def log(msg):
        print(msg, flush=True)

    t0 = time.time()
    if args.relabel:
        d = np.load(os.path.join(args.relabel, "sched_dataset.npz"))
        with open(os.path.join(args.relabel, "sched_mlp.json")) as f:
            info = json.load(f)["images"]
        X, Y, D, P = d["X"], d["Y"].copy(), d["D"], d["P"]
        pos = 0
        for im in info:
            n = im["tiles"]
            tr = {"final_mean": np.zeros(n), "maxval": 1.0, "dirty": D[pos:pos + n],
                  "pixels": P[pos:pos + n]}
            Y[pos:pos + n, 0] = labels(tr, im["ms_by_rounds"])[:, 0]
            im["k_hist"] = np.bincount(Y[pos:pos + n, 0].astype(int),
                                       minlength=MAX_ROUNDS + 1).tolist()
            pos += n
        assert pos == X.shape[0], "dataset and timings disagree"
        log(f"relabelled {len(info)} images from {args.relabel}")
    else:
        cases = training_cases({"c4": args.c4})
        log(f"{len(cases)} training images generated in {time.time() - t0:.1f} s")
        X, Y, info, D, P = build_dataset(cases, log=lambda m: None if " c4_" in m else log(m))
    log(f"dataset: {X.shape[0]} tiles, k* histogram "
        f"{np.bincount(Y[:, 0].astype(int), minlength=MAX_ROUNDS + 1).tolist()}")
    candidates = []
    for seed in range(max(1, args.inits)):
        w, hist = train(X, Y, steps=args.steps, lr=args.lr, seed=seed)
        log(f"init {seed}: loss {hist[0][1]:.5f} -> {hist[-1][1]:.5f}")
        candidates.append((seed, w, hist))
    chosen, val = 0, None
    if len(candidates) > 1:
        vcases = validation_cases()
        scores = [validation_ms(w, vcases) for _, w, _ in candidates]
        for (seed, _, _), v in zip(candidates, scores):
            log(f"init {seed}: held-out reconstruction time {v:.3f} ms")
        chosen = int(np.argmin(scores))
        val = [round(v, 4) for v in scores]
    seed, w, hist = candidates[chosen]
    for s, L in hist[:: max(1, len(hist) // 10)] + hist[-1:]:
        log(f"step {s}: loss {L:.5f}")
    meta = {"params": MLP_PARAMS, "features": MLP_FEATURES, "feature_sub": FEATURE_SUB,
            "feature_set": FEATURE_SET,
            "max_rounds": MAX_ROUNDS, "samples": int(X.shape[0]), "images": info,
            "steps": args.steps, "lr": args.lr, "loss_first": hist[0][1],
            "loss_last": hist[-1][1],
            "label_k_histogram": np.bincount(Y[:, 0].astype(int),
                                             minlength=MAX_ROUNDS + 1).tolist(),
            "init_seed": seed, "validation_ms_per_init": val}
    save_model(w, meta, os.path.join(args.out, "sched_mlp.f32"))
    np.savez_compressed(os.path.join(args.out, "sched_dataset.npz"), X=X, Y=Y, D=D, P=P)
    log(f"model written to {args.out} ({time.time() - t0:.1f} s)")


if __name__ == "__main__":
    main()
