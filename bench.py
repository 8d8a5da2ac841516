#!/usr/bin/env python3
"""Throughput of bit-exact IWPP reconstruction by dilation on one MI355X.

A step = one full reconstruction of the workload's marker under its mask
(marker reset from a resident copy, then iwpp_reconstruct_ex), inputs already
in HBM.  value = image pixels / step time, in Mpixel/s.  The default workload
is the 32768^2 learned-schedule configuration: the step is timed with the
learned schedule (value) and with the plain heuristic schedule (reported
beside it).  The outputs of the timed steps are compared bit for bit with the
CPU oracle.

Prints ONE JSON line on stdout; progress goes to stderr.
    python bench.py [--gpus 1] [--steps K] [--warmup W] [--config c5]
                    [--schedule both|learned|plain] [--queue auto]
"""
import argparse
import fnmatch
import json
import os
import statistics
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
for _p in (ROOT, os.path.join(ROOT, "iwpp-mp_amd")):
    if _p not in sys.path:
        sys.path.insert(0, _p)

METRIC = "Mpixel/s bit-exact reconstruction, learned-schedule vs plain IWPP, 1 MI355X"
HBM_PEAK_GBS = 8000.0


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def profile_traffic(workload, patterns):
    """HBM bytes per launch of a kernel family from the committed rocprofv3
    summary of this workload and queue (profiles/<workload>_<queue>_kernels.md,
    or profiles/<workload>_kernels.md for the batch: FETCH_SIZE and
    WRITE_SIZE passes, calibrated to bytes there), averaged over the rows whose
    kernel name matches one of `patterns` (shell-style), weighted by their
    dispatches, from each row's median dispatch.  (None, reason) when there is
    no summary."""
    path = os.path.join(ROOT, "profiles", f"{workload}_kernels.md")
    if not os.path.exists(path):
        return None, f"no {os.path.relpath(path, ROOT)}"
    col = 7
    num = den = 0.0
    table = False  # inside the traffic table (the first one whose header names FETCH_SIZE)
    for line in open(path):
        cells = [c.strip() for c in line.strip().strip("|").split("|")]
        if not table:
            table = cells[0] == "kernel" and any("FETCH_SIZE" in c for c in cells)
            continue
        if not line.strip():
            break
        if len(cells) <= col or not cells[0].startswith("`"):
            continue
        name = cells[0].strip("`")
        if not any(fnmatch.fnmatchcase(name, p) for p in patterns) or cells[col] == "-":
            continue
        calls = float(cells[1])
        num += calls * float(cells[col]) * 1e6
        den += calls
    if den == 0:
        return None, f"no {' | '.join(patterns)} rows in {os.path.relpath(path, ROOT)}"
    return num / den, (f"{os.path.relpath(path, ROOT)}: calibrated FETCH_SIZE + WRITE_SIZE of the "
                       f"median dispatch of each kernel, weighted by dispatches over the kernels {' | '.join(patterns)}")


def cpu_model():
    """The host CPU's model name (as lscpu reports it), or None."""
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.lower().startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return None


def seeding_bytes(shapes, tsize, fused, block_rows, seeded):
    """Algorithmic bytes of the seeding phase over images of `shapes` (h, w).
    Without sweeps k_unstable_scan reads every pixel's marker and mask.  When
    the last sweep pass has set the megapixel flags (fused), k_seed_edges reads
    only what that pass could not test: the 4-byte border words of both sides of
    each inner strip edge (strips of 1024 bytes: 64 lanes of 16 bytes) for every
    row, and marker and mask of the two rows beside each block boundary.  Both
    then write one flag byte per megapixel, and k_seed_queue reads the flags and
    writes an 8-byte ring entry per seeded megapixel."""
    total = 0
    for h, w in shapes:
        nmp = -(-h // 64) * -(-w // 64)
        if fused:
            strips = -(-w * tsize // 1024)
            bounds = max(0, -(-h // block_rows) - 1)
            total += 8 * h * max(0, strips - 1) + bounds * 2 * 2 * tsize * w
        else:
            total += 2 * tsize * h * w
        total += 2 * nmp
    return total + 8 * seeded


def seam_bytes(shapes, tsize, block_rows, seam_rows):
    """Algorithmic bytes of one raster fix-up launch over images of `shapes`
    (h, w): 3*sizeof(T) per pixel of the windows beside the block boundaries
    (min(seam_rows, block_rows / 2) rows each; the downward windows, after a
    boundary, end at the image's last row)."""
    k = min(seam_rows, block_rows // 2)
    total = 0
    for h, w in shapes:
        rows = sum(min(k, h - yb) for yb in range(block_rows, h, block_rows))
        total += 3 * tsize * w * rows
    return total


def single_kernel_roofline(candidates, workload):
    """Roofline of the single kernel with the largest average launch time.
    candidates: name -> (total ms per reconstruction, launches, algorithmic
    bytes per launch, profile name patterns).  Returns (roofline dict, table)."""
    table = {}
    for name, (tot, launches, alg, patterns) in candidates.items():
        avg = tot / launches if launches else 0.0
        ach = alg / (avg * 1e-3) / 1e9 if avg > 0 else None
        table[name] = {"ms_per_reconstruction": tot, "launches": launches, "avg_launch_ms": avg,
                       "algorithmic_bytes_per_launch": alg, "achieved_GBps": ach,
                       "frac": ach / HBM_PEAK_GBS if ach else None, "patterns": patterns}
    dom = max(table, key=lambda k: table[k]["avg_launch_ms"])
    t = table[dom]
    traffic, traffic_src = profile_traffic(workload, t["patterns"])
    roofline = {
        "bound": "hbm",
        "kernel": dom,
        "achieved": t["achieved_GBps"],
        "peak": HBM_PEAK_GBS,
        "unit": "GB/s",
        "frac": t["frac"],
        "traffic": traffic,
        "algorithmic_bytes_per_launch": t["algorithmic_bytes_per_launch"],
        "avg_launch_ms": t["avg_launch_ms"],
        "basis": "the single kernel with the largest average launch time inside the timed steps; "
                 "algorithmic bytes per launch / that average (the library's HIP events on its "
                 "stream, around each pass)",
        "traffic_basis": traffic_src,
    }
    for v in table.values():
        v.pop("patterns")
    return roofline, table


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--config", default="c5", choices=["c1", "c2", "c3", "c4", "c5"])
    ap.add_argument("--queue", default="auto", choices=["mutex", "dq", "uq", "auto"])
    ap.add_argument("--schedule", default="both", choices=["both", "plain", "learned"],
                    help="both: value is the learned schedule, the plain one is timed beside it")
    ap.add_argument("--sweeps", type=int, default=-1, help="sweep rounds (-1: schedule)")
    ap.add_argument("--split-sweeps", action="store_true",
                    help="sweep rounds of a row pass and two column passes (default: two raster "
                         "passes)")
    ap.add_argument("--cpu-runs", type=int, default=5)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--kernel-iters", type=int, default=5)
    ap.add_argument("--uniform-steps", type=int, default=5,
                    help="timed steps of the plain schedule at each uniform sweep-round count "
                         "0..3 (0: skip)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (got --gpus {args.gpus}); the reconstruction "
              "path is single-GPU and no multi-GPU figure is produced.", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch

    import iwpp
    from iwpp import workloads as W

    if not torch.cuda.is_available():
        log("no HIP device")
        sys.exit(1)
    torch.cuda.set_device(0)
    if args.config == "c4":
        return bench_batch(args)

    t = time.time()
    case = W.CONFIGS[args.config]()
    log(f"generated {case.name} {case.mask.shape} {case.mask.dtype} in {time.time() - t:.1f}s")
    px = case.pixels
    tsize = case.mask.itemsize
    mask = iwpp.DeviceImage.from_numpy(case.mask)
    marker0 = iwpp.DeviceImage.from_numpy(case.marker)
    work = iwpp.DeviceImage(marker0.h, marker0.w, marker0.dtype, marker0.pitch)
    ws = iwpp.Workspace(px, marker0.dtype)
    split = 1 if args.split_sweeps else 0
    opt = iwpp.default_options(sweep_rounds=args.sweeps, split_sweeps=split)

    from iwpp import learned

    def run_schedule(schedule, o=opt, steps=args.steps, warmup=args.warmup, label=None):
        label = label or schedule
        if schedule == "learned":
            learned.install()
        else:
            iwpp.set_model(None)

        def step():
            return iwpp.reconstruct(work, mask, case.connectivity, case.tile, args.queue,
                                    schedule, o, ws)

        for i in range(warmup):
            work.buf.copy_(marker0.buf)
            st = step()
            log(f"{label} warmup {i}: {st}")
        # Each step reconstructs a fresh copy of the marker (the reconstruction
        # is in place); the copy is made before the step's timed bracket.
        stats = []
        tot = 0.0
        for _ in range(steps):
            work.buf.copy_(marker0.buf)
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            stats.append(step())
            torch.cuda.synchronize()
            tot += time.perf_counter() - t0
        ms = tot * 1e3 / steps
        log(f"{label}: timed {steps} steps: {ms:.3f} ms/step, "
            f"{px / (ms * 1e-3) / 1e6:.1f} Mpixel/s")
        return ms, stats, work.numpy()

    runs = {}
    runs_out_uniform = {}
    for sched in (("plain", "learned") if args.schedule == "both" else (args.schedule,)):
        runs[sched] = run_schedule(sched)
    main_sched = "learned" if "learned" in runs else "plain"
    ms_step, stats, result = runs[main_sched]
    value = px / (ms_step * 1e-3) / 1e6
    # The plain schedule with every uniform number of sweep rounds: the best
    # of these in hindsight is what a schedule has to beat.
    uniform = {}
    if args.uniform_steps > 0 and args.sweeps < 0:
        for k in range(0, 4):
            o = iwpp.default_options(sweep_rounds=k, split_sweeps=split)
            ms_k, st_k, out_k = run_schedule("plain", o, args.uniform_steps, 1,
                                             f"plain, {k} rounds")
            uniform[k] = {"ms_per_step": ms_k, "tile_visits": st_k[-1]["tile_visits"],
                          "ms_sweep": statistics.mean(x["ms_sweep"] for x in st_k),
                          "ms_queue": statistics.mean(x["ms_queue"] for x in st_k)}
            runs_out_uniform[k] = out_k
    iwpp.set_model(None)

    # --- per-kernel device times (HIP events on the stream the kernels run on)
    stream = torch.cuda.current_stream()
    ev = [torch.cuda.Event(enable_timing=True) for _ in range(2)]

    def timed(fn, reset):
        ts = []
        for _ in range(args.kernel_iters):
            reset()
            ev[0].record(stream)
            fn()
            ev[1].record(stream)
            torch.cuda.synchronize()
            ts.append(ev[0].elapsed_time(ev[1]))
        return statistics.median(ts)

    scratch = iwpp.DeviceImage(marker0.h, marker0.w, marker0.dtype, marker0.pitch)

    def reset_marker():
        work.buf.copy_(marker0.buf)
        iwpp.clip(work, mask)

    ms_row = timed(lambda: iwpp.sweep_rows(work, mask), reset_marker)
    ms_col = timed(lambda: iwpp.sweep_cols(work, scratch, mask, case.connectivity, 1,
                                           opt.col_block_rows), reset_marker)
    ms_raster = timed(lambda: iwpp.sweep_raster(work, scratch, mask, case.connectivity, 1,
                                                opt.col_block_rows), reset_marker)
    bytes_sweep = 3 * tsize * px  # algorithmic bytes of one directional pass
    phases = {k: statistics.mean(s[k] for s in stats) for k in ("ms_sweep", "ms_seed", "ms_queue")}
    counters = {k: stats[-1][k] for k in ("tile_visits", "seeded", "sweep_rounds", "queue_rounds",
                                          "restarts", "queue", "tile", "col_block_rows",
                                          "seam_rows")}
    used_queue = {0: "mutex", 1: "dq", 2: "uq"}[stats[-1]["queue"]]
    rounds = counters["sweep_rounds"]
    visits = statistics.mean(s["tile_visits"] for s in stats)
    # Per reconstruction: device time of each kernel family inside the timed
    # steps (the library's HIP events on its own stream), its launches, and
    # the algorithmic bytes one launch must move: 3*sizeof(T) per pixel a
    # sweep pass owns (read marker, read mask, write marker; raster, row and
    # column passes alike), 3*sizeof(T) per pixel of each megapixel visit
    # (64*64 pixels) in the queue phase, and for the seeding phase what
    # seeding_bytes() counts.  The rocprofv3 names of the in-place passes end in their
    # template mode: raster 5 (6: clipping, 7: seeding), column 2 (3: seeding).
    queue_name = "k_queue_" + ("mutex" if used_queue == "mutex" else f"bsp<{used_queue.upper()}>")
    ms_rows = statistics.mean(s_["ms_rows"] for s_ in stats)
    ms_cols = statistics.mean(s_["ms_cols"] for s_ in stats)
    ms_seams = statistics.mean(s_["ms_seams"] for s_ in stats)
    if split:
        sweeps = {
            "k_row_pass (forward + backward, one launch per round)":
                (ms_rows, rounds, bytes_sweep, ("k_row_pass*",)),
            "k_col_pass (in place, down and up: two launches per round)":
                (ms_cols, 2 * rounds, bytes_sweep, ("k_col_pass<*, [23]>",)),
        }
    else:
        # the fix-ups run as k_col_pass modes 5 (down) and 7 or 5 (up): their
        # profile rows are shared with the passes of those modes
        seam_launch = seam_bytes([case.marker.shape], tsize, stats[-1]["col_block_rows"],
                                 stats[-1]["seam_rows"])
        # (with one round the downward pass is the clipping one, mode 6 alone;
        # the upward fix-up of the last round shares mode 7 with the upward pass)
        down_pat = "k_col_pass<*, 1, 6>" if rounds == 1 else "k_col_pass<*, 1, [56]>"
        sweeps = {
            "k_col_pass raster, downward (in place, clipping in round 0; one launch per round)":
                (ms_rows, rounds, bytes_sweep, (down_pat,)),
            "k_col_pass raster fix-ups beside the block boundaries (down and up: two launches "
            "per round)":
                (ms_seams, 2 * rounds if stats[-1]["seam_rows"] > 0 else 0, seam_launch,
                 ("k_col_pass<*, 1, 5>",)),
            "k_col_pass raster, upward (in place, seeding flags in the last round; one launch "
            "per round)":
                (ms_cols, rounds, bytes_sweep, ("k_col_pass<*, -1, [57]>",)),
        }
    candidates = {
        **sweeps,
        queue_name: (phases["ms_queue"], 1, visits * 64 * 64 * 3 * tsize, ("k_queue_*",)),
        "seeding (k_seed_edges | k_unstable_scan, k_seed_queue)":
            (phases["ms_seed"], 1,
             seeding_bytes([case.marker.shape], tsize, rounds > 0, stats[-1]["col_block_rows"],
                           counters["seeded"]),
             ("k_seed_*", "k_unstable_scan*")),
    }
    roofline, per_kernel = single_kernel_roofline(candidates, f"{case.name}_{used_queue}")
    passes = (3 if split else 2) * rounds
    sweep_avg = (phases["ms_sweep"] - (0 if split else ms_seams)) / max(1, passes)
    roofline["sweep_family"] = {
        "kernels": ("k_row_pass*, k_col_pass (every pass of the sweep phase)" if split else
                    "k_col_pass raster (every full pass of the sweep phase; fix-ups excluded)"),
        "avg_launch_ms": sweep_avg,
        "achieved": bytes_sweep / (sweep_avg * 1e-3) / 1e9 if sweep_avg > 0 else None,
        "frac": bytes_sweep / (sweep_avg * 1e-3) / 1e9 / HBM_PEAK_GBS if sweep_avg > 0 else None,
    }
    roofline["end_to_end"] = {
        "basis": "3*sizeof(T)*pixels (every pixel's marker and mask read once and marker "
                 "written once) / step time",
        "achieved": 3 * tsize * px / (ms_step * 1e-3) / 1e9,
        "frac": 3 * tsize * px / (ms_step * 1e-3) / 1e9 / HBM_PEAK_GBS,
    }
    sweep_kernels = {
        "row_pass_ms": ms_row,
        "row_pass_GBps": bytes_sweep / (ms_row * 1e-3) / 1e9,
        "row_pass_frac": bytes_sweep / (ms_row * 1e-3) / 1e9 / HBM_PEAK_GBS,
        "col_pass_ms": ms_col,
        "col_pass_GBps": bytes_sweep / (ms_col * 1e-3) / 1e9,
        "col_pass_frac": bytes_sweep / (ms_col * 1e-3) / 1e9 / HBM_PEAK_GBS,
        "raster_pass_ms": ms_raster,
        "raster_pass_GBps": bytes_sweep / (ms_raster * 1e-3) / 1e9,
        "raster_pass_frac": bytes_sweep / (ms_raster * 1e-3) / 1e9 / HBM_PEAK_GBS,
        "per_kernel": per_kernel,
        "basis": "row_pass / col_pass / raster_pass: one isolated launch each over the clipped "
                 "marker (the column and raster passes out of place), HIP events",
    }

    # --- where the queue kernel's waves spend their time: one more (untimed)
    # reconstruction with the in-kernel phase clocks on (summed over waves)
    if main_sched == "learned":
        learned.install()
    work.buf.copy_(marker0.buf)
    st_tr = iwpp.reconstruct(work, mask, case.connectivity, case.tile, args.queue, main_sched,
                             iwpp.default_options(sweep_rounds=args.sweeps, trace_phases=1,
                                                  split_sweeps=split), ws)
    iwpp.set_model(None)
    keys = ("us_wait", "us_fetch", "us_converge", "us_store", "us_notify")
    tot_us = sum(st_tr[k] for k in keys)
    queue_split = {
        "basis": "queue kernel, one reconstruction with trace_phases: wave-microseconds per phase "
                 "(wait = taking work, fetch = loading the tile, converge = local sweeps, store = "
                 "write-back and its completion, notify = neighbour atomics, release and the next "
                 "tile's loads), summed over waves",
        **{k[3:]: st_tr[k] for k in keys},
        "fractions": {k[3:]: (st_tr[k] / tot_us if tot_us > 0 else None) for k in keys},
    }

    # --- CPU baseline: the oracle (a port of the sequential hybrid), one core
    cpu = None
    exact = None
    if not args.no_cpu:
        import oracle
        times = []
        want = None
        for _ in range(max(1, args.cpu_runs)):
            t = time.perf_counter()
            want = oracle.reconstruct(case.marker, case.mask, case.connectivity)
            times.append(time.perf_counter() - t)
            if sum(times) > 20.0:  # bounded sample: about 20 s of CPU work at most
                break
        med = statistics.median(times)
        cpu = {"value": px / med / 1e6, "unit": "Mpixel/s", "cores": 1, "kind": "port",
               "sample": f"full {case.name} image, median of {len(times)} runs "
                         f"({med * 1e3:.1f} ms each), oracle/iwpp_oracle.c -O3",
               "host_cpus": os.cpu_count(), "cpu_model": cpu_model()}
        exact = all(bool(np.array_equal(r[2], want)) for r in runs.values()) and \
            all(bool(np.array_equal(o, want)) for o in runs_out_uniform.values())
        log(f"cpu oracle {med * 1e3:.1f} ms; GPU outputs bit-exact: {exact}")

    out = {
        "metric": METRIC,
        "value": value,
        "unit": "Mpixel/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u8" if tsize == 1 else "u16",
        "data": "synthetic",
        "config": {"workload": case.name, "pixels": px, "connectivity": case.connectivity,
                   "tile": case.tile, "queue": args.queue, "queue_used": used_queue,
                   "schedule": main_sched, "sweeps": "split" if split else "raster",
                   **case.meta},
        "roofline": roofline,
        "cpu_baseline": cpu,
        "exact": exact,
        "phases_ms": phases,
        "queue_phase_split": queue_split,
        "counters": counters,
        "sweep_kernels": sweep_kernels,
        "schedules": {k: {"ms_per_step": v[0], "mpixel_s": px / (v[0] * 1e-3) / 1e6,
                          "sweep_rounds": v[1][-1]["sweep_rounds"],
                          "tiles_skipped": v[1][-1]["sweeps_skipped_tiles"],
                          "schedule_used": {0: "plain", 1: "learned"}[v[1][-1]["schedule"]]}
                      for k, v in runs.items()},
    }
    if len(runs) == 2:
        out["learned_vs_plain"] = runs["plain"][0] / runs["learned"][0]
    if uniform:
        best_k = min(uniform, key=lambda k: uniform[k]["ms_per_step"])
        out["uniform_rounds"] = uniform
        out["best_uniform_rounds"] = best_k
        out[f"{main_sched}_vs_best_uniform"] = uniform[best_k]["ms_per_step"] / ms_step
    ws.close()
    print(json.dumps(out), flush=True)


def _pool_images(arrays, torch, iwpp):
    """DeviceImages carved out of one device buffer (so one copy resets all)."""
    offs, total = [], 0
    for a in arrays:
        pitch = (a.shape[1] * a.itemsize + 255) // 256 * 256
        offs.append((total, pitch))
        total += pitch * a.shape[0]
    pool = torch.zeros(total, dtype=torch.uint8, device="cuda")
    host = torch.zeros(total, dtype=torch.uint8)
    imgs = []
    dt = iwpp.U8 if arrays[0].dtype.itemsize == 1 else iwpp.U16
    for a, (off, pitch) in zip(arrays, offs):
        h, w = a.shape
        view = host[off:off + h * pitch].view(h, pitch)
        view[:, :w * a.itemsize] = torch.from_numpy(a.view("uint8").reshape(h, w * a.itemsize))
        imgs.append(iwpp.DeviceImage(h, w, dt, pitch, buf=pool[off:off + h * pitch].view(h, pitch)))
    pool.copy_(host.to("cuda"))
    return pool, imgs


def bench_batch(args):
    """c4: 256 images reconstructed (a) one call per image, (b) in one batched
    call, (c) by replaying a captured hipGraph of the batched call."""
    import numpy as np
    import torch

    import iwpp
    from iwpp import workloads as W

    t = time.time()
    cases = W.c4()
    px = sum(c.pixels for c in cases)
    log(f"generated {len(cases)} images, {px} px in {time.time() - t:.1f}s")
    pool0, mk0 = _pool_images([c.marker for c in cases], torch, iwpp)
    poolw, mkw = _pool_images([c.marker for c in cases], torch, iwpp)
    _, msk = _pool_images([c.mask for c in cases], torch, iwpp)
    pairs = iwpp.Batch(zip(mkw, msk))  # descriptor array built once for every call below
    ws = iwpp.Workspace()

    def reset():
        poolw.copy_(pool0)

    def timed(fn, steps, warmup):
        for _ in range(warmup):
            reset()
            fn()
        torch.cuda.synchronize()
        tot = 0.0
        for _ in range(steps):
            reset()
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            fn()
            torch.cuda.synchronize()
            tot += time.perf_counter() - t0
        return tot * 1e3 / steps

    def loop():
        for mk, ms in pairs:
            iwpp.reconstruct(mk, ms, 8, 0, "auto", "plain", None, ws)

    ms_loop = timed(loop, max(1, args.steps // 2), 1)
    loop_out = [m.numpy() for m in mkw]
    log(f"per-image loop: {ms_loop:.2f} ms, {px / ms_loop / 1e3:.1f} Mpixel/s")
    reset()
    tuning = iwpp.autotune(pairs, 8, 2, None, ws)
    log(f"autotuned: {tuning}")
    stats = []
    ms_batch = timed(lambda: stats.append(iwpp.reconstruct_batch(pairs, 8, 0, "auto", "plain",
                                                                 None, ws)), args.steps,
                     args.warmup)
    batch_st = stats[-1]
    batch_same = all(np.array_equal(m.numpy(), o) for m, o in zip(mkw, loop_out))
    log(f"batched: {ms_batch:.2f} ms, {px / ms_batch / 1e3:.1f} Mpixel/s, same as loop: {batch_same}")
    from iwpp import learned
    learned.install()
    stats_l = []
    ms_learned = timed(lambda: stats_l.append(iwpp.reconstruct_batch(pairs, 8, 0, "auto",
                                                                     "learned", None, ws)),
                       args.steps, args.warmup)
    learned_st = stats_l[-1]
    learned_same = all(np.array_equal(m.numpy(), o) for m, o in zip(mkw, loop_out))
    iwpp.set_model(None)
    log(f"batched, learned schedule: {ms_learned:.2f} ms, {px / ms_learned / 1e3:.1f} Mpixel/s, "
        f"same as loop: {learned_same}")
    gws = iwpp.Workspace()
    reset()
    torch.cuda.synchronize()
    g = iwpp.BatchGraph(pairs, gws, 8, options=iwpp.default_options(
        sweep_rounds=tuning["sweep_rounds"]))
    ms_graph = timed(lambda: (g.launch(), g.stream.synchronize()), args.steps, args.warmup)
    graph_same = all(np.array_equal(m.numpy(), o) for m, o in zip(mkw, loop_out))
    graph_st = g.stats()
    log(f"graph: {ms_graph:.2f} ms, {px / ms_graph / 1e3:.1f} Mpixel/s, same as loop: {graph_same}")
    forms = {"graph": ms_graph, "batched call": ms_batch, "batched call, learned": ms_learned}
    form = min(forms, key=forms.get)
    best_ms = forms[form]

    cpu, exact = None, None
    if not args.no_cpu:
        import oracle
        t_cpu, px_cpu, ok = 0.0, 0, True
        for c, o in zip(cases, loop_out):
            t0 = time.perf_counter()
            want = oracle.reconstruct(c.marker, c.mask, 8)
            t_cpu += time.perf_counter() - t0
            px_cpu += c.pixels
            ok = ok and np.array_equal(o, want)
            if t_cpu > 20.0:
                break
        cpu = {"value": px_cpu / t_cpu / 1e6, "unit": "Mpixel/s", "cores": 1, "kind": "port",
               "sample": f"first {px_cpu} px of the c4 images in order, "
                         f"{t_cpu:.1f} s, oracle/iwpp_oracle.c -O3", "host_cpus": os.cpu_count(),
               "cpu_model": cpu_model()}
        exact = bool(ok and batch_same and graph_same and learned_same)
        log(f"oracle on the sample: {ok}")
    phases = {k: batch_st[k] for k in ("ms_sweep", "ms_seed", "ms_queue")}
    rounds = max(1, batch_st["sweep_rounds"])
    tsize = cases[0].mask.itemsize
    candidates = {
        "k_col_pass raster, downward (in place, clipping in round 0; one launch per round)":
            (batch_st["ms_rows"], rounds, 3 * tsize * px,
             ("k_col_pass<*, 1, 6>" if rounds == 1 else "k_col_pass<*, 1, [56]>",)),
        "k_col_pass raster, upward (in place, seeding flags in the last round; one launch "
        "per round)":
            (batch_st["ms_cols"], rounds, 3 * tsize * px, ("k_col_pass<*, -1, [57]>",)),
        "k_col_pass raster fix-ups beside the block boundaries (down and up: two launches "
        "per round)":
            (batch_st["ms_seams"], 2 * rounds if batch_st["seam_rows"] > 0 else 0,
             seam_bytes([c.mask.shape for c in cases], tsize, batch_st["col_block_rows"],
                        batch_st["seam_rows"]), ("k_col_pass<*, 1, 5>",)),
        "k_queue_" + ("mutex" if batch_st["queue"] == 0 else "bsp"):
            (phases["ms_queue"], 1, batch_st["tile_visits"] * 64 * 64 * 3 * tsize,
             ("k_queue_mutex<*, false>" if batch_st["queue"] == 0 else "k_queue_bsp<*, false>",)),
        "seeding (k_seed_edges | k_unstable_scan, k_seed_queue)":
            (phases["ms_seed"], 1,
             seeding_bytes([c.mask.shape for c in cases], tsize, batch_st["sweep_rounds"] > 0,
                           batch_st["col_block_rows"], batch_st["seeded"]),
             ("k_seed_*", "k_unstable_scan*")),
    }
    # Traffic from the median dispatch: the profile's dispatches of the batched
    # queue kernel (template flag false: several images) are the batched
    # calls' and the autotuner's trials; the median is a batched call's.
    roofline, per_kernel = single_kernel_roofline(candidates, "c4_batch_256_u8")
    roofline["end_to_end"] = {
        "basis": "3*sizeof(T)*pixels / time of the fastest batched form",
        "achieved": 3 * tsize * px / (best_ms * 1e-3) / 1e9,
        "frac": 3 * tsize * px / (best_ms * 1e-3) / 1e9 / HBM_PEAK_GBS,
    }
    out = {
        "metric": METRIC,
        "value": px / best_ms / 1e3,
        "unit": "Mpixel/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": best_ms,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u8",
        "data": "synthetic",
        "config": {"workload": "c4_batch_256_u8", "images": len(cases), "pixels": px,
                   "connectivity": 8, "form": form, "tuning": tuning},
        "roofline": roofline,
        "per_kernel": per_kernel,
        "cpu_baseline": cpu,
        "exact": exact,
        "per_image_loop": {"ms": ms_loop, "mpixel_s": px / ms_loop / 1e3},
        "batched_call": {"ms": ms_batch, "mpixel_s": px / ms_batch / 1e3, "counters": batch_st},
        "graph_replay": {"ms": ms_graph, "mpixel_s": px / ms_graph / 1e3,
                         "counters": graph_st},
        "batched_learned": {"ms": ms_learned, "mpixel_s": px / ms_learned / 1e3,
                            "counters": learned_st, "same_as_loop": learned_same},
        "speedup_vs_loop": ms_loop / best_ms,
        "learned_batch_vs_loop": ms_loop / ms_learned,
        "learned_batch_vs_plain_batch": ms_batch / ms_learned,
    }
    # release the graph before its workspace, both before the interpreter exits
    g.close()
    gws.close()
    ws.close()
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
