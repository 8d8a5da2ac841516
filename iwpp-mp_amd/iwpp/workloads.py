"""Synthetic workloads: the five benchmark configurations and scaled-down twins.

Seeds: base S0 = 1312 (the seed of the reference's generator,
IWPP MP/src/testCaseGenerator.cu:52-55); configuration k uses S0 + k.
Every generator is integer-only and deterministic (see include/iwpp_synth.h).
"""
from dataclasses import dataclass, field

import numpy as np

from . import (U8, U16, synth_batch_shapes, synth_mask, synth_marker_hdome,
               synth_marker_seeds)

S0 = 1312


@dataclass
class Case:
    name: str
    mask: np.ndarray
    marker: np.ndarray
    connectivity: int
    tile: int
    queues: tuple
    meta: dict = field(default_factory=dict)

    @property
    def pixels(self):
        return int(self.mask.size)


def seeds_count(n_pixels, per_million):
    return (n_pixels * per_million) // 1_000_000


def c1(side=1024, seed=S0 + 1):
    """1024^2 uint8, blurred noise (sigma 4), h-dome h=30, 8-conn, one tile, mutex."""
    mask = synth_mask(side, side, U8, 255, 4.0, 0, 0, 0, seed)
    marker = synth_marker_hdome(mask, 30)
    return Case("c1_hdome_1024_u8", mask, marker, 8, max(64, side // 64 * 64), ("mutex",),
                {"seed": seed, "sigma": 4, "h": 30})


def c2(side=4096, seed=S0 + 2):
    """4096^2 uint8, sigma 8, marker = 0.1% random seeds, 8-conn, 1024 tiles, mutex/DQ/UQ."""
    mask = synth_mask(side, side, U8, 255, 8.0, 0, 0, 0, seed)
    marker = synth_marker_seeds(mask, seeds_count(mask.size, 1000), seed)
    return Case("c2_seeds_4096_u8", mask, marker, 8, 1024, ("mutex", "dq", "uq"),
                {"seed": seed, "sigma": 8, "seed_ppm": 1000})


def c3(side=16384, seed=S0 + 3, ndisks=200):
    """16384^2 uint16 (12-bit), sigma 16 + 200 disks r in [50,400], h-dome h=50,
    4-conn, 2048 tiles, UQ."""
    mask = synth_mask(side, side, U16, 4095, 16.0, ndisks, 50, 400, seed)
    marker = synth_marker_hdome(mask, 50)
    return Case("c3_hdome_16384_u16", mask, marker, 4, 2048, ("uq",),
                {"seed": seed, "sigma": 16, "disks": ndisks, "h": 50})


def c4(n=256, seed=S0 + 4, max_side=None):
    """256 uint8 images, sides log-uniform in [256, 4096], aspect 1:1..1:4; half
    h-dome (h in [10,60]), half 0.1% seeds; 8-conn.  Mask sigma 8."""
    ws, hs, kinds, hvals = synth_batch_shapes(n, seed)
    cases = []
    for i in range(n):
        w, h = int(ws[i]), int(hs[i])
        if max_side:
            w, h = max(1, min(w, max_side)), max(1, min(h, max_side))
        s = seed * 1000 + i
        mask = synth_mask(h, w, U8, 255, 8.0, 0, 0, 0, s)
        if kinds[i] == 0:
            marker = synth_marker_hdome(mask, int(hvals[i]))
        else:
            marker = synth_marker_seeds(mask, seeds_count(mask.size, 1000), s)
        cases.append(Case(f"c4_{i:03d}_{w}x{h}", mask, marker, 8, 0, ("auto",),
                          {"seed": s, "kind": "hdome" if kinds[i] == 0 else "seeds",
                           "h": int(hvals[i])}))
    return cases


def c5(side=32768, seed=S0 + 5, ndisks=1000):
    """32768^2 uint8, sigma 32 + 1000 disks r in [50,400], 0.01% seeds, 8-conn."""
    mask = synth_mask(side, side, U8, 255, 32.0, ndisks, 50, 400, seed)
    marker = synth_marker_seeds(mask, seeds_count(mask.size, 100), seed)
    return Case("c5_seeds_32768_u8", mask, marker, 8, 2048, ("uq",),
                {"seed": seed, "sigma": 32, "disks": ndisks, "seed_ppm": 100})


CONFIGS = {"c1": c1, "c2": c2, "c3": c3, "c4": c4, "c5": c5}
