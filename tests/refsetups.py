"""Inputs of the reference's own collab tests, regenerated from the formulas in their source.

Each function restates a generator from the reference's test helpers (cited per function) in
numpy with the same u32 wrapping and f32 rounding, so the digests the reference recorded for
those inputs can be checked here.  These are data generators, not code from the reference.
"""
from __future__ import annotations

import numpy as np

U32 = np.uint64(0xFFFFFFFF)


def _hash(idx: np.ndarray) -> np.ndarray:
    """idx*2654435761 + 0x9E3779B9, ^>>15, *0x85EBCA6B, ^>>13 in wrapping u32."""
    h = (idx.astype(np.uint64) * np.uint64(2654435761) + np.uint64(0x9E3779B9)) & U32
    h ^= h >> np.uint64(15)
    h = (h * np.uint64(0x85EBCA6B)) & U32
    h ^= h >> np.uint64(13)
    return h


def _unit(h: np.ndarray) -> np.ndarray:
    return h.astype(np.float32) / np.float32(4294967295.0)


def make_unique_frame(width: int, height: int) -> np.ndarray:
    """collab/tests/helpers.rs:36-52: x*10 plus a per-pixel hash offset times 10."""
    y, x = np.mgrid[0:height, 0:width]
    idx = (y * width + x).astype(np.uint64)
    off = _unit(_hash(idx))
    return (x.astype(np.float32) * np.float32(10.0) + off * np.float32(10.0)).astype(np.float32)


def unique_frame(width: int, height: int) -> np.ndarray:
    """collab/tests/fused/mod.rs:64-68: make_unique_frame over its peak."""
    raw = make_unique_frame(width, height)
    return (raw / raw.max()).astype(np.float32)


def noisy_flat_field(width: int, height: int, base: float, sigma: float) -> np.ndarray:
    """collab/tests/helpers.rs:13-30: base plus rescaled Irwin-Hall noise of four hash uniforms."""
    unit_std = np.sqrt(np.float32(1.0) / np.float32(3.0)).astype(np.float32)
    idx = np.arange(width * height, dtype=np.uint64)
    acc = np.zeros(width * height, np.float32)
    for k in range(4):
        acc = (acc + (_unit(_hash(idx * np.uint64(4) + np.uint64(k))) - np.float32(0.5))).astype(np.float32)
    out = np.float32(base) + (acc / unit_std) * np.float32(sigma)
    return out.astype(np.float32).reshape(height, width)


def cross_frame_fields(width: int, height: int, radius: int, blk_step: int = 8):
    """collab/tests/fused/mod.rs:586-641: per-block vectors and confidences straddling c_min."""
    bx = (width + blk_step - 1) // blk_step
    by = (height + blk_step - 1) // blk_step
    blocks = bx * by
    mv = np.zeros(2 * radius * blocks * 2, np.int32)
    conf = np.zeros(2 * radius * blocks, np.float32)
    for t in range(2 * radius):
        for b in range(blocks):
            mv[(t * blocks + b) * 2] = (b % 11) - 5 + t
            mv[(t * blocks + b) * 2 + 1] = 4 - (b % 9) - t
            conf[t * blocks + b] = np.float32(((b * 7 + t * 3) % 11)) / np.float32(10.0)
    slots = []
    for t in range(radius):
        slots.append(radius - 1 - t)
        slots.append(radius + 1 + t)
    return mv, conf, slots, bx, by


def jittered_ring(width: int, height: int) -> np.ndarray:
    """collab/tests/fused/mod.rs:685-726: the centre frame plus per-neighbour jitter, 5 slots."""
    frame = unique_frame(width, height).reshape(-1)
    jitter = noisy_flat_field(width, height * 4, 0.5, 0.002).reshape(-1)
    px = width * height
    ring = []
    for slot in range(5):
        if slot == 2:
            ring.append(frame)
            continue
        n = slot if slot < 2 else slot - 1
        ring.append(((frame + jitter[n * px:(n + 1) * px]) - np.float32(0.5)).astype(np.float32))
    return np.concatenate(ring).astype(np.float32)


def probe_index(i: int, length: int) -> int:
    return (i * 7919 + 1013) % length


def digest(accum: np.ndarray, wsum: np.ndarray, group_weight: np.ndarray):
    """collab/tests/fused/mod.rs:209-243 for a luma ring."""
    acc = accum.astype(np.float64)
    w = wsum.astype(np.float64)
    pix = np.where(wsum != 0, acc * 5.0 / np.where(wsum != 0, w, 1.0), 0.0)
    n = pix.size
    return {
        "covered": int(np.count_nonzero(wsum)),
        "pixel_mean": float(pix.sum() / n),
        "pixel_rms": float(np.sqrt((pix * pix).sum() / n)),
        "weight_mean": float(group_weight.astype(np.float64).sum() / group_weight.size),
        "probes": [float(pix[probe_index(i, n)]) for i in range(8)],
    }
