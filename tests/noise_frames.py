"""Test frames for the noise measurement, built as the reference's noise tests build theirs
(nlmeans/tests/helpers.rs: make_uniform_frame, make_noisy_gaussian_frame, make_gradient_frame)."""
from __future__ import annotations

import numpy as np


def uniform(width: int, height: int, channels: int, value: float) -> np.ndarray:
    return np.full(width * height * channels, value, np.float32)


def noisy_gaussian(width: int, height: int, channels: int, base: float, sigmas) -> np.ndarray:
    """base + sigma * (sum of four hashed uniforms in [-0.5, 0.5]) / sqrt(1/3), clamped to [0, 1];
    sample i's k-th uniform hashes 4 i + k (a multiply-xorshift mix in u32)."""
    n = width * height * channels
    idx = np.arange(n, dtype=np.uint64)
    unit_std = np.sqrt(np.float32(1.0) / np.float32(3.0))
    total = np.zeros(n, np.float32)
    for k in range(4):
        h = ((idx * 4 + k) & 0xFFFFFFFF).astype(np.uint32)
        h = (h.astype(np.uint64) * 2654435761 + 0x9E3779B9) & 0xFFFFFFFF
        h ^= h >> 15
        h = (h * 0x85EBCA6B) & 0xFFFFFFFF
        h ^= h >> 13
        u = h.astype(np.float32) / np.float32(4294967295.0)
        total = total + (u - np.float32(0.5))
    sig = np.array([sigmas[c % len(sigmas)] for c in range(channels)], np.float32)
    sigma = np.tile(sig, width * height)
    return np.clip(np.float32(base) + (total / unit_std) * sigma, 0.0, 1.0).astype(np.float32)


def gradient(width: int, height: int, low: float, high: float) -> np.ndarray:
    x = np.arange(width, dtype=np.float32) / np.float32(max(width - 1, 1))
    row = np.float32(low) + np.float32(high - low) * x
    return np.tile(row, height).astype(np.float32)
