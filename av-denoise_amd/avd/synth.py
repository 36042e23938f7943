"""Seeded synthetic clips for tests and benchmarks (the role of the reference CLI's ``synthesize``).

A scene is a sum of six 2-D sinusoids (periods 8..128 px, random orientation, phase and
amplitude) plus a linear gradient, normalised analytically to 0..1 and mapped to the legal code
range of the bit depth.  Moving clips translate the scene by a random integer step in [-6, 6] per
frame plus a 0.25 px sub-pixel shift, evaluated exactly on the continuous scene.  Gaussian noise
is added per plane and the codes are clipped.  Everything is drawn from seeded generators, so a
(seed, shape) pair always gives the same clip.
"""
from __future__ import annotations

import math

import numpy as np
import torch


def _scene_params(rng: np.random.Generator, components: int = 6):
    periods = rng.uniform(8.0, 128.0, components)
    theta = rng.uniform(0.0, math.pi, components)
    phase = rng.uniform(0.0, 2.0 * math.pi, components)
    amp = rng.uniform(0.5, 1.0, components)
    grad = rng.uniform(-1.0, 1.0, 2)
    return periods, theta, phase, amp, grad


def _render(params, width: int, height: int, ox: float, oy: float, scale: float, device) -> torch.Tensor:
    """The scene on a width x height grid sampled at (x*scale + ox, y*scale + oy), in 0..1."""
    periods, theta, phase, amp, grad = params
    ys = torch.arange(height, dtype=torch.float64, device=device).unsqueeze(1) * scale + oy
    xs = torch.arange(width, dtype=torch.float64, device=device).unsqueeze(0) * scale + ox
    acc = torch.zeros(height, width, dtype=torch.float64, device=device)
    for T, th, ph, a in zip(periods, theta, phase, amp):
        acc += a * torch.sin(2.0 * math.pi * (xs * math.cos(th) + ys * math.sin(th)) / T + ph)
    span_x, span_y = 1920.0, 1080.0
    acc += grad[0] * (xs / span_x - 0.5) * 2.0 + grad[1] * (ys / span_y - 0.5) * 2.0
    bound = float(np.sum(amp)) + abs(grad[0]) * 2.0 * max(1.0, (width * scale + abs(ox)) / span_x) + \
        abs(grad[1]) * 2.0 * max(1.0, (height * scale + abs(oy)) / span_y)
    return ((acc / bound) + 1.0) * 0.5


def legal_range(depth: int):
    k = float(1 << (depth - 8))
    return 16.0 * k, 235.0 * k


def clip(frames: int, width: int, height: int, depth: int, seed: int, chroma: str = "420", moving: bool = True,
         sigma_luma: float = 10.0, sigma_chroma: float | None = None, device="cpu", lo_hi=None):
    """Returns a list of frames, each a list of int32 code planes (Y[, U, V]) on ``device``.

    ``sigma_*`` are noise standard deviations in codes of the given depth; ``lo_hi`` clips codes
    (defaults to the full 0..2^depth-1 range).
    """
    rng = np.random.default_rng(seed)
    planes = 1 if chroma == "gray" else 3
    params = [_scene_params(np.random.default_rng(seed * 7919 + 17 + p)) for p in range(planes)]
    steps = rng.integers(-6, 7, size=(frames, 2)) if moving else np.zeros((frames, 2), np.int64)
    offsets = np.cumsum(steps, axis=0).astype(np.float64)
    gen = torch.Generator(device="cpu").manual_seed(seed)
    lo, hi = legal_range(depth)
    maxc = (1 << depth) - 1
    clo, chi = lo_hi if lo_hi is not None else (0, maxc)
    if sigma_chroma is None:
        sigma_chroma = sigma_luma
    out = []
    for t in range(frames):
        ox = offsets[t, 0] + (0.25 * t if moving else 0.0)
        oy = offsets[t, 1]
        frame = []
        for p in range(planes):
            if p == 0:
                w, h, sc = width, height, 1.0
            elif chroma == "420":
                w, h, sc = (width + 1) // 2, (height + 1) // 2, 2.0
            elif chroma == "422":
                w, h, sc = (width + 1) // 2, height, 2.0
            else:
                w, h, sc = width, height, 1.0
            base = _render(params[p], w, h, ox, oy, sc, device)
            codes = lo + base * (hi - lo)
            sig = sigma_luma if p == 0 else sigma_chroma
            noise = torch.randn(h, w, generator=gen, dtype=torch.float64).to(device) * sig
            codes = torch.clamp(torch.round(codes + noise), clo, chi).to(torch.int32)
            frame.append(codes)
        out.append(frame)
    return out


def pack_words(codes: torch.Tensor, depth: int) -> torch.Tensor:
    """Packs an (H, W) int code plane into little-endian u32 words as a uint8 byte tensor."""
    flat = codes.reshape(-1).to(torch.int64)
    spw = 4 if depth == 8 else 2
    bits = 32 // spw
    n = flat.numel()
    words = (n + spw - 1) // spw
    pad = torch.zeros(words * spw, dtype=torch.int64, device=codes.device)
    pad[:n] = flat
    pad = pad.view(words, spw)
    w = torch.zeros(words, dtype=torch.int64, device=codes.device)
    for lane in range(spw):
        w |= pad[:, lane] << (lane * bits)
    return _to_u32_bytes(w)


def _to_u32_bytes(w: torch.Tensor) -> torch.Tensor:
    w32 = (w & 0xFFFFFFFF)
    w32 = torch.where(w32 >= 2**31, w32 - 2**32, w32).to(torch.int32)
    return w32.view(torch.uint8).contiguous()


def unpack_words(buf: torch.Tensor, depth: int, width: int, height: int) -> torch.Tensor:
    """Inverse of ``pack_words``: a uint8 byte tensor of u32 words back to an (H, W) int32 plane."""
    w = buf.view(torch.int32).to(torch.int64) & 0xFFFFFFFF
    spw = 4 if depth == 8 else 2
    bits = 32 // spw
    mask = (1 << bits) - 1
    lanes = [(w >> (l * bits)) & mask for l in range(spw)]
    flat = torch.stack(lanes, dim=1).reshape(-1)[: width * height]
    return flat.view(height, width).to(torch.int32)
