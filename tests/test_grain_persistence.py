"""Grain persistence and the vector save (grain_project, grain_persist, grain_persist_reduce,
nl4d/kernels/grain_persistence.rs:54-279; grain_save_vectors, nl4d/kernels/grain.rs:50-75).

On the CPU the oracle is pinned by the patterns' stated properties (nl4d/grain/consts.rs:70-75:
each ignores a cell's mean, slopes and curves and changes sign at least 4 times along each axis),
by closed forms (a frame against itself persists fully, cross sum equal to auto sum bit for bit;
independent grain does not persist and its auto sum reads sigma squared per pixel), by the flat
gates and by the lag-to-slot arithmetic of the 9-slot history.  On the GPU every kernel, called
through the C ABI, must equal the oracle bit for bit.
"""
from __future__ import annotations

import ctypes

import numpy as np
import pytest

import oracle as O

from avd import _lib as L


def frame(W, H, sigma, seed, base=0.5, C=1, ch=0):
    rng = np.random.default_rng(seed)
    f = np.zeros((H, W, C), np.float32)
    f[..., ch] = base + sigma * rng.standard_normal((H, W))
    return f


def project(ring, out, W, H, C=1, ch=0, slot=0, history=None, hslot=0, luma_band=True):
    cells = (W // 8) * (H // 8)
    history = np.zeros(9 * cells * 9, np.float32) if history is None else history
    O.grain_project(ring, out, W, H, C, ch, slot, history, hslot, luma_band)
    return history


def persist(history, now_slot, cells, lags):
    part = O.grain_persist(history, now_slot, cells, lags)
    blocks = -(-cells // 1024)
    return O.grain_persist_reduce(part, np.zeros(3 * lags, np.float32), 0, blocks, lags).reshape(lags, 3)


# ---- the oracle ---------------------------------------------------------------------------------------

def test_patterns_ignore_mean_slopes_and_curves():
    yy, xx = np.mgrid[0:8, 0:8]
    words = O.grain_pattern_words()
    for p in range(8):
        bits = (int(words[2 * p + 1]) << 32) | int(words[2 * p])
        sign = np.array([1 if (bits >> b) & 1 else -1 for b in range(64)]).reshape(8, 8)
        for f in (np.ones((8, 8)), xx, yy, xx ** 2, yy ** 2):
            assert (sign * f).sum() == 0, p
        assert (np.diff(sign, axis=1) != 0).sum(1).min() >= 4 and (np.diff(sign, axis=0) != 0).sum(0).min() >= 4


def test_projections_of_smooth_cells_vanish():
    W, H = 64, 48
    yy, xx = np.mgrid[0:H, 0:W]
    smooth = (0.4 + 0.001 * (xx % 8) + 0.0005 * (yy % 8) + 0.0001 * (xx % 8) ** 2).astype(np.float32)[..., None]
    hist = project(smooth[None], np.full((H, W, 1), 0.5, np.float32), W, H).reshape(9, -1, 9)
    assert np.max(np.abs(hist[0, :, :8])) < 1e-5
    assert np.all(hist[0, :, 8] == 1.0)


def test_a_frame_against_itself_persists_fully():
    W, H, sigma = 96, 64, 2 / 255
    f = frame(W, H, sigma, 1)
    out = np.full((H, W, 1), 0.5, np.float32)
    hist = project(f[None], out, W, H, hslot=3)
    project(f[None], out, W, H, history=hist, hslot=4)
    rec = persist(hist, 4, (W // 8) * (H // 8), 1)
    assert rec[0, 0] == (W // 8) * (H // 8) and rec[0, 1] == rec[0, 2]
    assert abs(rec[0, 1] / rec[0, 0] / sigma ** 2 - 1) < 0.15


def test_independent_grain_does_not_persist():
    W, H, sigma = 256, 128, 2 / 255
    out = np.full((H, W, 1), 0.5, np.float32)
    hist = project(frame(W, H, sigma, 1)[None], out, W, H, hslot=0)
    project(frame(W, H, sigma, 2)[None], out, W, H, history=hist, hslot=1)
    rec = persist(hist, 1, (W // 8) * (H // 8), 1)
    assert abs(rec[0, 1] / rec[0, 0] / sigma ** 2 - 1) < 0.05  # auto: one pixel's variance
    assert abs(rec[0, 2]) < 0.05 * rec[0, 1]


def test_flat_gates_and_the_luma_band():
    W, H = 64, 48
    f = frame(W, H, 1 / 255, 3)
    out = np.full((H, W, 1), 0.5, np.float32)
    out[3, 12] += 6 / 255  # cell (1, 0) of the denoised frame not flat
    f[20, 20, 0] = 3 / 255  # cell (2, 2) near black
    flags = project(f[None], out, W, H).reshape(9, -1, 9)[0, :, 8].reshape(H // 8, W // 8)
    assert flags[0, 1] == 0 and flags[2, 2] == 0 and flags.sum() == flags.size - 2
    dark = np.full((H, W, 1), 10 / 255, np.float32)
    assert not project(f[None], dark, W, H).reshape(9, -1, 9)[0, :, 8].any()
    assert project(f[None], dark, W, H, luma_band=False).reshape(9, -1, 9)[0, :, 8].sum() == flags.size - 1


def test_lag_n_reads_the_slot_n_frames_back():
    # grain_persistence.rs:168: lag n + 1 reads slot (now + 9 - 1 - n) % 9.
    W, H = 64, 48
    cells = (W // 8) * (H // 8)
    out = np.full((H, W, 1), 0.5, np.float32)
    f = frame(W, H, 2 / 255, 7)
    hist = np.zeros(9 * cells * 9, np.float32)
    for slot in range(9):
        project(frame(W, H, 2 / 255, 100 + slot)[None], out, W, H, history=hist, hslot=slot)
    project(f[None], out, W, H, history=hist, hslot=1)
    project(f[None], out, W, H, history=hist, hslot=6)  # 4 frames before slot 1, wrapping
    rec = persist(hist, 1, cells, 8)
    assert np.all(rec[:, 0] == cells)
    assert rec[3, 2] == rec[3, 1]  # lag 4: the same frame
    others = np.delete(np.arange(8), 3)
    assert np.all(np.abs(rec[others, 2]) < 0.3 * rec[others, 1])


def test_persist_reduce_adds_exact_integer_sums():
    lags, blocks = 3, 300
    rng = np.random.default_rng(1)
    part = rng.integers(-100, 101, lags * blocks * 3).astype(np.float32)
    record = np.arange(2 + 3 * lags, dtype=np.float32)
    got = O.grain_persist_reduce(part, record, 2, blocks, lags)
    want = record.astype(np.float64)
    want[2:] += part.reshape(lags, blocks, 3).astype(np.float64).sum(1).reshape(-1)
    assert np.array_equal(got, want.astype(np.float32))


def test_save_vectors_copies_one_neighbour():
    blocks = 37
    mv = np.arange(5 * blocks * 2, dtype=np.int32) - 100
    conf = np.linspace(0, 1, 5 * blocks).astype(np.float32)
    saved_mv = np.full(3 * blocks * 2, -9, np.int32)
    saved_conf = np.full(3 * blocks, -9.0, np.float32)
    O.grain_save_vectors(mv, conf, saved_mv, saved_conf, 2 * blocks * 2, 2 * blocks, 1, blocks)
    assert np.array_equal(saved_mv[2 * blocks:4 * blocks], mv[4 * blocks:6 * blocks])
    assert np.array_equal(saved_conf[blocks:2 * blocks], conf[2 * blocks:3 * blocks])
    assert np.all(saved_mv[:2 * blocks] == -9) and np.all(saved_conf[2 * blocks:] == -9.0)


def test_sites_reject_bad_arguments_without_a_gpu():
    import avd

    lib, b = L.load(), ctypes.c_void_p(16)
    assert lib.avd_site_grain_project(b, b, b, b, 64, 48, 1, 0, 0, 9, 1, None) == avd.InvalidOptions.status
    assert lib.avd_site_grain_project(b, b, b, b, 64, 48, 1, 1, 0, 0, 1, None) == avd.InvalidGeometry.status
    assert lib.avd_site_grain_persist(b, b, b, 0, 0, 48, 9, None) == avd.InvalidOptions.status
    assert lib.avd_site_grain_save_vectors(None, b, b, b, 0, 0, 0, 4, None) == avd.NullArgument.status
    assert lib.avd_grain_persist_partials_len(2049, 8) == 3 * 8 * 3


# ---- the HIP kernels against the oracle -----------------------------------------------------------------

def _ptr(t):
    return ctypes.c_void_p(t.data_ptr())


@pytest.mark.gpu
@pytest.mark.parametrize("shape", [(64, 48, 1, 0), (70, 45, 2, 1), (1920, 1080, 1, 0)], ids=str)
def test_site_grain_project_and_persist_are_bit_exact(shape):
    import torch

    from gpu_helpers import dev, host, require_gpu

    require_gpu()
    W, H, C, ch = shape
    print(f"grain project + persist {W}x{H} C{C}", flush=True)
    cells = (W // 8) * (H // 8)
    rng = np.random.default_rng(W)
    out = np.full((H, W, C), 0.5, np.float32)
    out[rng.integers(0, H, 20), rng.integers(0, W, 20), ch] += 6 / 255  # some cells not flat
    frames = [frame(W, H, 2 / 255, 10 + k, C=C, ch=ch) for k in range(3)]
    frames[2] = frames[0].copy()
    ring = np.stack(frames)
    words = O.grain_pattern_words()
    d_ring, d_out, d_words = dev(ring), dev(out), dev(words.view(np.int32))
    d_hist = torch.zeros(9 * cells * 9, dtype=torch.float32, device="cuda")
    want_hist = np.zeros(9 * cells * 9, np.float32)
    for k, slot in [(0, 7), (1, 8), (2, 0)]:
        L.check(L.load().avd_site_grain_project(_ptr(d_ring), _ptr(d_out), _ptr(d_words), _ptr(d_hist), W, H, C, ch,
                                                k, slot, 1, None))
        O.grain_project(ring, out, W, H, C, ch, k, want_hist, slot, True)
    got_hist = host(d_hist)
    assert np.array_equal(got_hist.view(np.uint32), want_hist.view(np.uint32))
    lags = 8
    n = L.load().avd_grain_persist_partials_len(cells, lags)
    d_part = torch.full((n,), -7.0, dtype=torch.float32, device="cuda")
    record = rng.standard_normal(5 + 3 * lags).astype(np.float32)
    d_rec = dev(record)
    L.check(L.load().avd_site_grain_persist(_ptr(d_hist), _ptr(d_part), _ptr(d_rec), 5, 0, cells, lags, None))
    want_part = O.grain_persist(want_hist, 0, cells, lags)
    assert np.array_equal(host(d_part).view(np.uint32), want_part.view(np.uint32))
    blocks = -(-cells // 1024)
    want_rec = O.grain_persist_reduce(want_part, record, 5, blocks, lags)
    assert np.array_equal(host(d_rec).view(np.uint32), want_rec.view(np.uint32))
    per_lag = want_part.reshape(lags, blocks, 3)
    assert np.array_equal(per_lag[1, :, 1], per_lag[1, :, 2]) and per_lag[1, :, 0].sum() > 0  # lag 2: slot 7, the same frame


@pytest.mark.gpu
def test_site_grain_save_vectors_copies_bit_exactly():
    import torch

    from gpu_helpers import dev, host, require_gpu

    require_gpu()
    blocks = 129600 // 4
    rng = np.random.default_rng(2)
    mv = rng.integers(-50, 50, 6 * blocks * 2).astype(np.int32)
    conf = rng.random(6 * blocks).astype(np.float32)
    d_mv, d_conf = dev(mv), dev(conf)
    d_smv = torch.full((4 * blocks * 2,), -9, dtype=torch.int32, device="cuda")
    d_sconf = torch.full((4 * blocks,), -9.0, dtype=torch.float32, device="cuda")
    L.check(L.load().avd_site_grain_save_vectors(_ptr(d_mv), _ptr(d_conf), _ptr(d_smv), _ptr(d_sconf),
                                                 3 * blocks * 2, 3 * blocks, 2, blocks, None))
    want_mv = np.full(4 * blocks * 2, -9, np.int32)
    want_conf = np.full(4 * blocks, -9.0, np.float32)
    O.grain_save_vectors(mv, conf, want_mv, want_conf, 3 * blocks * 2, 3 * blocks, 2, blocks)
    assert np.array_equal(host(d_smv), want_mv) and np.array_equal(host(d_sconf), want_conf)
