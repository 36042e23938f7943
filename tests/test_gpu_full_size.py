"""The grouping search and the whole NL4D engine at the benchmark's full size.

3840x2160 10-bit luma and its 1920x1080 4:2:0 chroma, temporal radius 3 (seven ring slots), spatial
radius 12, stride 3, groups of up to 16: the frame sizes at which the workgroup grids are remapped
over the eight XCDs, the trajectory's raw buffer loads address the whole ring (seven 16.6 MB slots)
through 32-bit byte offsets, and the matrix-core centre search runs with its redo list.  The
motion fields come from the motion site (its parity is tested on its own); the oracle recomputes
the members of about 1700 references -- the corners, the last row and column, the grid's edges and
a seeded spread -- from the same ring, fields and confidences, and they must match bit for bit.

The engine test runs the FP32 form over a two-frame 3840x2160 luma stream at temporal radius 2
(volume grids of 4 frames), stride 3, groups of up to 16, and compares every output frame with the
oracle's multi-threaded FP32 restatement bit for bit: motion, both searches, the filter, the binned
aggregation (tile lists of a 120 x 68 tile grid per slot) and the normalise at size.
"""
from __future__ import annotations

import ctypes

import numpy as np
import pytest
import torch

import avd
from avd import _lib as L

import oracle as O
from gpu_helpers import (assert_members_equal, collab_pass, dev, gray_geometry, host, ptr, require_gpu, run_engine,
                         site_collab_search, words_dev)
from test_gpu_k16 import _opts, _spec

pytestmark = pytest.mark.gpu


@pytest.fixture(autouse=True)
def _gpu():
    require_gpu()


def _sample_refs(refs_x, refs_y, n_random, seed):
    ids = set()
    for iy in (0, 1, refs_y // 2, refs_y - 2, refs_y - 1):
        for ix in range(refs_x):
            if ix < 3 or ix >= refs_x - 3 or ix % 97 == 0:
                ids.add(iy * refs_x + ix)
    for ix in (0, 1, refs_x // 2, refs_x - 2, refs_x - 1):
        for iy in range(refs_y):
            if iy < 3 or iy >= refs_y - 3 or iy % 53 == 0:
                ids.add(iy * refs_x + ix)
    rng = np.random.default_rng(seed)
    ids.update(int(v) for v in rng.integers(0, refs_x * refs_y, n_random))
    return np.asarray(sorted(ids), np.uint32)


@pytest.mark.parametrize("plane", ["luma", "chroma"])
def test_full_size_search_matches_the_oracle_on_sampled_references(plane):
    W0, H0, depth, r, S, step, K = 3840, 2160, 10, 3, 12, 3, 16
    frames = 2 * r + 1
    print(f"full-size search: {plane}", flush=True)
    clip = avd.synth.clip(frames, W0, H0, depth, 0xA5D00002, chroma="420", sigma_luma=16.0, sigma_chroma=8.0,
                          device="cuda")
    if plane == "luma":
        W, H, C, sigma = W0, H0, 1, 16.0 / 1023
        ring = torch.stack([f[0].reshape(-1) for f in clip]).to(torch.int16)
    else:
        W, H, C, sigma = W0 // 2, H0 // 2, 2, 8.0 / 1023
        ring = torch.stack([torch.stack([f[1].reshape(-1), f[2].reshape(-1)], 1).reshape(-1) for f in clip])
        ring = ring.to(torch.int16)
    ring = ring.contiguous()
    geom = O.MotionGeom(W, H, 16, 8, 4, 2)
    centre = r
    slots = [s for s in range(frames) if s != centre]
    floor, th = float(O.sad_noise_floor(16, sigma)), float(O.thsad(16, 1.0))
    mp = L.MotionPass()
    mp.width, mp.height, mp.channels, mp.blksize, mp.overlap = W, H, C, 16, 8
    mp.search_radius, mp.pyramid_levels, mp.centre_slot, mp.neighbours = 4, 2, centre, len(slots)
    mp.max_value, mp.sad_floor, mp.thsad = 1023.0, floor, th
    ns = dev(np.asarray(slots, np.uint32).view(np.int32))
    mv = torch.zeros(len(slots) * geom.blocks * 2, dtype=torch.int32, device="cuda")
    conf = torch.zeros(len(slots) * geom.blocks, dtype=torch.float32, device="cuda")
    L.check(avd.load().avd_site_motion(ctypes.byref(mp), ptr(ring), ptr(ns), ptr(mv), ptr(conf), None))
    mv_h, conf_h = host(mv), host(conf)
    p = collab_pass(W, H, C, r, S, 2, 0.05, 2.7, sigma, 2.0, centre, frames, geom, code_bits=10, step=step,
                    group_size=K)
    got = site_collab_search(p, codes=ring, nslots=slots, mv=mv_h, conf=conf_h)
    refs_x, refs_y = O.refs_along(W, step), O.refs_along(H, step)
    assert got.size == refs_x * refs_y * (K + 1)
    ids = _sample_refs(refs_x, refs_y, 1500, 7 + C)
    ring_h = host(ring).view(np.uint16)
    op = O.collab_params(W, H, C, r, S, 2, 0.05, 2.7, sigma, 2.0, geom, step=step, group_size=K)
    O.set_threads(min(16, O.max_threads()))
    want = O.collab_search_refs(ring_h, op, centre, slots, mv_h, conf_h, ids)
    got_sub = got.reshape(-1, K + 1)[ids.astype(np.int64)].reshape(-1)
    print(f"{plane}: {ids.size} references compared of {refs_x * refs_y}", flush=True)
    assert_members_equal(got_sub, want, f"{plane} {W}x{H}", group_size=K)
    # The grid groups (volumes filled from every frame) are among them.
    words = got.reshape(-1, K + 1)[:, K]
    assert np.any(words & 0x100)


def test_full_size_engine_fp32_stream_matches_the_oracle():
    W, H, depth, r, S, step = 3840, 2160, 10, 2, 12, 3
    frames, sigma = 2, 16.0 / 1023
    print("full-size engine: 3840x2160 luma, r=2, stride 3, groups of 16, FP32 form", flush=True)
    clip = avd.synth.clip(frames, W, H, depth, 0xA5D00004, chroma="gray", sigma_luma=16.0)
    codes = [f[0].numpy() for f in clip]
    eng = avd.build_engine("nl4d", gray_geometry(W, H, depth), luma=_opts(r, S, sigma, step, 16))
    got = run_engine(eng, [[words_dev(c, depth)] for c in codes], depth, [(W, H)])
    O.set_threads(min(16, O.max_threads()))
    want, _ = O.run_stream(O.Nl4dPass(_spec(r, S, sigma, step, 16), W, H, 1, depth, depth),
                           [[O.pack_plane(c, depth)] for c in codes])
    assert len(got) == len(want) == frames
    for i, (g, w) in enumerate(zip(got, want)):
        assert np.array_equal(g[0], w[0]), f"frame {i}"
