"""The chroma pass's motion from the luma pass, on the GPU, against the oracle.

The field map (avd_site_map_luma_fields) and the seeded refine (avd_site_motion_refine) through the
C ABI, bit-exact with the oracle's restatements; then whole NL4D and NLMeans engines cleaning luma
and chroma, whose chroma passes refine the luma pass's mapped motion (engine/spec.rs:350-365,
nlmeans/dispatch/motion.rs:226-313): NL4D output words bit-exact with the linked oracle passes at
4:2:0, 4:2:2 and 4:4:4, NLMeans within one code.
"""
from __future__ import annotations

import ctypes

import numpy as np
import pytest
import torch

import avd
from avd import _lib as L

import oracle as O
from gpu_helpers import dev, host, ptr, require_gpu, run_engine, words_dev

pytestmark = pytest.mark.gpu


@pytest.fixture(autouse=True)
def _gpu():
    require_gpu()


def _layout(g: O.MotionGeom, pad=0):
    lay = L.FieldLayout()
    lay.blocks_x, lay.blocks_y, lay.step, lay.blksize = g.blocks_x, g.blocks_y, g.step, g.blksize
    lay.stride = g.blocks * 2 + pad
    return lay


@pytest.mark.parametrize("chroma,W,H,blk,ov", [("420", 70, 45, 16, 8), ("422", 64, 33, 16, 8), ("444", 41, 29, 16, 8),
                                               ("420", 9, 9, 16, 8), ("420", 99, 61, 12, 8), ("422", 50, 40, 8, 4)])
def test_map_matches_the_oracle(chroma, W, H, blk, ov):
    print(f"map {chroma} {W}x{H} blk {blk}", flush=True)
    sx, sy = O.chroma_shifts(chroma)
    luma = O.MotionGeom(W, H, blk, ov, 4, 2)
    cg = O.MotionGeom((W + sx) >> sx, (H + sy) >> sy, blk, ov, 4, 2)
    fields = 4
    rng = np.random.default_rng(W * 7 + H)
    values = rng.integers(-60, 61, fields * luma.blocks * 2).astype(np.int32)
    want = O.map_luma_fields(values, luma, cg, sx, sy, fields)
    d_in = dev(values)
    d_out = torch.full((fields * cg.blocks * 2,), 12345, dtype=torch.int32, device="cuda")
    L.check(avd.load().avd_site_map_luma_fields(ctypes.byref(_layout(luma)), ptr(d_in), ctypes.byref(_layout(cg)),
                                                ptr(d_out), sx, sy, fields, None))
    assert np.array_equal(host(d_out), want)


def test_padded_fields_map_like_packed_ones():
    """pair_fields_map_like_window_fields (motion/tests/chroma_map.rs:183-226): padding between
    fields, filled with sentinels, changes nothing."""
    luma, cg = O.MotionGeom(64, 64, 16, 8, 4, 2), O.MotionGeom(32, 32, 16, 8, 4, 2)
    rng = np.random.default_rng(2)
    values = rng.integers(-30, 31, 2 * luma.blocks * 2).astype(np.int32)
    pad = 8
    padded = np.concatenate([np.concatenate([values[f * luma.blocks * 2:(f + 1) * luma.blocks * 2],
                                             np.full(pad, 9999, np.int32)]) for f in range(2)])
    d_in = dev(padded)
    d_out = torch.zeros(2 * (cg.blocks * 2 + pad), dtype=torch.int32, device="cuda")
    L.check(avd.load().avd_site_map_luma_fields(ctypes.byref(_layout(luma, pad)), ptr(d_in),
                                                ctypes.byref(_layout(cg, pad)), ptr(d_out), 1, 1, 2, None))
    got = host(d_out).reshape(2, -1)[:, :cg.blocks * 2].reshape(-1)
    assert np.array_equal(got, O.map_luma_fields(values, luma, cg, 1, 1, 2))


@pytest.mark.parametrize("blk,ov,R,C", [(16, 8, 1, 1), (16, 8, 1, 2), (12, 6, 1, 1), (16, 8, 2, 1), (8, 4, 1, 2)])
def test_refine_matches_the_oracle(blk, ov, R, C):
    print(f"refine blk {blk} R {R} C {C}", flush=True)
    W, H, depth = 80, 56, 10
    clip = avd.synth.clip(5, W, H, depth, 13 + blk, chroma="444" if C == 2 else "gray", sigma_luma=6.0)
    if C == 2:
        codes = [np.stack([f[1].numpy(), f[2].numpy()], axis=-1).astype(np.uint16) for f in clip]
    else:
        codes = [f[0].numpy().astype(np.uint16)[..., None] for f in clip]
    ring = np.stack([c.reshape(-1) for c in codes])
    g = O.MotionGeom(W, H, blk, ov, R, 1)
    centre, slots = 2, [0, 1, 3, 4]
    rng = np.random.default_rng(blk + R)
    seeds = rng.integers(-6, 7, len(slots) * g.blocks * 2).astype(np.int32)
    floor, th = float(O.sad_noise_floor(blk, 6.0 / 1023)), float(O.thsad(blk, 1.0))
    mp = L.MotionPass()
    mp.width, mp.height, mp.channels, mp.blksize, mp.overlap = W, H, C, blk, ov
    mp.search_radius, mp.pyramid_levels, mp.centre_slot, mp.neighbours = 0, 1, centre, len(slots)
    mp.max_value, mp.sad_floor, mp.thsad = 1023.0, floor, th
    d_ring = dev(ring.view(np.int16))
    ns = dev(np.asarray(slots, np.uint32).view(np.int32))
    d_seeds = dev(seeds)
    mv = torch.zeros(len(slots) * g.blocks * 2, dtype=torch.int32, device="cuda")
    conf = torch.zeros(len(slots) * g.blocks, dtype=torch.float32, device="cuda")
    L.check(avd.load().avd_site_motion_refine(ctypes.byref(mp), ptr(d_ring), ptr(ns), ptr(d_seeds), R, ptr(mv),
                                              ptr(conf), None))
    got_mv, got_conf = host(mv), host(conf)
    lane0 = [c[..., 0].reshape(-1).astype(np.int32) for c in codes]
    per = g.blocks * 2
    for i, s in enumerate(slots):
        want_mv, want_conf = O.motion_refine(g, [lane0[centre]], [lane0[s]], seeds[i * per:(i + 1) * per], 1023.0,
                                             floor, th, radius=R)
        assert np.array_equal(got_mv[i * per:(i + 1) * per], want_mv), i
        assert np.array_equal(got_conf[i * g.blocks:(i + 1) * g.blocks], want_conf), i
    # In place: the seeds buffer is also the output.
    L.check(avd.load().avd_site_motion_refine(ctypes.byref(mp), ptr(d_ring), ptr(ns), ptr(d_seeds), R, ptr(d_seeds),
                                              ptr(conf), None))
    assert np.array_equal(host(d_seeds), got_mv)


LAYOUT = {"420": 1, "422": 2, "444": 3}


def _nl4d_spec(r, S, sigma):
    return O.Nl4dSpec(temporal_radius=r, refine=2, spatial_radius=S, lambda_ht=2.7, c_min=0.05, kaiser_beta=2.0,
                      sigma=sigma, motion=O.MotionGeom(0, 0, 16, 8, 4, 2))


@pytest.mark.parametrize("chroma,W,H,r,frames", [("420", 64, 48, 2, 6), ("422", 56, 40, 1, 5), ("444", 48, 40, 1, 4),
                                                 ("420", 70, 46, 1, 3)])
def test_nl4d_engine_chroma_refines_luma_motion_bit_exact(chroma, W, H, r, frames):
    print(f"nl4d {chroma} {W}x{H} r {r} frames {frames}", flush=True)
    depth, S = 10, 6
    clip = avd.synth.clip(frames, W, H, depth, 61 + W, chroma=chroma, sigma_luma=16.0, sigma_chroma=8.0)
    geom = L.Geometry(W, H, LAYOUT[chroma], L.SampleFormat.u16(10), L.SampleFormat.u16(10), 0)
    common = dict(temporal_radius=r, spatial_radius=S, lambda_ht=2.7, pooled_threshold=0, noise_map=0,
                  motion={"estimation": 1}, field_lambda=0.0)
    eng = avd.build_engine("nl4d", geom, luma=avd.nl4d_options(sigma=16 / 1023, **common),
                           chroma=avd.nl4d_options(sigma=8 / 1023, **common))
    cw, ch = avd.chroma_dims(LAYOUT[chroma], W, H)
    dims = [(W, H), (cw, ch), (cw, ch)]
    got = run_engine(eng, [[words_dev(p.numpy(), depth) for p in f] for f in clip], depth, dims)
    pl = O.Nl4dPass(_nl4d_spec(r, S, 16 / 1023), W, H, 1, depth, depth)
    pc = O.Nl4dPass(_nl4d_spec(r, S, 8 / 1023), cw, ch, 2, depth, depth)
    O.link_motion(pl, pc, chroma)
    want_l, _ = O.run_stream(pl, [[O.pack_plane(f[0].numpy(), depth)] for f in clip])
    want_c, _ = O.run_stream(pc, [[O.pack_plane(f[1].numpy(), depth), O.pack_plane(f[2].numpy(), depth)] for f in clip])
    assert len(got) == len(want_l) == len(want_c) == frames
    for i in range(frames):
        assert np.array_equal(got[i][0], want_l[i][0]), f"luma {i}"
        assert np.array_equal(got[i][1], want_c[i][0]), f"U {i}"
        assert np.array_equal(got[i][2], want_c[i][1]), f"V {i}"


def _max_code_diff(a_words, b_words, depth, pixels):
    return int(np.max(np.abs(O.unpack_plane(a_words, depth, pixels) - O.unpack_plane(b_words, depth, pixels))))


def test_nlm_engine_chroma_refines_luma_motion():
    W, H, depth, frames, r = 80, 64, 10, 6, 2
    print(f"nlm 420 {W}x{H}", flush=True)
    sigma = 12.0 / 1023
    clip = avd.synth.clip(frames, W, H, depth, 19, chroma="420", sigma_luma=12.0, sigma_chroma=8.0)
    geom = L.Geometry(W, H, 1, L.SampleFormat.u16(10), L.SampleFormat.u16(10), 0)
    opt = avd.nlmeans_options(hq=1, temporal=1, temporal_radius=r, search_radius=3, patch_radius=2, sigma=sigma,
                              motion_compensation=1, motion={"estimation": 1})
    eng = avd.build_engine("nlmeans", geom, luma=opt, chroma=opt)
    dims = [(W, H), (W // 2, H // 2), (W // 2, H // 2)]
    got = run_engine(eng, [[words_dev(p.numpy(), depth) for p in f] for f in clip], depth, dims)

    def spec(strength):
        return O.NlmSpec(r, 3, 2, strength, 1.0, hq=True, sigma=sigma, motion=O.MotionGeom(0, 0, 16, 8, 4, 2))

    pl = O.NlmPass(spec(O.HQ_STRENGTH_LUMA[r]), W, H, 1, depth, depth)
    pc = O.NlmPass(spec(O.HQ_STRENGTH_CHROMA[r]), W // 2, H // 2, 2, depth, depth)
    O.link_motion(pl, pc, "420")
    want_l, _ = O.run_stream(pl, [[O.pack_plane(f[0].numpy(), depth)] for f in clip])
    want_c, _ = O.run_stream(pc, [[O.pack_plane(f[1].numpy(), depth), O.pack_plane(f[2].numpy(), depth)] for f in clip])
    assert len(got) == len(want_l) == len(want_c) == frames
    for i in range(frames):
        assert _max_code_diff(got[i][0], want_l[i][0], depth, W * H) <= 1, f"luma {i}"
        assert _max_code_diff(got[i][1], want_c[i][0], depth, W * H // 4) <= 1, f"U {i}"
        assert _max_code_diff(got[i][2], want_c[i][1], depth, W * H // 4) <= 1, f"V {i}"
