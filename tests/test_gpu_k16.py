"""Groups of 16 members and other grid strides on the GPU, against the oracle.

The reference groups 8 members on a stride-4 grid (collab/mod.rs:16-20); groups of 16 (volume
grids of 4 frames x 4 volumes, 2 x 8 at temporal radius 1, stacks of up to 16) at any stride are
this library's extension, restated in the oracle and pinned there by mathematics
(tests/test_oracle_k16.py).  Here the HIP search (matrix-core and vector centre searches, the
trajectory at one and two references per wave, the f32-ring search), the filter (direct scatter
and binned) and the engine stream must match the oracle bit for bit; the mixed transform form stays
within 0.05 dB of the FP32 form.  The binned aggregation's overflow path is forced with a small
tile list and must leave the output unchanged.
"""
from __future__ import annotations

import numpy as np
import pytest

import avd
from avd import _lib as L
from gpu_helpers import (assert_members_equal, collab_pass, dev, gray_geometry, run_engine, site_collab_filter,
                         site_collab_search, words_dev)
from test_gpu_parity import _codes_ring, _oracle_fields, _psnr

import oracle as O

pytestmark = pytest.mark.gpu


def _scene_planes(frames, W, H, C, depth, seed, content="scene"):
    top = (1 << depth) - 1
    if content == "scene":
        clip = avd.synth.clip(frames, W, H, depth, seed, chroma="444" if C == 2 else "gray", sigma_luma=12.0)
        return [[f[0].numpy()] if C == 1 else [f[1].numpy(), f[2].numpy()] for f in clip]
    rng = np.random.default_rng(seed)
    if content == "stripes":  # every column constant: a column's shifts tie, so one lane holds the best
        return [[np.broadcast_to(rng.integers(0, top + 1, (1, W)), (H, W)).copy() for _ in range(C)]
                for _ in range(frames)]
    return [[rng.integers(0, top + 1, (H, W)) for _ in range(C)] for _ in range(frames)]


SEARCH_CASES = [
    # (name, W, H, C, depth, radius, S, step, code_bits, content, blksize, overlap)
    ("stride 3 luma, matrix-core centre", 80, 56, 1, 10, 3, 12, 3, 10, "scene", 16, 8),
    ("stride 3 chroma", 60, 44, 2, 10, 2, 8, 3, 10, "scene", 16, 8),
    ("stride 4 luma r1 (2 x 8 grid)", 72, 48, 1, 10, 1, 9, 4, 10, "scene", 16, 8),
    ("stride 3 r0 (stack of 16)", 64, 40, 1, 10, 0, 6, 3, 10, "scene", 16, 8),
    ("12-bit, vector centre", 64, 48, 1, 12, 2, 9, 3, 12, "scene", 16, 8),
    ("radius 14, vector centre, wide keys", 64, 48, 1, 10, 2, 14, 3, 0, "scene", 16, 8),
    ("full-range noise, capped keys", 40, 32, 1, 10, 2, 3, 3, 10, "noise", 16, 8),
    ("vertical stripes, one lane's list runs short", 64, 48, 1, 10, 2, 8, 3, 10, "stripes", 16, 8),
    ("vertical stripes, chroma", 48, 40, 2, 10, 2, 8, 3, 10, "stripes", 16, 8),
    ("16 covering blocks, one reference per wave", 64, 48, 1, 10, 6, 6, 3, 10, "scene", 32, 24),
    ("short frame", 40, 11, 1, 10, 2, 5, 3, 10, "scene", 16, 8),
    ("tiny frame", 9, 8, 1, 10, 1, 3, 3, 10, "scene", 16, 8),
]


@pytest.mark.parametrize("case", SEARCH_CASES, ids=[c[0] for c in SEARCH_CASES])
def test_sixteen_member_codes_search_is_bit_exact(case):
    name, W, H, C, depth, radius, S, step, code_bits, content, blksize, overlap = case
    print(f"k16 search case: {name}", flush=True)
    frames = 2 * radius + 1
    codes = _codes_ring(_scene_planes(frames, W, H, C, depth, 61 + W + radius, content), C)
    geom = O.MotionGeom(W, H, blksize, overlap, 4, 2)
    slots = [s for s in range(frames) if s != radius]
    sigma = 12.0 / 1023
    mv, conf = _oracle_fields(codes, W, H, C, radius, slots, sigma, float((1 << depth) - 1), geom)
    p = collab_pass(W, H, C, radius, S, 2, 0.05, 2.7, sigma, 2.0, radius, frames, geom, code_bits=code_bits,
                    step=step, group_size=16)
    op = O.collab_params(W, H, C, radius, S, 2, 0.05, 2.7, sigma, 2.0, geom, step=step, group_size=16)
    want = O.collab_search(codes, op, radius, slots, mv, conf)
    got = site_collab_search(p, codes=dev(codes.view(np.int16)), nslots=slots, mv=mv, conf=conf)
    assert_members_equal(got, want, name, group_size=16)
    if radius >= 1 and content == "scene" and H >= 32:
        assert np.any(got.reshape(-1, 17)[:, 16] & 0x100), "some groups build a 16-member grid"


@pytest.mark.parametrize("C,radius,S,pooled", [(1, 2, 9, False), (2, 1, 6, False), (1, 3, 12, False),
                                                (1, 0, 5, False), (1, 2, 9, True), (2, 2, 6, True)])
def test_sixteen_member_search_and_filter_are_bit_exact(C, radius, S, pooled):
    W, H, depth, step = 72, 52, 10, 3
    print(f"k16 filter C={C} r={radius} S={S} pooled={pooled}", flush=True)
    frames = 2 * radius + 1
    codes = _codes_ring(_scene_planes(frames, W, H, C, depth, 9 + C + radius), C)
    values = (codes.astype(np.float32) / np.float32(1023.0)).astype(np.float32)
    geom = O.MotionGeom(W, H, 16, 8, 4, 2)
    slots = [s for s in range(frames) if s != radius]
    sigma = 12.0 / 1023
    mv, conf = _oracle_fields(codes, W, H, C, radius, slots, sigma, 1023.0, geom)
    p = collab_pass(W, H, C, radius, S, 2, 0.05, 2.7, sigma, 2.0, radius, frames, geom, code_bits=depth,
                    pooled=pooled, step=step, group_size=16)
    op = O.collab_params(W, H, C, radius, S, 2, 0.05, 2.7, sigma, 2.0, geom, step=step, pooled=pooled,
                         group_size=16)
    want = O.collab_search(codes, op, radius, slots, mv, conf)
    got = site_collab_search(p, codes=dev(codes.view(np.int16)), nslots=slots, mv=mv, conf=conf)
    assert_members_equal(got, want, "search", group_size=16)
    acc, ws, gw = site_collab_filter(p, dev(values), slots, got, frames)
    wacc = np.zeros(frames * W * H * C, np.int32)
    wws = np.zeros(frames * W * H, np.int32)
    wgw = O.collab_filter(values, op, radius, slots, want, wacc, wws, "f32")
    assert np.array_equal(gw, wgw)
    assert np.array_equal(acc, wacc)
    assert np.array_equal(ws, wws)


def test_sixteen_member_f32_ring_search_is_bit_exact():
    """The f32-ring form of the search (the reference's F32Search distances) at 16 members."""
    W, H, C, radius, S, step = 56, 40, 1, 2, 6, 3
    frames = 2 * radius + 1
    codes = _codes_ring(_scene_planes(frames, W, H, C, 10, 44), C)
    values = (codes.astype(np.float32) / np.float32(1023.0)).astype(np.float32)
    geom = O.MotionGeom(W, H, 16, 8, 4, 2)
    slots = [s for s in range(frames) if s != radius]
    sigma = 12.0 / 1023
    mv, conf = _oracle_fields(codes, W, H, C, radius, slots, sigma, 1023.0, geom)
    p = collab_pass(W, H, C, radius, S, 2, 0.05, 2.7, sigma, 2.0, radius, frames, geom, step=step, group_size=16)
    op = O.collab_params(W, H, C, radius, S, 2, 0.05, 2.7, sigma, 2.0, geom, step=step, group_size=16)
    want = O.collab_search(values, op, radius, slots, mv, conf)
    got = site_collab_search(p, values=dev(values), nslots=slots, mv=mv, conf=conf)
    assert_members_equal(got, want, "f32 ring", group_size=16)


def _spec(r, S, sigma, step, group_size, lam=2.7):
    return O.Nl4dSpec(temporal_radius=r, refine=2, spatial_radius=S, lambda_ht=lam, c_min=0.05, kaiser_beta=2.0,
                      sigma=sigma, motion=O.MotionGeom(0, 0, 16, 8, 4, 2), step=step, group_size=group_size)


def _opts(r, S, sigma, step, group_size, form=0):
    return avd.nl4d_options(temporal_radius=r, spatial_radius=S, sigma=sigma, lambda_ht=2.7, pooled_threshold=0,
                            noise_map=0, motion={"estimation": 1}, field_lambda=0.0, step=step,
                            group_size=group_size, transform_form=form)


@pytest.mark.parametrize("frames,r,step", [(6, 3, 3), (7, 2, 3), (5, 1, 3), (3, 3, 3), (6, 2, 2)])
def test_sixteen_member_engine_stream_is_bit_exact(frames, r, step):
    W, H, depth, S = 72, 56, 10, 8
    sigma = 16.0 / 1023
    print(f"k16 engine frames={frames} r={r} step={step}", flush=True)
    clip = avd.synth.clip(frames, W, H, depth, 200 + r, chroma="gray", sigma_luma=16.0)
    codes = [f[0].numpy() for f in clip]
    eng = avd.build_engine("nl4d", gray_geometry(W, H, depth), luma=_opts(r, S, sigma, step, 16))
    got = run_engine(eng, [[words_dev(c, depth)] for c in codes], depth, [(W, H)])
    want, _ = O.run_stream(O.Nl4dPass(_spec(r, S, sigma, step, 16), W, H, 1, depth, depth),
                           [[O.pack_plane(c, depth)] for c in codes])
    assert len(got) == len(want) == frames
    for i, (g, w) in enumerate(zip(got, want)):
        assert np.array_equal(g[0], w[0]), f"frame {i}"


def test_sixteen_member_engine_420_is_bit_exact_and_mixed_within_0_05_db():
    W, H, depth, r, S, step = 64, 48, 10, 3, 8, 3
    frames = 6
    sig_y, sig_c = 16.0, 8.0
    clip = avd.synth.clip(frames, W, H, depth, 303, chroma="420", sigma_luma=sig_y, sigma_chroma=sig_c)
    clean = avd.synth.clip(frames, W, H, depth, 303, chroma="420", sigma_luma=0.0, sigma_chroma=0.0)
    geom = L.Geometry(W, H, 1, L.SampleFormat.u16(10), L.SampleFormat.u16(10), 0)
    dims = [(W, H), (W // 2, H // 2), (W // 2, H // 2)]
    outs = {}
    for form in (0, 1):
        eng = avd.build_engine("nl4d", geom, luma=_opts(r, S, sig_y / 1023, step, 16, form),
                               chroma=_opts(r, S, sig_c / 1023, step, 16, form))
        outs[form] = run_engine(eng, [[words_dev(p.numpy(), depth) for p in f] for f in clip], depth, dims)
    pl = O.Nl4dPass(_spec(r, S, sig_y / 1023, step, 16), W, H, 1, depth, depth)
    pc = O.Nl4dPass(_spec(r, S, sig_c / 1023, step, 16), W // 2, H // 2, 2, depth, depth)
    O.link_motion(pl, pc, "420")
    want_l, _ = O.run_stream(pl, [[O.pack_plane(f[0].numpy(), depth)] for f in clip])
    want_c, _ = O.run_stream(pc, [[O.pack_plane(f[1].numpy(), depth), O.pack_plane(f[2].numpy(), depth)] for f in clip])
    for i in range(frames):
        assert np.array_equal(outs[0][i][0], want_l[i][0]), f"luma {i}"
        assert np.array_equal(outs[0][i][1], want_c[i][0]), f"U {i}"
        assert np.array_equal(outs[0][i][2], want_c[i][1]), f"V {i}"
    peak = 1023.0
    for k, (w, h) in enumerate(dims):
        ref = np.concatenate([clean[i][k].numpy().reshape(-1) for i in range(frames)])
        dec = {f: np.concatenate([O.unpack_plane(outs[f][i][k], depth, w * h) for i in range(frames)]) for f in (0, 1)}
        p32, pmx = _psnr(dec[0], ref, peak), _psnr(dec[1], ref, peak)
        print(f"plane {k}: FP32 {p32:.3f} dB, mixed {pmx:.3f} dB")
        assert abs(pmx - p32) <= 0.05, (k, p32, pmx)


@pytest.mark.parametrize("group_size,chroma", [(8, "gray"), (16, "gray"), (8, "420"), (16, "420")])
def test_overflowing_tile_lists_leave_the_output_unchanged(group_size, chroma):
    """A tile list of 256 entries overflows on these frames: the (member, tile) pairs past it take
    the global-atomic path, whose sums must equal the binned ones (the FP32 form then still matches
    the oracle bit for bit, and the mixed form matches its own run without overflow)."""
    W, H, depth, r, S, step = 96, 72, 10, 2, 8, 3 if group_size == 16 else 2
    frames = 5
    print(f"overflow case K={group_size} {chroma}", flush=True)
    clip = avd.synth.clip(frames, W, H, depth, 404 + group_size, chroma=chroma, sigma_luma=16.0, sigma_chroma=8.0)
    geom = L.Geometry(W, H, 0 if chroma == "gray" else 1, L.SampleFormat.u16(10), L.SampleFormat.u16(10), 0)
    dims = [(W, H)] if chroma == "gray" else [(W, H), (W // 2, H // 2), (W // 2, H // 2)]
    planes = [[words_dev(p.numpy(), depth) for p in f[:len(dims)]] for f in clip]
    sigma = 16.0 / 1023
    results = {}
    for form in (0, 1):
        for cap in (None, 256):
            eng = avd.build_engine("nl4d", geom, luma=_opts(r, S, sigma, step, group_size, form),
                                   chroma=_opts(r, S, sigma, step, group_size, form) if chroma != "gray" else None)
            if cap:
                eng.set_tile_cap_for_test(cap)
            results[(form, cap)] = run_engine(eng, planes, depth, dims)
            if cap:
                assert eng.tile_overflows_for_test() > 0, "the small lists overflowed"
    for form in (0, 1):
        for i in range(frames):
            for k in range(len(dims)):
                assert np.array_equal(results[(form, None)][i][k], results[(form, 256)][i][k]), (form, i, k)
    pl = O.Nl4dPass(_spec(r, S, sigma, step, group_size), W, H, 1, depth, depth)
    want_l, _ = O.run_stream(pl, [[O.pack_plane(f[0].numpy(), depth)] for f in clip])
    for i in range(frames):
        assert np.array_equal(results[(0, 256)][i][0], want_l[i][0]), f"luma {i}"
