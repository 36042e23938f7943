"""GPU parity: the HIP path through the C ABI against the oracle on the same seeded inputs.

Bar: integer and index work (codes, block-match vectors, group members, fixed-point
accumulators, NL4D output words) bit-exact; NLMeans sums within 1e-4 relative of the FP64
oracle and outputs within 1 LSB of the FP32 oracle.
"""
import ctypes

import numpy as np
import pytest
import torch

import avd
from avd import _lib as L

import oracle as O
import refsetups as R
from gpu_helpers import (assert_members_equal, collab_pass, dev, gray_geometry, host, ptr, require_gpu, run_engine,
                         site_collab_filter, site_collab_search, words_dev, words_host)

pytestmark = pytest.mark.gpu


@pytest.fixture(autouse=True)
def _gpu():
    require_gpu()


# ---- plane I/O ----------------------------------------------------------------------------

@pytest.mark.parametrize("depth,C,pixels", [(8, 1, 1001), (10, 1, 777), (12, 2, 515), (8, 2, 64), (16, 1, 33)])
def test_ingest_egress_match_the_oracle(depth, C, pixels):
    rng = np.random.default_rng(depth * 10 + C)
    codes = [rng.integers(0, 1 << depth, pixels) for _ in range(C)]
    words = [O.pack_plane(c, depth) for c in codes]
    fmt = L.SampleFormat.for_depth(depth)
    planes = [dev(w.view(np.uint8)) for w in words]
    slot = torch.zeros(pixels * C, dtype=torch.float32, device="cuda")
    cod = torch.zeros(pixels * C, dtype=torch.int16, device="cuda")
    L.check(avd.load().avd_site_ingest(ptr(planes[0]), ptr(planes[-1]), ptr(slot), ptr(cod), pixels, C, fmt, None))
    want = O.ingest(words, pixels, C, depth)
    assert np.array_equal(host(slot), want)
    assert np.array_equal(host(cod).view(np.uint16).astype(np.int64), np.stack(codes, 1).reshape(-1))
    outs = [torch.zeros_like(p) for p in planes]
    L.check(avd.load().avd_site_egress(ptr(slot), ptr(outs[0]), ptr(outs[-1]), pixels, C, fmt, None))
    for o, w in zip(outs, words):
        assert np.array_equal(words_host(o), w)


# ---- NLMeans window kernels ----------------------------------------------------------------

def _window(W, H, C, P, S, h2, off, step=8, bx=1, by=1):
    w = L.NlmWindow()
    w.width, w.height, w.channels, w.patch_radius, w.search_radius = W, H, C, P, S
    w.h2_inv_norm, w.noise_offset, w.step, w.blocks_x, w.blocks_y = h2, off, step, bx, by
    return w


@pytest.mark.parametrize("W,H,C,P,S", [(70, 45, 1, 2, 3), (64, 40, 2, 3, 2), (37, 29, 1, 0, 1), (90, 70, 1, 4, 2)])
def test_nlm_single_matches_the_oracle(W, H, C, P, S):
    rng = np.random.default_rng(W + C)
    f = rng.uniform(0, 1, W * H * C).astype(np.float32)
    h2 = float(O.h2_inv_norm(1.2, P))
    off = 0.01
    lut = O.spatial_offset_lut(S, off)
    acc = [np.zeros(W * H * C), np.zeros(W * H), np.zeros(W * H)]
    O.nlm_single(f, W, H, C, P, S, h2, lut, *acc, precision="f64")
    d_acc = torch.zeros(W * H * C, device="cuda")
    d_ws = torch.zeros(W * H, device="cuda")
    d_mw = torch.zeros(W * H, device="cuda")
    d_f, d_lut = dev(f), dev(lut)  # held for the launch
    L.check(avd.load().avd_site_nlm_single(ctypes.byref(_window(W, H, C, P, S, h2, off)), ptr(d_f), ptr(d_lut),
                                           ptr(d_acc), ptr(d_ws), ptr(d_mw), None))
    for got, want in zip((host(d_acc), host(d_ws), host(d_mw)), acc):
        assert np.allclose(got, want, rtol=1e-4, atol=1e-6), np.max(np.abs(got - want))


@pytest.mark.parametrize("S", [3, 1, 0])
@pytest.mark.parametrize("conf", [False, True])
@pytest.mark.parametrize("C", [1, 2])
def test_nlm_pair_matches_the_oracle(conf, C, S):
    # Search radius 0 and 1 leave some of a workgroup's offset parts without a row of offsets.
    W, H, P = 66, 50, 3
    rng = np.random.default_rng(3 + C)
    frames = [rng.uniform(0, 1, W * H * C).astype(np.float32) for _ in range(3)]
    h2 = float(O.h2_inv_norm(0.8, P))
    off = 0.02
    bx, by = (W + 7) // 8, (H + 7) // 8
    cf = rng.uniform(0, 1, bx * by).astype(np.float32) if conf else None
    cb = rng.uniform(0, 1, bx * by).astype(np.float32) if conf else None
    acc = [np.zeros(W * H * C), np.zeros(W * H), np.zeros(W * H)]
    O.nlm_pair(frames[0], frames[1], frames[2], W, H, C, P, S, h2, off, cf, cb, 8, bx, by, *acc, precision="f64")
    d = [torch.zeros(W * H * C, device="cuda"), torch.zeros(W * H, device="cuda"), torch.zeros(W * H, device="cuda")]
    keep = [dev(x) for x in frames] + ([dev(cf), dev(cb)] if conf else [None, None])  # held for the launch
    L.check(avd.load().avd_site_nlm_pair(ctypes.byref(_window(W, H, C, P, S, h2, off, 8, bx, by)), ptr(keep[0]),
                                         ptr(keep[1]), ptr(keep[2]), ptr(keep[3]), ptr(keep[4]), ptr(d[0]), ptr(d[1]),
                                         ptr(d[2]), None))
    for got, want in zip((host(x) for x in d), acc):
        assert np.allclose(got, want, rtol=1e-4, atol=1e-6), np.max(np.abs(got - want))


# ---- motion ------------------------------------------------------------------------------

def _moving_codes(W, H, depth, frames, seed):
    clip = avd.synth.clip(frames, W, H, depth, seed, chroma="gray", sigma_luma=4.0)
    return [f[0].numpy().astype(np.int32) for f in clip]


@pytest.mark.parametrize("levels,R", [(2, 4), (1, 3), (2, 8)])
def test_motion_matches_the_oracle(levels, R):
    W, H, depth = 96, 72, 10
    codes = _moving_codes(W, H, depth, 5, 11)
    ring = np.stack([c.reshape(-1) for c in codes]).astype(np.uint16)
    geom = O.MotionGeom(W, H, 16, 8, R, levels)
    centre, slots = 2, [0, 1, 3, 4]
    sigma = 4.0 / 1023
    floor, th = float(O.sad_noise_floor(16, sigma)), float(O.thsad(16, 1.0))
    mp = L.MotionPass()
    mp.width, mp.height, mp.channels, mp.blksize, mp.overlap = W, H, 1, 16, 8
    mp.search_radius, mp.pyramid_levels, mp.centre_slot, mp.neighbours = R, levels, centre, 4
    mp.max_value, mp.sad_floor, mp.thsad = 1023.0, floor, th
    blocks = geom.blocks
    mv = torch.zeros(4 * blocks * 2, dtype=torch.int32, device="cuda")
    conf = torch.zeros(4 * blocks, dtype=torch.float32, device="cuda")
    ns = dev(np.asarray(slots, np.uint32).view(np.int32))
    d_ring = dev(ring.view(np.int16))  # held for the launch
    L.check(avd.load().avd_site_motion(ctypes.byref(mp), ptr(d_ring), ptr(ns), ptr(mv), ptr(conf), None))
    got_mv, got_conf = host(mv), host(conf)
    pyr = [[c.reshape(-1), O.downscale_sum(c.reshape(-1), W, H)] for c in codes]
    for i, s in enumerate(slots):
        want_mv, want_conf = O.motion_search(geom, pyr[centre], pyr[s], 1023.0, floor, th)
        assert np.array_equal(got_mv[i * blocks * 2:(i + 1) * blocks * 2], want_mv), i
        assert np.array_equal(got_conf[i * blocks:(i + 1) * blocks], want_conf), i
    assert np.any(got_mv != 0)


# ---- NL4D grouping and filter ---------------------------------------------------------------

def test_collab_matches_the_recorded_reference_setups():
    """The reference's recorded f32 setups (collab/tests/fused/recorded.rs) through the HIP sites."""
    W = H = 64
    mv, conf, slots, bx, by = R.cross_frame_fields(W, H, 2)
    ring = R.unique_frame(W, H * 5).reshape(-1)
    geom = O.MotionGeom(W, H, 16, 8)
    p = collab_pass(W, H, 1, 2, 4, 2, 0.5, 5.3, 0.02, 0.0, 2, 5, geom)
    op = O.collab_params(W, H, 1, 2, 4, 2, 0.5, 5.3, 0.02, 0.0, geom)
    want = O.collab_search(ring, op, 2, slots, mv, conf)
    got = site_collab_search(p, values=dev(ring), nslots=slots, mv=mv, conf=conf)
    assert np.array_equal(got, want)
    acc, ws, gw = site_collab_filter(p, dev(ring), slots, got, 5)
    wacc = np.zeros(5 * W * H, np.int32)
    wws = np.zeros(5 * W * H, np.int32)
    wgw = O.collab_filter(ring, op, 2, slots, want, wacc, wws, "f32")
    assert np.array_equal(acc, wacc) and np.array_equal(ws, wws) and np.array_equal(gw, wgw)
    d = R.digest(acc, ws, gw)
    assert d["covered"] == 16453


def _codes_ring(planes, C):
    """Frames of (H, W) code planes as one interleaved u16 ring of shape (frames, H*W*C)."""
    return np.stack([np.stack([p.reshape(-1) for p in fr], 1).reshape(-1) for fr in planes]).astype(np.uint16)


def _oracle_fields(codes, W, H, C, centre, slots, sigma, max_value, geom):
    pyr = {s: [codes[s].reshape(-1, C)[:, 0].astype(np.int32),
               O.downscale_sum(codes[s].reshape(-1, C)[:, 0].astype(np.int32), W, H)] for s in [centre] + slots}
    floor, th = float(O.sad_noise_floor(geom.blksize, sigma)), float(O.thsad(geom.blksize, 1.0))
    mvs, confs = [], []
    for s in slots:
        m, c = O.motion_search(geom, pyr[centre], pyr[s], max_value, floor, th)
        mvs.append(m)
        confs.append(c)
    mv = np.concatenate(mvs) if mvs else np.zeros(2, np.int32)
    conf = np.concatenate(confs) if confs else np.ones(1, np.float32)
    return mv, conf


@pytest.mark.parametrize("case", [
    # (name, W, H, C, depth, radius, S, code_bits, content)
    ("odd width, narrow keys", 81, 45, 1, 10, 2, 9, 10, "scene"),
    ("odd width, wide keys", 81, 45, 1, 10, 2, 9, 0, "scene"),
    ("odd width chroma", 43, 37, 2, 10, 1, 7, 10, "scene"),
    ("radius 16 (run length 33)", 72, 64, 1, 10, 2, 16, 10, "scene"),
    ("radius 16 chroma", 70, 50, 2, 8, 1, 16, 8, "scene"),
    ("12-bit", 64, 48, 1, 12, 3, 12, 12, "scene"),
    ("12-bit chroma", 48, 40, 2, 12, 2, 6, 12, "scene"),
    ("flat with steps: ties", 64, 40, 1, 10, 2, 8, 10, "flat"),
    ("tiny frame", 9, 8, 1, 10, 1, 3, 10, "scene"),
    ("narrow frame", 8, 40, 2, 10, 2, 5, 10, "scene"),
    ("8-bit full range noise", 56, 48, 1, 8, 3, 12, 8, "noise"),
    # Radius 14 leaves the matrix-core centre search (two grid rows of candidates exceed 32 rows).
    ("radius 14, vector centre search", 64, 48, 1, 10, 2, 14, 10, "scene"),
    # Full-range 10-bit noise puts the best distances past the 32-bit key range of the matrix-core
    # centre search, whose references are then searched again with exact keys.
    ("10-bit full range noise, capped keys", 40, 32, 1, 10, 2, 2, 10, "noise"),
    ("10-bit chroma noise, capped keys", 36, 28, 2, 10, 1, 2, 10, "noise"),
    # 32-pixel blocks at step 8 cover an anchor with up to 16 blocks; at these radii the rectangle
    # tables only fit two references per wave.
    ("deep window, 16 covering blocks", 64, 48, 1, 10, 6, 6, 10, "scene", 32, 24),
    ("deep window chroma, 16 covering blocks", 48, 40, 2, 10, 5, 5, 10, "scene", 32, 24),
    ("radius 4, 9 covering blocks", 64, 48, 1, 10, 4, 7, 10, "scene", 24, 16),
])
def test_codes_search_edge_cases_are_bit_exact(case):
    """Member lists of the integer search against the oracle over the shapes and contents that
    exercise its paths: odd rows (unaligned loads), 32- and 64-bit keys, run lengths past the
    unrolled forms, clipped rectangles, equal distances resolved by visit order, and rectangle
    tables of every size (four or two references per wave)."""
    name, W, H, C, depth, radius, S, code_bits, content = case[:9]
    blksize, overlap = case[9:] if len(case) > 9 else (16, 8)
    print(f"search case: {name}", flush=True)
    frames = 2 * radius + 1
    rng = np.random.default_rng(W * 1000 + H + depth)
    top = (1 << depth) - 1
    if content == "scene":
        clip = avd.synth.clip(frames, W, H, depth, 31 + W, chroma="444" if C == 2 else "gray", sigma_luma=12.0)
        planes = [[f[0].numpy()] if C == 1 else [f[1].numpy(), f[2].numpy()] for f in clip]
    elif content == "flat":
        base = (np.arange(W)[None, :] // 16 * 37 + np.arange(H)[:, None] // 8 * 11) % top
        planes = [[np.broadcast_to(base, (H, W)).astype(np.int64)] for _ in range(frames)]
    else:
        planes = [[rng.integers(0, top + 1, (H, W)) for _ in range(C)] for _ in range(frames)]
    codes = _codes_ring(planes, C)
    geom = O.MotionGeom(W, H, blksize, overlap, 4, 2)
    centre = radius
    slots = [s for s in range(frames) if s != centre]
    sigma = 12.0 / 1023
    mv, conf = _oracle_fields(codes, W, H, C, centre, slots, sigma, float(top), geom)
    p = collab_pass(W, H, C, radius, S, 2, 0.05, 2.7, sigma, 2.0, centre, frames, geom, code_bits=code_bits)
    op = O.collab_params(W, H, C, radius, S, 2, 0.05, 2.7, sigma, 2.0, geom)
    want = O.collab_search(codes, op, centre, slots, mv, conf)
    d_codes = dev(codes.view(np.int16))
    got = site_collab_search(p, codes=d_codes, nslots=slots, mv=mv, conf=conf)
    assert_members_equal(got, want, name)


@pytest.mark.parametrize("C,step,S", [(1, 3, 12), (1, 8, 9), (2, 2, 6), (1, 1, 13)])
def test_codes_search_other_grid_steps_are_bit_exact(C, step, S):
    """Reference grids at other strides through the integer search (the matrix-core centre search
    takes every stride whose two candidate rows of references fit 32 rows)."""
    W, H, depth, radius = 72, 52, 10, 2
    print(f"search step {step} C {C} S {S}", flush=True)
    frames = 2 * radius + 1
    clip = avd.synth.clip(frames, W, H, depth, 17 + step, chroma="444" if C == 2 else "gray", sigma_luma=12.0)
    planes = [[f[0].numpy()] if C == 1 else [f[1].numpy(), f[2].numpy()] for f in clip]
    codes = _codes_ring(planes, C)
    geom = O.MotionGeom(W, H, 16, 8, 4, 2)
    slots = [s for s in range(frames) if s != radius]
    sigma = 12.0 / 1023
    mv, conf = _oracle_fields(codes, W, H, C, radius, slots, sigma, 1023.0, geom)
    p = collab_pass(W, H, C, radius, S, 2, 0.05, 2.7, sigma, 2.0, radius, frames, geom, code_bits=depth, step=step)
    op = O.collab_params(W, H, C, radius, S, 2, 0.05, 2.7, sigma, 2.0, geom, step=step)
    want = O.collab_search(codes, op, radius, slots, mv, conf)
    got = site_collab_search(p, codes=dev(codes.view(np.int16)), nslots=slots, mv=mv, conf=conf)
    assert_members_equal(got, want, f"step {step}")


@pytest.mark.parametrize("C,radius,S,pooled", [(1, 2, 9, False), (2, 1, 6, False), (1, 3, 12, False), (1, 0, 5, False),
                                                (1, 2, 9, True), (2, 2, 6, True), (1, 0, 5, True)])
def test_collab_codes_search_and_filter_are_bit_exact(C, radius, S, pooled):
    W, H, depth = 80, 56, 10
    frames = 2 * radius + 1
    clip = avd.synth.clip(frames, W, H, depth, 5 + C, chroma="444" if C == 2 else "gray", sigma_luma=12.0)
    planes = [[f[0].numpy()] if C == 1 else [f[1].numpy(), f[2].numpy()] for f in clip]
    codes = np.stack([np.stack([p.reshape(-1) for p in fr], 1).reshape(-1) for fr in planes]).astype(np.uint16)
    values = (codes.astype(np.float32) / np.float32(1023.0)).astype(np.float32)
    geom = O.MotionGeom(W, H, 16, 8, 4, 2)
    centre = radius
    slots = [s for s in range(frames) if s != centre]
    pyr = [[codes[s].reshape(-1, C)[:, 0].astype(np.int32),
            O.downscale_sum(codes[s].reshape(-1, C)[:, 0].astype(np.int32), W, H)] for s in range(frames)]
    sigma = 12.0 / 1023
    floor, th = float(O.sad_noise_floor(16, sigma)), float(O.thsad(16, 1.0))
    mvs, confs = [], []
    for s in slots:
        m, c = O.motion_search(geom, pyr[centre], pyr[s], 1023.0, floor, th)
        mvs.append(m)
        confs.append(c)
    mv = np.concatenate(mvs) if mvs else np.zeros(2, np.int32)
    conf = np.concatenate(confs) if confs else np.ones(1, np.float32)
    lam = 2.7
    p = collab_pass(W, H, C, radius, S, 2, 0.05, lam, sigma, 2.0, centre, frames, geom, code_bits=depth,
                    pooled=pooled)
    op = O.collab_params(W, H, C, radius, S, 2, 0.05, lam, sigma, 2.0, geom, pooled=pooled)
    want = O.collab_search(codes, op, centre, slots, mv, conf)
    d_codes = dev(codes.view(np.int16))
    got = site_collab_search(p, codes=d_codes, nslots=slots, mv=mv, conf=conf)
    assert np.array_equal(got, want)
    acc, ws, gw = site_collab_filter(p, dev(values), slots, got, frames)
    wacc = np.zeros(frames * W * H * C, np.int32)
    wws = np.zeros(frames * W * H, np.int32)
    wgw = O.collab_filter(values, op, centre, slots, want, wacc, wws, "f32")
    assert np.array_equal(acc, wacc)
    assert np.array_equal(ws, wws)
    assert np.array_equal(gw, wgw)
    if radius >= 2:
        assert np.any(got.reshape(-1, 9)[:, 8] & 0x100), "some groups build a volume grid"


# ---- engines ------------------------------------------------------------------------------

def _nl4d_spec(r, S, sigma, lam=2.7, pooled=False):
    return O.Nl4dSpec(temporal_radius=r, refine=2, spatial_radius=S, lambda_ht=lam, c_min=0.05, kaiser_beta=2.0,
                      sigma=sigma, motion=O.MotionGeom(0, 0, 16, 8, 4, 2), pooled=pooled)


@pytest.mark.parametrize("frames,r", [(7, 1), (9, 2), (2, 2), (5, 3)])
def test_nl4d_engine_stream_is_bit_exact(frames, r):
    W, H, depth = 72, 56, 10
    S = 7
    sigma = 16.0 / 1023
    clip = avd.synth.clip(frames, W, H, depth, 100 + r, chroma="gray", sigma_luma=16.0)
    codes = [f[0].numpy() for f in clip]
    opt = avd.nl4d_options(temporal_radius=r, spatial_radius=S, sigma=sigma, lambda_ht=2.7, pooled_threshold=0,
                           noise_map=0, motion={"estimation": 1}, field_lambda=0.0)
    eng = avd.build_engine("nl4d", gray_geometry(W, H, depth), luma=opt)
    got = run_engine(eng, [[words_dev(c, depth)] for c in codes], depth, [(W, H)])
    p = O.Nl4dPass(_nl4d_spec(r, S, sigma), W, H, 1, depth, depth)
    want, _ = O.run_stream(p, [[O.pack_plane(c, depth)] for c in codes])
    assert len(got) == len(want) == frames
    for i, (g, w) in enumerate(zip(got, want)):
        assert np.array_equal(g[0], w[0]), f"frame {i}"


def test_nl4d_engine_420_luma_and_chroma_are_bit_exact():
    W, H, depth, r, S = 64, 48, 10, 2, 6
    sigma = 16.0 / 1023
    frames = 6
    clip = avd.synth.clip(frames, W, H, depth, 77, chroma="420", sigma_luma=16.0, sigma_chroma=8.0)
    geom = L.Geometry(W, H, 1, L.SampleFormat.u16(10), L.SampleFormat.u16(10), 0)
    opt = avd.nl4d_options(temporal_radius=r, spatial_radius=S, sigma=sigma, lambda_ht=2.7, pooled_threshold=0,
                           noise_map=0, motion={"estimation": 1}, field_lambda=0.0)
    eng = avd.build_engine("nl4d", geom, luma=opt, chroma=opt)
    dims = [(W, H), (W // 2, H // 2), (W // 2, H // 2)]
    got = run_engine(eng, [[words_dev(p.numpy(), depth) for p in f] for f in clip], depth, dims)
    pl = O.Nl4dPass(_nl4d_spec(r, S, sigma), W, H, 1, depth, depth)
    pc = O.Nl4dPass(_nl4d_spec(r, S, sigma), W // 2, H // 2, 2, depth, depth)
    O.link_motion(pl, pc, "420")  # the chroma pass refines the luma pass's motion
    want_l, _ = O.run_stream(pl, [[O.pack_plane(f[0].numpy(), depth)] for f in clip])
    want_c, _ = O.run_stream(pc, [[O.pack_plane(f[1].numpy(), depth), O.pack_plane(f[2].numpy(), depth)] for f in clip])
    assert len(got) == frames
    for i in range(frames):
        assert np.array_equal(got[i][0], want_l[i][0]), f"luma {i}"
        assert np.array_equal(got[i][1], want_c[i][0]), f"U {i}"
        assert np.array_equal(got[i][2], want_c[i][1]), f"V {i}"


@pytest.mark.parametrize("chroma", ["gray", "420"])
def test_nl4d_engine_with_pooled_threshold_is_bit_exact(chroma):
    """The reference's default threshold (pooled) through the engine, luma and 4:2:0 chroma."""
    W, H, depth, r, S = 64, 48, 10, 2, 6
    sigma = 16.0 / 1023
    frames = 6
    clip = avd.synth.clip(frames, W, H, depth, 91, chroma=chroma, sigma_luma=16.0, sigma_chroma=8.0)
    geom = L.Geometry(W, H, 0 if chroma == "gray" else 1, L.SampleFormat.u16(10), L.SampleFormat.u16(10), 0)
    opt = avd.nl4d_options(temporal_radius=r, spatial_radius=S, sigma=sigma, pooled_threshold=1, noise_map=0,
                           motion={"estimation": 1}, field_lambda=0.0)
    eng = avd.build_engine("nl4d", geom, luma=opt, chroma=opt if chroma != "gray" else None)
    dims = [(W, H)] if chroma == "gray" else [(W, H), (W // 2, H // 2), (W // 2, H // 2)]
    got = run_engine(eng, [[words_dev(p.numpy(), depth) for p in f[:len(dims)]] for f in clip], depth, dims)
    pl = O.Nl4dPass(_nl4d_spec(r, S, sigma, lam=4.158, pooled=True), W, H, 1, depth, depth)
    pc = O.Nl4dPass(_nl4d_spec(r, S, sigma, lam=3.234, pooled=True), W // 2, H // 2, 2, depth, depth)
    if chroma != "gray":
        O.link_motion(pl, pc, chroma)
    want_l, _ = O.run_stream(pl, [[O.pack_plane(f[0].numpy(), depth)] for f in clip])
    for i in range(frames):
        assert np.array_equal(got[i][0], want_l[i][0]), f"luma {i}"
    if chroma != "gray":
        want_c, _ = O.run_stream(pc, [[O.pack_plane(f[1].numpy(), depth), O.pack_plane(f[2].numpy(), depth)]
                                      for f in clip])
        for i in range(frames):
            assert np.array_equal(got[i][1], want_c[i][0]), f"U {i}"
            assert np.array_equal(got[i][2], want_c[i][1]), f"V {i}"


def _psnr(a, b, peak):
    mse = float(np.mean((a.astype(np.float64) - b.astype(np.float64)) ** 2))
    return 99.0 if mse == 0 else 10.0 * np.log10(peak * peak / mse)


MIXED_CASES = [
    # name, W, H, depth, temporal radius, S, frames, seed, luma / chroma sigma in codes, chroma
    ("r2_96x64_10bit", 96, 64, 10, 2, 8, 7, 123, 16.0, 8.0, "420"),  # grid of 4 frames per volume
    ("r1_70x46_8bit", 70, 46, 8, 1, 6, 6, 77, 5.0, 3.0, "420"),  # grid of 2, ragged odd chroma, 8-bit
    ("r3_80x56_10bit", 80, 56, 10, 3, 8, 8, 31, 12.0, 6.0, "420"),
    ("r2_64x48_422", 64, 48, 10, 2, 7, 6, 5, 14.0, 7.0, "422"),
    ("r1_56x40_444", 56, 40, 10, 1, 7, 5, 6, 14.0, 7.0, "444"),
    ("r2_72x56_gray", 72, 56, 10, 2, 8, 6, 9, 16.0, 0.0, "gray"),
    ("r2_64x48_12bit", 64, 48, 12, 2, 8, 6, 41, 64.0, 32.0, "420"),  # codes above 11 bits round to f16
]
LAYOUT = {"gray": 0, "420": 1, "422": 2, "444": 3}


@pytest.mark.parametrize("case", MIXED_CASES, ids=[c[0] for c in MIXED_CASES])
def test_nl4d_mixed_transform_form_stays_within_0_05_db(case):
    """The mixed form (f16 matrix-core transforms, f32 accumulation) against the FP32 form on the
    same 4:2:0 clip: PSNR against the clean scene within 0.05 dB on every plane, under 1 % of the
    samples more than 2 codes apart and none more than 32 (codes in 10-bit units at 12 bits), and the FP32 form still bit-exact with
    the oracle.  (A coefficient within the DCT table's f16 rounding of the threshold can flip, which
    also moves a sparse group's weight: at frame borders, where one to four groups cover a pixel,
    that shows as a few codes; without the hi/lo split of the Haar inputs 1.2 % of the r = 3 chroma
    samples moved more than 2 codes, with it 0.3 %.)"""
    name, W, H, depth, r, S, frames, seed, sig_y, sig_c, chroma = case
    print(f"mixed-form case {name}", flush=True)
    peak = float((1 << depth) - 1)
    noisy = avd.synth.clip(frames, W, H, depth, seed, chroma=chroma, sigma_luma=sig_y, sigma_chroma=sig_c)
    clean = avd.synth.clip(frames, W, H, depth, seed, chroma=chroma, sigma_luma=0.0, sigma_chroma=0.0)
    fmt = L.SampleFormat.u8() if depth == 8 else L.SampleFormat.u16(depth)
    geom = L.Geometry(W, H, LAYOUT[chroma], fmt, fmt, 0)
    cd = avd.chroma_dims(LAYOUT[chroma], W, H)
    dims = [(W, H)] + ([cd, cd] if cd else [])
    outs = {}
    for form in (0, 1):
        common = dict(temporal_radius=r, spatial_radius=S, lambda_ht=2.7, pooled_threshold=0, noise_map=0,
                      motion={"estimation": 1}, field_lambda=0.0, transform_form=form)
        eng = avd.build_engine("nl4d", geom, luma=avd.nl4d_options(sigma=sig_y / peak, **common),
                               chroma=avd.nl4d_options(sigma=sig_c / peak, **common) if cd else None)
        outs[form] = run_engine(eng, [[words_dev(p.numpy(), depth) for p in f] for f in noisy], depth, dims)
    pl = O.Nl4dPass(_nl4d_spec(r, S, sig_y / peak), W, H, 1, depth, depth)
    want_l, _ = O.run_stream(pl, [[O.pack_plane(f[0].numpy(), depth)] for f in noisy])
    for i in range(frames):
        assert np.array_equal(outs[0][i][0], want_l[i][0]), f"FP32 form, luma {i}"
    for k, (w, h) in enumerate(dims):
        ref = np.concatenate([clean[i][k].numpy().reshape(-1) for i in range(frames)])
        dec = {f: np.concatenate([O.unpack_plane(outs[f][i][k], depth, w * h) for i in range(frames)])
               for f in (0, 1)}
        p32, pmx = _psnr(dec[0], ref, peak), _psnr(dec[1], ref, peak)
        diff = np.abs(dec[0] - dec[1])
        # Code distances in 10-bit units: a 12-bit code step is a quarter of a 10-bit one.
        diff = diff / float(max(1 << (depth - 10), 1)) if depth > 10 else diff
        far = float(np.mean(diff > 2))
        print(f"plane {k}: FP32 {p32:.3f} dB, mixed {pmx:.3f} dB, max |diff| {float(diff.max())}, "
              f"share > 2 codes {far:.5f}")
        assert abs(pmx - p32) <= 0.05, (k, p32, pmx)
        assert far < 1e-2 and float(diff.max()) <= 32, (k, far, float(diff.max()))
        if depth > 8:
            assert not np.array_equal(dec[0], dec[1]) or k > 0, "the mixed form ran"


def test_nl4d_chroma_passes_through_when_only_luma_is_cleaned():
    W, H, depth = 48, 40, 8
    clip = avd.synth.clip(5, W, H, depth, 3, chroma="420", sigma_luma=6.0)
    geom = L.Geometry(W, H, 1, L.SampleFormat.u8(), L.SampleFormat.u8(), 0)
    opt = avd.nl4d_options(temporal_radius=1, spatial_radius=5, sigma=6 / 255, pooled_threshold=0, noise_map=0,
                           motion={"estimation": 1}, field_lambda=0.0)
    eng = avd.build_engine("nl4d", geom, luma=opt)
    dims = [(W, H), (W // 2, H // 2), (W // 2, H // 2)]
    got = run_engine(eng, [[words_dev(p.numpy(), depth) for p in f] for f in clip], depth, dims)
    for i, f in enumerate(clip):
        assert np.array_equal(got[i][1], O.pack_plane(f[1].numpy(), depth))
        assert np.array_equal(got[i][2], O.pack_plane(f[2].numpy(), depth))


def _max_lsb(a_words, b_words, depth, pixels):
    a = O.unpack_plane(a_words, depth, pixels)
    b = O.unpack_plane(b_words, depth, pixels)
    return int(np.max(np.abs(a - b)))


@pytest.mark.parametrize("case", ["spatial8", "temporal10", "hq_motion10", "hq_confidence_only10"])
def test_nlm_engine_stream_within_one_lsb(case):
    if case == "spatial8":
        W, H, depth, frames = 80, 48, 8, 3
        opt = avd.nlmeans_options(search_radius=7, patch_radius=2)
        spec = O.NlmSpec(0, 7, 2, 1.2, 1.0)
        sig = 10.0
    elif case == "temporal10":
        W, H, depth, frames = 72, 40, 10, 5
        opt = avd.nlmeans_options(temporal=1, temporal_radius=2, search_radius=4, patch_radius=3)
        spec = O.NlmSpec(2, 4, 3, 1.2, 1.0)
        sig = 12.0
    elif case == "hq_confidence_only10":
        # HQ confidence with motion compensation off: blocks scored where they stand, frames unwarped.
        W, H, depth, frames = 72, 56, 10, 5
        sigma = 12.0 / 1023
        opt = avd.nlmeans_options(hq=1, temporal=1, temporal_radius=2, search_radius=3, patch_radius=2, sigma=sigma,
                                  motion_compensation=0)
        spec = O.NlmSpec(2, 3, 2, O.HQ_STRENGTH_LUMA[2], 1.0, hq=True, sigma=sigma)
        sig = 12.0
    else:
        W, H, depth, frames = 80, 64, 10, 6
        sigma = 12.0 / 1023
        opt = avd.nlmeans_options(hq=1, temporal=1, temporal_radius=2, search_radius=3, patch_radius=2, sigma=sigma,
                                  motion_compensation=1, motion={"estimation": 1})
        spec = O.NlmSpec(2, 3, 2, O.HQ_STRENGTH_LUMA[2], 1.0, hq=True, sigma=sigma,
                         motion=O.MotionGeom(0, 0, 16, 8, 4, 2))
        sig = 12.0
    clip = avd.synth.clip(frames, W, H, depth, 9, chroma="gray", sigma_luma=sig)
    codes = [f[0].numpy() for f in clip]
    eng = avd.build_engine("nlmeans", gray_geometry(W, H, depth), luma=opt)
    got = run_engine(eng, [[words_dev(c, depth)] for c in codes], depth, [(W, H)])
    want, _ = O.run_stream(O.NlmPass(spec, W, H, 1, depth, depth), [[O.pack_plane(c, depth)] for c in codes])
    assert len(got) == len(want) == frames
    for i in range(frames):
        assert _max_lsb(got[i][0], want[i][0], depth, W * H) <= 1, i


def test_engine_protocol_errors():
    W, H, depth = 32, 24, 8
    opt = avd.nlmeans_options()
    eng = avd.build_engine("nlmeans", gray_geometry(W, H, depth), luma=opt)
    plane = words_dev(np.full((H, W), 100), depth)
    out = torch.zeros_like(plane)
    with pytest.raises(avd.NothingToEmit):
        eng.emit_into([out])
    assert eng.push([plane]) == 1
    with pytest.raises(avd.OutputsPending):
        eng.push([plane])
    eng.emit_into([out])
    with pytest.raises(avd.ContextAfterPush):
        eng.push_context([plane])
    with pytest.raises(avd.PlaneMismatch):
        eng.push([torch.zeros(8, dtype=torch.uint8, device="cuda")])
    eng.fail_for_test()
    with pytest.raises(avd.NeedsReset):
        eng.push([plane])
    eng.reset()
    assert eng.push([plane]) == 1
    eng.emit_into([out])
    assert np.array_equal(words_host(out), words_host(plane))
