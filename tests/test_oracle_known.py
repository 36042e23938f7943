"""Known answers, closed forms and invariants that pin the oracle where no recorded digest does.

Each test names the reference test or formula it follows.  A plausible slip in the oracle (a
wrong tap, a swapped Haar operand, a dropped tie rule) fails at least one of them.
"""
import math

import numpy as np
import pytest

import oracle as O


# --- geometry (collab/geometry.rs:6-29, tests at :91-159) ------------------------------------

def test_refs_cover_1080p_exactly():
    assert O.refs_along(1920) == 479
    assert O.refs_along(1080) == 269


def test_last_ref_clamps_inside_the_frame():
    n = O.refs_along(21)
    assert O.ref_pos(n - 1, 21) == 13
    assert all(O.ref_pos(i, 21) + 8 <= 21 for i in range(n))


@pytest.mark.parametrize("dim", [8, 9, 16, 21, 64])
def test_every_pixel_covered_one_to_three_times(dim):
    n = O.refs_along(dim)
    for x in range(dim):
        cover = sum(1 for i in range(n) if O.ref_pos(i, dim) <= x < O.ref_pos(i, dim) + 8)
        assert 1 <= cover <= 3


# --- aggregation constants (collab/kernels/aggregate.rs:305-444) ------------------------------

def test_to_fixed_rounds_rather_than_truncating():
    L = O.lib()
    L.avdo_to_fixed_f32.restype = L.avdo_to_fixed_weight_f32.restype = __import__("ctypes").c_int32
    import ctypes
    f = lambda v, s: L.avdo_to_fixed_f32(ctypes.c_float(v), ctypes.c_float(s))
    fw = lambda v, s: L.avdo_to_fixed_weight_f32(ctypes.c_float(v), ctypes.c_float(s))
    assert (f(0.6, 1.0), f(0.4, 1.0), f(-0.6, 1.0)) == (1, 0, -1)
    assert f(0.3 / 65536.0, 65536.0) == 0 and fw(0.3 / 65536.0, 65536.0) == 2
    assert fw(4.0, 1024.0) == fw(1.0, 1024.0) and fw(-1.0, 1024.0) == 0


def test_cross_frame_scale_keeps_every_radius_under_budget():
    budget = (2**31 - 1) / 2.0
    for S in range(1, 17):
        for r in range(1, 9):
            per_axis = ((8 - 1) + 2 * S) // 4 + 1
            worst = per_axis * per_axis * 8 * (4 * r + 1) * 5.0 * float(O.cross_frame_accum_scale(S, r))
            assert 0 < worst <= budget
    assert O.cross_frame_accum_scale(9, 2) >= 32768.0
    assert (((8 - 1) + 2 * 9) // 4 + 1) ** 2 * 8 == 392


def test_kaiser_window_known_answers():
    assert np.array_equal(O.kaiser_window(0.0), np.ones(8, np.float32))
    for beta in (1.0, 2.0, 4.0, 8.0):
        w = O.kaiser_window(beta)
        for i in range(4):
            assert abs(w[i] - w[7 - i]) < 1e-6
            if i < 3:
                assert w[i + 1] > w[i]
        assert np.all((w > 0) & (w <= 1))
    for beta in (1.0, 2.0, 4.0):
        i0 = sum(((beta / 2) ** k / math.factorial(k)) ** 2 for k in range(32))
        assert abs(O.kaiser_window(beta)[0] - 1.0 / i0) < 1e-6
    assert abs(O.kaiser_window(2.0)[0] - 0.4388) < 1e-3


# --- transforms (collab/kernels/transforms.rs, collab/tests/transforms.rs) --------------------

def _basis():
    return np.array([[O.dct8_basis(r, c) for c in range(8)] for r in range(8)])


def test_dct8_matches_the_f64_basis_and_inverts():
    import ctypes
    rng = np.random.default_rng(1)
    B = _basis()
    for _ in range(20):
        x = rng.uniform(-1, 1, 8).astype(np.float32)
        line = x.copy()
        O.lib().avdo_dct8_fwd_f32(line.ctypes.data_as(ctypes.POINTER(ctypes.c_float)))
        assert np.max(np.abs(line - B @ x.astype(np.float64))) < 1e-6
        O.lib().avdo_dct8_inv_f32(line.ctypes.data_as(ctypes.POINTER(ctypes.c_float)))
        assert np.max(np.abs(line - x)) < 1e-6


def test_dct_noise_profile_known_answers():
    assert np.array_equal(O.dct_noise_profile(0.0), np.ones(8, np.float32))
    assert np.array_equal(O.dct_noise_profile(-0.1), np.ones(8, np.float32))
    for rho in (0.05, 0.3, 0.5, 0.67, 0.8, 0.85, 0.86, 0.95, 0.99):
        g = O.dct_noise_profile(rho)
        assert abs(float(g.sum()) - 8.0) < 1e-3
        assert np.all(g[:-1] > g[1:])


def _ladder(sig2, k):
    """The reference's reordering variance ladder (transforms.rs:280-297 host mirror)."""
    v = list(sig2)
    n = k
    while n > 1:
        half = n // 2
        snap = v[:n]
        for p in range(half):
            m = (snap[2 * p] + snap[2 * p + 1]) / 2.0
            v[p] = m
            v[half + p] = m
        n = half
    return v


def test_variance_ladder_known_answer_and_uniform_case():
    got = _ladder([1.0, 3.0, 2.0, 2.0, 5.0, 1.0, 4.0, 0.0], 8)
    assert np.allclose(got, [2.25, 2.25, 2.0, 2.5, 2.0, 2.0, 3.0, 2.0], atol=1e-6)
    for k in (1, 2, 4, 8):
        assert np.allclose(_ladder([0.3] * k, k), [0.3] * k, atol=1e-7)


def _ref_stack_haar(stack, k):
    """haar_reg_fwd_level over len = k, k/2, .. 2 with the reference's reordering (transforms.rs:122-147)."""
    s = stack.copy()
    n = k
    f = np.float32(0.70710678118654752440)
    while n >= 2:
        half = n // 2
        snap = s[:n].copy()
        for p in range(half):
            s[p] = (snap[2 * p] + snap[2 * p + 1]) * f
            s[half + p] = (snap[2 * p] - snap[2 * p + 1]) * f
        n = half
    return s


def _ref_grid(stack, gf):
    """grid_fwd (fused/grid.rs:119-152) with the reference's strided reordering."""
    s = stack.copy()
    f = np.float32(0.70710678118654752440)
    vols = 8 // gf

    def strided(start, stride, length):
        levels = [lv for lv in (4, 2) if lv <= length]
        for lv in levels:
            half = lv // 2
            snap = [s[start + k * stride].copy() for k in range(lv)]
            for p in range(half):
                s[start + p * stride] = (snap[2 * p] + snap[2 * p + 1]) * f
                s[start + (half + p) * stride] = (snap[2 * p] - snap[2 * p + 1]) * f

    for v in range(vols):
        strided(v * gf, 1, gf)
    for fr in range(gf):
        strided(fr, gf, vols)
    return s


@pytest.mark.parametrize("mode", [("stack", 2), ("stack", 4), ("stack", 8), ("grid", 2), ("grid", 4)])
def test_in_place_group_haar_is_the_reference_transform_permuted(mode):
    """The oracle's in-place butterflies give the reference's coefficients bit for bit, slot 0 alike."""
    import ctypes
    kind, n = mode
    rng = np.random.default_rng(7)
    stk = rng.uniform(-1, 1, (8, 64)).astype(np.float32)
    mine = stk.copy()
    if kind == "stack":
        O.lib().avdo_group_fwd_f32(mine.ctypes.data_as(ctypes.POINTER(ctypes.c_float)), n, 0, 1)
        ref = _ref_stack_haar(stk, n)
        rows = n
    else:
        O.lib().avdo_group_fwd_f32(mine.ctypes.data_as(ctypes.POINTER(ctypes.c_float)), 8, 1, n)
        ref = _ref_grid(stk, n)
        rows = 8
    assert np.array_equal(mine[0], ref[0])
    mine_rows = sorted(tuple(r) for r in mine[:rows])
    ref_rows = sorted(tuple(r) for r in ref[:rows])
    assert mine_rows == ref_rows
    back = mine.copy()
    if kind == "stack":
        O.lib().avdo_group_inv_f32(back.ctypes.data_as(ctypes.POINTER(ctypes.c_float)), n, 0, 1)
    else:
        O.lib().avdo_group_inv_f32(back.ctypes.data_as(ctypes.POINTER(ctypes.c_float)), 8, 1, n)
    assert np.max(np.abs(back[:rows] - stk[:rows])) < 1e-6


# --- motion (motion/confidence.rs tests, motion/kernels/tests/argmin.rs rules) ----------------

def test_confidence_closed_forms():
    sigma = 4.0 / 255.0
    assert abs(float(O.sad_noise_floor(16, sigma)) - 256 * 2.0 * sigma / math.sqrt(math.pi)) < 1e-6
    assert float(O.sad_noise_floor(16, 0.0)) == 0.0
    assert abs(float(O.thsad(16, 2.0)) - 2.0 * 256 * 0.02) < 1e-6
    import ctypes
    O.lib().avdo_match_confidence.restype = ctypes.c_float
    mc = lambda s, f, t: O.lib().avdo_match_confidence(ctypes.c_float(s), ctypes.c_float(f), ctypes.c_float(t))
    assert mc(1.0, 2.0, 5.12) == 1.0       # below the floor: full confidence
    assert mc(2.0 + 5.12, 2.0, 5.12) == 0.0  # excess == thsad: zero
    assert abs(mc(3.0, 2.0, 2.0) - 0.6) < 1e-6


def _serial_block_sads(cen, nb, ox, oy, blk, R, cdx, cdy):
    h, w = cen.shape
    side = 2 * R + 1
    sads = np.zeros(side * side, np.int64)
    for dy in range(side):
        for dx in range(side):
            s = 0
            for py in range(blk):
                for px in range(blk):
                    cx, cy = min(max(ox + px, 0), w - 1), min(max(oy + py, 0), h - 1)
                    nx = min(max(ox + cdx - R + dx + px, 0), w - 1)
                    ny = min(max(oy + cdy - R + dy + py, 0), h - 1)
                    s += abs(int(cen[cy, cx]) - int(nb[ny, nx]))
            sads[dy * side + dx] = s
    return sads


def _expected(sads, R, seed):
    side = 2 * R + 1
    idx = int(np.argmin(sads))  # first index of the minimum
    best = (seed[0] + idx % side - R, seed[1] + idx // side - R)
    return seed if sads[R * side + R] <= sads[idx] else best


@pytest.mark.parametrize("content", ["flat", "periodic", "textured"])
def test_fine_block_match_follows_the_serial_scan(content):
    W, H, R = 40, 27, 4
    rng = np.random.default_rng(3)
    if content == "flat":
        cen = np.full((H, W), 200, np.int32)
        nb = np.full((H, W), 90, np.int32)
    elif content == "periodic":
        tile = rng.integers(0, 1024, (4, 4))
        cen = np.tile(tile, (H // 4 + 2, W // 4 + 2))[:H, :W].astype(np.int32)
        nb = np.roll(np.roll(np.tile(tile, (H // 4 + 2, W // 4 + 2)), 1, 0), 2, 1)[:H, :W].astype(np.int32)
    else:
        cen = rng.integers(0, 1024, (H, W)).astype(np.int32)
        nb = np.roll(np.roll(cen, -2, 0), 3, 1) + rng.integers(-2, 3, (H, W)).astype(np.int32)
    bx, by = (W + 7) // 8, (H + 7) // 8
    mv = np.zeros(bx * by * 2, np.int32)
    O.lib().avdo_block_match_fine(cen.ctypes.data_as(O.I32P), nb.ctypes.data_as(O.I32P), W, H, 16, 8, R, bx, by, 0,
                                  mv.ctypes.data_as(O.I32P), None, __import__("ctypes").c_float(1023.0),
                                  __import__("ctypes").c_float(0.0), __import__("ctypes").c_float(5.12))
    for b in range(bx * by):
        sads = _serial_block_sads(cen, nb, (b % bx) * 8, (b // bx) * 8, 16, R, 0, 0)
        assert tuple(mv[2 * b:2 * b + 2]) == _expected(sads, R, (0, 0)), b
    if content == "flat":
        assert not mv.any()
    if content == "periodic":
        assert any(np.sum(_serial_block_sads(cen, nb, 8, 8, 16, R, 0, 0) == 0) > 1 for _ in [0])


# --- plane I/O and invariants -----------------------------------------------------------------

@pytest.mark.parametrize("depth,pixels", [(8, 37), (10, 33), (12, 64), (16, 7)])
def test_ingest_egress_round_trip(depth, pixels):
    rng = np.random.default_rng(depth)
    codes = rng.integers(0, 1 << depth, pixels)
    words = O.pack_plane(codes, depth)
    f = O.ingest([words], pixels, 1, depth)
    assert np.array_equal(f, (codes.astype(np.float32) / np.float32(O.max_value(depth))))
    back = O.egress(f, pixels, 1, depth)[0]
    assert np.array_equal(back, words)
    assert np.array_equal(O.unpack_plane(back, depth, pixels), codes)


def test_nlm_uniform_frame_and_zero_radius_pass_through():
    W, H = 24, 17
    f = np.full(W * H, 0.37, np.float32)
    for S in (0, 2):
        acc = np.zeros(W * H, np.float32)
        ws = np.zeros(W * H, np.float32)
        mw = np.zeros(W * H, np.float32)
        O.nlm_single(f, W, H, 1, 2, S, float(O.h2_inv_norm(1.2, 2)), O.spatial_offset_lut(S, 0.0), acc, ws, mw)
        out = O.nlm_finish(f, acc, ws, mw, 1.0, W * H, 1)
        assert np.allclose(out, f, atol=1e-6)


def test_nlm_weights_are_symmetric_in_the_window():
    """fused.rs:349-357: a frame against itself weighs offset q and -q alike, so a mirrored frame
    gives the mirrored result."""
    rng = np.random.default_rng(5)
    W, H = 20, 14
    f = rng.uniform(0, 1, (H, W)).astype(np.float32)
    h2 = float(O.h2_inv_norm(1.2, 1))

    def run(img):
        acc = np.zeros(W * H, np.float64)
        ws = np.zeros(W * H, np.float64)
        mw = np.zeros(W * H, np.float64)
        flat = np.ascontiguousarray(img).reshape(-1)
        O.nlm_single(flat, W, H, 1, 1, 2, h2, O.spatial_offset_lut(2, 0.0), acc, ws, mw, "f64")
        return O.nlm_finish(flat, acc, ws, mw, 1.0, W * H, 1, "f64").reshape(H, W)

    a = run(f)
    b = run(f[:, ::-1])[:, ::-1]
    assert np.max(np.abs(a - b)) < 1e-12


def test_temporal_nlm_with_identical_frames_matches_the_pair_count():
    """With every ring frame equal, a pair pass scores the same distances as the single window, so the
    weights gathered per pixel are exactly (2 * pairs + 1) times the spatial set plus the q=0 terms."""
    rng = np.random.default_rng(9)
    W, H = 16, 12
    f = rng.uniform(0, 1, W * H).astype(np.float32)
    h2 = float(O.h2_inv_norm(1.2, 1))
    z = lambda: (np.zeros(W * H), np.zeros(W * H), np.zeros(W * H))
    a1, w1, m1 = z()
    O.nlm_single(f, W, H, 1, 1, 2, h2, O.spatial_offset_lut(2, 0.0), a1, w1, m1, "f64")
    a2, w2, m2 = z()
    O.nlm_pair(f, f, f, W, H, 1, 1, 2, h2, 0.0, None, None, 8, 2, 2, a2, w2, m2, "f64")
    # The pair window includes q = 0 (weight 1) twice and every other offset in both directions.
    assert np.allclose(w2, 2 * w1 + 2.0, atol=1e-9)


def test_nl4d_uniform_codes_pass_through():
    W, H, depth = 40, 32, 10
    spec = O.Nl4dSpec(temporal_radius=1, spatial_radius=4, sigma=8 / 1023.0, lambda_ht=2.7)
    p = O.Nl4dPass(spec, W, H, 1, depth, depth)
    codes = np.full((H, W), 517, np.int32)
    words = O.pack_plane(codes, depth)
    outs, _ = O.run_stream(p, [[words]] * 4)
    assert len(outs) == 4
    for w in outs:
        assert np.array_equal(w[0], words)
