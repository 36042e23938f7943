"""The oracle's luma-to-chroma motion: the field map, the seeded refine and the linked passes.

The map cases are the reference's known answers (motion/tests/chroma_map.rs:94-260: 16x16 blocks
at step 8, identity at 4:4:4, halving, the footprint's per-component lower median, rounding half
away from zero, padded strides, ragged edges, an even count of live vectors).  Random fields are
checked against a numpy restatement that sorts instead of ranking.  The refine is pinned against
a brute-force SAD argmin around each seed.  The linked passes follow nl4d/tests/luma_motion.rs: on
a panning luma clip with flat (noise-only) chroma, the chroma pass's own search misses the pan
while the refined luma motion follows it on every block that stays inside the frame.
"""
from __future__ import annotations

import numpy as np
import pytest

import oracle as O

STEP, BLK = 8, 16


def _geom(w, h):
    return O.MotionGeom(w, h, BLK, BLK - STEP, 4, 2)


def _fields(g: O.MotionGeom, fields, vector_of):
    out = []
    for f in range(fields):
        for row in range(g.blocks_y):
            for col in range(g.blocks_x):
                out += list(vector_of(f, col, row))
    return np.asarray(out, np.int32)


def _at(mapped, g: O.MotionGeom, col, row, field=0):
    i = field * g.blocks * 2 + (row * g.blocks_x + col) * 2
    return int(mapped[i]), int(mapped[i + 1])


def test_four_four_four_is_the_identity():
    luma = _geom(48, 40)
    values = _fields(luma, 1, lambda f, c, r: (c * 3 - 7, r * -5 + 2))
    assert np.array_equal(O.map_luma_fields(values, luma, luma, 0, 0, 1), values)


def test_four_two_zero_halves_both_axes():
    luma, chroma = _geom(64, 64), _geom(32, 32)
    mapped = O.map_luma_fields(_fields(luma, 1, lambda f, c, r: (12, -20)), luma, chroma, 1, 1, 1)
    assert np.array_equal(mapped.reshape(-1, 2), np.tile([6, -10], (chroma.blocks, 1)))


def test_four_two_two_halves_x_only():
    luma, chroma = _geom(64, 64), _geom(32, 64)
    mapped = O.map_luma_fields(_fields(luma, 1, lambda f, c, r: (12, r * 10 - 30)), luma, chroma, 1, 0, 1)
    for row in range(chroma.blocks_y):
        for col in range(chroma.blocks_x):
            assert _at(mapped, chroma, col, row) == (6, row * 10 - 30)


def test_the_footprint_median_is_per_component():
    luma, chroma = _geom(64, 64), _geom(32, 32)
    fx = [[2, 4, 6], [8, 10, 12], [14, 16, 18]]
    fy = [[-40, 4, 8], [-8, 100, 0], [12, -4, 6]]
    values = _fields(luma, 1, lambda f, c, r: (fx[r][c], fy[r][c]) if c < 3 and r < 3 else (500, 500))
    assert _at(O.map_luma_fields(values, luma, chroma, 1, 1, 1), chroma, 0, 0) == (5, 2)


@pytest.mark.parametrize("vector,want", [((1, -1), (1, -1)), ((3, -3), (2, -2)), ((5, -7), (3, -4)),
                                         ((0, 2), (0, 1))])
def test_odd_vectors_round_half_away_from_zero(vector, want):
    luma, chroma = _geom(64, 64), _geom(32, 32)
    mapped = O.map_luma_fields(_fields(luma, 1, lambda f, c, r: vector), luma, chroma, 1, 1, 1)
    assert _at(mapped, chroma, 1, 1) == want


def test_ragged_edges_take_in_bounds_luma_vectors():
    luma, chroma = _geom(37, 35), _geom(19, 18)
    assert (luma.blocks_x, luma.blocks_y, chroma.blocks_x, chroma.blocks_y) == (5, 5, 3, 3)
    mapped = O.map_luma_fields(_fields(luma, 1, lambda f, c, r: (c * 20, r * 40)), luma, chroma, 1, 1, 1)
    assert _at(mapped, chroma, 2, 2) == (40, 80)
    assert _at(mapped, chroma, 1, 1) == (30, 60)


def test_an_even_count_takes_the_lower_median():
    luma, chroma = _geom(29, 16), _geom(15, 8)
    assert (luma.blocks_x, luma.blocks_y, chroma.blocks_x, chroma.blocks_y) == (4, 2, 2, 1)
    mapped = O.map_luma_fields(_fields(luma, 1, lambda f, c, r: (c * 20, r * 40)), luma, chroma, 1, 1, 1)
    assert _at(mapped, chroma, 1, 0) == (20, 0)


def _numpy_map(values, luma, chroma, sx, sy, fields):
    """The same map by sorting each footprint (np.sort) instead of ranking."""
    cx = (max((BLK << sx) - BLK, 0)) // STEP + 1
    cy = (max((BLK << sy) - BLK, 0)) // STEP + 1
    v = values.reshape(fields, luma.blocks_y, luma.blocks_x, 2)
    out = np.zeros((fields, chroma.blocks_y, chroma.blocks_x, 2), np.int64)

    def down(x, s):
        return int(np.sign(x) * ((abs(int(x)) + (1 << s >> 1)) >> s)) if s else int(x)

    for f in range(fields):
        for row in range(chroma.blocks_y):
            for col in range(chroma.blocks_x):
                c0, r0 = ((col * STEP) << sx) // STEP, ((row * STEP) << sy) // STEP
                c1, r1 = min(c0 + cx - 1, luma.blocks_x - 1), min(r0 + cy - 1, luma.blocks_y - 1)
                c0, r0 = min(c0, c1), min(r0, r1)
                foot = v[f, r0:r1 + 1, c0:c1 + 1].reshape(-1, 2)
                srt = np.sort(foot, axis=0)
                mid = (len(foot) - 1) // 2
                out[f, row, col] = (down(srt[mid, 0], sx), down(srt[mid, 1], sy))
    return out.reshape(-1).astype(np.int32)


@pytest.mark.parametrize("chroma_layout,W,H", [("420", 70, 45), ("422", 64, 33), ("444", 41, 29), ("420", 9, 9)])
def test_random_fields_match_a_sorting_median(chroma_layout, W, H):
    sx, sy = O.chroma_shifts(chroma_layout)
    luma, chroma = _geom(W, H), _geom((W + sx) >> sx, (H + sy) >> sy)
    rng = np.random.default_rng(W * H)
    values = rng.integers(-40, 41, luma.blocks * 2 * 3).astype(np.int32)
    got = O.map_luma_fields(values, luma, chroma, sx, sy, 3)
    assert np.array_equal(got, _numpy_map(values, luma, chroma, sx, sy, 3))


def _brute_refine(cen, nb, W, H, g, seeds, R):
    """argmin of the clamped SAD over the (2R+1)^2 window around each seed: lowest SAD, then the
    lowest row-major offset, the window centre keeping a tie with that pick."""
    out = np.zeros_like(seeds)
    ys, xs = np.arange(g.blksize), np.arange(g.blksize)
    for br in range(g.blocks_y):
        for bc in range(g.blocks_x):
            b = br * g.blocks_x + bc
            sx, sy = int(seeds[2 * b]), int(seeds[2 * b + 1])
            oy, ox = br * g.step, bc * g.step
            blk = cen[np.clip(oy + ys, 0, H - 1)][:, np.clip(ox + xs, 0, W - 1)].astype(np.int64)
            best, pick, centre = None, None, None
            for dy in range(-R, R + 1):
                for dx in range(-R, R + 1):
                    win = nb[np.clip(oy + sy + dy + ys, 0, H - 1)][:, np.clip(ox + sx + dx + xs, 0, W - 1)]
                    sad = int(np.abs(blk - win).sum())
                    if dx == 0 and dy == 0:
                        centre = sad
                    if best is None or sad < best:
                        best, pick = sad, (sx + dx, sy + dy)
            out[2 * b:2 * b + 2] = (sx, sy) if centre <= best else pick
    return out


@pytest.mark.parametrize("R", [1, 2])
def test_the_refine_is_the_argmin_around_each_seed(R):
    W, H = 48, 40
    rng = np.random.default_rng(5)
    base = rng.integers(0, 1024, (H + 16, W + 16)).astype(np.int32)
    cen = base[4:4 + H, 4:4 + W].copy()
    nb = base[6:6 + H, 1:1 + W].copy()  # the content moves by (+3, -2)
    g = O.MotionGeom(W, H, 16, 8, R, 1)
    seeds = rng.integers(-4, 5, g.blocks * 2).astype(np.int32)
    mv, conf = O.motion_refine(g, [cen.reshape(-1)], [nb.reshape(-1)], seeds, 1023.0, 0.0, 5.0, radius=R)
    assert np.array_equal(mv, _brute_refine(cen, nb, W, H, g, seeds, R))
    assert conf.shape == (g.blocks,) and np.all((conf >= 0) & (conf <= 1))
    # A seed one pixel from the true motion is pulled onto it wherever the block is inside.
    near = np.tile(np.asarray([3 - 1, -2 + 1], np.int32), g.blocks)
    mv, _ = O.motion_refine(g, [cen.reshape(-1)], [nb.reshape(-1)], near, 1023.0, 0.0, 5.0, radius=R)
    inner = [(br * g.blocks_x + bc) for br in range(1, g.blocks_y - 2) for bc in range(1, g.blocks_x - 2)]
    assert all(tuple(mv[2 * b:2 * b + 2]) == (3, -2) for b in inner)


def _pan_clip(W, H, frames, pan, seed):
    """A smooth random texture moving `pan` luma pixels right per frame, and noise-only chroma."""
    rng = np.random.default_rng(seed)
    tex = rng.normal(0, 1, (H + 8, W + pan * frames + 8))
    k = np.ones(5) / 5
    tex = np.apply_along_axis(lambda r: np.convolve(r, k, "same"), 1, tex)
    tex = np.apply_along_axis(lambda c: np.convolve(c, k, "same"), 0, tex)
    tex = (tex - tex.min()) / (tex.max() - tex.min()) * 800 + 100
    luma, chroma = [], []
    for t in range(frames):
        x0 = pan * (frames - t)  # frame t shows the texture `pan * t` further right
        y = tex[4:4 + H, x0:x0 + W] + rng.normal(0, 1.0, (H, W))
        luma.append(np.clip(np.rint(y), 0, 1023).astype(np.uint16))
        uv = [np.clip(np.rint(512 + rng.normal(0, 2.0, (H // 2, W // 2))), 0, 1023).astype(np.uint16) for _ in range(2)]
        chroma.append(uv)
    return luma, chroma


def _spec(r):
    return O.Nl4dSpec(temporal_radius=r, refine=2, spatial_radius=3, lambda_ht=2.7, sigma=2.0 / 1023,
                      motion=O.MotionGeom(0, 0, 16, 8, 4, 2), step=4)


def test_the_linked_chroma_pass_follows_the_luma_pan():
    W, H, r, pan, depth, frames = 96, 64, 1, 4, 10, 4
    luma, chroma = _pan_clip(W, H, frames, pan, 3)
    results = {}
    for linked in (False, True):
        pl = O.Nl4dPass(_spec(r), W, H, 1, depth, depth)
        pc = O.Nl4dPass(_spec(r), W // 2, H // 2, 2, depth, depth)
        if linked:
            O.link_motion(pl, pc, "420")
        for y, uv in zip(luma, chroma):
            pl.push([O.pack_plane(y, depth)])
            pc.push([O.pack_plane(uv[0], depth), O.pack_plane(uv[1], depth)])
        results[linked] = (pl.last_fields[0], pc.last_fields[0], pc.geom)
    # The last pass centres on frame 2 with neighbours 1 and 3 (offsets -1, +1).
    lmv, cmv, cg = results[True]
    lg = O.MotionGeom(W, H, 16, 8, 4, 2)

    def share(mv, g, per_px, margin, tol):
        fields = mv.reshape(2, g.blocks_y, g.blocks_x, 2)
        shares = []
        for i, k in enumerate((-1, 1)):
            want, hit, n = per_px * k, 0, 0
            for row in range(g.blocks_y):
                for col in range(g.blocks_x):
                    x0 = col * g.step
                    if x0 + want < margin or x0 + g.blksize + want > g.W - margin or row * g.step + g.blksize > g.H:
                        continue
                    n += 1
                    dx, dy = fields[i, row, col]
                    hit += abs(int(dx) - want) <= tol and abs(int(dy)) <= tol
            shares.append(hit / n)
        return shares

    # Within a chroma pixel of the mapped pan counts as following it (matched_flags,
    # nlmeans/tests/models.rs:207-216); luma must find the pan exactly.
    assert min(share(lmv, lg, pan, 2, 0)) == 1.0, "premise: luma finds the pan"
    assert min(share(results[False][1], cg, pan // 2, 2, 1)) < 0.5, "premise: a chroma search on flat chroma misses"
    assert share(cmv, cg, pan // 2, 2, 1) == [1.0, 1.0]
