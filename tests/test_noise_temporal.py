"""Temporal block statistics (nlm_temporal_noise_stats, noise/kernels/stats.rs:142-604, record layout
noise/temporal.rs:18-131).

On the CPU the oracle is pinned by the reference's own expectations (noise/tests/temporal.rs:67-78,
noise/tests/stats.rs:312-705: known pans cancel, zero vectors equal the raw records, off-frame
shifts zero the record, the moved lane, the covering-block rule, ramp textures, the gain rule on
pure noise), by closed forms on constant residuals and by an f64 brute force of every lane, so a
dropped term, a wrong lane or a transposed index fails.  On the GPU the HIP kernel, called
through the C ABI, must equal the oracle bit for bit.
"""
from __future__ import annotations

import ctypes

import numpy as np
import pytest

import oracle as O

from avd import _lib as L

U32 = np.uint64(0xFFFFFFFF)
DECOY = (1000, 1000)  # a vector far outside every test frame (noise/tests/stats.rs:104)


# ---- the reference tests' frame builders (tune/tests/fixtures.rs:54-102) ------------------------

def hashed_unit(index, seed: int) -> np.ndarray:
    h = (np.asarray(index, np.uint64) * np.uint64(0x9E3779B9)) & U32
    h ^= np.uint64((seed * 0x85EBCA6B) & 0xFFFFFFFF)
    h ^= h >> np.uint64(15)
    h = (h * np.uint64(0x2C1B3C6D)) & U32
    h ^= h >> np.uint64(12)
    unit = h.astype(np.float32) / np.float32(4294967295.0)
    return (np.float32(0.2) + np.float32(0.6) * unit).astype(np.float32)


def hashed_ring(width, height, channels, frames, seed):
    return hashed_unit(np.arange(width * height * channels * frames), seed)


def unit_gaussian(idx, seed: int) -> np.ndarray:
    idx = np.asarray(idx, np.uint64)
    total = np.zeros(idx.shape, np.float32)
    for k in range(4):
        h = ((idx * np.uint64(4) + np.uint64(k)) * np.uint64(2654435761)) & U32
        h = (h + np.uint64((seed * 0x9E3779B9 + k) & 0xFFFFFFFF)) & U32
        h ^= h >> np.uint64(15)
        h = (h * np.uint64(0x85EBCA6B)) & U32
        h ^= h >> np.uint64(13)
        total = total + (h.astype(np.float32) / np.float32(4294967295.0) - np.float32(0.5))
    return total / np.sqrt(np.float32(1.0) / np.float32(3.0))


def noisy_over(clean, width, height, sigma, seed):
    sample = unit_gaussian(np.arange(width * height), seed)
    return np.clip(clean + sample * np.float32(sigma), 0.0, 1.0).astype(np.float32)


def texture_frame(width, height, shift_x, shift_y):
    field = width + 16
    y, x = np.mgrid[0:height, 0:width]
    return hashed_unit(((y + shift_y) * field + x + shift_x).reshape(-1), 11)


def motion_grid(width, height, step, vector_at):
    bx, by = -(-width // step), -(-height // step)
    return np.array([vector_at(c, r) for r in range(by) for c in range(bx)], np.int32).reshape(-1), step, bx, by


def stats(previous, new, width, height, channels, luma_fields, grid=None):
    """The oracle on a two-frame ring (previous in slot 0, new in slot 1), the motion grid uploaded
    behind as many decoy vectors, so a reader that ignores the offset zeroes every record."""
    ring = np.concatenate([np.asarray(previous, np.float32), np.asarray(new, np.float32)])
    motion = None
    if grid is not None:
        field, step, bx, by = grid
        behind = np.tile(np.array(DECOY, np.int32), field.size // 2)
        motion = (np.concatenate([behind, field]), behind.size, step, bx, by)
    return O.temporal_stats(ring, width, height, channels, 1, 0, luma_fields, motion)


def record(records, width, height, channels, bx, by):
    nbx, _ = O.temporal_blocks(width, height)
    n = O.temporal_record_len(channels)
    start = (by * nbx + bx) * n
    return records[start:start + n]


def extent(width, height, bx, by):
    return min(16, width - 16 * bx), min(16, height - 16 * by)


def host_variance(prev, new, width, height, bx, by, vector):
    bw, bh = extent(width, height, bx, by)
    vals = []
    for ly in range(bh):
        for lx in range(bw):
            x, y = 16 * bx + lx, 16 * by + ly
            px, py = x + vector[0], y + vector[1]
            if 0 <= px < width and 0 <= py < height:
                vals.append(float(new[y * width + x]) - float(prev[py * width + px]))
    if not vals:
        return None
    v = np.array(vals)
    return float(np.mean(v * v) - np.mean(v) ** 2)


def host_gain(prev, new, width, height, bx, by, vector):
    zero = host_variance(prev, new, width, height, bx, by, (0, 0))
    moved = host_variance(prev, new, width, height, bx, by, vector)
    return None if zero is None or moved is None else moved / zero


def host_sum_d2(prev, new, width, height, bx, by, vector):
    bw, bh = extent(width, height, bx, by)
    total = 0.0
    for ly in range(bh):
        for lx in range(bw):
            x, y = 16 * bx + lx, 16 * by + ly
            d = float(new[y * width + x]) - float(prev[(y + vector[1]) * width + x + vector[0]])
            total += d * d
    return total


# ---- layout and closed forms ---------------------------------------------------------------------

def test_blocks_record_and_slot_lengths():
    # noise/tests/temporal.rs:67-78.
    assert O.temporal_blocks(32, 16) == (2, 1)
    assert O.temporal_blocks(33, 17) == (3, 2)
    assert O.temporal_record_len(1) == 29 and O.temporal_record_len(2) == 32
    assert O.temporal_slot_len(32, 16, 1) == 58
    lib = L.load()
    for w, h, c in [(32, 16, 1), (33, 17, 2), (3840, 2160, 1), (1920, 1080, 2), (1, 1, 1)]:
        assert lib.avd_noise_temporal_stats_len(w, h, c) == O.temporal_slot_len(w, h, c)
    assert O.temporal_texture_lane(1, 0) == 3 and O.temporal_texture_lane(2, 0) == 5
    assert O.temporal_texture_lane(2, 1) == 31


def test_constant_residual_closed_forms():
    # 33x17: a full block, a block 16 wide and 1 high, and a 1x16 and a 1x1 block.  new 0.5 over
    # previous 0.25: every in-frame residual is 0.25 and every sum below is exact in f32.
    w, h = 33, 17
    r = stats(np.full(w * h, 0.25, np.float32), np.full(w * h, 0.5, np.float32), w, h, 1, True)
    for bx, by in [(0, 0), (1, 0), (2, 0), (0, 1), (2, 1)]:
        bw, bh = extent(w, h, bx, by)
        rec = record(r, w, h, 1, bx, by)
        assert rec[0] == 0.25 * bw * bh and rec[1] == 0.0625 * bw * bh, (bx, by)
        assert rec[2] == 0.0625 * (bw - 1) * bh, (bx, by)
        assert rec[3] == 0.0 and rec[4] == 0.0  # flat content, nothing moved
        for q in range(4):
            f = rec[5 + 6 * q: 11 + 6 * q]
            qw = max(0, min(8, bw - 8 * (q % 2)))
            qh = max(0, min(8, bh - 8 * (q // 2)))
            n = qw * qh
            assert f[0] == 0.25 * n and f[1] == 0.0625 * n and f[2] == 0.5 * n, (bx, by, q)
            assert f[3] == (0.0 if qw == 8 and qh == 8 else O.UNMEASURED_FLATNESS), (bx, by, q)
            assert f[4] == (np.float32(0.5) if n else np.float32(1.0e30)), (bx, by, q)
            assert f[5] == (np.float32(0.5) if n else -np.float32(1.0e30)), (bx, by, q)


def test_without_luma_fields_every_quarter_lane_is_zero():
    w, h = 45, 29
    ring = hashed_ring(w, h, 2, 2, 5)
    off = stats(ring[:w * h * 2], ring[w * h * 2:], w, h, 2, False)
    on = stats(ring[:w * h * 2], ring[w * h * 2:], w, h, 2, True)
    n = O.temporal_record_len(2)
    quarters = slice(7, 7 + 24)
    for b in range(off.size // n):
        assert not np.any(off[b * n:][quarters]), b
        keep = np.r_[0:7, 31:32]
        assert np.array_equal(off[b * n:(b + 1) * n][keep], on[b * n:(b + 1) * n][keep]), b


@pytest.mark.parametrize("channels", [1, 2])
def test_every_lane_against_an_f64_brute_force(channels):
    w, h = 45, 29
    ring = hashed_ring(w, h, channels, 2, 7 + channels).reshape(2, h, w, channels).astype(np.float64)
    got = stats(ring[0].reshape(-1), ring[1].reshape(-1), w, h, channels, True)
    nbx, nby = O.temporal_blocks(w, h)
    for by in range(nby):
        for bx in range(nbx):
            bw, bh = extent(w, h, bx, by)
            new = ring[1, 16 * by:16 * by + bh, 16 * bx:16 * bx + bw]
            prev = ring[0, 16 * by:16 * by + bh, 16 * bx:16 * bx + bw]
            d = new - prev
            rec = record(got, w, h, channels, bx, by)
            want = list(d.sum((0, 1))) + list((d * d).sum((0, 1))) + [(d[:, :-1, 0] * d[:, 1:, 0]).sum()]
            assert np.allclose(rec[:2 * channels + 1], want, rtol=2e-5, atol=1e-6), (bx, by)
            for c in range(channels):
                tile = np.zeros((16, 16))
                tile[:bh, :bw] = new[:, :, c]
                cells = tile.reshape(4, 4, 4, 4).mean((1, 3))
                cols, rows = bw // 4, bh // 4
                steps = [cells[i, j] - cells[i, j + 1] for i in range(rows) for j in range(cols - 1)]
                steps += [cells[i, j] - cells[i + 1, j] for i in range(rows - 1) for j in range(cols)]
                texture = np.mean(np.square(steps)) if steps else 0.0
                assert np.isclose(rec[O.temporal_texture_lane(channels, c)], texture, rtol=1e-4, atol=1e-9), (bx, by, c)
            mean = np.zeros((16, 16))
            mean[:bh, :bw] = 0.5 * (new[:, :, 0] + prev[:, :, 0])
            grid = mean.reshape(8, 2, 8, 2).mean((1, 3))
            for q in range(4):
                qy, qx = 8 * (q // 2), 8 * (q % 2)
                qd, qn = d[qy:qy + 8, qx:qx + 8, 0], new[qy:qy + 8, qx:qx + 8, 0]
                f = rec[2 * channels + 3 + 6 * q: 2 * channels + 9 + 6 * q]
                assert np.allclose(f[:3], [qd.sum(), (qd * qd).sum(), qn.sum()], rtol=2e-5, atol=1e-6), (bx, by, q)
                if qn.size:
                    assert f[4] == np.float32(qn.min()) and f[5] == np.float32(qn.max()), (bx, by, q)
                g = grid[4 * (q // 2):4 * (q // 2) + 4, 4 * (q % 2):4 * (q % 2) + 4]
                e = np.sum(np.square(np.diff(g, axis=1))) + np.sum(np.square(np.diff(g, axis=0)))
                full = bw >= qx + 8 and bh >= qy + 8
                if full:
                    assert np.isclose(f[3], e / 24.0, rtol=1e-4, atol=1e-10), (bx, by, q)
                else:
                    assert f[3] == O.UNMEASURED_FLATNESS, (bx, by, q)


# ---- the reference's kernel expectations (noise/tests/stats.rs) ---------------------------------

def test_texture_lane_reads_mean_cell_energy():
    # noise/tests/stats.rs:415-455: a ramp across; the ragged last column holds 10 pairs.
    w, h = 56, 32
    slope = np.float32(3.0 / 1024.0)
    ramp = ((np.arange(w * h) % w).astype(np.float32) * slope).astype(np.float32)
    step = 4.0 * float(slope)
    for luma_fields in (False, True):
        r = stats(np.full(w * h, 0.5, np.float32), ramp, w, h, 1, luma_fields)
        nbx, nby = O.temporal_blocks(w, h)
        for by in range(nby):
            for bx in range(nbx):
                want = (4.0 * step * step / 10.0) if bx + 1 == nbx else (12.0 * step * step / 24.0)
                got = record(r, w, h, 1, bx, by)[3]
                assert abs(got - want) / want <= 1e-5, (bx, by)


def test_each_channel_texture_lane_reads_its_own_cell_energy():
    # noise/tests/stats.rs:457-516, with and without a still motion grid.
    w, h = 56, 32
    idx = np.arange(w * h)
    new = np.stack([(idx % w).astype(np.float32) * np.float32(3.0 / 1024.0),
                    (idx // w).astype(np.float32) * np.float32(5.0 / 1024.0)], 1).reshape(-1).astype(np.float32)
    a, d = 4.0 * 3.0 / 1024.0, 4.0 * 5.0 / 1024.0
    full = [12.0 * a * a / 24.0, 12.0 * d * d / 24.0]
    ragged = [4.0 * a * a / 10.0, 6.0 * d * d / 10.0]
    still = motion_grid(w, h, 16, lambda c, r: (0, 0))
    for grid in (None, still):
        r = stats(np.full(new.size, 0.5, np.float32), new, w, h, 2, False, grid)
        nbx, nby = O.temporal_blocks(w, h)
        for by in range(nby):
            for bx in range(nbx):
                want = ragged if bx + 1 == nbx else full
                rec = record(r, w, h, 2, bx, by)
                for c in range(2):
                    got = rec[O.temporal_texture_lane(2, c)]
                    assert abs(got - want[c]) / want[c] <= 1e-5, (bx, by, c)


def test_compensated_stats_cancel_a_known_pan():
    # noise/tests/stats.rs:312-344: new(x, y) matches previous(x + 3, y - 2).
    w, h, sigma = 128, 96, 2.0 / 255.0
    prev = noisy_over(texture_frame(w, h, 0, 4), w, h, sigma, 1)
    new = noisy_over(texture_frame(w, h, 3, 2), w, h, sigma, 2)

    def median_sigma(r):
        nbx, nby = O.temporal_blocks(w, h)
        out = []
        for by in range(nby):
            for bx in range(nbx):
                rec = record(r, w, h, 1, bx, by)
                if rec[1] == 0.0:
                    continue
                bw, bh = extent(w, h, bx, by)
                m = rec[0] / (bw * bh)
                out.append(np.sqrt(max(rec[1] / (bw * bh) - m * m, 0.0)) / np.sqrt(2.0))
        return sorted(out)[len(out) // 2]

    comp = median_sigma(stats(prev, new, w, h, 1, True, motion_grid(w, h, 8, lambda c, r: (3, -2))))
    raw = median_sigma(stats(prev, new, w, h, 1, True))
    assert abs(comp - sigma) / sigma <= 0.05, comp
    assert raw > 3.0 * sigma, raw


@pytest.mark.parametrize("channels", [1, 2])
def test_zero_vector_compensated_stats_equal_raw_stats(channels):
    # noise/tests/stats.rs:346-381.
    w, h = 45, 29
    ring = hashed_ring(w, h, channels, 2, 5)
    n = w * h * channels
    for luma_fields in (False, True):
        comp = stats(ring[:n], ring[n:], w, h, channels, luma_fields, motion_grid(w, h, 8, lambda c, r: (0, 0)))
        raw = stats(ring[:n], ring[n:], w, h, channels, luma_fields)
        assert np.array_equal(comp.view(np.uint32), raw.view(np.uint32)), luma_fields
        moved = 2 * channels + O.BLOCK_MOVED_LANE
        assert not np.any(raw.reshape(-1, O.temporal_record_len(channels))[:, moved])


def test_off_frame_displacement_zeroes_the_record():
    # noise/tests/stats.rs:383-413: every block follows (20, 0); the last two block columns leave.
    w, h = 96, 48
    prev, new = texture_frame(w, h, 0, 0), texture_frame(w, h, 20, 0)
    nbx, nby = O.temporal_blocks(w, h)
    for luma_fields in (False, True):
        r = stats(prev, new, w, h, 1, luma_fields, motion_grid(w, h, 8, lambda c, r: (20, 0)))
        for by in range(nby):
            for bx in range(nbx):
                zeroed = not np.any(record(r, w, h, 1, bx, by))
                assert zeroed == (bx + 2 >= nbx), (bx, by, luma_fields)


def test_moved_lane_flags_non_zero_vectors():
    # noise/tests/stats.rs:518-544: odd block columns follow (1, 0).
    w, h = 80, 32
    prev, new = texture_frame(w, h, 0, 0), texture_frame(w, h, 1, 0)
    r = stats(prev, new, w, h, 1, True, motion_grid(w, h, 16, lambda c, r: (0, 0) if c % 2 == 0 else (1, 0)))
    nbx, nby = O.temporal_blocks(w, h)
    for by in range(nby):
        for bx in range(nbx):
            assert record(r, w, h, 1, bx, by)[2 + O.BLOCK_MOVED_LANE] == float(bx % 2), (bx, by)


@pytest.mark.parametrize("step", [4, 8, 32])
def test_covering_block_rule_picks_the_origin_block(step):
    # noise/tests/stats.rs:546-627: each block reads the vector of the motion block covering its
    # origin; every other motion block holds the decoy.
    w = h = 64
    mbx, mby = -(-w // step), -(-h // step)
    nbx, nby = O.temporal_blocks(w, h)
    cov_col = lambda bx: min(bx * 16 // step, mbx - 1)
    cov_row = lambda by: min(by * 16 // step, mby - 1)

    def motion_vector(col, row):
        rx, ry = 1 + (col + 2 * row) % 3, (col * row) % 3
        return (rx if col * step < w // 2 else -rx, ry if row * step < h // 2 else -ry)

    block_vector = lambda bx, by: motion_vector(cov_col(bx), cov_row(by))
    ring = hashed_ring(w, h, 1, 2, 13)
    prev, new = ring[:w * h].copy(), ring[w * h:].copy()
    for by in range(nby):  # plant each block's match at its shifted position
        for bx in range(nbx):
            vx, vy = block_vector(bx, by)
            for ly in range(16):
                for lx in range(16):
                    x, y = 16 * bx + lx, 16 * by + ly
                    prev[(y + vy) * w + x + vx] = new[y * w + x]
    new = noisy_over(new, w, h, 0.01, 14)

    def at(col, row):
        origin = any(cov_col(bx) == col for bx in range(nbx)) and any(cov_row(by) == row for by in range(nby))
        return motion_vector(col, row) if origin else DECOY

    r = stats(prev, new, w, h, 1, True, motion_grid(w, h, step, at))
    for by in range(nby):
        for bx in range(nbx):
            v = block_vector(bx, by)
            assert host_gain(prev, new, w, h, bx, by, v) <= 0.5
            want = host_sum_d2(prev, new, w, h, bx, by, v)
            assert abs(record(r, w, h, 1, bx, by)[1] - want) / want <= 1e-4, (step, bx, by)


def test_noise_fitted_vector_keeps_the_zero_vector_residual():
    # noise/tests/stats.rs:655-705: on independent noise most blocks keep the zero vector; a block
    # follows exactly when the host gain rule says so.
    w, h, sigma = 64, 48, 2.0 / 255.0
    clean = np.full(w * h, 0.5, np.float32)
    prev, new = noisy_over(clean, w, h, sigma, 31), noisy_over(clean, w, h, sigma, 32)
    nbx, nby = O.temporal_blocks(w, h)
    for luma_fields in (False, True):
        comp = stats(prev, new, w, h, 1, luma_fields, motion_grid(w, h, 8, lambda c, r: (1, 0)))
        raw = stats(prev, new, w, h, 1, luma_fields)
        kept = 0
        for by in range(nby):
            for bx in range(nbx):
                follows = host_gain(prev, new, w, h, bx, by, (1, 0)) <= 0.5
                rec = record(comp, w, h, 1, bx, by)
                assert rec[2 + O.BLOCK_MOVED_LANE] == (1.0 if follows else 0.0), (bx, by)
                if not follows:
                    assert np.array_equal(rec.view(np.uint32), record(raw, w, h, 1, bx, by).view(np.uint32))
                    kept += 1
        assert kept * 4 >= nbx * nby * 3


# ---- the boundary ---------------------------------------------------------------------------------

def test_site_rejects_bad_arguments_without_a_gpu():
    lib = L.load()
    buf = ctypes.c_void_p(16)
    import avd

    assert lib.avd_site_noise_temporal_stats(None, 64, 48, 1, 1, 0, None, 0, 8, 8, 6, 1, buf, None) == \
        avd.NullArgument.status
    assert lib.avd_site_noise_temporal_stats(buf, 64, 48, 3, 1, 0, None, 0, 8, 8, 6, 1, buf, None) == \
        avd.InvalidGeometry.status
    assert lib.avd_site_noise_temporal_stats(buf, 0, 48, 1, 1, 0, None, 0, 8, 8, 6, 1, buf, None) == \
        avd.InvalidGeometry.status
    assert lib.avd_site_noise_temporal_stats(buf, 64, 48, 1, 1, 0, buf, 0, 0, 8, 6, 1, buf, None) == \
        avd.InvalidOptions.status
    assert lib.avd_site_noise_temporal_zero(None, 4, None) == avd.NullArgument.status
    # 2^20 x 2^20 is within the side limit but its records pass 32-bit offsets.
    assert lib.avd_site_noise_temporal_stats(buf, 1 << 20, 1 << 20, 1, 1, 0, None, 0, 8, 8, 6, 1, buf, None) == \
        avd.InvalidGeometry.status


# ---- the HIP kernel against the oracle ------------------------------------------------------------

def _site(ring, w, h, c, slot_new, slot_prev, luma_fields, motion):
    import torch

    from gpu_helpers import dev, host, require_gpu

    require_gpu()
    d_ring = dev(ring)
    d_stats = torch.full((O.temporal_slot_len(w, h, c),), -7.0, dtype=torch.float32, device="cuda")
    if motion is None:
        d_mv, args = None, (0, 0, 0, 0)
    else:
        d_mv, args = dev(np.ascontiguousarray(motion[0], np.int32)), motion[1:]
    rc = L.load().avd_site_noise_temporal_stats(
        ctypes.c_void_p(d_ring.data_ptr()), w, h, c, slot_new, slot_prev,
        None if d_mv is None else ctypes.c_void_p(d_mv.data_ptr()), *args, 1 if luma_fields else 0,
        ctypes.c_void_p(d_stats.data_ptr()), None)
    L.check(rc)
    return host(d_stats)


def _mixed_field(w, h, step, seed):
    """Vectors of every kind: zero, small in-frame shifts, shifts that leave the frame at the
    edges, and decoys whose match never lies inside it."""
    rng = np.random.default_rng(seed)
    bx, by = -(-w // step), -(-h // step)
    kinds = rng.integers(0, 6, bx * by)
    v = rng.integers(-3, 4, (bx * by, 2))
    v[kinds == 0] = 0
    v[kinds == 1] = (w + 5, -h - 5)
    v[kinds == 2] = rng.integers(-40, 41, (int((kinds == 2).sum()), 2))
    # Three decoy vectors sit in front of the grid, so the field starts at offset 6.
    field = np.concatenate([np.tile(np.array(DECOY, np.int32), 3), v.astype(np.int32).reshape(-1)])
    return field, 6, step, bx, by


GPU_CASES = [
    # name, W, H, C, luma_fields, motion step (0: none), frames in the ring, slot_new, slot_prev
    ("luma45x29", 45, 29, 1, True, 0, 2, 1, 0),
    ("luma45x29_plain", 45, 29, 1, False, 0, 2, 1, 0),
    ("chroma45x29", 45, 29, 2, True, 0, 2, 1, 0),
    ("chroma33x17_plain", 33, 17, 2, False, 0, 2, 1, 0),
    ("luma_comp96x72", 96, 72, 1, True, 8, 3, 2, 0),
    ("chroma_comp96x72_plain", 96, 72, 2, False, 16, 3, 0, 2),
    ("chroma_comp61x37", 61, 37, 2, True, 4, 2, 1, 0),
    ("tiny1x1", 1, 1, 1, True, 0, 2, 1, 0),
    ("tiny7x3_comp", 7, 3, 1, True, 8, 2, 1, 0),
    ("hd1080_comp", 1920, 1080, 1, True, 8, 2, 1, 0),
    ("chroma960x540_comp", 960, 540, 2, True, 8, 2, 1, 0),
    ("uhd_luma", 3840, 2160, 1, True, 0, 2, 1, 0),
]


@pytest.mark.gpu
@pytest.mark.parametrize("case", GPU_CASES, ids=[c[0] for c in GPU_CASES])
def test_site_noise_temporal_stats_is_bit_exact(case):
    name, w, h, c, luma_fields, step, frames, slot_new, slot_prev = case
    print(f"temporal stats case {name}", flush=True)
    clean = texture_frame(w, h, 0, 0)
    ring = []
    for f in range(frames):
        base = np.repeat(clean, c) if c > 1 else clean
        ring.append(noisy_over(base, w * c, h, 3.0 / 255.0, 40 + f).reshape(-1))
    ring = np.concatenate(ring).astype(np.float32)
    motion = _mixed_field(w, h, step, 7 + w) if step else None
    got = _site(ring, w, h, c, slot_new, slot_prev, luma_fields, motion)
    want = O.temporal_stats(ring, w, h, c, slot_new, slot_prev, luma_fields, motion)
    assert got.size == want.size
    bad = np.flatnonzero(got.view(np.uint32) != want.view(np.uint32))
    assert bad.size == 0, (name, bad[:8], got[bad[:8]], want[bad[:8]])


@pytest.mark.gpu
def test_site_noise_temporal_zero_fills_the_slot():
    import torch

    from gpu_helpers import host, require_gpu

    require_gpu()
    d = torch.full((1000,), -7.0, dtype=torch.float32, device="cuda")
    L.check(L.load().avd_site_noise_temporal_zero(ctypes.c_void_p(d.data_ptr() + 4 * 10), 977, None))
    got = host(d)
    assert np.all(got[10:987] == 0.0) and np.all(got[:10] == -7.0) and np.all(got[987:] == -7.0)
