"""Film-grain measurement (grain_measure, nl4d/kernels/grain.rs:243-513).

On the CPU the oracle is pinned by the closed forms and expectations of the reference's kernel
tests (nl4d/grain/tests/kernels.rs:573-760: flat grain fills the source histogram at its std and
only interior cells count, a constant cell reads bucket 0, a checkerboard's and a ramp's high pass,
the one-frame histogram by the colour grid's luma and none without it) and by an f64 brute force of
the 46 lag sums, so a wrong lag offset, a missing mean removal or a dropped gate fails.  On the GPU
the HIP kernel, called through the C ABI, must give the same histogram counts and bit-identical
partials.
"""
from __future__ import annotations

import ctypes

import numpy as np
import pytest

import oracle as O

from avd import _lib as L

EDGES = O.grain_bucket_edges()


def scene(W, H, sigma=2 / 255, base=0.5, shift=(0, 0), seed=1, C=1, ch=0, texture=None):
    """Two ring slots of a flat (or textured) field with independent white grain; slot 1 is slot 0's
    scene moved by `shift` (slot_next(x + dx, y + dy) = scene(x, y)); the denoised frame is flat."""
    rng = np.random.default_rng(seed)
    pad = 16
    field = np.full((H + 2 * pad, W + 2 * pad), base, np.float64)
    if texture is not None:
        field += texture * rng.standard_normal(field.shape)
    dx, dy = shift
    t = field[pad:pad + H, pad:pad + W]
    n = field[pad - dy:pad - dy + H, pad - dx:pad - dx + W]
    slots = np.stack([t + sigma * rng.standard_normal((H, W)), n + sigma * rng.standard_normal((H, W))])
    ring = np.zeros((2, H, W, C), np.float32)
    ring[..., ch] = slots
    out_t = np.zeros((H, W, C), np.float32)
    out_t[..., ch] = base
    return ring, out_t


def grid(W, H, step=8, vector=(0, 0), conf=1.0):
    bx, by = -(-W // step), -(-H // step)
    return np.tile(np.array(vector, np.int32), bx * by), np.full(bx * by, conf, np.float32), bx, by, step


def measure(ring, out_t, W, H, C=1, ch=0, g=None, **kw):
    mv, conf, bx, by, step = g if g is not None else grid(W, H)
    return O.grain_measure(ring, out_t, W, H, C, ch, 0, 1, mv, conf, bx, by, step, **kw)


def cells(part):
    return part.reshape(-1, O.GRAIN_PARTIAL)


def bucket_of(s):
    return max(i for i in range(64) if i == 0 or s >= EDGES[i])


# ---- constants -------------------------------------------------------------------------------------

def test_lag_table_and_bucket_edges():
    lags = O.grain_lags()
    assert len(lags) == O.GRAIN_LAGS and lags[:7] == [(0, d) for d in range(7)]
    assert lags[7] == (1, -6) and lags[-1] == (3, 6) and len(set(lags)) == 46
    assert EDGES.size == 65 and EDGES[0] == np.float32(np.float32(0.05) / np.float32(255.0))
    assert abs(EDGES[-1] / np.float32(32 / 255) - 1) < 1e-6
    ratio = EDGES[1:].astype(np.float64) / EDGES[:-1]
    assert np.allclose(ratio, (32 / 0.05) ** (1 / 64), rtol=1e-6)
    assert L.load().avd_grain_partials_len(70, 45) == 8 * 5 * 48


# ---- the reference's expectations ------------------------------------------------------------------

def test_flat_static_grain_counts_interior_cells_at_its_std():
    W, H, sigma = 96, 64, 2 / 255
    ring, out_t = scene(W, H, sigma)
    hist, part = measure(ring, out_t, W, H)
    p = cells(part)
    cx, cy = W // 8, H // 8
    accepted = p[:, 46].reshape(cy, cx) > 0
    want = np.zeros((cy, cx), bool)
    want[:cy - 1, 1:cx - 1] = True  # off the left, right and bottom edges
    assert np.array_equal(accepted, want)
    src = hist[:1024].reshape(16, 64)
    assert src.sum() == want.sum() and src[8].sum() == want.sum()  # mean 0.5 -> brightness bin 8
    counts = src[8]
    median_bucket = int(np.searchsorted(np.cumsum(counts), counts.sum() / 2))
    assert abs(median_bucket - bucket_of(sigma)) <= 1
    assert hist[1024:].sum() == cx * cy  # every usable cell counts one-frame grain
    lag0 = p[accepted.reshape(-1), 0].sum() / p[accepted.reshape(-1), 46].sum()
    assert abs(lag0 / sigma ** 2 - 1) < 0.1
    assert np.all(np.abs(p[accepted.reshape(-1), 1:46].sum(0) / p[accepted.reshape(-1), 46].sum()) < 0.15 * sigma ** 2)


def test_lag_sums_against_an_f64_brute_force():
    W, H = 64, 48
    ring, out_t = scene(W, H, 3 / 255, shift=(2, -1), seed=4)
    mv, conf, bx, by, step = grid(W, H, 8, (2, -1))
    mv = mv.reshape(-1, 2)
    mv[5] = (-3, 2)  # one block with its own vector: halo pixels read their own cell's vector
    mv = mv.reshape(-1)
    _, part = measure(ring, out_t, W, H, g=(mv, conf, bx, by, step))
    p = cells(part)
    t, n = ring[0, ..., 0].astype(np.float64), ring[1, ..., 0].astype(np.float64)

    def grain(x, y):
        x, y = min(max(x, 0), W - 1), min(max(y, 0), H - 1)
        b = min(y // 8 * 8 // step, by - 1) * bx + min(x // 8 * 8 // step, bx - 1)
        vx, vy = mv[2 * b], mv[2 * b + 1]
        return (n[min(max(y + vy, 0), H - 1), min(max(x + vx, 0), W - 1)] - t[y, x]) / np.sqrt(2)

    checked = 0
    for c in range(p.shape[0]):
        if p[c, 46] == 0:
            assert not np.any(p[c, :46])
            continue
        ox, oy = c % (W // 8) * 8, c // (W // 8) * 8
        cell = np.array([[grain(ox + i, oy + j) for i in range(8)] for j in range(8)])
        mean = cell.mean()
        for k, (dy, dx) in enumerate(O.grain_lags()):
            want = sum((cell[j, i] - mean) * (grain(ox + i + dx, oy + j + dy) - mean)
                       for j in range(8) for i in range(8))
            assert abs(p[c, k] - want) <= 2e-4 * abs(p[c, 0]) + 1e-9, (c, k)
        checked += 1
    assert checked >= 20


def test_a_constant_cell_reads_bucket_0():
    # nl4d/grain/tests/kernels.rs:675-686.
    W, H = 64, 48
    ring, out_t = scene(W, H, 0.0)
    hist, _ = measure(ring, out_t, W, H)
    frame = hist[1024:].reshape(16, 64)
    assert frame[8, 0] == (W // 8) * (H // 8) and frame.sum() == frame[8, 0]
    assert hist[:1024].sum() == 0  # zero grain std is below the floor


@pytest.mark.parametrize("amp", [0.5 / 255, 2 / 255, 6 / 255])
def test_one_frame_high_pass_of_a_checkerboard_and_a_ramp(amp):
    # nl4d/grain/tests/kernels.rs:688-716: a checkerboard of +-a steps by 2a on both axes, so the
    # high pass is a sqrt 2; a ramp of slope s across reads s / 2.
    W, H = 64, 48
    ring, out_t = scene(W, H, 0.0)
    yy, xx = np.mgrid[0:H, 0:W]
    ring[0, ..., 0] = 0.5 + amp * np.where((xx + yy) % 2 == 0, 1.0, -1.0)
    hist, _ = measure(ring, out_t, W, H)
    assert hist[1024:].reshape(16, 64)[8, bucket_of(amp * np.sqrt(2))] == (W // 8) * (H // 8)
    ring[0, ..., 0] = 0.4 + amp * (xx % 8)
    hist, _ = measure(ring, out_t, W, H)
    assert hist[1024:].reshape(16, 64)[8, bucket_of(amp / 2)] == (W // 8) * (H // 8)


def test_gates_reject_what_the_reference_rejects():
    W, H = 64, 48
    ring, out_t = scene(W, H, 2 / 255)
    base_hist, _ = measure(ring, out_t, W, H)
    n_src, n_frame = base_hist[:1024].sum(), base_hist[1024:].sum()
    hist, part = measure(ring, out_t, W, H, has_source=False)  # no source rejects every cell
    assert hist.sum() == 0 and not np.any(part)
    hist, _ = measure(ring, out_t, W, H, g=grid(W, H, conf=0.79))  # confidence gates source grain only
    assert hist[:1024].sum() == 0 and hist[1024:].sum() == n_frame
    dark = out_t.copy()
    dark[...] = 10 / 255  # below the luma band
    assert measure(ring, dark, W, H)[0].sum() == 0
    assert measure(ring, dark, W, H, luma_band=False)[0][:1024].sum() == n_src
    rough = out_t.copy()
    rough[3, 12] += 6 / 255  # cell (1, 0) of the denoised frame is no longer flat: one cell fewer
    hist, _ = measure(ring, rough, W, H)
    assert hist[:1024].sum() == n_src - 1 and hist[1024:].sum() == n_frame - 1
    clipped = ring.copy()
    clipped[1, 20, 20, 0] = 3 / 255  # a source sample near black rejects its cell
    hist, _ = measure(clipped, out_t, W, H)
    assert hist[:1024].sum() == n_src - 1 and hist[1024:].sum() == n_frame - 1


def test_a_pan_followed_by_its_vector_reads_the_grain():
    W, H, sigma = 96, 64, 2 / 255
    ring, out_t = scene(W, H, sigma, shift=(3, -2), seed=9, texture=0.02)
    def median_std(hist):
        counts = hist[:1024].reshape(16, 64).sum(0)
        return int(np.searchsorted(np.cumsum(counts), counts.sum() / 2)) if counts.sum() else -1
    followed = median_std(measure(ring, out_t, W, H, g=grid(W, H, vector=(3, -2)))[0])
    still = median_std(measure(ring, out_t, W, H)[0])
    assert abs(followed - bucket_of(sigma)) <= 1
    assert still > bucket_of(2 * sigma) and still > followed + 4  # the still vector reads texture


def test_one_frame_cells_bin_by_the_colour_grid():
    # nl4d/grain/tests/kernels.rs:718-752: with a colour grid a one-frame cell bins by its 16x16
    # block's mean luma on the 8-bit scale, and counts nothing until the grid is held.
    W, H = 64, 48
    ring, out_t = scene(W, H, 1 / 255, C=2, ch=1)
    mbx = W // 16
    means = np.zeros((mbx * (H // 16), 3), np.float32)
    means[:, 0] = (np.arange(means.shape[0]) % 16) * 16 + 3
    kw = dict(C=2, ch=1, luma_means=means.reshape(-1), means_blocks_x=mbx)
    hist, _ = measure(ring, out_t, W, H, has_luma_means=True, **kw)
    frame = hist[1024:].reshape(16, 64).sum(1)
    for b in range(means.shape[0]):
        assert frame[b % 16] >= 4  # four cells per 16x16 block
    assert frame.sum() == (W // 8) * (H // 8)
    hist, _ = measure(ring, out_t, W, H, has_luma_means=False, **kw)
    assert hist[1024:].sum() == 0


def test_site_rejects_bad_arguments_without_a_gpu():
    import avd

    lib = L.load()
    b = ctypes.c_void_p(16)
    args = [b, b, 64, 48, 1, 0, 0, 1, b, b, 8, 6, 8, 0, 1, None, 0, 0, 1, b, b, b, None]
    assert lib.avd_site_grain_measure(*([None] + args[1:])) == avd.NullArgument.status
    bad = list(args); bad[5] = 1
    assert lib.avd_site_grain_measure(*bad) == avd.InvalidGeometry.status
    bad = list(args); bad[12] = 0
    assert lib.avd_site_grain_measure(*bad) == avd.InvalidOptions.status
    bad = list(args); bad[17] = 4
    assert lib.avd_site_grain_measure(*bad) == avd.NullArgument.status


# ---- the HIP kernel against the oracle --------------------------------------------------------------

def _site(ring, out_t, W, H, C, ch, g, entry=0, has_source=True, luma_means=None, has_luma_means=False,
          means_bx=0, luma_band=True):
    import torch

    from gpu_helpers import dev, host, require_gpu

    require_gpu()
    mv, conf, bx, by, step = g
    d = {k: dev(np.ascontiguousarray(v)) for k, v in dict(ring=ring, out=out_t, mv=mv, conf=conf, edges=EDGES).items()}
    means = dev(np.ascontiguousarray(luma_means, np.float32)) if luma_means is not None else None
    hist = torch.zeros(2 * O.GRAIN_HIST, dtype=torch.int32, device="cuda")
    part = torch.full((L.load().avd_grain_partials_len(W, H),), -7.0, dtype=torch.float32, device="cuda")
    p = lambda t: None if t is None else ctypes.c_void_p(t.data_ptr())
    L.check(L.load().avd_site_grain_measure(p(d["ring"]), p(d["out"]), W, H, C, ch, 0, 1, p(d["mv"]), p(d["conf"]),
                                            bx, by, step, entry, int(has_source), p(means), int(has_luma_means),
                                            means_bx, int(luma_band), p(d["edges"]), p(hist), p(part), None))
    return host(hist), host(part)


GPU_CASES = [
    # name, W, H, C, channel, sigma, shift, texture, step
    ("flat96x64", 96, 64, 1, 0, 2 / 255, (0, 0), None, 8),
    ("ragged70x45", 70, 45, 1, 0, 2 / 255, (0, 0), None, 8),
    ("pan96x64", 96, 64, 1, 0, 2 / 255, (3, -2), 0.02, 8),
    ("pan_step16_textured", 128, 72, 1, 0, 1 / 255, (-4, 3), 0.004, 16),
    ("chroma_ch1", 64, 48, 2, 1, 1 / 255, (1, 1), None, 8),
    ("hd1080", 1920, 1080, 1, 0, 1.5 / 255, (2, 0), 0.003, 8),
    ("uhd", 3840, 2160, 1, 0, 1.5 / 255, (0, 0), 0.003, 8),
]


@pytest.mark.gpu
@pytest.mark.parametrize("case", GPU_CASES, ids=[c[0] for c in GPU_CASES])
def test_site_grain_measure_is_bit_exact(case):
    name, W, H, C, ch, sigma, shift, texture, step = case
    print(f"grain measure case {name}", flush=True)
    ring, out_t = scene(W, H, sigma, shift=shift, seed=W + H, C=C, ch=ch, texture=texture)
    mv, conf, bx, by, _ = grid(W, H, step, shift)
    rng = np.random.default_rng(W)
    mv = mv.reshape(-1, 2)
    jitter = rng.random(len(mv)) < 0.1
    mv[jitter] = rng.integers(-9, 10, (int(jitter.sum()), 2))
    conf = np.where(rng.random(len(conf)) < 0.1, np.float32(0.5), conf).astype(np.float32)
    g = (mv.reshape(-1).astype(np.int32), conf, bx, by, step)
    extra = {}
    if C == 2:
        mbx = -(-W // 16)
        means = np.zeros((mbx * -(-H // 16), 3), np.float32)
        means[:, 0] = rng.uniform(0, 255, len(means))
        extra = dict(luma_means=means.reshape(-1), has_luma_means=True, means_bx=mbx)
    hist, part = _site(ring, out_t, W, H, C, ch, g, **extra)
    kw = {}
    if extra:
        kw = dict(luma_means=extra["luma_means"], has_luma_means=True, means_blocks_x=extra["means_bx"])
    want_hist, want_part = O.grain_measure(ring, out_t, W, H, C, ch, 0, 1, g[0], g[1], bx, by, step, **kw)
    assert np.array_equal(hist, want_hist), name
    assert part.size == want_part.size
    bad = np.flatnonzero(part.view(np.uint32) != want_part.view(np.uint32))
    assert bad.size == 0, (name, bad[:8], part[bad[:8]], want_part[bad[:8]])
    assert want_hist[:1024].sum() > 0, name  # the case measures something


# ---- the partial reduce (grain_reduce_partials, nl4d/kernels/grain.rs:515-560) ----------------------

def random_partials(cells, seed, integer=False):
    rng = np.random.default_rng(seed)
    p = np.zeros((cells, O.GRAIN_PARTIAL), np.float32)
    p[:, :47] = rng.integers(-50, 51, (cells, 47)) if integer else rng.standard_normal((cells, 47)) * 1e-3
    # strength groups 0..15 plus values the clamp must handle: past the last group, negative, fractional
    p[:, 47] = rng.choice([0, 1, 3, 7, 15, 16, 99, -2, 4.7], cells)
    return p


def test_reduce_sums_each_group_and_keeps_earlier_sums():
    cells = 1000
    p = random_partials(cells, 3, integer=True)  # integer partials: every f32 sum is exact
    chunk = np.arange(10 + 16 * 47, dtype=np.float32)
    got = O.grain_reduce(p, chunk, record_offset=10)
    groups = np.minimum(np.maximum(np.floor(p[:, 47]), 0), 15).astype(int)  # Rust's saturating cast
    want = chunk.astype(np.float64).copy()
    for g in range(16):
        want[10 + 47 * g:10 + 47 * (g + 1)] += p[groups == g, :47].astype(np.float64).sum(0)
    assert np.array_equal(got, want.astype(np.float32))
    assert np.array_equal(got[:10], chunk[:10])


def test_reduce_of_measured_partials_counts_the_accepted_pixels():
    W, H = 96, 64
    ring, out_t = scene(W, H, 2 / 255)
    hist, part = measure(ring, out_t, W, H)
    rec = O.grain_reduce(part).reshape(16, 47)
    # The count lane totals 64 pixels per accepted cell, in the group of its std bucket.
    assert rec[:, 46].sum() == 64 * hist[:1024].sum()
    by_group = hist[:1024].reshape(16, 16, 4).sum((0, 2))
    assert np.array_equal(rec[:, 46], 64 * by_group.astype(np.float32))
    lag0 = rec[:, 0].sum() / rec[:, 46].sum()
    assert abs(lag0 / (2 / 255) ** 2 - 1) < 0.1


def _site_reduce(part, chunk, record_offset):
    import torch

    from gpu_helpers import dev, host, require_gpu

    require_gpu()
    d_part, d_chunk = dev(part), dev(chunk)
    L.check(L.load().avd_site_grain_reduce(ctypes.c_void_p(d_part.data_ptr()), part.size // O.GRAIN_PARTIAL,
                                           ctypes.c_void_p(d_chunk.data_ptr()), record_offset, None))
    return host(d_chunk)


REDUCE_CASES = [("cells1", 1, 0), ("cells127", 127, 5), ("cells1000", 1000, 0), ("cells12345", 12345, 47),
                ("uhd_cells", 480 * 270, 16 * 47)]


@pytest.mark.gpu
@pytest.mark.parametrize("case", REDUCE_CASES, ids=[c[0] for c in REDUCE_CASES])
def test_site_grain_reduce_is_bit_exact(case):
    name, cells, offset = case
    print(f"grain reduce case {name}", flush=True)
    part = random_partials(cells, cells).reshape(-1)
    chunk = np.random.default_rng(1).standard_normal(offset + 16 * 47).astype(np.float32)
    got = _site_reduce(part, chunk, offset)
    want = O.grain_reduce(part, chunk, offset)
    bad = np.flatnonzero(got.view(np.uint32) != want.view(np.uint32))
    assert bad.size == 0, (name, bad[:8], got[bad[:8]], want[bad[:8]])


@pytest.mark.gpu
def test_site_measure_then_reduce_matches_the_oracle_at_4k():
    W, H = 3840, 2160
    print("grain measure + reduce 3840x2160", flush=True)
    ring, out_t = scene(W, H, 1.5 / 255, shift=(1, 0), seed=5, texture=0.002)
    g = grid(W, H, 8, (1, 0))
    _, part = _site(ring, out_t, W, H, 1, 0, g)
    chunk = np.zeros(3 * 16 * 47, np.float32)
    got = _site_reduce(part, chunk, 16 * 47)
    _, want_part = O.grain_measure(ring, out_t, W, H, 1, 0, 0, 1, g[0], g[1], g[2], g[3], g[4])
    want = O.grain_reduce(want_part, chunk, 16 * 47)
    assert np.array_equal(got.view(np.uint32), want.view(np.uint32))
    assert want[16 * 47:2 * 16 * 47].reshape(16, 47)[:, 46].sum() > 0
