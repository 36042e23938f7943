"""The spatial noise measurement (the Immerkaer estimate, noise/spatial.rs and
noise/kernels/stats.rs:24-140): the oracle against the reference's own expectations and closed
forms on the CPU, then the HIP stages against the oracle, bit for bit, on the GPU."""
from __future__ import annotations

import ctypes

import numpy as np
import pytest

import noise_frames as F
import oracle as O

from avd import _lib as L

# Interior areas of a 70x20 frame's ragged 3x3 tile grid, summing to 1224 (noise/tests/spatial.rs).
RAGGED_AREAS = [[217.0, 224.0, 35.0], [248.0, 256.0, 40.0], [93.0, 96.0, 15.0]]


def total_sums(frame, width, height, channels):
    return O.noise_reduce(O.noise_partial(frame, width, height, channels))


# ---- host closed forms and the library's host helpers --------------------------------------------

def test_partials_len_matches_tile_grid():
    for (w, h), want in [((32, 8), 2), ((33, 9), 8), ((1920, 1080), 60 * 135 * 2)]:
        assert O.noise_partials_len(w, h) == want
        assert L.load().avd_noise_partials_len(w, h) == want


def test_sigma_from_abs_sum_below_three_samples_is_nan():
    # No interior sample of the 3x3 mask exists there (the reference's version underflows).
    for w, h in [(2, 2), (1, 64), (64, 2), (0, 0)]:
        assert np.isnan(L.load().avd_noise_sigma_from_abs_sum(1.0, w, h)), (w, h)
    assert np.isfinite(L.load().avd_noise_sigma_from_abs_sum(1.0, 3, 3))


def test_sigma_from_abs_sum_zero_and_formula():
    assert O.sigma_from_abs_sum(0.0, 64, 64) == 0.0
    assert L.load().avd_noise_sigma_from_abs_sum(0.0, 64, 64) == 0.0
    # sqrt(pi / 2) s / (6 (w - 2)(h - 2)), the library's f32 value equal to the oracle's.
    for s, w, h in [(123.25, 256, 256), (1.0e4, 3840, 2160), (7.5, 70, 20)]:
        want = O.sigma_from_abs_sum(s, w, h)
        assert abs(float(want) - np.sqrt(np.pi / 2) * s / (6 * (w - 2) * (h - 2))) <= 1e-6 * float(want)
        assert np.float32(L.load().avd_noise_sigma_from_abs_sum(s, w, h)) == want


@pytest.mark.parametrize("values,want", [([1, 2, 3, 4, 5], 2.0), ([1, 2, 3, 4], 1.75), ([10, 20, 30], 15.0),
                                         ([42], 42.0)])
def test_lower_quartile(values, want):
    assert abs(float(O.lower_quartile(values)) - want) < 1e-6


def test_block_p25_uniform_response_matches_frame_wide():
    partials = np.zeros(9 * 2, np.float32)
    total = np.float32(0.0)
    for ty, row in enumerate(RAGGED_AREAS):
        for tx, area in enumerate(row):
            s = np.float32(0.02) * np.float32(area)
            partials[(ty * 3 + tx) * 2] = s
            total += s
    got = O.sigma_block_p25_from_partials(partials, 1, 70, 20)[0]
    want = O.sigma_from_abs_sum(total, 70, 20)
    assert abs(got - want) < want * 1e-4


def test_block_p25_picks_the_third_smallest_of_nine():
    targets = [5.0, 2.0, 8.0, 1.0, 9.0, 3.0, 7.0, 4.0, 6.0]
    partials = np.zeros(9 * 2, np.float32)
    for ty, row in enumerate(RAGGED_AREAS):
        for tx, area in enumerate(row):
            k = ty * 3 + tx
            partials[k * 2] = np.float32(targets[k] / 255.0) * 6.0 * area / np.sqrt(np.float32(np.pi / 2))
    got = O.sigma_block_p25_from_partials(partials, 1, 70, 20)[0]
    assert abs(got - 3.0 / 255.0) < 1e-4


# ---- the mask, pinned by closed forms ----------------------------------------------------------

def test_an_impulse_sums_to_sixteen_times_its_height():
    """The mask [1 -2 1; -2 4 -2; 1 -2 1] around an interior impulse of height a: |4a| + 4 |-2a| +
    4 |a| = 16 a; a dropped term or a wrong weight changes it."""
    w, h, a = 40, 24, 0.25
    for (x, y) in [(2, 2), (17, 9), (37, 21), (31, 7), (32, 8)]:
        frame = np.zeros(w * h, np.float32)
        frame[y * w + x] = a
        tot = total_sums(frame, w, h, 1)
        assert tot[0] == np.float32(16 * a) and tot[1] == 0.0, (x, y, tot)


def test_affine_content_cancels_exactly():
    """The mask is orthogonal to affine content: exact (zero) on a plane of dyadic values."""
    w, h = 70, 20
    yy, xx = np.mgrid[0:h, 0:w]
    for c in (1, 2):
        planes = [(0.125 + xx / 256.0 + yy / 512.0), (0.75 - xx / 1024.0 + yy / 128.0)][:c]
        frame = np.stack(planes, axis=-1).astype(np.float32).reshape(-1)
        assert not total_sums(frame, w, h, c).any()


def test_a_two_channel_frame_sums_each_channel_in_its_lane():
    w, h = 48, 16
    one = F.noisy_gaussian(w, h, 1, 0.5, [8 / 255])
    two = np.stack([one, np.zeros_like(one)], -1).reshape(-1)
    t1, t2 = total_sums(one, w, h, 1), total_sums(two, w, h, 2)
    assert t2[0] == t1[0] and t2[1] == 0.0 and t1[1] == 0.0


# ---- the estimate's behaviour, with the reference's tolerances (nlmeans/tests/noise.rs) ---------

def test_recovers_a_known_sigma():
    w = h = 256
    true = 8.0 / 255.0
    est = O.sigma_from_abs_sum(total_sums(F.noisy_gaussian(w, h, 1, 0.5, [true]), w, h, 1)[0], w, h)
    assert abs(est - true) / true <= 0.20


def test_zero_on_uniform():
    w = h = 128
    est = O.sigma_from_abs_sum(total_sums(F.uniform(w, h, 1, 0.5), w, h, 1)[0], w, h)
    assert est < 0.2 / 255.0


def test_per_channel():
    w = h = 256
    su, sv = 8.0 / 255.0, 2.0 / 255.0
    tot = total_sums(F.noisy_gaussian(w, h, 2, 0.5, [su, sv]), w, h, 2)
    assert abs(O.sigma_from_abs_sum(tot[0], w, h) - su) / su <= 0.25
    assert abs(O.sigma_from_abs_sum(tot[1], w, h) - sv) / sv <= 0.25


def test_gradient_bias_bounded():
    w = h = 256
    true = 6.0 / 255.0
    noise = F.noisy_gaussian(w, h, 1, 0.5, [true])
    frame = np.clip(F.gradient(w, h, 0.2, 0.8) + (noise - np.float32(0.5)), 0.0, 1.0).astype(np.float32)
    est = O.sigma_from_abs_sum(total_sums(frame, w, h, 1)[0], w, h)
    assert abs(est - true) / true <= 0.30


# ---- the HIP stages against the oracle -----------------------------------------------------------

def _site(ring, w, h, c, frame, slots, slot):
    import torch

    from gpu_helpers import dev, host, require_gpu

    require_gpu()
    d_ring = dev(ring)
    d_part = torch.full((O.noise_partials_len(w, h),), -1.0, dtype=torch.float32, device="cuda")
    d_res = torch.full((2 * slots,), -7.0, dtype=torch.float32, device="cuda")
    rc = L.load().avd_site_noise_spatial(ctypes.c_void_p(d_ring.data_ptr()), w, h, c, frame,
                                        ctypes.c_void_p(d_part.data_ptr()), ctypes.c_void_p(d_res.data_ptr()),
                                        slot, None)
    L.check(rc)
    return host(d_part), host(d_res)


CASES = [
    ("ragged70x20", 70, 20, 1, 1, 0, 1, 0),
    ("odd33x9", 33, 9, 2, 1, 0, 1, 0),
    ("noise256", 256, 256, 1, 1, 0, 1, 0),
    ("chroma256", 256, 256, 2, 1, 0, 1, 0),
    ("ring3_frame1_slot2", 96, 40, 2, 3, 1, 3, 2),
    ("hd1080", 1920, 1080, 1, 1, 0, 1, 0),
    ("tiny3x3", 3, 3, 1, 1, 0, 1, 0),
]


@pytest.mark.gpu
@pytest.mark.parametrize("case", CASES, ids=[c[0] for c in CASES])
def test_site_noise_spatial_is_bit_exact(case):
    name, w, h, c, frames, frame, slots, slot = case
    ring = np.concatenate([F.noisy_gaussian(w, h, c, 0.25 + 0.2 * f, [6 / 255, 3 / 255]) +
                           np.float32(0.001 * f) for f in range(frames)]).astype(np.float32)
    print(f"noise site case {name}", flush=True)
    part, res = _site(ring, w, h, c, frame, slots, slot)
    want_part = O.noise_partial(ring, w, h, c, frame)
    assert np.array_equal(part, want_part), name
    want = O.noise_reduce(want_part, np.full(2 * slots, -7.0, np.float32), slot)
    assert np.array_equal(res, want), name
