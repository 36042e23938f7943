"""Host side of the grain table (nl4d/grain/fit.rs: bucket_edges, hist_median, scaling points,
quantise_ar), checked against the reference's known answers (nl4d/grain/tests/fit.rs:61-235) and
closed forms; no GPU."""
from __future__ import annotations

import ctypes

import numpy as np
import pytest

import oracle as O

from avd import _lib as L


def lib():
    return L.load()


def edges():
    e = np.zeros(65, np.float32)
    lib().avd_grain_bucket_edges(ctypes.c_void_p(e.ctypes.data))
    return e


def hist_median(counts):
    c = np.ascontiguousarray(counts, np.uint32)
    e = edges()
    found = ctypes.c_int32(-1)
    v = lib().avd_grain_hist_median(ctypes.c_void_p(c.ctypes.data), ctypes.c_void_p(e.ctypes.data),
                                     ctypes.byref(found))
    return v if found.value else None


def scaling_points(points, template):
    luma = np.array([p[0] for p in points], np.uint8)
    sigma = np.array([p[1] for p in points], np.float64)
    out = np.zeros(len(points), np.uint8)
    shift = ctypes.c_uint32(0)
    assert lib().avd_grain_scaling_points(ctypes.c_void_p(luma.ctypes.data), ctypes.c_void_p(sigma.ctypes.data),
                                          len(points), template, ctypes.c_void_p(out.ctypes.data),
                                          ctypes.byref(shift)) == 0
    return [(int(l), int(v)) for l, v in zip(luma, out)], shift.value


def test_bucket_edges_are_log_spaced_between_the_limits():
    # nl4d/grain/tests/fit.rs:61-71; the same values as the edges the measurement tests use.
    e = edges()
    assert np.array_equal(e, O.grain_bucket_edges())
    assert abs(float(e[0]) - 0.05 / 255) < 1e-9 and abs(float(e[64]) - 32 / 255) < 1e-7
    assert abs(e[1] / e[0] - e[64] / e[63]) < 1e-4


def test_histogram_median():
    # nl4d/grain/tests/fit.rs:85-113: within one bucket of the exact median; none when empty.
    e = edges()
    rng = np.random.default_rng(7)
    values = (2.0 + np.abs(rng.standard_normal(4000))) / 255.0
    counts = np.zeros(64, np.uint32)
    for v in values.astype(np.float32):
        counts[max(i for i in range(64) if i == 0 or v >= e[i])] += 1
    exact = float(np.median(values))
    est = hist_median(counts)
    ratio = float(e[1] / e[0])
    assert est / exact < ratio and exact / est < ratio
    assert hist_median(np.zeros(64)) is None
    one = np.zeros(64, np.uint32)
    one[20] = 10  # a single bucket: the geometric middle of its edges
    assert abs(hist_median(one) / np.sqrt(float(e[20]) * float(e[21])) - 1) < 1e-12


def test_finest_scaling_shift_fits_the_peak():
    # nl4d/grain/tests/fit.rs:198-204.
    assert [lib().avd_grain_finest_scaling_shift(p) for p in (0.1, 0.4, 0.9, 2.0)] == [11, 9, 8, 8]


def test_scaling_points_use_the_finest_shift_under_255():
    # nl4d/grain/tests/fit.rs:225-233.
    scaled, shift = scaling_points([(24, 1 / 255), (120, 2 / 255)], 40.0)
    assert shift == 11 and scaled == [(24, 51), (120, 102)]


@pytest.mark.parametrize("points", [[(24, 1 / 255), (120, 2 / 255)], [(40, 6 / 255), (200, 9 / 255)],
                                    [(16, 20 / 255), (80, 30 / 255), (240, 3 / 255)], [(128, 90 / 255)]])
def test_scaling_points_match_their_parts(points):
    # nl4d/grain/tests/fit.rs:176-196: the whole equals peak -> finest shift -> points at that shift.
    scaled, shift = scaling_points(points, 40.0)
    raw = [s * 255.0 / 40.0 for _, s in points]
    assert shift == lib().avd_grain_finest_scaling_shift(max(raw))
    assert scaled == [(l, int(min(np.round(r * (1 << shift)), 255.0))) for (l, _), r in zip(points, raw)]


def test_quantise_ar_matches_the_oracle_and_the_known_answers():
    w = np.zeros(24)
    w[23] = 0.555
    q = np.zeros(24, np.int32)
    s = ctypes.c_uint32(0)
    lib().avd_grain_quantise_ar(ctypes.c_void_p(w.ctypes.data), ctypes.c_void_p(q.ctypes.data), ctypes.byref(s))
    assert (s.value, q[23]) == (7, 71)  # nl4d/grain/tests/fit.rs:135-149
    rng = np.random.default_rng(3)
    for scale in (0.1, 0.3, 0.6, 1.2, 4.0):
        w = rng.uniform(-scale, scale, 24)
        lib().avd_grain_quantise_ar(ctypes.c_void_p(w.ctypes.data), ctypes.c_void_p(q.ctypes.data), ctypes.byref(s))
        want_q, want_s = O.grain_quantise(w)
        assert s.value == want_s and np.array_equal(q, want_q), scale


def pooled(hist, records):
    h = np.ascontiguousarray(hist, np.uint32)
    r = np.ascontiguousarray(records, np.float64)
    e = edges()
    out = np.full(47, -1.0)
    ok = lib().avd_grain_pooled_record(ctypes.c_void_p(h.ctypes.data), ctypes.c_void_p(r.ctypes.data),
                                       ctypes.c_void_p(e.ctypes.data), ctypes.c_void_p(out.ctypes.data))
    return out if ok else None


def test_pooled_record_sums_the_groups_near_the_median():
    # Every block in bucket 30 of brightness bins 3 and 9: the median is bucket 30's geometric
    # middle, and half to twice it, a 4x span of 1.106x buckets, reaches buckets 23.6 to 37.4, so
    # strength groups 5 to 9 (buckets 20-23 ... 36-39) pool.
    hist = np.zeros((16, 64), np.uint32)
    hist[3, 30], hist[9, 30] = 100, 200
    records = np.zeros((16, 47))
    records[:, :46] = np.arange(16)[:, None] + 1.0
    records[:, 46] = 20000.0
    rec = pooled(hist, records)
    assert rec is not None
    assert np.array_equal(rec[:46], np.full(46, float(sum(range(6, 11)))))
    assert rec[46] == 5 * 20000.0
    records[:, 46] = 9000.0  # 45000 pixels in the band: below the 50000 a texture needs
    assert pooled(hist, records) is None
    assert pooled(np.zeros((16, 64)), records) is None


def test_take_out_fine_noise_caps_the_floor_share():
    c = np.array([4.0] + [1.0] * 46)
    lib().avd_grain_take_out_fine_noise(ctypes.c_void_p(c.ctypes.data), 0.5, 0.5)
    assert c[0] == 4.0 - 0.25 * 4.0 and np.all(c[1:] == 1.0)
    c[0] = 4.0
    lib().avd_grain_take_out_fine_noise(ctypes.c_void_p(c.ctypes.data), 1.0, 3.0)  # at most 0.95 of the floor
    assert c[0] == 4.0 - 0.95 * 4.0


def floor_fraction(weights):
    w = np.ascontiguousarray(weights, np.float64)
    return lib().avd_grain_floor_fraction(ctypes.c_void_p(w.ctypes.data))


def test_floor_fraction_of_white_and_coarse_grain():
    # White grain (no weights) has a flat spectrum: the floor is all of it.  Strong positive
    # neighbour weights make coarse grain with almost no power at the finest frequencies
    # (nl4d/grain/fit.rs:218-224).  The closed form: a left weight a alone has power
    # 1 / (1 - 2a cos fx + a^2), lowest (1 + a)^-2 at fx = pi, and over the 64 frequencies a mean of
    # (1 + a^64) / ((1 - a^2)(1 - a^64)).
    assert floor_fraction(np.zeros(24)) == 1.0
    offsets = O.ar_offsets()
    for a in (0.2, 0.5, 0.8):
        w = np.zeros(24)
        w[offsets.index((0, -1))] = a
        want = (1 - a * a) / (1 + a) ** 2 * (1 - a ** 64) / (1 + a ** 64)
        assert abs(floor_fraction(w) / want - 1) < 1e-9, a
    coarse = np.zeros(24)
    coarse[offsets.index((0, -1))], coarse[offsets.index((-1, 0))] = 0.45, 0.45
    single = np.zeros(24)
    single[offsets.index((0, -1))] = 0.45
    assert floor_fraction(coarse) < floor_fraction(single) < floor_fraction(np.zeros(24))  # coarser, smaller floor
    assert 0.0 <= floor_fraction(np.full(24, 1e6)) <= 1.0
