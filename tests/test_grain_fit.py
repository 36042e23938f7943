"""The AR(3) grain fit in f64 (undo_mean_removal, yule_walker, solve and quantise_ar,
nl4d/grain/fit.rs:79-305).

On the CPU the oracle is pinned by the reference's own expectations (nl4d/grain/tests/fit.rs:115-345):
the quantisation known answers, Yule-Walker recovering known weights from a field's autocovariance,
an empty record rejected, and the mean-removal correction giving white grain zero weights and
recovering an AR field's correlation.  The records of the last two come from the oracle's grain
measurement and reduce, so the whole chain from samples to weights is pinned.  On the GPU the fit,
called through the C ABI, must equal the oracle bit for bit in f64.
"""
from __future__ import annotations

import ctypes

import numpy as np
import pytest
from scipy.signal import lfilter

import oracle as O

from avd import _lib as L

OFFSETS = O.ar_offsets()


def weights_with(**named):
    w = np.zeros(24)
    for (dy, dx), v in named.values():
        w[OFFSETS.index((dy, dx))] = v
    return w


SCENE_EIGHT = weights_with(a=((0, -1), 0.555), b=((-1, 0), 0.477), c=((0, -2), -0.172), d=((-2, 0), -0.148),
                           e=((-1, -1), -0.164), f=((-1, 1), 0.117))


def ar_field(weights, width, height, seed):
    """A causal AR field x = sum w(dy, dx) x[y + dy][x + dx] + e over unit white noise: each row is an
    IIR filter along x (the dy = 0 weights) of the innovation plus the rows above."""
    rng = np.random.default_rng(seed)
    pad = 32
    h, w = height + pad, width + 2 * pad
    x = np.zeros((h, w))
    row_taps = [-weights[OFFSETS.index((0, dx))] for dx in (-1, -2, -3)]
    for y in range(h):
        rhs = rng.standard_normal(w)
        for (dy, dx), c in zip(OFFSETS, weights):
            if dy < 0 and c != 0.0 and y + dy >= 0:
                rhs += c * np.roll(x[y + dy], -dx)
        x[y] = lfilter([1.0], [1.0] + row_taps, rhs)
    return x[pad:, pad:pad + width]


def autocov_of(field):
    """The 46 lag sums over every pixel pair inside the field, and the pixel count."""
    H, W = field.shape
    out = []
    for dy, dx in O.grain_lags():
        a = field[0:H - dy, max(0, -dx):W - max(0, dx)]
        b = field[dy:H, max(0, dx):W - max(0, -dx)]
        out.append(float((a * b).sum()))
    return np.array(out + [float(H * W)])


def measured_record(field, sigma=2 / 255):
    """The field as source grain (slot t flat, slot t + 1 the field times sqrt 2 over it), measured
    and reduced by the oracle, its 16 strength-group records pooled in f64."""
    H, W = field.shape
    g = field / field.std() * sigma
    ring = np.stack([np.full((H, W), 0.5), 0.5 + np.sqrt(2.0) * g]).astype(np.float32)
    out_t = np.full((H, W), 0.5, np.float32)
    bx, by = -(-W // 8), -(-H // 8)
    _, part = O.grain_measure(ring, out_t, W, H, 1, 0, 0, 1, np.zeros(2 * bx * by, np.int32),
                              np.ones(bx * by, np.float32), bx, by, 8)
    return O.grain_reduce(part).reshape(16, 47).astype(np.float64).sum(0)


def fit(record):
    corrected, weights, quantised, shift, solved = O.grain_fit(record)
    return corrected[0], weights[0], quantised[0], int(shift[0]), bool(solved[0])


# ---- the reference's expectations ------------------------------------------------------------------

def test_quantise_known_answers():
    # nl4d/grain/tests/fit.rs:135-174.
    w = np.zeros(24)
    w[23] = 0.555
    q, s = O.grain_quantise(w)
    assert (s, q[23]) == (7, 71)
    w[23] = 0.2
    q, s = O.grain_quantise(w)
    assert (s, q[23]) == (9, 102)
    for value, shift in [(0.2, 9), (0.4, 8), (0.7, 7), (1.5, 6), (3.0, 6)]:
        w = np.zeros(24)
        w[0], w[23] = -0.1, value
        q, s = O.grain_quantise(w)
        assert s == shift, value
        assert q[23] == min(127, int(np.floor(value * (1 << shift) + 0.5))) and q[0] == -round(0.1 * (1 << shift))
    w = np.zeros(24)
    w[0], w[1] = 3.0, -3.0
    q, _ = O.grain_quantise(w)
    assert (q[0], q[1]) == (127, -128)


def test_yule_walker_recovers_known_weights():
    # nl4d/grain/tests/fit.rs:115-125: the solve alone on a field's autocovariance.
    field = ar_field(SCENE_EIGHT, 400, 400, 11)
    weights = O.grain_yule_walker(autocov_of(field))
    assert weights is not None and np.max(np.abs(weights - SCENE_EIGHT)) < 0.01
    assert O.grain_yule_walker(np.zeros(47)) is None


def test_an_empty_record_is_rejected():
    # nl4d/grain/tests/fit.rs:127-133.
    corrected, weights, quantised, shift, solved = fit(np.zeros(47))
    assert not solved and not np.any(weights) and not np.any(quantised) and shift == 0
    assert not np.any(corrected)


def test_mean_correction_gives_white_grain_zero_weights():
    # nl4d/grain/tests/fit.rs:299-312, the record measured by the grain kernels.
    rng = np.random.default_rng(21)
    record = measured_record(rng.standard_normal((600, 600)))
    corrected, weights, _, _, solved = fit(record)
    assert solved and abs(weights.sum()) < 0.03
    raw = O.grain_yule_walker(record)  # without the correction the weights read the bias
    assert raw.sum() < -0.15
    measured_rho1 = record[1] / record[0]
    assert measured_rho1 < -0.01  # the cell mean's bias the correction undoes
    assert abs(corrected[1] / corrected[0]) < abs(measured_rho1) / 3


def test_mean_correction_recovers_an_ar_fields_correlation():
    # nl4d/grain/tests/fit.rs:314-345 with the record measured by the grain kernels.
    weights = weights_with(a=((0, -1), 0.3), b=((-1, 0), 0.3))
    field = ar_field(weights, 1016, 1016, 3)
    full = autocov_of(field)
    record = measured_record(field)
    corrected, got, _, _, solved = fit(record)
    truth, measured, recovered = full[:46] / full[0], record[:46] / record[0], corrected[:46] / corrected[0]
    assert np.max(np.abs(recovered - truth)) < 0.02
    assert truth[1] - measured[1] > 0.03
    _, full_weights, _, _, _ = fit(full)
    assert solved and np.max(np.abs(got - full_weights)) < 0.02


def test_site_rejects_null_arguments_without_a_gpu():
    import avd

    b = ctypes.c_void_p(16)
    assert L.load().avd_site_grain_fit(None, 1, b, b, b, b, b, None) == avd.NullArgument.status


# ---- the GPU fit against the oracle ------------------------------------------------------------------

def _site_fit(records):
    import torch

    from gpu_helpers import dev, host, require_gpu

    require_gpu()
    n = records.shape[0]
    d_rec = dev(np.ascontiguousarray(records, np.float64))
    outs = [torch.full((n, 47), -7.0, dtype=torch.float64, device="cuda"),
            torch.full((n, 24), -7.0, dtype=torch.float64, device="cuda"),
            torch.full((n, 24), -7, dtype=torch.int32, device="cuda"),
            torch.full((n,), 77, dtype=torch.int32, device="cuda"),
            torch.full((n,), -7, dtype=torch.int32, device="cuda")]
    L.check(L.load().avd_site_grain_fit(ctypes.c_void_p(d_rec.data_ptr()), n,
                                        *[ctypes.c_void_p(t.data_ptr()) for t in outs], None))
    return [host(t) for t in outs]


@pytest.mark.gpu
def test_site_grain_fit_is_bit_exact():
    print("grain fit: measured, synthetic, empty and singular records", flush=True)
    rng = np.random.default_rng(5)
    records = [measured_record(rng.standard_normal((200, 200))),
               measured_record(ar_field(SCENE_EIGHT, 256, 256, 2)),
               autocov_of(ar_field(SCENE_EIGHT, 128, 128, 4)),
               np.zeros(47),                                   # empty: no solve
               np.r_[np.ones(46), 64.0],                       # every lag equal (the correction breaks the tie)
               np.r_[np.zeros(46), 640.0]]                     # zero lag 0: no correction, singular
    for k in range(10):                                        # random decaying correlations
        rho = np.r_[1.0, 0.4 * rng.uniform(-1, 1, 45) * np.exp(-0.3 * np.arange(1, 46))]
        records.append(np.r_[rho * rng.uniform(0.5, 2.0) * 1e-3, 64.0 * rng.integers(20, 2000)])
    records = np.array(records)
    got = _site_fit(records)
    want = O.grain_fit(records)
    names = ["corrected", "weights", "quantised", "shifts", "solved"]
    for name, g, w in zip(names, got, want):
        g = np.asarray(g)
        w = np.asarray(w).astype(g.dtype)
        assert np.array_equal(g.view(np.uint8), w.view(np.uint8)), (name, g, w)
    assert want[4][:3].all() and not want[4][3] and not want[4][5]  # empty, and zero lag 0: unsolved
