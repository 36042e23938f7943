"""Known answers for the oracle's NLM window sums (single, pair with confidence, finish).

The NLM arithmetic the oracle restates (nlmeans/kernels/helpers.rs:36-72, fused.rs:138-530,
accumulate.rs:54-101, params.rs:229-268) is pinned here by closed forms that do not share its
structure: on a flat frame with one impulse of height delta, the patch distance of pixel x at
offset q is 3 delta^2 (luma channel scale) times the number of the two patches (at x and at x+q)
that hold the impulse, so every weight is exp(-h2 max(3 delta^2 n - offset, 0)) with n in {0, 1, 2}
read off the geometry alone.  Each check below fails under a named mistake planted in a copy of
the oracle's C source (``test_each_planted_mistake_is_caught``): a swapped channel scale, a
missing noise-offset floor, a self weight without the maximum weight, a backward neighbour at +q,
confidence applied to the wrong direction, and a pair maximum over one direction only.
"""
from __future__ import annotations

import ctypes
import os
import shutil
import subprocess

import numpy as np
import pytest

import oracle as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _impulse(W, H, a, z, delta):
    f = np.full((H, W), a, np.float64)
    f[z[1], z[0]] += delta
    return f.astype(np.float32)


def _inside(x, z, P):
    return max(abs(z[0] - x[0]), abs(z[1] - x[1])) <= P


def _zeros(n):
    return np.zeros(n), np.zeros(n), np.zeros(n)


# ---- single window --------------------------------------------------------------------------

W1, H1, P1, S1 = 17, 15, 1, 2
A1, DELTA1, Z1 = 0.4, 0.02, (8, 7)


def _single_closed_form(x, f, h2, offset):
    """wsum, accum, maxw of interior pixel x over q != 0, from the impulse count alone."""
    ws = acc = 0.0
    mw = 0.0
    d_unit = 3.0 * (float(f[Z1[1], Z1[0]]) - float(f[0, 0])) ** 2  # the f32 frame's own step
    for qy in range(-S1, S1 + 1):
        for qx in range(-S1, S1 + 1):
            if qx == 0 and qy == 0:
                continue
            n = int(_inside(x, Z1, P1)) + int(_inside((x[0] + qx, x[1] + qy), Z1, P1))
            w = np.exp(-h2 * max(d_unit * n - offset, 0.0))
            ws += w
            acc += w * float(f[x[1] + qy, x[0] + qx])
            mw = max(mw, w)
    return ws, acc, mw


def check_single_impulse(offset_ratio=0.0):
    """Single window, luma, an impulse on a flat frame; offset = offset_ratio * 3 delta^2."""
    f = _impulse(W1, H1, A1, Z1, DELTA1)
    h2 = float(O.h2_inv_norm(1.2, P1))
    d_unit = 3.0 * (float(f[Z1[1], Z1[0]]) - float(f[0, 0])) ** 2
    offset = offset_ratio * d_unit
    lut = O.spatial_offset_lut(S1, offset)
    acc, ws, mw = _zeros(W1 * H1)
    O.nlm_single(f.reshape(-1), W1, H1, 1, P1, S1, h2, lut, acc, ws, mw, "f64")
    m = P1 + S1
    for y in range(m, H1 - m):
        for x in range(m, W1 - m):
            want_ws, want_acc, want_mw = _single_closed_form((x, y), f, h2, float(np.float32(offset)))
            p = y * W1 + x
            assert abs(ws[p] - want_ws) <= 1e-12 * want_ws, ((x, y), ws[p], want_ws)
            assert abs(acc[p] - want_acc) <= 1e-12 * want_acc, ((x, y), acc[p], want_acc)
            assert abs(mw[p] - want_mw) <= 1e-15, ((x, y), mw[p], want_mw)
    return f, acc, ws, mw


def check_offset_floor():
    """With 3 delta^2 < offset < 6 delta^2, a patch pair holding the impulse once weighs exactly 1
    (max(d - offset, 0) = 0) and one holding it twice exp(-h2 (6 delta^2 - offset))."""
    f, acc, ws, mw = check_single_impulse(offset_ratio=1.5)
    p = Z1[1] * W1 + Z1[0]
    # At the impulse itself every q != 0 pairs two patches that both hold it (n = 2) when |q| <= P,
    # and one (n = 1) otherwise: 8 offsets at n = 2, 16 at n = 1.
    h2 = float(O.h2_inv_norm(1.2, P1))
    d_unit = 3.0 * (float(f[Z1[1], Z1[0]]) - float(f[0, 0])) ** 2
    w2 = np.exp(-h2 * (2 * d_unit - float(np.float32(1.5 * d_unit))))
    assert abs(ws[p] - (16.0 + 8.0 * w2)) <= 1e-12 * ws[p]
    assert mw[p] == 1.0


def check_two_channel_scale():
    """Two equal chroma channels scale by 1.5 each (helpers.rs:51-62): 2 d^2 1.5 = 3 d^2, the luma
    distance, so the weights of a U = V frame equal the luma frame's exactly."""
    rng = np.random.default_rng(11)
    W, H, P, S = 14, 11, 1, 2
    f = rng.uniform(0.2, 0.8, (H, W)).astype(np.float32)
    h2 = float(O.h2_inv_norm(1.2, P))
    lut = O.spatial_offset_lut(S, 0.0)
    a1, w1, m1 = _zeros(W * H)
    O.nlm_single(f.reshape(-1), W, H, 1, P, S, h2, lut, a1, w1, m1, "f64")
    f2 = np.repeat(f.reshape(-1), 2).astype(np.float32)
    a2, w2, m2 = np.zeros(W * H * 2), np.zeros(W * H), np.zeros(W * H)
    O.nlm_single(f2, W, H, 2, P, S, h2, lut, a2, w2, m2, "f64")
    assert np.array_equal(w1, w2)
    assert np.array_equal(m1, m2)
    assert np.array_equal(a2[0::2], a1) and np.array_equal(a2[1::2], a1)


# ---- pair window with confidence ------------------------------------------------------------

W2, H2, P2, S2 = 19, 17, 1, 2
ZF, ZB, DF, DB = (9, 8), (11, 6), 0.03, -0.025
CF, CB = 0.25, 0.5  # the backward direction holds the larger weight, so the maximum reads it
STEP2 = 4


def check_pair_impulses(offset_ratio=0.0):
    """A flat centre; the forward frame holds an impulse at ZF, the backward one at ZB.  The
    forward distance at q counts the impulse in the patch at x + q, the backward one in the patch at
    x - q; weights scale by their direction's block confidence and include q = 0."""
    a = 0.5
    c = np.full((H2, W2), a, np.float32)
    fw = _impulse(W2, H2, a, ZF, DF)
    bw = _impulse(W2, H2, a, ZB, DB)
    h2 = float(O.h2_inv_norm(1.0, P2))
    df_unit = 3.0 * (float(fw[ZF[1], ZF[0]]) - a) ** 2
    db_unit = 3.0 * (float(bw[ZB[1], ZB[0]]) - a) ** 2
    offset = float(np.float32(offset_ratio * min(df_unit, db_unit)))
    bx, by = (W2 + STEP2 - 1) // STEP2, (H2 + STEP2 - 1) // STEP2
    conf_f = np.full(bx * by, CF, np.float32)
    conf_b = np.full(bx * by, CB, np.float32)
    acc, ws, mw = _zeros(W2 * H2)
    O.nlm_pair(c.reshape(-1), fw.reshape(-1), bw.reshape(-1), W2, H2, 1, P2, S2, h2, offset, conf_f, conf_b, STEP2,
               bx, by, acc, ws, mw, "f64")
    m = P2 + S2
    for y in range(m, H2 - m):
        for x in range(m, W2 - m):
            want_ws = want_acc = want_mw = 0.0
            for qy in range(-S2, S2 + 1):
                for qx in range(-S2, S2 + 1):
                    nf = int(_inside((x + qx, y + qy), ZF, P2))
                    nb = int(_inside((x - qx, y - qy), ZB, P2))
                    wf = CF * np.exp(-h2 * max(df_unit * nf - offset, 0.0))
                    wb = CB * np.exp(-h2 * max(db_unit * nb - offset, 0.0))
                    want_ws += wf + wb
                    want_acc += wf * float(fw[y + qy, x + qx]) + wb * float(bw[y - qy, x - qx])
                    want_mw = max(want_mw, wf, wb)
            p = y * W2 + x
            assert abs(ws[p] - want_ws) <= 1e-12 * want_ws, ((x, y), ws[p], want_ws)
            assert abs(acc[p] - want_acc) <= 1e-12 * want_acc, ((x, y), acc[p], want_acc)
            assert abs(mw[p] - want_mw) <= 1e-15, ((x, y), mw[p], want_mw)


# ---- finish ---------------------------------------------------------------------------------

def check_finish_by_hand():
    """accumulate.rs:54-101: sw = self_weight * max_w, out = (x sw + accum) / (sw + wsum), and the
    input itself where the denominator is at most 1e-30."""
    frame = np.array([0.5, 0.25, 0.75], np.float32)
    accum = np.array([1.2, 0.0, 2.0])
    wsum = np.array([3.0, 0.0, 4.0])
    maxw = np.array([0.8, 0.0, 0.5])
    out = O.nlm_finish(frame, accum, wsum, maxw, 1.5, 3, 1, "f64")
    # pixel 0: sw = 1.2, (0.6 + 1.2) / 4.2; pixel 1: nothing gathered, passthrough;
    # pixel 2: sw = 0.75, (0.5625 + 2.0) / 4.75.
    assert abs(float(out[0]) - 1.8 / 4.2) <= 1e-7
    assert float(out[1]) == 0.25
    assert abs(float(out[2]) - 2.5625 / 4.75) <= 1e-7


# ---- host closed forms: the reference's known answers (nlmeans/tests/params.rs:17-67) -------

def test_noise_offset_known_answers():
    sigma = 4.0 / 255.0
    assert abs(float(O.noise_offset(sigma, 1, 4)) - 6.0 * sigma * sigma * 81.0) < 1e-6
    # Chroma: 2 * 1.5 * (sigma^2 + sigma^2) * area, the same 6 sigma^2 area as luma.
    assert abs(float(O.noise_offset(sigma, 2, 4)) - 6.0 * sigma * sigma * 81.0) < 1e-6


def test_h2_inv_norm_known_answers():
    sigma = 8.0 / 255.0
    eff = 1.0 * sigma * 255.0  # HQ auto strength: strength * sigma * 255
    assert abs(float(O.h2_inv_norm(eff, 4)) - 65025.0 / (3.0 * eff * eff * 81.0)) < 1e-6
    # Literal values: strength 1.2 over 5x5 and 9x9 patches.
    assert abs(float(O.h2_inv_norm(1.2, 2)) - 602.0833333) < 1e-3
    assert abs(float(O.h2_inv_norm(1.2, 4)) - 185.8281893) < 1e-3


def test_hq_strength_table_known_answers():
    # nlmeans/params.rs:13-16, read at radii 0, 2, 4, 8.
    assert [O.HQ_STRENGTH_LUMA[r] for r in (0, 2, 4, 8)] == [0.45, 0.42, 0.35, 0.30]
    assert [O.HQ_STRENGTH_CHROMA[r] for r in (0, 1, 2, 8)] == [1.00, 0.85, 0.70, 0.70]


# ---- the checks on the built oracle ---------------------------------------------------------

def test_single_window_impulse_closed_form():
    check_single_impulse()


def test_noise_offset_floor_gives_weight_one():
    check_offset_floor()


def test_two_equal_chroma_channels_weigh_as_luma():
    check_two_channel_scale()


def test_pair_window_impulses_and_confidence_closed_form():
    check_pair_impulses()
    check_pair_impulses(offset_ratio=0.5)


def test_finish_hand_computed():
    check_finish_by_hand()


# ---- planted mistakes -----------------------------------------------------------------------

MUTATIONS = {
    # name: (source text, replacement, checks that must fail)
    "swapped_channel_scale": ("R scale = (C == 1) ? (R)3.0 : (R)1.5;", "R scale = (C == 1) ? (R)1.5 : (R)3.0;",
                              [check_single_impulse, check_two_channel_scale]),
    "no_offset_floor": ("if (e < 0) e = 0;", "", [check_offset_floor]),
    "self_weight_without_max_w": ("R sw = (R)wref * maxw[p];", "R sw = (R)wref;", [check_finish_by_hand]),
    "backward_at_plus_q": ("nlm_patch_dist##SFX(cf, bf, W, H, C, P, x, y, -qx, -qy, scale)",
                           "nlm_patch_dist##SFX(cf, bf, W, H, C, P, x, y, qx, qy, scale)", [check_pair_impulses]),
    "confidence_swapped": ("wf *= (R)conf_fwd[bidx];", "wf *= (R)conf_bwd[bidx];", [check_pair_impulses]),
    "pair_max_forward_only": ("R m2 = wf > wb ? wf : wb;", "R m2 = wf;", [check_pair_impulses]),
}


@pytest.mark.skipif(shutil.which("gcc") is None, reason="needs gcc to build the mutated oracle")
@pytest.mark.parametrize("name", sorted(MUTATIONS))
def test_each_planted_mistake_is_caught(name, tmp_path, monkeypatch):
    old, new, checks = MUTATIONS[name]
    src = open(os.path.join(ROOT, "oracle", "avd_oracle.c")).read()
    assert src.count(old) >= 1, f"mutation {name} no longer applies"
    mutated = tmp_path / "mutated.c"
    mutated.write_text(src.replace(old, new))
    so = tmp_path / "libmutated.so"
    subprocess.run(["gcc", "-O1", "-std=gnu11", "-fPIC", "-fopenmp", "-ffp-contract=off", "-mfma", "-mavx2",
                    "-shared", "-o", str(so), str(mutated), "-lm"], check=True)
    lib = ctypes.CDLL(str(so))
    lib.avdo_match_confidence.restype = ctypes.c_float
    lib.avdo_max_threads.restype = ctypes.c_int
    monkeypatch.setattr(O, "_LIB", lib)
    for check in checks:
        with pytest.raises(AssertionError):
            check()
