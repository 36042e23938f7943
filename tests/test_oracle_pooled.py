"""The oracle's pooled threshold (collab/kernels/fused/pooled.rs) pinned by the reference's own
tests (collab/tests/fused/pooled.rs): its host restatement of the keep rule on its seeded groups,
and its known answers.  Groups are indexed [v][j][u] (the reference's [sub][j][i])."""
import numpy as np
import pytest

import oracle as O

F = np.float32
SIDE = 8
FLOOR = F(1.0e-20)
UNIT = np.ones(SIDE, F)


def seeded_group(seed: int) -> np.ndarray:
    """seeded_group (pooled.rs:163-182): xorshift32 from seed * 2654435761 + 1, one step per element
    in (v, j, u) order, value (state % 10000 / 10000 - 0.5) * 6."""
    g = np.zeros((SIDE, SIDE, SIDE), F)
    state = (seed * 2654435761 + 1) & 0xFFFFFFFF
    for v in range(SIDE):
        for j in range(SIDE):
            for u in range(SIDE):
                state ^= (state << 13) & 0xFFFFFFFF
                state ^= state >> 17
                state ^= (state << 5) & 0xFFFFFFFF
                unit = F(state % 10000) / F(10000.0)
                g[v, j, u] = (unit - F(0.5)) * F(6.0)
    return g


def host_reference(group, variances, profile, k_use, threshold, dc_lambda):
    """The reference test's host copy of the keep rule (pooled.rs:113-160), in f32."""
    variances, profile = np.asarray(variances, F), np.asarray(profile, F)
    thr, dcl = F(threshold), F(dc_lambda)

    def variance(v, j, u):
        return variances[j] * profile[u] * profile[v]

    def energy(v, j, u):
        c = group[v, j, u]
        row_inv = F(1.0) / max(variances[j] * profile[v], FLOOR)
        col_inv = F(1.0) / max(profile[u], FLOOR)
        return c * c * row_inv * col_inv

    result = group.copy()
    retained = F(0.0)
    for v in range(SIDE):
        for j in range(k_use):
            for u in range(SIDE):
                if v == 0 and u == 0:
                    keep = abs(group[v, j, u]) >= dcl * np.sqrt(variance(v, j, u)) or j == 0
                else:
                    s, count = energy(v, j, u), F(1.0)
                    for rv, cu in ((v, u - 1), (v, u + 1), (v - 1, u), (v + 1, u)):
                        if 0 <= rv < SIDE and 0 <= cu < SIDE and not (rv == 0 and cu == 0):
                            s = s + energy(rv, j, cu)
                            count += F(1.0)
                    keep = s >= thr * thr * count
                if keep:
                    retained = retained + variance(v, j, u)
                else:
                    result[v, j, u] = 0.0
    return result, retained


@pytest.mark.parametrize("seed,k_use", [(1, 8), (2, 8), (3, 4), (4, 2)])
def test_oracle_matches_the_reference_host_rule(seed, k_use):
    profile = np.array([1.3, 1.1, 1.0, 0.95, 0.9, 0.9, 0.9, 0.95], F)
    variances = np.array([1.0, 1.0, 1.0, 1.0, 1.2, 1.2, 1.2, 1.2], F)
    group = seeded_group(seed)
    got, lanes = O.pooled_threshold(group, variances, profile, k_use, 1.6, 2.7)
    want, want_retained = host_reference(group, variances, profile, k_use, 1.6, 2.7)
    assert np.array_equal(got, want)
    assert abs(float(lanes.sum()) - float(want_retained)) <= float(want_retained) * 1e-5
    assert np.count_nonzero(got != group) > 0, "some coefficients are zeroed"


def _single(cells, k_use=8, threshold=1.0, dc_lambda=1.0, variances=UNIT):
    group = np.zeros((SIDE, SIDE, SIDE), F)
    for (v, j, u), val in cells.items():
        group[v, j, u] = val
    got, ret = O.pooled_threshold(group, variances, UNIT, k_use, threshold, dc_lambda)
    return group, got, ret


def test_a_weak_coefficient_with_strong_neighbours_is_kept():
    _, got, _ = _single({(3, 1, 3): 0.5, (3, 1, 2): 2.0, (3, 1, 4): 2.0, (2, 1, 3): 2.0, (4, 1, 3): 2.0})
    assert got[3, 1, 3] == F(0.5)


def test_a_lone_strong_coefficient_is_zeroed():
    _, got, _ = _single({(3, 1, 3): 1.5})
    assert got[3, 1, 3] == 0.0


def test_a_corner_averages_over_its_two_in_range_neighbours():
    _, got, _ = _single({(7, 1, 7): 1.0, (7, 1, 6): 1.0, (6, 1, 7): 1.0})
    assert got[7, 1, 7] == 1.0


def test_the_spatial_dc_keeps_its_own_test_and_is_never_a_neighbour():
    _, got, _ = _single({(0, 1, 0): 0.5, (0, 1, 1): 0.9, (0, 2, 1): 0.9, (0, 2, 0): 100.0, (1, 2, 0): 0.9})
    assert got[0, 1, 0] == 0.0
    assert got[0, 2, 0] == 100.0
    assert got[0, 2, 1] == 0.0
    assert got[1, 2, 0] == 0.0


def test_the_group_dc_is_always_kept():
    _, got, _ = _single({(0, 0, 0): 1.0e-6})
    assert got[0, 0, 0] == F(1.0e-6)


def test_planes_past_k_use_are_left_untouched():
    group = seeded_group(7)
    got, _ = O.pooled_threshold(group, UNIT, UNIT, 2, 5.0, 5.0)
    assert np.array_equal(got[:, 2:, :], group[:, 2:, :])


def test_zero_noise_variance_stays_finite():
    group = seeded_group(9)
    got, ret = O.pooled_threshold(group, np.zeros(SIDE, F), UNIT, 8, 1.0, 1.0)
    assert np.all(np.isfinite(got)) and np.all(np.isfinite(ret))


def test_pool_ratio_is_the_calibrated_threshold_over_the_default_lambda():
    assert O.pool_ratio(1) == float(F(2.42) / F(4.158))
    assert O.pool_ratio(2) == float(F(2.42) / F(3.234))


def test_a_noiseless_flat_frame_passes_through_with_pooling_on():
    """pooled.rs:394-410: a flat 0.4 frame at sigma 1e-6 passes through (within 1e-3)."""
    W, H = 64, 48
    frame = np.full(W * H, 0.4, np.float32)
    geom = O.MotionGeom(W, H, 16, 8)
    p = O.collab_params(W, H, 1, 0, 4, 2, 0.5, 2.7, 1.0e-6, 2.0, geom, pooled=True)
    members = O.collab_search(frame, p, 0, [], np.zeros(2, np.int32), np.ones(1, np.float32))
    acc = np.zeros(W * H, np.int32)
    ws = np.zeros(W * H, np.int32)
    gw = O.collab_filter(frame, p, 0, [], members, acc, ws)
    assert np.all(np.isfinite(gw))
    out = O.collab_normalise(acc, ws, W * H, 1)
    assert np.max(np.abs(out - 0.4)) < 1e-3


def test_pooling_changes_the_output():
    import refsetups as R
    W, H = 64, 48
    mv, conf, slots, _, _ = R.cross_frame_fields(W, H, 2)
    ring = R.unique_frame(W, H * 5).reshape(-1)
    geom = O.MotionGeom(W, H, 16, 8)
    outs = []
    for pooled in (False, True):
        p = O.collab_params(W, H, 1, 2, 4, 2, 0.5, 5.3, 0.02, 0.0, geom, pooled=pooled)
        members = O.collab_search(ring, p, 2, slots, mv, conf)
        acc = np.zeros(5 * W * H, np.int32)
        ws = np.zeros(5 * W * H, np.int32)
        O.collab_filter(ring, p, 2, slots, members, acc, ws)
        outs.append(acc)
    assert not np.array_equal(outs[0], outs[1])
