"""The oracle against the digests and known answers the reference's own tests record.

Inputs are regenerated from the reference's test generators (``refsetups``); expected numbers are
the literals in ``collab/tests/fused/recorded.rs`` with that file's tolerances (2e-5 relative on
whole-run statistics, 1e-3 absolute per probe pixel, ``fused/mod.rs:262-281``).
"""
import numpy as np
import pytest

import oracle as O
import refsetups as R

SPATIAL_RADIUS, K_MAX, BLKSIZE, BLK_STEP, SIGMA, LAMBDA = 4, 8, 16, 8, 0.02, 5.3

RECORDED = {
    "unique": (12288, 0.500102660422, 0.577471243275, 1250.000000000,
               [0.917905456141, 0.787302672863, 0.650475382805, 0.517024146186,
                0.386674649788, 0.255359411240, 0.120508321126, 0.989773918601]),
    "transposed": (12288, 0.500026936557, 0.577366054447, 1238.896681776,
                   [0.075242505755, 0.724634047477, 0.367116374354, 0.014184951782,
                    0.659262769363, 0.307459000618, 0.951007338131, 0.589665272066]),
    "ragged": (9984, 0.500121022858, 0.577469311374, 1240.579711065,
               [0.745903455294, 0.891958951950, 0.032306798299, 0.177212221869,
                0.322843606131, 0.468612211367, 0.608005691977, 0.754309082031]),
    "short": (4096, 0.500333883408, 0.577536371630, 768.518540988,
              [0.836218530965, 0.571090123027, 0.304096429037, 0.033846737369,
               0.774293684286, 0.508250150663, 0.242787978384, 0.977499961853]),
    "noise": (4096, 0.500598531425, 0.500781913699, 168.165750156,
              [0.473047106911, 0.525827771943, 0.505852930189, 0.501965226326,
               0.472216666744, 0.498205827272, 0.493595121410, 0.515348414403]),
    "rho": (4096, 0.500501375321, 0.502136468679, 54.170715162,
            [0.439152209001, 0.572103197408, 0.475816598569, 0.509686441252,
             0.405882571403, 0.498373582524, 0.493889111273, 0.547764034977]),
    "cross": (16453, 0.441598762473, 0.546715666545, 1045.608471951,
              [0.837928771973, 0.574595237938, 0.300403234153, 0.000000000000,
               0.775989927049, 0.000000000000, 0.237124125163, 0.979660034180]),
    "jitter": (19992, 0.488864379137, 0.570867709963, 1233.333334961,
               [0.836363474528, 0.574619293213, 0.300458908081, 0.034561157227,
                0.774574279785, 0.513134002686, 0.238787333171, 0.979087829590]),
}


def run_case(ring, W, H, radius=0, S=SPATIAL_RADIUS, c_min=0.0, refine=0, mv=None, conf=None, slots=(0,),
             centre=0, bx=1, by=1, rho=0.0):
    frames = ring.size // (W * H)
    geom = O.MotionGeom(W, H, BLKSIZE, BLKSIZE - BLK_STEP)
    p = O.collab_params(W, H, 1, radius, S, refine, c_min, LAMBDA, SIGMA, 0.0, geom,
                        profile=O.dct_noise_profile(rho))
    p.blocks_x, p.blocks_y = bx, by
    mv = np.zeros(2, np.int32) if mv is None else mv
    conf = np.ones(1, np.float32) if conf is None else conf
    members = O.collab_search(ring.astype(np.float32), p, centre, list(slots), mv, conf)
    accum = np.zeros(frames * W * H, np.int32)
    wsum = np.zeros(frames * W * H, np.int32)
    gw = O.collab_filter(ring.astype(np.float32), p, centre, list(slots), members, accum, wsum, "f32")
    return R.digest(accum, wsum, gw), members


def check(name, got):
    covered, mean, rms, wmean, probes = RECORDED[name]
    assert got["covered"] == covered
    for have, want in ((got["pixel_mean"], mean), (got["pixel_rms"], rms), (got["weight_mean"], wmean)):
        assert abs(have - want) / max(abs(want), 1e-30) < 2e-5, (name, have, want)
    for i, (have, want) in enumerate(zip(got["probes"], probes)):
        assert abs(have - want) < 1e-3, (name, i, have, want)


def test_recorded_unique_content():
    check("unique", run_case(R.unique_frame(128, 96).reshape(-1), 128, 96)[0])


def test_recorded_transposed_ramp():
    src = R.unique_frame(96, 128)
    check("transposed", run_case(np.ascontiguousarray(src.T).reshape(-1), 128, 96)[0])


def test_recorded_ragged_reference_row():
    check("ragged", run_case(R.unique_frame(104, 96).reshape(-1), 104, 96)[0])


def test_recorded_short_search_space():
    check("short", run_case(R.unique_frame(64, 64).reshape(-1), 64, 64, S=1)[0])


def test_recorded_noise():
    check("noise", run_case(R.noisy_flat_field(64, 64, 0.5, 0.05).reshape(-1), 64, 64)[0])


def test_recorded_correlation_shaping():
    check("rho", run_case(R.noisy_flat_field(64, 64, 0.5, 0.05).reshape(-1), 64, 64, rho=0.86)[0])


def test_recorded_across_frames():
    W = H = 64
    mv, conf, slots, bx, by = R.cross_frame_fields(W, H, 2)
    ring = R.unique_frame(W, H * 5).reshape(-1)
    got, members = run_case(ring, W, H, radius=2, c_min=0.5, refine=2, mv=mv, conf=conf, slots=slots, centre=2,
                            bx=bx, by=by)
    words = members.reshape(-1, 9)[:, 8]
    assert np.any(words & 0x100) and np.any((words & 0x100) == 0), "both grid and fallback groups occur"
    check("cross", got)


def test_recorded_jittered_copies_scatter_into_every_frame():
    W = H = 64
    bx = by = 8
    ring = R.jittered_ring(W, H)
    mv = np.zeros(4 * bx * by * 2, np.int32)
    conf = np.ones(4 * bx * by, np.float32)
    frames = 5
    geom = O.MotionGeom(W, H, BLKSIZE, BLKSIZE - BLK_STEP)
    p = O.collab_params(W, H, 1, 2, SPATIAL_RADIUS, 0, 0.0, LAMBDA, SIGMA, 0.0, geom)
    members = O.collab_search(ring, p, 2, [0, 1, 3, 4], mv, conf)
    accum = np.zeros(frames * W * H, np.int32)
    wsum = np.zeros(frames * W * H, np.int32)
    gw = O.collab_filter(ring, p, 2, [0, 1, 3, 4], members, accum, wsum, "f32")
    for slot in range(5):
        assert wsum[slot * W * H:(slot + 1) * W * H].astype(np.int64).sum() > 0
    check("jitter", R.digest(accum, wsum, gw))
