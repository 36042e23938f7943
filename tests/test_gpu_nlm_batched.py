"""The NLMeans batched form on the GPU: every window of a frame and the finish in one launch.

Its outputs must equal the plain form's (one launch per temporal pair, the single window and the
finish, nlmeans/dispatch/mod.rs:665-679) bit for bit on whole streams: spatial, temporal, HQ with
motion compensation and per-neighbour confidence (luma and 4:2:0 chroma, whose chroma pass refines
the luma motion), confidence alone, and HQ at temporal radius 8.  The radius-8 stream is also held
against the oracle, within one code.
"""
from __future__ import annotations

import numpy as np
import pytest

import avd
from avd import _lib as L

import oracle as O
from gpu_helpers import gray_geometry, require_gpu, run_engine, words_dev

pytestmark = pytest.mark.gpu

CASES = {
    # name: (W, H, depth, chroma, frames, options)
    "spatial_8bit": (80, 48, 8, "gray", 3, dict(search_radius=7, patch_radius=2)),
    "temporal_420": (72, 40, 10, "420", 6, dict(temporal=1, temporal_radius=2, search_radius=4, patch_radius=3)),
    "hq_motion_420": (80, 64, 10, "420", 6, dict(hq=1, temporal=1, temporal_radius=2, search_radius=3, patch_radius=2,
                                                 sigma=12 / 1023, motion_compensation=1, motion={"estimation": 1})),
    "hq_confidence_only": (72, 56, 10, "gray", 5, dict(hq=1, temporal=1, temporal_radius=2, search_radius=3,
                                                       patch_radius=2, sigma=12 / 1023, motion_compensation=0)),
    "hq_radius8_motion": (64, 48, 10, "gray", 19, dict(hq=1, temporal=1, temporal_radius=8, search_radius=2,
                                                       patch_radius=2, sigma=10 / 1023, motion_compensation=1,
                                                       motion={"estimation": 1})),
}
LAYOUT = {"gray": 0, "420": 1}


def _run(case, batched):
    W, H, depth, chroma, frames, opts = CASES[case]
    clip = avd.synth.clip(frames, W, H, depth, 17 + W, chroma=chroma, sigma_luma=12.0, sigma_chroma=8.0)
    fmt = L.SampleFormat.u8() if depth == 8 else L.SampleFormat.u16(depth)
    geom = L.Geometry(W, H, LAYOUT[chroma], fmt, fmt, 0)
    o = avd.nlmeans_options(**opts)
    eng = avd.build_engine("nlmeans", geom, luma=o, chroma=o if chroma != "gray" else None)
    eng.set_nlm_form("batched" if batched else "plain")
    cd = avd.chroma_dims(LAYOUT[chroma], W, H)
    dims = [(W, H)] + ([cd, cd] if cd else [])
    return clip, run_engine(eng, [[words_dev(p.numpy(), depth) for p in f] for f in clip], depth, dims)


@pytest.mark.parametrize("case", list(CASES))
def test_batched_form_equals_the_plain_form(case):
    require_gpu()
    print(f"nlm batched vs plain: {case}", flush=True)
    _, plain = _run(case, False)
    _, batched = _run(case, True)
    assert len(plain) == len(batched) == CASES[case][4]
    for i, (p, b) in enumerate(zip(plain, batched)):
        for k, (x, y) in enumerate(zip(p, b)):
            assert np.array_equal(x, y), f"frame {i} plane {k}"


def test_radius8_hq_stream_matches_the_oracle():
    require_gpu()
    W, H, depth, _, frames, opts = CASES["hq_radius8_motion"]
    clip, got = _run("hq_radius8_motion", True)
    spec = O.NlmSpec(8, 2, 2, O.HQ_STRENGTH_LUMA[8], 1.0, hq=True, sigma=opts["sigma"],
                     motion=O.MotionGeom(0, 0, 16, 8, 4, 2))
    want, _ = O.run_stream(O.NlmPass(spec, W, H, 1, depth, depth), [[O.pack_plane(f[0].numpy(), depth)] for f in clip])
    assert len(got) == len(want) == frames
    for i in range(frames):
        a = O.unpack_plane(got[i][0], depth, W * H)
        b = O.unpack_plane(want[i][0], depth, W * H)
        assert int(np.max(np.abs(a - b))) <= 1, i
