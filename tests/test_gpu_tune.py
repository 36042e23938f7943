"""The launch-site tuner on the GPU: its candidates compute the baseline's bits, so an engine gives
the same output with tuning on or off, and each tuned site records every candidate's time."""
from __future__ import annotations

import numpy as np
import pytest

import avd
from avd import _lib as L
from gpu_helpers import require_gpu, run_engine, words_dev

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("chroma,r,P,S", [("gray", 0, 2, 7), ("420", 2, 3, 4), ("gray", 1, 1, 3)])
def test_tuned_nlm_windows_match_the_baseline(chroma, r, P, S):
    require_gpu()
    W, H, depth = 200, 72, 10
    frames = 4
    clip = avd.synth.clip(frames, W, H, depth, 5 + r, chroma=chroma, sigma_luma=12.0, sigma_chroma=8.0)
    layout = 0 if chroma == "gray" else 1
    geom = L.Geometry(W, H, layout, L.SampleFormat.u16(depth), L.SampleFormat.u16(depth), 0)
    dims = avd.plane_dims(geom)
    opts = avd.nlmeans_options(temporal=1 if r else 0, temporal_radius=r, patch_radius=P, search_radius=S)
    outs = {}
    for tuned in (False, True):
        avd.set_tuning(tuned)
        eng = avd.build_engine("nlmeans", geom, luma=opts, chroma=opts if layout else None)
        outs[tuned] = run_engine(eng, [[words_dev(p.numpy(), depth) for p in f] for f in clip], depth, dims)
    avd.set_tuning(True)
    for i in range(frames):
        for k in range(len(dims)):
            assert np.array_equal(outs[True][i][k], outs[False][i][k]), (i, k)
    key = f"P{P} S{S} {W}x{H}"
    entries = [e for e in avd.tuning_entries() if key in e["key"]]
    assert entries, avd.tuning_entries()
    for e in entries:
        print(e)
        # every candidate (16, 8 or 4 rows per wave; 8 or 4 for the batched window) matched the baseline
        assert len(e["ms"]) == (2 if e["site"] == "nlm_window" else 3) and all(t >= 0 for t in e["ms"]), e
        assert e["ms"][e["choice"]] == min(e["ms"])
