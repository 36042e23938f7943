"""The host wire API (HostDenoiser: push / recv / flush over host bytes) on the GPU.

Its frames must be the device engine's frames byte for byte; the queue holds at most two frames
in flight (QueueFull past that, with no state change), a wrong plane length is refused without
poisoning the stream, and a y4m file denoised through it matches the engine too.
"""
from __future__ import annotations

import io
import os

import numpy as np
import pytest

import avd
from avd import _lib as L
from avd import y4m
from gpu_helpers import require_gpu, run_engine, words_dev

pytestmark = pytest.mark.gpu


def _wire(plane: np.ndarray, depth: int) -> bytes:
    return plane.astype(np.uint8 if depth == 8 else "<u2").tobytes()


def _nl4d(r=2):
    return avd.nl4d_options(temporal_radius=r, spatial_radius=6, sigma=16 / 1023, lambda_ht=2.7, pooled_threshold=0,
                            noise_map=0, motion={"estimation": 1}, field_lambda=0.0, step=3, group_size=16)


def _engine_frames(clip, geom, kind, luma, chroma, depth, dims):
    eng = avd.build_engine(kind, geom, luma=luma, chroma=chroma)
    outs = run_engine(eng, [[words_dev(p.numpy(), depth) for p in f] for f in clip], depth, dims)
    return [[w.view(np.uint8)[: (dw * dh) * (1 if depth == 8 else 2)].tobytes() for w, (dw, dh) in zip(o, dims)]
            for o in outs]


def test_host_denoiser_matches_the_engine_and_keeps_its_queue():
    require_gpu()
    W, H, depth, frames = 64, 48, 10, 7
    clip = avd.synth.clip(frames, W, H, depth, 9, chroma="420", sigma_luma=16.0, sigma_chroma=8.0)
    geom = L.Geometry(W, H, 1, L.SampleFormat.u16(depth), L.SampleFormat.u16(depth), 0)
    dims = [(W, H), (W // 2, H // 2), (W // 2, H // 2)]
    want = _engine_frames(clip, geom, "nl4d", _nl4d(), _nl4d(), depth, dims)
    den = avd.HostDenoiser("nl4d", geom, luma=_nl4d(), chroma=_nl4d())
    assert den.input_lengths == [W * H * 2, W * H // 2, W * H // 2]
    got = []
    queue_full_seen = False
    for f in clip:
        planes = [_wire(p.numpy(), depth) for p in f]
        if den.pending() < avd.HostDenoiser.MAX_PENDING:  # (a full queue is reported first)
            with pytest.raises(avd.PlaneMismatch):  # a short plane changes nothing
                den.push([planes[0][:-2]] + planes[1:])
        while True:
            try:
                den.push(planes)
                break
            except avd.QueueFull:
                queue_full_seen = True
                assert den.pending() == avd.HostDenoiser.MAX_PENDING
                got.append(den.recv())
    got += den.flush()
    assert queue_full_seen
    assert den.pending() == 0 and den.recv() is None
    assert len(got) == len(want) == frames
    for i, (g, w) in enumerate(zip(got, want)):
        assert g == w, f"frame {i}"
    # The denoiser takes a fresh stream after flush: the same clip again gives the same frames.
    again = []
    for f in clip:
        den.push([_wire(p.numpy(), depth) for p in f])
        out = den.recv()
        if out is not None:
            again.append(out)
    again += den.flush()
    assert again == want


def test_host_denoiser_nlmeans_gray_8bit_and_reset():
    require_gpu()
    W, H, depth, frames = 80, 40, 8, 5
    clip = avd.synth.clip(frames, W, H, depth, 21, chroma="gray", sigma_luma=10.0)
    geom = L.Geometry(W, H, 0, L.SampleFormat.u8(), L.SampleFormat.u8(), 0)
    opts = avd.nlmeans_options(temporal=1, temporal_radius=1, patch_radius=2, search_radius=3)
    want = _engine_frames(clip, geom, "nlmeans", opts, None, depth, [(W, H)])
    den = avd.HostDenoiser("nlmeans", geom, luma=opts)
    den.push([_wire(clip[0][0].numpy(), depth)])
    den.reset_stream()  # drops the stream: the next push starts a new one
    got = []
    for f in clip:
        den.push([_wire(f[0].numpy(), depth)])
        out = den.recv()
        if out is not None:
            got.append(out)
    got += den.flush()
    assert got == want


def test_y4m_file_through_the_host_denoiser(tmp_path):
    require_gpu()
    W, H, depth, frames = 48, 32, 10, 6
    clip = avd.synth.clip(frames, W, H, depth, 33, chroma="420", sigma_luma=16.0, sigma_chroma=8.0)
    header = y4m.Header(W, H, y4m.colorspace_for(1, depth), extensions=["COLORRANGE=LIMITED"])
    src, dst = os.path.join(tmp_path, "in.y4m"), os.path.join(tmp_path, "out.y4m")
    with open(src, "wb") as f:
        y4m.write(f, header, [[_wire(p.numpy(), depth) for p in fr] for fr in clip])
    assert y4m.denoise_file(src, dst, "nl4d", luma=_nl4d(1), chroma=_nl4d(1)) == frames
    geom = L.Geometry(W, H, 1, L.SampleFormat.u16(depth), L.SampleFormat.u16(depth), 0)
    want = _engine_frames(clip, geom, "nl4d", _nl4d(1), _nl4d(1), depth, header.plane_dims())
    with open(dst, "rb") as f:
        h2, it = y4m.read(io.BytesIO(f.read()))
        got = list(it)
    assert h2 == header
    assert got == want
