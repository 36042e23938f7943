"""The y4m stream format and the host wire API's checks, without a GPU."""
from __future__ import annotations

import io

import numpy as np
import pytest

import avd
from avd import y4m


@pytest.mark.parametrize("chroma,depth", [(c, d) for c in (1, 2, 3) for d in (8, 10, 12)])
def test_colorspace_round_trips_every_supported_combination(chroma, depth):
    # y4m_format.rs:101-125: every layout and depth maps to a tag and back.
    tag = y4m.colorspace_for(chroma, depth)
    assert y4m.COLORSPACES[tag][:2] == (chroma, depth)


def test_only_eight_bit_420_carries_a_siting():
    assert [y4m.colorspace_for(1, 8, s) for s in ("left", "center", "top_left", "unknown")] == \
        ["420mpeg2", "420jpeg", "420paldv", "420"]
    assert y4m.colorspace_for(1, 10, "left") == "420p10"
    with pytest.raises(ValueError):
        y4m.colorspace_for(1, 16)


def test_header_parses_and_writes_back_with_vendor_extensions():
    line = b"YUV4MPEG2 W1920 H1080 F24000:1001 Ip A1:1 C420p10 XCOLORRANGE=LIMITED\n"
    h = y4m.parse_header(line)
    assert (h.width, h.height, h.colorspace, h.fps, h.depth, h.chroma) == (1920, 1080, "420p10", "24000:1001", 10, 1)
    assert h.extensions == ["COLORRANGE=LIMITED"]
    assert y4m.parse_header(h.line()) == h
    assert h.plane_lengths() == [1920 * 1080 * 2, 960 * 540 * 2, 960 * 540 * 2]


@pytest.mark.parametrize("bad", [b"YUV4MPEG2 W16 H16 Cmono\n", b"YUV4MPEG2 W16 H16 C444p16\n", b"YUV4MPEG2 H16\n",
                                 b"RIFF W16 H16\n"])
def test_unsupported_headers_are_refused(bad):
    with pytest.raises(ValueError):
        y4m.parse_header(bad)


def test_stream_round_trip_and_truncation():
    h = y4m.Header(17, 9, "422p12")
    rng = np.random.default_rng(3)
    frames = [[rng.integers(0, 4096, n // 2).astype("<u2").tobytes() for n in h.plane_lengths()] for _ in range(3)]
    buf = io.BytesIO()
    assert y4m.write(buf, h, frames) == 3
    buf.seek(0)
    h2, it = y4m.read(buf)
    assert h2 == h and list(it) == frames
    cut = io.BytesIO(buf.getvalue()[:-5])
    _, it = y4m.read(cut)
    with pytest.raises(ValueError):
        list(it)


def test_host_denoiser_rejects_float_planes_before_touching_the_device():
    geom = avd.Geometry(64, 48, 0, avd.SampleFormat(2, 0), avd.SampleFormat(2, 0), 0)
    with pytest.raises((avd.InvalidGeometry, avd.Gpu)):
        avd.HostDenoiser("nlmeans", geom, luma=avd.nlmeans_options())
