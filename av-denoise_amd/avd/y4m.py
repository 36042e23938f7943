"""YUV4MPEG2 streams in and out, and a whole-file denoise through the host wire API.

The colourspace mapping follows the reference CLI's (av-denoise/src/bin/y4m_format.rs:15-85):
4:2:0, 4:2:2 and 4:4:4 at 8, 10 and 12 bits; 8-bit 4:2:0 carries its chroma siting (C420mpeg2
left, C420jpeg centre, C420paldv top-left, C420 unknown); float, 16-bit and gray streams are
refused; ``X``-prefixed vendor extensions (such as ``XCOLORRANGE=LIMITED``) are kept and written
back.  Frames are lists of plane byte strings (Y, U, V), 1 byte per sample at 8 bits and 2 bytes
little-endian above, as the host wire API takes them.
"""
from __future__ import annotations

from dataclasses import dataclass, field

MAGIC = b"YUV4MPEG2"

# colourspace tag -> (chroma layout: 1 4:2:0, 2 4:2:2, 3 4:4:4; bit depth; siting)
COLORSPACES = {
    "420": (1, 8, "unknown"), "420jpeg": (1, 8, "center"), "420mpeg2": (1, 8, "left"),
    "420paldv": (1, 8, "top_left"), "420p10": (1, 10, None), "420p12": (1, 12, None),
    "422": (2, 8, None), "422p10": (2, 10, None), "422p12": (2, 12, None),
    "444": (3, 8, None), "444p10": (3, 10, None), "444p12": (3, 12, None),
}


@dataclass
class Header:
    width: int
    height: int
    colorspace: str = "420"
    fps: str = "25:1"
    interlace: str = "p"
    aspect: str = "1:1"
    extensions: list = field(default_factory=list)  # vendor extensions without their leading X

    @property
    def chroma(self) -> int:
        return COLORSPACES[self.colorspace][0]

    @property
    def depth(self) -> int:
        return COLORSPACES[self.colorspace][1]

    def plane_dims(self):
        w, h = self.width, self.height
        cw, ch = {1: ((w + 1) // 2, (h + 1) // 2), 2: ((w + 1) // 2, h), 3: (w, h)}[self.chroma]
        return [(w, h), (cw, ch), (cw, ch)]

    def plane_lengths(self):
        bps = 1 if self.depth == 8 else 2
        return [w * h * bps for (w, h) in self.plane_dims()]

    def line(self) -> bytes:
        parts = [MAGIC.decode(), f"W{self.width}", f"H{self.height}", f"F{self.fps}", f"I{self.interlace}",
                 f"A{self.aspect}", f"C{self.colorspace}"] + [f"X{e}" for e in self.extensions]
        return (" ".join(parts) + "\n").encode()


def colorspace_for(chroma: int, depth: int, siting: str = "unknown") -> str:
    """subsampling_to_y4m (y4m_format.rs:30-55): the tag of a layout and depth; only 8-bit 4:2:0
    carries a siting."""
    if depth not in (8, 10, 12):
        raise ValueError("y4m output supports 8, 10 and 12-bit samples")
    if chroma == 1 and depth == 8:
        return {"left": "420mpeg2", "center": "420jpeg", "top_left": "420paldv"}.get(siting, "420")
    base = {1: "420", 2: "422", 3: "444"}[chroma]
    return base if depth == 8 else f"{base}p{depth}"


def parse_header(line: bytes) -> Header:
    tokens = line.strip().split(b" ")
    if not tokens or tokens[0] != MAGIC:
        raise ValueError("not a YUV4MPEG2 stream")
    h = Header(0, 0)
    for t in tokens[1:]:
        if not t:
            continue
        key, val = chr(t[0]), t[1:].decode()
        if key == "W":
            h.width = int(val)
        elif key == "H":
            h.height = int(val)
        elif key == "C":
            if val not in COLORSPACES:
                raise ValueError(f"unsupported y4m colorspace C{val}, need 4:2:0, 4:2:2, or 4:4:4 at 8, 10 or 12 bits")
            h.colorspace = val
        elif key == "F":
            h.fps = val
        elif key == "I":
            h.interlace = val
        elif key == "A":
            h.aspect = val
        elif key == "X":
            h.extensions.append(val)
    if h.width <= 0 or h.height <= 0:
        raise ValueError("y4m header without a frame size")
    return h


def read(stream):
    """(header, frame iterator) of a binary stream; each frame is [Y, U, V] bytes."""
    line = stream.readline()
    header = parse_header(line)
    lengths = header.plane_lengths()

    def frames():
        while True:
            tag = stream.readline()
            if not tag:
                return
            if not tag.startswith(b"FRAME"):
                raise ValueError("y4m frame header expected")
            planes = [stream.read(n) for n in lengths]
            if any(len(p) != n for p, n in zip(planes, lengths)):
                raise ValueError("truncated y4m frame")
            yield planes

    return header, frames()


def write(stream, header: Header, frames) -> int:
    stream.write(header.line())
    n = 0
    for planes in frames:
        stream.write(b"FRAME\n")
        for p in planes:
            stream.write(bytes(p))
        n += 1
    return n


def denoise_file(src_path: str, dst_path: str, kind: str = "nl4d", luma=None, chroma=None) -> int:
    """Denoises a y4m file into another through the host wire API, frame by frame (push, recv when
    the queue holds its limit, flush at the end); returns the frames written."""
    from . import Geometry, HostDenoiser, QueueFull, SampleFormat

    with open(src_path, "rb") as src, open(dst_path, "wb") as dst:
        header, frames = read(src)
        fmt = SampleFormat.for_depth(header.depth)
        geom = Geometry(header.width, header.height, header.chroma, fmt, fmt, 0)
        den = HostDenoiser(kind, geom, luma=luma, chroma=chroma)
        dst.write(header.line())
        written = 0

        def put(planes):
            nonlocal written
            dst.write(b"FRAME\n")
            for p in planes:
                dst.write(p)
            written += 1

        for planes in frames:
            while True:
                try:
                    den.push(planes)
                    break
                except QueueFull:
                    put(den.recv())
        for planes in den.flush():
            put(planes)
        den.close()
        return written
