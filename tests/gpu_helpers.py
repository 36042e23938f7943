"""Shared helpers of the GPU tests: device buffers, site calls and oracle-side counterparts."""
from __future__ import annotations

import ctypes

import numpy as np
import pytest
import torch

import avd
from avd import _lib as L
from avd import synth

import oracle as O


def require_gpu():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")


def dev(a: np.ndarray) -> torch.Tensor:
    return torch.from_numpy(np.ascontiguousarray(a)).to("cuda")


def host(t: torch.Tensor) -> np.ndarray:
    torch.cuda.synchronize()
    return t.cpu().numpy()


def ptr(t) -> int:
    return None if t is None else t.data_ptr()


def words_dev(codes: np.ndarray, depth: int) -> torch.Tensor:
    """An (H, W) code plane as a device uint8 tensor of packed u32 words."""
    w = O.pack_plane(codes, depth)
    return dev(w.view(np.uint8))


def words_host(t: torch.Tensor) -> np.ndarray:
    return host(t).view(np.uint32)


def collab_pass(W, H, C, radius, S, refine, c_min, lam, sigma, kaiser_beta, centre, slots, geom, profile=None,
                code_bits=0, pooled=False, step=4, group_size=8):
    p = L.CollabPass()
    p.code_bits = code_bits
    p.pool_ratio = O.pool_ratio(C) if pooled else 0.0
    p.width, p.height, p.channels = W, H, C
    p.step, p.spatial_radius, p.temporal_radius, p.refine, p.group_size = step, S, radius, refine, group_size
    p.blk_step, p.blksize, p.blocks_x, p.blocks_y = geom.step, geom.blksize, geom.blocks_x, geom.blocks_y
    p.centre_slot, p.slots = centre, slots
    p.c_min = c_min
    p.lam[0] = p.lam[1] = lam
    p.sigma[0] = p.sigma[1] = sigma
    prof = O.dct_noise_profile(0.0) if profile is None else profile
    p.weight_scale = float(O.weight_scale(sigma, float(prof.max())))
    p.accum_scale = float(O.cross_frame_accum_scale(S, radius, step, group_size))
    k = O.kaiser_window(kaiser_beta)
    for i in range(8):
        p.kaiser[i] = float(k[i])
        p.profile[i] = float(prof[i])
    return p


def site_collab_search(p, codes=None, values=None, nslots=None, mv=None, conf=None):
    refs = O.refs_along(p.width, p.step) * O.refs_along(p.height, p.step)
    # the kernels write group_size member words and the group word per reference
    members = torch.zeros(refs * (p.group_size + 1), dtype=torch.int32, device="cuda")
    ns = dev(np.asarray(nslots if nslots else [0], np.uint32).view(np.int32))
    mvd = dev(mv if mv is not None else np.zeros(2, np.int32))
    cfd = dev(conf if conf is not None else np.ones(1, np.float32))
    L.check(avd.load().avd_site_collab_search(ctypes.byref(p), ptr(codes), ptr(values), ptr(ns), ptr(mvd), ptr(cfd),
                                              ptr(members), None))
    return host(members).view(np.uint32)


def unpack_aggregate(agg: np.ndarray, second: np.ndarray | None, C: int):
    """The filter's packed accumulators as the reference's (accum [px*C] i32, wsum [px] i32)."""
    a = (agg & 0xFFFFFFFF).astype(np.uint32).view(np.int32)
    w = ((agg >> 32) & 0xFFFFFFFF).astype(np.uint32).view(np.int32) + (a < 0).astype(np.int32)
    accum = a if C == 1 else np.stack([a, second], 1).reshape(-1)
    return accum, w


def site_collab_filter(p, values, nslots, members, frames):
    refs = O.refs_along(p.width, p.step) * O.refs_along(p.height, p.step)
    px = p.width * p.height
    agg = torch.zeros(frames * px, dtype=torch.int64, device="cuda")
    second = torch.zeros(frames * px, dtype=torch.int32, device="cuda") if p.channels == 2 else None
    gw = torch.zeros(refs, dtype=torch.float32, device="cuda")
    ns = dev(np.asarray(nslots if nslots else [0], np.uint32).view(np.int32))
    md = dev(members.view(np.int32))
    L.check(avd.load().avd_site_collab_filter(ctypes.byref(p), ptr(values), ptr(ns), ptr(md), ptr(agg),
                                              ptr(second), ptr(gw), None))
    accum, wsum = unpack_aggregate(host(agg).view(np.uint64), host(second) if second is not None else None,
                                   p.channels)
    return accum, wsum, host(gw)


def gray_geometry(W, H, depth_in, depth_out=None):
    fmt_in = L.SampleFormat.for_depth(depth_in)
    fmt_out = L.SampleFormat.for_depth(depth_out or depth_in)
    return L.Geometry(W, H, 0, fmt_in, fmt_out, 0)


def run_engine(engine, frames_planes, depth_out, dims):
    """Pushes every frame (list of code planes per frame), emits each ready frame, finishes and
    emits the tail.  Returns the emitted frames as lists of u32 word arrays."""
    outs = []

    def out_bufs():
        return [torch.zeros(avd.plane_bytes(L.SampleFormat.for_depth(depth_out), w, h), dtype=torch.uint8,
                            device="cuda") for (w, h) in dims]

    for planes in frames_planes:
        ready = engine.push(planes)
        for _ in range(ready):
            bufs = out_bufs()
            engine.emit_into(bufs)
            outs.append([words_host(b) for b in bufs])
    tail = engine.finish()
    for _ in range(tail):
        bufs = out_bufs()
        engine.emit_into(bufs)
        outs.append([words_host(b) for b in bufs])
    return outs


def assert_members_equal(got: np.ndarray, want: np.ndarray, what: str = "", group_size: int = 8) -> None:
    """Member words match; on a mismatch names the first differing reference and its two groups."""
    g, w = got.reshape(-1, group_size + 1), want.reshape(-1, group_size + 1)
    bad = np.nonzero(np.any(g != w, axis=1))[0]
    if len(bad):
        i = int(bad[0])

        def fmt(row):
            return [f"({v & 0x1FFF},{(v >> 13) & 0x1FFF},t{v >> 26})" for v in row[:-1]] + [hex(int(row[-1]))]
        raise AssertionError(f"{what}: {len(bad)} references differ; first {i}: got {fmt(g[i])} want {fmt(w[i])}")
