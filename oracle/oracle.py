"""CPU oracle for av-denoise's denoising core -- TEST INFRASTRUCTURE ONLY.

Only ``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of ``bench.py`` import
this module, and only to check results against.  The product (``av-denoise_amd``) never imports it.

It wraps ``liboracle.so`` (``avd_oracle.c``) and adds, in plain Python/numpy:

* the host-side closed forms the reference computes once per engine
  (``nlmeans/params.rs:229-268``, ``noise/correlation.rs:43-88``, ``collab/kernels/aggregate.rs:36-125``,
  ``collab/geometry.rs:6-29``, ``motion/confidence.rs:17-29``);
* the spatial noise measurement's host closed forms (``noise/spatial.rs:31-200``,
  ``noise/stats.rs:20-38``) around the C restatement of its two kernels;
* stream drivers that replay the reference's ring and pass scheduling:
  NLMeans with repeated edges (``nlmeans/denoiser/ring.rs``, ``nlmeans/engine.rs:201-262``,
  ``nlmeans/dispatch/mod.rs:653-679``) and NL4D with shifted edges and head/middle/tail passes
  (``nl4d/denoiser.rs:370-600``, ``nl4d/engine.rs:264-360``, ``nl4d/collab_pass.rs:50-200``).

Each function cites the reference file:line it restates; the tests under ``tests/`` pin them
against the reference's golden vectors and known answers.
"""
from __future__ import annotations

import ctypes
import math
import os
from dataclasses import dataclass, field

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

F32P = ctypes.POINTER(ctypes.c_float)
F64P = ctypes.POINTER(ctypes.c_double)
U32P = ctypes.POINTER(ctypes.c_uint32)
I32P = ctypes.POINTER(ctypes.c_int32)
U16P = ctypes.POINTER(ctypes.c_uint16)


class CollabParams(ctypes.Structure):
    _fields_ = [
        ("W", ctypes.c_int), ("H", ctypes.c_int), ("C", ctypes.c_int),
        ("step", ctypes.c_int), ("refs_x", ctypes.c_int), ("refs_y", ctypes.c_int),
        ("S", ctypes.c_int), ("radius", ctypes.c_int), ("grid_frames", ctypes.c_int),
        ("refine", ctypes.c_int), ("k_max", ctypes.c_int),
        ("blk_step", ctypes.c_int), ("blksize", ctypes.c_int),
        ("blocks_x", ctypes.c_int), ("blocks_y", ctypes.c_int),
        ("c_min", ctypes.c_float), ("lam", ctypes.c_float * 2), ("sigma", ctypes.c_float * 2),
        ("weight_scale", ctypes.c_float), ("accum_scale", ctypes.c_float),
        ("kaiser", ctypes.c_float * 8), ("profile", ctypes.c_float * 8),
        ("pool_ratio", ctypes.c_float),
    ]


def lib():
    """The compiled oracle; build it with ``make -C oracle`` (``__graft_entry__.build`` does)."""
    global _LIB
    if _LIB is None:
        path = os.path.join(HERE, "liboracle.so")
        if not os.path.exists(path):
            raise RuntimeError(f"{path} is missing; run `make -C oracle` first")
        _LIB = ctypes.CDLL(path)
        _LIB.avdo_match_confidence.restype = ctypes.c_float
        _LIB.avdo_max_threads.restype = ctypes.c_int
    return _LIB


def _p(a, t):
    return a.ctypes.data_as(t)


def set_threads(n: int) -> None:
    lib().avdo_set_threads(int(n))


def max_threads() -> int:
    return int(lib().avdo_max_threads())


# ----------------------------------------------------------------------------------------------
# Sample formats and plane I/O (engine/plane.rs:8-52, engine/kernels.rs:25-158)
# ----------------------------------------------------------------------------------------------

def samples_per_word(depth: int) -> int:
    return 4 if depth == 8 else 2


def max_value(depth: int) -> float:
    return float((1 << depth) - 1)


def pack_plane(codes: np.ndarray, depth: int) -> np.ndarray:
    """Pack integer codes (H, W) into little-endian u32 words, tail lanes zero."""
    flat = codes.reshape(-1).astype(np.uint32)
    spw = samples_per_word(depth)
    words = (flat.size + spw - 1) // spw
    padded = np.zeros(words * spw, np.uint32)
    padded[: flat.size] = flat
    padded = padded.reshape(words, spw)
    bits = 32 // spw
    out = np.zeros(words, np.uint32)
    for lane in range(spw):
        out |= padded[:, lane] << np.uint32(lane * bits)
    return out


def unpack_plane(words: np.ndarray, depth: int, pixels: int) -> np.ndarray:
    spw = samples_per_word(depth)
    bits = 32 // spw
    mask = np.uint32((1 << bits) - 1)
    lanes = [(words >> np.uint32(l * bits)) & mask for l in range(spw)]
    return np.stack(lanes, axis=1).reshape(-1)[:pixels].astype(np.int32)


def ingest(planes_words, pixels: int, channels: int, depth: int) -> np.ndarray:
    out = np.zeros(pixels * channels, np.float32)
    p0 = np.ascontiguousarray(planes_words[0], np.uint32)
    p1 = np.ascontiguousarray(planes_words[1] if channels > 1 else planes_words[0], np.uint32)
    lib().avdo_ingest_words(_p(p0, U32P), _p(p1, U32P), _p(out, F32P), ctypes.c_uint32(pixels),
                            ctypes.c_uint32(channels), ctypes.c_uint32(samples_per_word(depth)),
                            ctypes.c_float(max_value(depth)))
    return out


def egress(frame: np.ndarray, pixels: int, channels: int, depth: int):
    spw = samples_per_word(depth)
    words = (pixels + spw - 1) // spw
    p0 = np.zeros(words, np.uint32)
    p1 = np.zeros(words, np.uint32)
    f = np.ascontiguousarray(frame, np.float32)
    lib().avdo_egress_words(_p(f, F32P), _p(p0, U32P), _p(p1, U32P), ctypes.c_uint32(pixels),
                            ctypes.c_uint32(channels), ctypes.c_uint32(spw), ctypes.c_float(max_value(depth)))
    return [p0, p1][:channels]


# ----------------------------------------------------------------------------------------------
# Closed forms
# ----------------------------------------------------------------------------------------------

HQ_STRENGTH_LUMA = [0.45, 0.45, 0.42, 0.42, 0.35, 0.35, 0.35, 0.30, 0.30]    # nlmeans/params.rs:13
HQ_STRENGTH_CHROMA = [1.00, 0.85, 0.70, 0.70, 0.70, 0.70, 0.70, 0.70, 0.70]  # nlmeans/params.rs:16


def channel_scale(channels: int) -> float:
    return 3.0 if channels == 1 else 1.5  # nlmeans/kernels/helpers.rs:51-62


def h2_inv_norm(strength: float, patch_radius: int) -> np.float32:
    """NLM_NORM / (NLM_LEGACY * h^2 * area) in f32 (nlmeans/params.rs:238-243)."""
    area = np.float32((2 * patch_radius + 1) ** 2)
    s = np.float32(strength)
    return np.float32(np.float32(255.0 * 255.0) / (np.float32(3.0) * s * s * area))


def noise_offset(sigma: float, channels: int, patch_radius: int) -> np.float32:
    """2 * scale * sum(sigma^2) * area in f32 (nlmeans/params.rs:250-262)."""
    area = np.float32((2 * patch_radius + 1) ** 2)
    s = np.float32(sigma)
    sum_sq = np.float32(0.0)
    for _ in range(channels):
        sum_sq = np.float32(sum_sq + s * s)
    return np.float32(np.float32(2.0) * np.float32(channel_scale(channels)) * sum_sq * area)


def spatial_offset_lut(search_radius: int, offset: float) -> np.ndarray:
    """noise/correlation.rs:75-88 at rho = 0: the flat offset everywhere but the centre."""
    side = 2 * search_radius + 1
    lut = np.full(side * side, np.float32(offset), np.float32)
    lut[search_radius * side + search_radius] = 0.0
    return lut


def sad_noise_floor(blksize: int, sigma: float) -> np.float32:
    area = np.float32(blksize * blksize)
    return np.float32(area * np.float32(2.0) * np.float32(sigma) / np.float32(math.sqrt(math.pi)))


def thsad(blksize: int, scale: float) -> np.float32:
    return np.float32(np.float32(scale) * np.float32(blksize * blksize) * np.float32(0.02))


def kaiser_window(beta: float) -> np.ndarray:
    """aggregate.rs:92-125: I0(beta*sqrt(1-(2i/7-1)^2)) / I0(beta), f64 then f32."""
    def i0(x):
        half, term, s = x / 2.0, 1.0, 1.0
        for k in range(1, 32):
            term *= half / k
            s += term * term
        return s
    denom = i0(float(np.float32(beta)))
    out = np.zeros(8, np.float32)
    for i in range(8):
        pos = 2.0 * i / 7.0 - 1.0
        out[i] = np.float32(i0(float(np.float32(beta)) * math.sqrt(1.0 - pos * pos)) / denom)
    return out


def cross_frame_accum_scale(spatial_radius: int, temporal_radius: int, step: int = 4,
                            group_size: int = 8) -> np.float32:
    """aggregate.rs:36-55: the largest power of two keeping the worst case under i32::MAX/2.  The
    reference's STEP (4) and MAX_K (8) enter as the grid stride and the group size, so a stride-3,
    16-member grid gets the bound of its own worst case."""
    refs_per_axis = ((8 - 1) + 2 * spatial_radius) // step + 1
    contribs = float(refs_per_axis) * float(refs_per_axis) * float(group_size)
    passes = float(4 * temporal_radius + 1)
    max_raw = contribs * passes * 5.0
    budget = float(2**31 - 1) / 2.0
    exponent = math.floor(math.log2(budget / max_raw))
    return np.float32(2.0 ** exponent)


def dct8_basis(row: int, col: int) -> float:
    """transforms.rs:39-43, in double."""
    scale = 1.0 / math.sqrt(8.0) if row == 0 else 0.5
    return scale * math.cos(math.pi * (2.0 * col + 1.0) * row / 16.0)


def dct_noise_profile(rho: float) -> np.ndarray:
    """transforms.rs:196-221: g(u) = sum_ij B_u(i) B_u(j) rho^|i-j|, exactly ones at rho <= 0."""
    if rho <= 0.0:
        return np.ones(8, np.float32)
    r = float(np.float32(rho))
    out = np.zeros(8, np.float32)
    for u in range(8):
        acc = 0.0
        for i in range(8):
            for j in range(8):
                acc += dct8_basis(u, i) * dct8_basis(u, j) * r ** abs(i - j)
        out[u] = np.float32(acc)
    return out


def weight_scale(sigma: float, g_max: float = 1.0) -> np.float32:
    norm = np.float32(np.float32(sigma) * np.float32(sigma) * np.float32(g_max) * np.float32(g_max))
    return norm if (np.isfinite(norm) and norm > np.float32(1e-12)) else np.float32(1e-12)


def refs_along(dim: int, step: int = 4) -> int:
    return (dim - 8 + step - 1) // step + 1


def ref_pos(index: int, dim: int, step: int = 4) -> int:
    return min(index * step, dim - 8)


def grid_frames(temporal_radius: int) -> int:
    return 1 if temporal_radius == 0 else (2 if temporal_radius == 1 else 4)


# ----------------------------------------------------------------------------------------------
# Pass-level restatements (numpy in, numpy out)
# ----------------------------------------------------------------------------------------------

def nlm_single(frame, W, H, C, P, S, h2, lut, accum, wsum, maxw, precision="f32"):
    fn = lib().avdo_nlm_single_f32 if precision == "f32" else lib().avdo_nlm_single_f64
    pt = F32P if precision == "f32" else F64P
    fn(_p(frame, F32P), W, H, C, P, S, ctypes.c_float(h2), _p(lut, F32P), _p(accum, pt), _p(wsum, pt), _p(maxw, pt))


def nlm_pair(centre, fwd, bwd, W, H, C, P, S, h2, offset, conf_f, conf_b, step, bx, by, accum, wsum, maxw,
             precision="f32"):
    fn = lib().avdo_nlm_pair_f32 if precision == "f32" else lib().avdo_nlm_pair_f64
    pt = F32P if precision == "f32" else F64P
    cf = _p(conf_f, F32P) if conf_f is not None else None
    cb = _p(conf_b, F32P) if conf_b is not None else None
    fn(_p(centre, F32P), _p(fwd, F32P), _p(bwd, F32P), W, H, C, P, S, ctypes.c_float(h2), ctypes.c_float(offset),
       cf, cb, step, bx, by, _p(accum, pt), _p(wsum, pt), _p(maxw, pt))


def nlm_finish(frame, accum, wsum, maxw, wref, pixels, C, precision="f32"):
    out = np.zeros(pixels * C, np.float32)
    fn = lib().avdo_nlm_finish_f32 if precision == "f32" else lib().avdo_nlm_finish_f64
    pt = F32P if precision == "f32" else F64P
    fn(_p(frame, F32P), _p(accum, pt), _p(wsum, pt), _p(maxw, pt), ctypes.c_float(wref), pixels, C, _p(out, F32P))
    return out


def downscale_sum(level0: np.ndarray, w: int, h: int) -> np.ndarray:
    dw, dh = max(w // 2, 1), max(h // 2, 1)
    out = np.zeros(dw * dh, np.int32)
    src = np.ascontiguousarray(level0, np.int32)
    lib().avdo_downscale_sum(_p(src, I32P), w, h, _p(out, I32P), dw, dh)
    return out


@dataclass
class MotionGeom:
    """MotionCtx for an Mvtools mode (motion/context.rs:53-90) with the coarse level of analyse.rs:112-145."""
    W: int
    H: int
    blksize: int = 16
    overlap: int = 8
    search_radius: int = 4
    pyramid_levels: int = 2

    @property
    def step(self):
        return self.blksize - self.overlap

    @property
    def blocks_x(self):
        return max((self.W + self.step - 1) // self.step, 1)

    @property
    def blocks_y(self):
        return max((self.H + self.step - 1) // self.step, 1)

    @property
    def blocks(self):
        return self.blocks_x * self.blocks_y


def motion_search(geom: MotionGeom, pyr_centre, pyr_nb, maxv, sad_floor, th, want_conf=True):
    """Coarse (level 1) then seeded fine (level 0) block match of one neighbour against the centre."""
    mv = np.zeros(geom.blocks * 2, np.int32)
    conf = np.zeros(geom.blocks, np.float32) if want_conf else None
    L = lib()
    use_seed = 0
    if geom.pyramid_levels >= 2:
        lw, lh = max(geom.W // 2, 1), max(geom.H // 2, 1)
        cblk = max(geom.blksize // 2, 2)
        cstep = max(geom.step // 2, 1)
        L.avdo_block_match_coarse(_p(pyr_centre[1], I32P), _p(pyr_nb[1], I32P), lw, lh, cblk, cstep,
                                  geom.search_radius, 2, geom.blocks_x, geom.blocks_y, geom.step, _p(mv, I32P))
        use_seed = 1
    L.avdo_block_match_fine(_p(pyr_centre[0], I32P), _p(pyr_nb[0], I32P), geom.W, geom.H, geom.blksize, geom.step,
                            geom.search_radius, geom.blocks_x, geom.blocks_y, use_seed, _p(mv, I32P),
                            _p(conf, F32P) if want_conf else None, ctypes.c_float(maxv), ctypes.c_float(sad_floor),
                            ctypes.c_float(th))
    return mv, conf


MAPPED_REFINE_RADIUS = 1  # nlmeans/dispatch/motion.rs:42-43


def map_luma_fields(luma_mv, luma: MotionGeom, chroma: MotionGeom, shift_x: int, shift_y: int, fields: int):
    """map_luma_fields (motion/chroma_map.rs:56-98, kernels/chroma_map.rs:57-123): ``fields`` luma
    fields, one after another, onto the chroma block grid."""
    src = np.ascontiguousarray(luma_mv, np.int32)
    lstride, cstride = luma.blocks * 2, chroma.blocks * 2
    assert src.size >= fields * lstride
    out = np.zeros(max(fields * cstride, 2), np.int32)
    lib().avdo_map_luma_fields(_p(src, I32P), luma.blocks_x, luma.blocks_y, luma.step, luma.blksize, lstride,
                               _p(out, I32P), chroma.blocks_x, chroma.blocks_y, chroma.step, chroma.blksize, cstride,
                               shift_x, shift_y, fields)
    return out[: fields * cstride]


def motion_refine(geom: MotionGeom, pyr_centre, pyr_nb, seeds, maxv, sad_floor, th, radius=MAPPED_REFINE_RADIUS):
    """run_seeded_refine (motion/analyse.rs:321-367) of one neighbour at the single-level refine
    geometry (MotionCtx::single_level, motion/context.rs:108-116): the fine match on level 0,
    ``radius`` around each seed, with its confidence."""
    mv = np.array(seeds, np.int32, copy=True)
    conf = np.zeros(geom.blocks, np.float32)
    lib().avdo_block_match_fine(_p(pyr_centre[0], I32P), _p(pyr_nb[0], I32P), geom.W, geom.H, geom.blksize,
                                geom.step, radius, geom.blocks_x, geom.blocks_y, 1, _p(mv, I32P), _p(conf, F32P),
                                ctypes.c_float(maxv), ctypes.c_float(sad_floor), ctypes.c_float(th))
    return mv, conf


class LumaMotionShare:
    """LumaMotionRing (motion/luma_ring.rs:10-150) between a luma pass and the chroma pass of the
    same stream: the luma pass leaves each window's fields under the centre's stream frame
    (share_window_motion, nlmeans/dispatch/motion.rs:226-252) and the chroma pass of that frame maps
    them onto its block grid and refines them instead of searching (refine_luma_motion, :254-313).
    ``shift_x``/``shift_y`` are the chroma subsampling (1, 1 for 4:2:0; 1, 0 for 4:2:2; 0, 0 for
    4:4:4).  Frames are kept rather than rotated through 2r+1 slots: the luma stream may be run to
    its end before the chroma stream starts."""

    def __init__(self, shift_x: int, shift_y: int):
        self.shift_x, self.shift_y = shift_x, shift_y
        self.luma_geom = None
        self.fields = {}

    def put(self, frame: int, geom: MotionGeom, mv) -> None:
        self.luma_geom = geom
        self.fields[frame] = np.array(mv, np.int32, copy=True)

    def mapped(self, frame: int, chroma: MotionGeom, neighbours: int):
        if frame not in self.fields:
            raise RuntimeError(f"the luma pass has not shared the motion of stream frame {frame}")
        return map_luma_fields(self.fields[frame], self.luma_geom, chroma, self.shift_x, self.shift_y, neighbours)


def warp(src, mv, W, H, C, geom: MotionGeom):
    dst = np.zeros_like(src)
    lib().avdo_warp(_p(src, F32P), _p(dst, F32P), _p(mv, I32P), W, H, C, geom.step, geom.blocks_x, geom.blocks_y)
    return dst


POOLED_THRESHOLD = 2.42  # nl4d/options.rs:109


def default_lambda_ht(C: int) -> float:
    """nl4d_default_lambda_ht (nl4d/options.rs:101-106): luma 4.158, chroma 3.234."""
    return 4.158 if C == 1 else 3.234


def pool_ratio(C: int) -> float:
    """nl4d_pool_ratio (nl4d/options.rs:115-117): the calibrated pooled threshold over the plane's
    default lambda, in f32."""
    return float(np.float32(POOLED_THRESHOLD) / np.float32(default_lambda_ht(C)))


def collab_params(W, H, C, radius, S, refine, c_min, lam, sigma, kaiser_beta, geom: MotionGeom, step=4,
                  profile=None, sigmas=None, pooled=False, group_size=8):
    p = CollabParams()
    p.pool_ratio = pool_ratio(C) if pooled else 0.0
    p.W, p.H, p.C = W, H, C
    p.step = step
    p.refs_x, p.refs_y = refs_along(W, step), refs_along(H, step)
    p.S, p.radius, p.grid_frames, p.refine, p.k_max = S, radius, grid_frames(radius), refine, group_size
    p.blk_step, p.blksize, p.blocks_x, p.blocks_y = geom.step, geom.blksize, geom.blocks_x, geom.blocks_y
    p.c_min = c_min
    sig = list(sigmas) if sigmas is not None else [sigma, sigma]
    prof = dct_noise_profile(0.0) if profile is None else np.asarray(profile, np.float32)
    for c in range(2):
        p.lam[c] = float(lam)
        p.sigma[c] = float(sig[min(c, len(sig) - 1)])
    p.weight_scale = float(weight_scale(sig[0], float(prof.max())))
    for i in range(8):
        p.profile[i] = float(prof[i])
    p.accum_scale = float(cross_frame_accum_scale(S, radius, step, group_size))
    k = kaiser_window(kaiser_beta)
    for i in range(8):
        p.kaiser[i] = float(k[i])
    return p


def collab_search(ring, params: CollabParams, centre_slot, nslots, mv, conf):
    """search_group over every reference; ``ring`` is u16 codes (exact SSD) or f32 values."""
    refs = params.refs_x * params.refs_y
    members = np.zeros(refs * (params.k_max + 1), np.uint32)  # k_max member words, then the group word
    ns = np.ascontiguousarray(nslots, np.uint32) if len(nslots) else np.zeros(1, np.uint32)
    mvv = np.ascontiguousarray(mv, np.int32) if mv is not None and mv.size else np.zeros(2, np.int32)
    cf = np.ascontiguousarray(conf, np.float32) if conf is not None and conf.size else np.zeros(1, np.float32)
    if ring.dtype == np.uint16:
        fn, rp = lib().avdo_collab_search_u16, _p(ring, U16P)
    else:
        fn, rp = lib().avdo_collab_search_f32, _p(np.ascontiguousarray(ring, np.float32), F32P)
    fn(rp, ctypes.byref(params), centre_slot, _p(ns, U32P), _p(mvv, I32P), _p(cf, F32P), _p(members, U32P))
    return members


def collab_search_refs(ring, params: CollabParams, centre_slot, nslots, mv, conf, refs):
    """collab_search for the listed references only (row-major ids): members[len(refs)][k_max + 1]."""
    ids = np.ascontiguousarray(refs, np.uint32)
    members = np.zeros(max(ids.size, 1) * (params.k_max + 1), np.uint32)
    ns = np.ascontiguousarray(nslots, np.uint32) if len(nslots) else np.zeros(1, np.uint32)
    mvv = np.ascontiguousarray(mv, np.int32) if mv is not None and mv.size else np.zeros(2, np.int32)
    cf = np.ascontiguousarray(conf, np.float32) if conf is not None and conf.size else np.zeros(1, np.float32)
    if ring.dtype == np.uint16:
        fn, rp = lib().avdo_collab_search_refs_u16, _p(ring, U16P)
    else:
        fn, rp = lib().avdo_collab_search_refs_f32, _p(np.ascontiguousarray(ring, np.float32), F32P)
    fn(rp, ctypes.byref(params), centre_slot, _p(ns, U32P), _p(mvv, I32P), _p(cf, F32P), _p(ids, U32P), int(ids.size),
       _p(members, U32P))
    return members[: ids.size * (params.k_max + 1)]


def collab_filter(values_ring, params: CollabParams, centre_slot, nslots, members, accum, wsum, precision="f32"):
    refs = params.refs_x * params.refs_y
    gw = np.zeros(refs, np.float32)
    ns = np.ascontiguousarray(nslots, np.uint32) if len(nslots) else np.zeros(1, np.uint32)
    fn = lib().avdo_collab_filter_f32 if precision == "f32" else lib().avdo_collab_filter_f64
    fn(_p(values_ring, F32P), ctypes.byref(params), centre_slot, _p(ns, U32P), _p(members, U32P), _p(accum, I32P),
       _p(wsum, I32P), _p(gw, F32P))
    return gw


def pooled_threshold(group, variances, profile, k_use, threshold, dc_lambda, precision="f32"):
    """pooled_threshold (fused/pooled.rs:24-147) over one group given as [v][j][u] (the
    reference's [sub][j][i]); returns the thresholded group in that layout and the per-lane
    retained variances."""
    dt = np.float32 if precision == "f32" else np.float64
    ptr = F32P if precision == "f32" else ctypes.POINTER(ctypes.c_double)
    stk = np.ascontiguousarray(np.transpose(np.asarray(group, dt), (1, 0, 2)))  # [j][v][u]
    var = np.ascontiguousarray(variances, dt)
    prof = np.ascontiguousarray(profile, np.float32)
    ret = np.zeros(8, dt)
    fn = lib().avdo_pooled_threshold_f32 if precision == "f32" else lib().avdo_pooled_threshold_f64
    cast = ctypes.c_float if precision == "f32" else ctypes.c_double
    fn(_p(stk, ptr), _p(var, ptr), _p(prof, F32P), int(k_use), cast(threshold), cast(dc_lambda), _p(ret, ptr))
    return np.transpose(stk, (1, 0, 2)).copy(), ret


def collab_normalise(accum_region, wsum_region, pixels, C):
    out = np.zeros(pixels * C, np.float32)
    lib().avdo_collab_normalise(_p(accum_region, I32P), _p(wsum_region, I32P), _p(out, F32P),
                                ctypes.c_size_t(pixels), C)
    return out


# ----------------------------------------------------------------------------------------------
# Stream drivers
# ----------------------------------------------------------------------------------------------

@dataclass
class NlmSpec:
    """Resolved NLM parameters of one pass (nlmeans/options.rs:106-137, nlmeans/params.rs:172-262)."""
    temporal_radius: int = 0
    search_radius: int = 2
    patch_radius: int = 4
    strength: float = 1.2
    self_weight: float = 1.0
    hq: bool = False
    sigma: float | None = None            # HQ pinned sigma (normalised)
    auto_strength: bool = True
    noise_floor: bool = True
    temporal_confidence: bool = True
    thsad_scale: float = 1.0
    motion: MotionGeom | None = None      # Mvtools block geometry (W/H filled per pass)

    def h2(self):
        strength = self.strength
        if self.hq and self.auto_strength and self.sigma is not None:
            strength = float(np.float32(np.float32(self.strength) * np.float32(self.sigma) * np.float32(255.0)))
        return h2_inv_norm(strength, self.patch_radius)

    def offset(self, channels):
        if self.hq and self.noise_floor and self.sigma is not None:
            return noise_offset(self.sigma, channels, self.patch_radius)
        return np.float32(0.0)


class _Ring:
    """A ring of 2r+1 ingested slots with their motion pyramids (nlmeans/denoiser/ring.rs)."""

    def __init__(self, W, H, C, depth, total):
        self.W, self.H, self.C, self.depth, self.total = W, H, C, depth, total
        self.values = np.zeros((total, H * W * C), np.float32)
        self.codes = np.zeros((total, H * W * C), np.uint16)
        self.pyr = [None] * total
        self.slot_frames = [0] * total
        self.head = 0
        self.loaded = 0
        self.real = 0

    def phys(self, logical):
        start = self.head % self.total
        return (start + (logical % self.total)) % self.total

    def _build_pyramid(self, slot):
        l0 = self.codes[slot].reshape(self.H * self.W, self.C)[:, 0].astype(np.int32)
        self.pyr[slot] = [l0, downscale_sum(l0, self.W, self.H)]

    def write(self, words):
        slot = self.head % self.total
        px = self.W * self.H
        self.values[slot] = ingest(words, px, self.C, self.depth)
        codes = [unpack_plane(w, self.depth, px) for w in words[: self.C]]
        self.codes[slot] = np.stack(codes, axis=1).reshape(-1).astype(np.uint16)
        self._build_pyramid(slot)
        self.slot_frames[slot] = self.real
        self.head += 1
        if self.loaded < self.total:
            self.loaded += 1
        self.real += 1

    def duplicate_last(self):
        last = (self.head - 1) % self.total
        nxt = self.head % self.total
        self.values[nxt] = self.values[last]
        self.codes[nxt] = self.codes[last]
        self.pyr[nxt] = self.pyr[last]
        self.slot_frames[nxt] = self.slot_frames[last]
        self.head += 1

    def ready(self):
        return self.loaded >= self.total


class NlmPass:
    """One NLMeans plane group over a stream (Repeat edges)."""

    def __init__(self, spec: NlmSpec, W, H, C, depth_in, depth_out, precision="f32", share=None, reads_motion=False):
        self.s, self.W, self.H, self.C = spec, W, H, C
        self.depth_out = depth_out
        self.precision = precision
        self.share, self.reads_motion = share, reads_motion
        self.r = spec.temporal_radius
        self.ring = _Ring(W, H, C, depth_in, 2 * self.r + 1)
        self.out = []
        if spec.motion is not None and self.r > 0:
            spec.motion = MotionGeom(W, H, spec.motion.blksize, spec.motion.overlap, spec.motion.search_radius,
                                     spec.motion.pyramid_levels)
        self.maxv = max_value(depth_in)

    def push(self, words):
        self.ring.write(words)
        if self.ring.real == 1 and self.r > 0:
            for _ in range(self.r):
                self.ring.duplicate_last()
                self.ring.loaded += 1
        if self.ring.ready():
            self.out.append(self._denoise())
        return len(self.out)

    def finish(self):
        target = min(self.ring.real, self.r)
        produced = 0
        while produced < target:
            self.ring.duplicate_last()
            if self.ring.loaded < self.ring.total:
                self.ring.loaded += 1
            if self.ring.ready():
                self.out.append(self._denoise())
                produced += 1
        self.ring = _Ring(self.W, self.H, self.C, self.ring.depth, self.ring.total)
        return target

    def _denoise(self):
        s, W, H, C, r = self.s, self.W, self.H, self.C, self.r
        px = W * H
        dt = np.float32 if self.precision == "f32" else np.float64
        accum = np.zeros(px * C, dt)
        wsum = np.zeros(px, dt)
        maxw = np.zeros(px, dt)
        h2 = float(s.h2())
        off = float(s.offset(C))
        centre = self.ring.phys(r)
        cf = self.ring.values[centre]
        frames = {}
        confs = {}
        geom = s.motion
        conf_only = r > 0 and geom is None and s.hq and s.temporal_confidence
        if conf_only:
            # run_confidence_pass (nlmeans/dispatch/motion.rs:554-612) with the confidence_only block
            # geometry (motion/context.rs:182-200): 16x16 blocks at step 8 scored where they stand
            # (radius 0, one level); the neighbours stay unwarped.
            geom = MotionGeom(W, H, 16, 8, 0, 1)
            sigma = s.sigma if s.sigma is not None else 0.0
            floor = float(sad_noise_floor(geom.blksize, sigma))
            th = float(thsad(geom.blksize, s.thsad_scale))
            for k in list(range(-r, 0)) + list(range(1, r + 1)):
                slot = self.ring.phys(r + k)
                _, conf = motion_search(geom, self.ring.pyr[centre], self.ring.pyr[slot], self.maxv, floor, th)
                frames[k] = self.ring.values[slot]
                confs[k] = conf
        elif r > 0 and geom is not None:
            sigma = s.sigma if s.sigma is not None else 0.0
            floor = float(sad_noise_floor(geom.blksize, sigma))
            th = float(thsad(geom.blksize, s.thsad_scale))
            frame = self.ring.slot_frames[centre]
            ks = list(range(-r, 0)) + list(range(1, r + 1))
            per = geom.blocks * 2
            mapped = self.share.mapped(frame, geom, len(ks)) if self.reads_motion else None
            shared = []
            for i, k in enumerate(ks):
                slot = self.ring.phys(r + k)
                if mapped is not None:
                    mv, conf = motion_refine(geom, self.ring.pyr[centre], self.ring.pyr[slot],
                                             mapped[i * per:(i + 1) * per], self.maxv, floor, th)
                else:
                    mv, conf = motion_search(geom, self.ring.pyr[centre], self.ring.pyr[slot], self.maxv, floor, th)
                shared.append(mv)
                frames[k] = warp(self.ring.values[slot], mv, W, H, C, geom)
                confs[k] = conf if (s.hq and s.temporal_confidence) else None
            if self.share is not None and not self.reads_motion:
                self.share.put(frame, geom, np.concatenate(shared))
        else:
            for k in range(-r, r + 1):
                if k:
                    frames[k] = self.ring.values[self.ring.phys(r + k)]
                    confs[k] = None
        for qk in range(-r, 0):
            fwd, bwd = frames[qk], frames[-qk]
            gb = geom if geom is not None else MotionGeom(W, H)
            nlm_pair(cf, fwd, bwd, W, H, C, s.patch_radius, s.search_radius, h2, off, confs[qk], confs[-qk],
                     gb.step, gb.blocks_x, gb.blocks_y, accum, wsum, maxw, self.precision)
        lut = spatial_offset_lut(s.search_radius, off)
        nlm_single(cf, W, H, C, s.patch_radius, s.search_radius, h2, lut, accum, wsum, maxw, self.precision)
        res = nlm_finish(cf, accum, wsum, maxw, s.self_weight, px, C, self.precision)
        return res


@dataclass
class Nl4dSpec:
    """Resolved NL4D parameters of one pass (nl4d/options.rs:163-185, nl4d/params.rs:27-198)."""
    temporal_radius: int = 2
    refine: int = 2
    spatial_radius: int = 9
    lambda_ht: float = 4.158
    c_min: float = 0.05
    kaiser_beta: float = 2.0
    sigma: float = 0.02
    thsad_scale: float = 1.0
    motion: MotionGeom = field(default_factory=lambda: MotionGeom(0, 0))
    step: int = 4
    pooled: bool = False
    group_size: int = 8


class Nl4dPass:
    """One NL4D plane group over a stream (Shifted edges, head/middle/tail passes)."""

    def __init__(self, spec: Nl4dSpec, W, H, C, depth_in, depth_out, precision="f32", share=None, reads_motion=False):
        self.s, self.W, self.H, self.C = spec, W, H, C
        self.depth_out = depth_out
        self.precision = precision
        self.share, self.reads_motion = share, reads_motion
        self.r = spec.temporal_radius
        self.N = 2 * self.r + 1
        self.ring = _Ring(W, H, C, depth_in, self.N)
        self.geom = MotionGeom(W, H, spec.motion.blksize, spec.motion.overlap, spec.motion.search_radius,
                               spec.motion.pyramid_levels)
        self.maxv = max_value(depth_in)
        self.params = collab_params(W, H, C, self.r, spec.spatial_radius, spec.refine, spec.c_min, spec.lambda_ht,
                                    spec.sigma, spec.kaiser_beta, self.geom, spec.step, pooled=spec.pooled,
                                    group_size=spec.group_size)
        px = W * H
        self.accum = np.zeros(self.N * px * C, np.int32)
        self.wsum = np.zeros(self.N * px, np.int32)
        self.passes = 0
        self.out = []
        self.last_members = None

    # nl4d/collab_pass.rs:50-200 and nlmeans/denoiser/machinery.rs:45-101
    def _run_pass(self, centre_logical, clear):
        px = self.W * self.H
        C = self.C
        centre = self.ring.phys(centre_logical)
        if clear == "whole":
            self.accum[:] = 0
            self.wsum[:] = 0
        elif clear == "newest":
            newest = (centre + self.r) % self.N
            self.accum[newest * px * C:(newest + 1) * px * C] = 0
            self.wsum[newest * px:(newest + 1) * px] = 0
        nslots = [self.ring.phys(l) for l in range(self.N) if l != centre_logical]
        sig = float(self.s.sigma)
        floor = float(sad_noise_floor(self.geom.blksize, sig))
        th = float(thsad(self.geom.blksize, self.s.thsad_scale))
        mvs, confs = [], []
        frame = self.ring.slot_frames[centre]
        per = self.geom.blocks * 2
        mapped = self.share.mapped(frame, self.geom, len(nslots)) if self.reads_motion else None
        for i, slot in enumerate(nslots):
            if mapped is not None:
                mv, conf = motion_refine(self.geom, self.ring.pyr[centre], self.ring.pyr[slot],
                                         mapped[i * per:(i + 1) * per], self.maxv, floor, th)
            else:
                mv, conf = motion_search(self.geom, self.ring.pyr[centre], self.ring.pyr[slot], self.maxv, floor, th)
            mvs.append(mv)
            confs.append(conf)
        if self.share is not None and not self.reads_motion and mvs:
            self.share.put(frame, self.geom, np.concatenate(mvs))
        mv = np.concatenate(mvs) if mvs else np.zeros(2, np.int32)
        conf = np.concatenate(confs) if confs else np.zeros(1, np.float32)
        members = collab_search(self.ring.codes, self.params, centre, nslots, mv, conf)
        self.last_members = members
        self.last_fields = (mv, conf)
        collab_filter(self.ring.values, self.params, centre, nslots, members, self.accum, self.wsum, self.precision)
        self.passes += 1

    def _read_region(self, slot):
        px = self.W * self.H
        C = self.C
        a = self.accum[slot * px * C:(slot + 1) * px * C]
        w = self.wsum[slot * px:(slot + 1) * px]
        out = collab_normalise(a, w, px, C)
        self.accum[slot * px * C:(slot + 1) * px * C] = 0
        self.wsum[slot * px:(slot + 1) * px] = 0
        return out

    def push(self, words):
        self.ring.write(words)
        if not self.ring.ready():
            return len(self.out)
        r = self.r
        clear = "whole" if self.passes == 0 else "newest"
        if self.passes == 0:
            for c in range(r):
                self._run_pass(c, clear)
                clear = "none"
        self._run_pass(r, clear)
        if self.passes > r:
            centre = self.ring.phys(r)
            slot = (centre + self.N - r) % self.N
            self.out.append(self._read_region(slot))
        return len(self.out)

    def finish(self):
        real = self.ring.real
        emit = min(real, 2 * self.r)
        if emit == 0:
            return 0
        short = real < self.N
        if short:
            while not self.ring.ready():
                self.ring.duplicate_last()
                self.ring.loaded += 1
            centres, last_real = list(range(emit)), emit - 1
        else:
            centres, last_real = list(range(self.r + 1, 2 * self.r + 1)), 2 * self.r
        clear = "whole" if (short or self.passes == 0) else "none"
        for c in centres:
            self._run_pass(c, clear)
            clear = "none"
        first = last_real + 1 - emit
        for logical in range(first, last_real + 1):
            self.out.append(self._read_region(self.ring.phys(logical)))
        self.ring = _Ring(self.W, self.H, self.C, self.ring.depth, self.N)
        self.passes = 0
        return emit


def chroma_shifts(chroma: str):
    """The subsampling shifts of a chroma layout name ("420", "422", "444")."""
    return {"420": (1, 1), "422": (1, 0), "444": (0, 0)}[chroma]


def link_motion(luma_pass, chroma_pass, chroma: str = "420"):
    """Has ``chroma_pass`` refine ``luma_pass``'s motion instead of searching its own, as the
    reference's planar engine does whenever both passes search motion (engine/spec.rs:350-365,
    :474-492).  Run the luma stream before (or interleaved ahead of) the chroma stream."""
    share = LumaMotionShare(*chroma_shifts(chroma))
    luma_pass.share, luma_pass.reads_motion = share, False
    chroma_pass.share, chroma_pass.reads_motion = share, True
    return share


def run_stream(pass_obj, frames_words):
    """Push every frame, then finish; returns the egressed output words per emitted frame."""
    outputs = []
    for words in frames_words:
        pass_obj.push(words)
        while pass_obj.out:
            outputs.append(pass_obj.out.pop(0))
    pass_obj.finish()
    while pass_obj.out:
        outputs.append(pass_obj.out.pop(0))
    px = pass_obj.W * pass_obj.H
    return [egress(f, px, pass_obj.C, pass_obj.depth_out) for f in outputs], outputs


# ---------------------------------------------------------------------------------------------
# Spatial noise measurement (the Immerkaer estimate)
# ---------------------------------------------------------------------------------------------

IMMERKAER_BLOCK_X, IMMERKAER_BLOCK_Y, IMMERKAER_LANES = 32, 8, 2  # noise/mod.rs:114-122


def noise_partials_len(width: int, height: int) -> int:
    """partials_len (noise/spatial.rs:31-34): two floats per 32x8 tile."""
    return -(-width // IMMERKAER_BLOCK_X) * -(-height // IMMERKAER_BLOCK_Y) * IMMERKAER_LANES


def noise_partial(ring: np.ndarray, width: int, height: int, channels: int, frame: int = 0) -> np.ndarray:
    """nlm_noise_partial (noise/kernels/stats.rs:31-99) over frame ``frame`` of an f32 ring
    [frame][H][W][C], in the reference's summation order (avdo_noise_partial)."""
    ring = np.ascontiguousarray(ring, np.float32).reshape(-1)
    out = np.zeros(noise_partials_len(width, height), np.float32)
    lib().avdo_noise_partial(_p(ring, F32P), ctypes.c_uint32(width), ctypes.c_uint32(height),
                             ctypes.c_uint32(channels), ctypes.c_uint32(frame), _p(out, F32P))
    return out


def noise_reduce(partials: np.ndarray, results: np.ndarray | None = None, slot: int = 0) -> np.ndarray:
    """nlm_noise_reduce (noise/kernels/stats.rs:101-140): the two per-channel totals into row
    ``slot`` of ``results`` (avdo_noise_reduce)."""
    partials = np.ascontiguousarray(partials, np.float32)
    if results is None:
        results = np.zeros(2 * (slot + 1), np.float32)
    lib().avdo_noise_reduce(_p(partials, F32P), ctypes.c_uint32(partials.size // 2), _p(results, F32P),
                            ctypes.c_uint32(slot))
    return results


def immerkaer_sigma_scale() -> np.float32:
    """sqrt(pi / 2) in f32 (noise/spatial.rs:127-130)."""
    return np.sqrt(np.float32(math.pi / 2))


def sigma_from_abs_sum(abs_sum: float, width: int, height: int) -> np.float32:
    """sigma_from_abs_sum (noise/spatial.rs:132-138), f32 arithmetic."""
    interior = np.float32(width - 2) * np.float32(height - 2)
    return immerkaer_sigma_scale() * np.float32(abs_sum) / (np.float32(6.0) * interior)


def lower_quartile(sorted_values) -> np.float32:
    """lower_quartile (noise/stats.rs:20-38) of a sorted, non-empty sequence: the quarter point,
    interpolated between its neighbours, in f32."""
    v = [np.float32(x) for x in sorted_values]
    if len(v) == 1:
        return v[0]
    position = np.float32(0.25) * np.float32(len(v) - 1)
    lo, hi = int(math.floor(position)), int(math.ceil(position))
    if lo == hi:
        return v[lo]
    fraction = position - np.float32(lo)
    return v[lo] + fraction * (v[hi] - v[lo])


def sigma_block_p25_from_partials(partials, channels: int, width: int, height: int) -> np.ndarray:
    """sigma_block_p25_from_partials (noise/spatial.rs:140-200): each tile's sigma over its overlap
    with the frame interior (tiles with none skipped), then the lower quartile per channel."""
    cx, cy = -(-width // IMMERKAER_BLOCK_X), -(-height // IMMERKAER_BLOCK_Y)
    scale = immerkaer_sigma_scale()
    sigmas = [[] for _ in range(channels)]
    for ty in range(cy):
        y0, y1 = ty * IMMERKAER_BLOCK_Y, min((ty + 1) * IMMERKAER_BLOCK_Y, height)
        oy0, oy1 = max(y0, 1), min(y1, height - 1)
        for tx in range(cx):
            x0, x1 = tx * IMMERKAER_BLOCK_X, min((tx + 1) * IMMERKAER_BLOCK_X, width)
            ox0, ox1 = max(x0, 1), min(x1, width - 1)
            if ox1 <= ox0 or oy1 <= oy0:
                continue
            area = np.float32((ox1 - ox0) * (oy1 - oy0))
            base = (ty * cx + tx) * IMMERKAER_LANES
            for c in range(channels):
                sigmas[c].append(scale * np.float32(partials[base + c]) / (np.float32(6.0) * area))
    out = np.zeros(IMMERKAER_LANES, np.float32)
    for c in range(channels):
        if sigmas[c]:
            out[c] = lower_quartile(sorted(sigmas[c]))
    return out


# ---- temporal block statistics (noise/temporal.rs:18-131, noise/kernels/stats.rs:142-660) --------

TEMPORAL_BLOCK = 16
TEMPORAL_QUARTER_BASE, TEMPORAL_QUARTER_FIELDS = 3, 6  # noise/temporal.rs:36-38
BLOCK_TEXTURE_LANE, BLOCK_MOVED_LANE = 1, 2            # noise/temporal.rs:27-31
QUARTER_SUM_D, QUARTER_SUM_D2, QUARTER_LUMA_SUM, QUARTER_FLATNESS, QUARTER_LUMA_MIN, QUARTER_LUMA_MAX = range(6)
UNMEASURED_FLATNESS = np.float32(3.0e38)               # noise/temporal.rs:62
COMPENSATION_GAIN_MAX = np.float32(0.5)                # noise/temporal.rs:44


def temporal_record_len(channels: int) -> int:
    """temporal_stats_record_len (noise/temporal.rs:117-119)."""
    return 2 * channels + TEMPORAL_QUARTER_BASE + 4 * TEMPORAL_QUARTER_FIELDS + channels - 1


def temporal_texture_lane(channels: int, channel: int) -> int:
    """texture_lane (noise/temporal.rs:125-131)."""
    if channel == 0:
        return 2 * channels + BLOCK_TEXTURE_LANE
    return 2 * channels + TEMPORAL_QUARTER_BASE + 4 * TEMPORAL_QUARTER_FIELDS + channel - 1


def temporal_blocks(width: int, height: int) -> tuple[int, int]:
    """temporal_stats_blocks (noise/temporal.rs:134-139): ragged edges truncated."""
    return -(-width // TEMPORAL_BLOCK), -(-height // TEMPORAL_BLOCK)


def temporal_slot_len(width: int, height: int, channels: int) -> int:
    """temporal_stats_slot_len (noise/temporal.rs:141-144)."""
    bx, by = temporal_blocks(width, height)
    return bx * by * temporal_record_len(channels)


def temporal_stats(ring: np.ndarray, width: int, height: int, channels: int, slot_new: int, slot_prev: int,
                   luma_fields: bool, motion=None) -> np.ndarray:
    """nlm_temporal_noise_stats (noise/kernels/stats.rs:191-604) over an f32 ring [slot][H][W][C]:
    one record per 16x16 block, in the reference's summation orders (avdo_temporal_stats).
    ``motion`` is None or (field i32, offset, step, blocks_x, blocks_y) with vectors (x, y)."""
    ring = np.ascontiguousarray(ring, np.float32).reshape(-1)
    out = np.full(temporal_slot_len(width, height, channels), np.nan, np.float32)
    if motion is None:
        field, args = None, (0, 1, 1, 1)
    else:
        field = np.ascontiguousarray(motion[0], np.int32)
        args = tuple(int(v) for v in motion[1:])
    lib().avdo_temporal_stats(_p(ring, F32P), ctypes.c_uint32(width), ctypes.c_uint32(height),
                              ctypes.c_uint32(channels), ctypes.c_uint32(slot_new), ctypes.c_uint32(slot_prev),
                              None if field is None else _p(field, I32P), *[ctypes.c_uint32(v) for v in args],
                              ctypes.c_int(1 if luma_fields else 0), _p(out, F32P))
    return out


# ---- film-grain measurement (nl4d/kernels/grain.rs:243-513, nl4d/grain/consts.rs) ------------------

GRAIN_CELL, GRAIN_LAGS, GRAIN_PARTIAL, GRAIN_BUCKETS, GRAIN_BINS = 8, 46, 48, 64, 16
GRAIN_HIST = GRAIN_BINS * GRAIN_BUCKETS
GRAIN_STD_MIN, GRAIN_STD_MAX = np.float32(0.05 / 255.0), np.float32(32.0 / 255.0)


def grain_lags() -> list[tuple[int, int]]:
    """LAGS (nl4d/grain/consts.rs:27, lag_table): (dy, dx) for dy 0..3, dx -6..6, the half plane."""
    return [(dy, dx) for dy in range(4) for dx in range(-6, 7) if dy > 0 or dx >= 0]


def grain_bucket_edges() -> np.ndarray:
    """bucket_edges (nl4d/grain/fit.rs:28-38): 65 edges log-spaced from STD_MIN to STD_MAX, computed
    in f64 from the f32 constants and stored as f32."""
    lo, hi = float(np.float32(0.05) / np.float32(255.0)), float(np.float32(32.0) / np.float32(255.0))
    return np.array([lo * (hi / lo) ** (i / GRAIN_BUCKETS) for i in range(GRAIN_BUCKETS + 1)], np.float32)


def grain_measure(ring, out_t, width, height, channels, channel, slot_t, slot_next, saved_mv, saved_conf,
                  blocks_x, blocks_y, step, entry=0, has_source=True, luma_means=None, has_luma_means=False,
                  means_blocks_x=0, luma_band=True, edges=None, hist=None):
    """grain_measure (nl4d/kernels/grain.rs:290-513) over every 8x8 cell (avdo_grain_measure): returns
    (hist, partials), hist the 2 x 16 x 64 i32 counts (source grain, then one-frame grain) added
    into ``hist`` when given, partials 48 floats per cell."""
    ring = np.ascontiguousarray(ring, np.float32).reshape(-1)
    out_t = np.ascontiguousarray(out_t, np.float32).reshape(-1)
    mv = np.ascontiguousarray(saved_mv, np.int32).reshape(-1)
    conf = np.ascontiguousarray(saved_conf, np.float32).reshape(-1)
    means = np.ascontiguousarray(luma_means if luma_means is not None else np.zeros(3, np.float32), np.float32)
    edges = np.ascontiguousarray(grain_bucket_edges() if edges is None else edges, np.float32)
    hist = np.zeros(2 * GRAIN_HIST, np.int32) if hist is None else np.ascontiguousarray(hist, np.int32)
    partials = np.full((width // GRAIN_CELL) * (height // GRAIN_CELL) * GRAIN_PARTIAL, np.nan, np.float32)
    u = ctypes.c_uint32
    lib().avdo_grain_measure(_p(ring, F32P), _p(out_t, F32P), u(width), u(height), u(channels), u(channel),
                             u(slot_t), u(slot_next), _p(mv, I32P), _p(conf, F32P), u(blocks_x), u(blocks_y),
                             u(step), u(entry), ctypes.c_int(int(bool(has_source))), _p(means, F32P),
                             ctypes.c_int(int(bool(has_luma_means))), u(means_blocks_x),
                             ctypes.c_int(int(bool(luma_band))), _p(edges, F32P), _p(hist, I32P), _p(partials, F32P))
    return hist, partials


def grain_reduce(partials, chunk=None, record_offset=0):
    """grain_reduce_partials (nl4d/kernels/grain.rs:515-560): every cell's partial added into its
    strength group's 47-float record of ``chunk`` (16 records from ``record_offset``, earlier sums
    kept), in the reference's order (avdo_grain_reduce)."""
    partials = np.ascontiguousarray(partials, np.float32).reshape(-1)
    cells = partials.size // GRAIN_PARTIAL
    if chunk is None:
        chunk = np.zeros(record_offset + 16 * 47, np.float32)
    chunk = np.ascontiguousarray(chunk, np.float32).copy()
    lib().avdo_grain_reduce(_p(partials, F32P), ctypes.c_uint32(cells), _p(chunk, F32P), ctypes.c_uint32(record_offset))
    return chunk


def grain_fit(records):
    """The AR(3) fit of each 47-value f64 record (undo_mean_removal, yule_walker and quantise_ar,
    nl4d/grain/fit.rs:120-300; avdo_grain_fit): returns (corrected records, weights [n][24] in
    AR_OFFSETS order, AV1 integers [n][24], shifts [n], solved [n])."""
    rec = np.ascontiguousarray(records, np.float64).reshape(-1, 47)
    n = rec.shape[0]
    corrected = np.zeros((n, 47), np.float64)
    weights = np.zeros((n, 24), np.float64)
    quantised = np.zeros((n, 24), np.int32)
    shift = np.zeros(n, np.uint32)
    solved = np.zeros(n, np.int32)
    lib().avdo_grain_fit(_p(rec, F64P), ctypes.c_uint32(n), _p(corrected, F64P), _p(weights, F64P),
                         _p(quantised, I32P), _p(shift, U32P), _p(solved, I32P))
    return corrected, weights, quantised, shift, solved.astype(bool)


def ar_offsets() -> list[tuple[int, int]]:
    """AR_OFFSETS (nl4d/grain/consts.rs:30, ar_offset_table): AV1's 24 causal lag-3 offsets (dy, dx)."""
    return [(dy, dx) for dy in range(-3, 1) for dx in range(-3, 4) if dy < 0 or dx < 0]


def grain_quantise(weights):
    """quantise_ar (nl4d/grain/fit.rs:272-305) of 24 weights: (AV1 integers, shift)."""
    w = np.ascontiguousarray(weights, np.float64).reshape(24)
    q = np.zeros(24, np.int32)
    shift = np.zeros(1, np.uint32)
    lib().avdo_grain_quantise(_p(w, F64P), _p(q, I32P), _p(shift, U32P))
    return q, int(shift[0])


def grain_yule_walker(record):
    """yule_walker and solve (nl4d/grain/fit.rs:160-216) of a 47-value f64 record, without the
    mean-removal correction: the 24 weights, or None."""
    rec = np.ascontiguousarray(record, np.float64).reshape(47)
    sol = np.zeros(24, np.float64)
    ok = lib().avdo_grain_yule_walker(_p(rec, F64P), _p(sol, F64P))
    return sol if ok else None


# ---- grain persistence (nl4d/kernels/grain_persistence.rs, nl4d/grain/consts.rs:46-85) -----------

GRAIN_PATTERNS = [0x99666699_99666699, 0x96696996_69969669, 0xA55AA55A_5AA55AA5, 0xAA55AA55_AA55AA55,
                  0xA55A5AA5_A55A5AA5, 0x99669966_66996699, 0xAA5555AA_55AAAA55, 0x96699669_96699669]
HISTORY_FRAMES, HISTORY_ENTRY = 9, 9


def grain_pattern_words() -> np.ndarray:
    """PATTERNS (nl4d/grain/consts.rs:76-85) as two u32 words each, the low word first."""
    return np.array([w for p in GRAIN_PATTERNS for w in (p & 0xFFFFFFFF, p >> 32)], np.uint32)


def grain_save_vectors(mv_field, confidence, saved_mv, saved_conf, mv_offset, conf_offset, entry, blocks):
    """grain_save_vectors (nl4d/kernels/grain.rs:50-75): copies into saved_mv / saved_conf in place."""
    lib().avdo_grain_save_vectors(_p(np.ascontiguousarray(mv_field, np.int32), I32P),
                                  _p(np.ascontiguousarray(confidence, np.float32), F32P), _p(saved_mv, I32P),
                                  _p(saved_conf, F32P), ctypes.c_uint32(mv_offset), ctypes.c_uint32(conf_offset),
                                  ctypes.c_uint32(entry), ctypes.c_uint32(blocks))


def grain_project(ring, out, width, height, channels, channel, frame_slot, history, history_slot, luma_band=True):
    """grain_project (nl4d/kernels/grain_persistence.rs:54-153): writes history slot ``history_slot``
    (9 floats per cell: 8 projections, the flat flag) into ``history`` in place."""
    ring = np.ascontiguousarray(ring, np.float32).reshape(-1)
    out = np.ascontiguousarray(out, np.float32).reshape(-1)
    u = ctypes.c_uint32
    lib().avdo_grain_project(_p(ring, F32P), _p(out, F32P), _p(grain_pattern_words(), U32P), _p(history, F32P),
                             u(width), u(height), u(channels), u(channel), u(frame_slot), u(history_slot),
                             ctypes.c_int(int(bool(luma_band))))


def grain_persist(history, now_slot, cells, lags):
    """grain_persist (nl4d/kernels/grain_persistence.rs:155-230): [lags][blocks][3] partials."""
    blocks = -(-cells // 1024)
    partials = np.full(lags * blocks * 3, np.nan, np.float32)
    lib().avdo_grain_persist(_p(np.ascontiguousarray(history, np.float32), F32P), _p(partials, F32P),
                             ctypes.c_uint32(now_slot), ctypes.c_uint32(cells), ctypes.c_uint32(lags))
    return partials


def grain_persist_reduce(partials, record, record_offset, blocks, lags):
    """grain_persist_reduce (nl4d/kernels/grain_persistence.rs:232-279): adds into ``record``."""
    record = np.ascontiguousarray(record, np.float32).copy()
    lib().avdo_grain_persist_reduce(_p(np.ascontiguousarray(partials, np.float32), F32P), _p(record, F32P),
                                    ctypes.c_uint32(record_offset), ctypes.c_uint32(blocks), ctypes.c_uint32(lags))
    return record
