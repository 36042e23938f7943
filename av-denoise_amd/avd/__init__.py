"""MI355X-native av-denoise denoising core: the Python face of ``libavd.so``.

The API mirrors av-denoise-core's engine boundary: ``build_engine`` (tune/setup.rs:79) returns an
``Engine`` with ``push`` / ``push_context`` / ``emit_into`` / ``finish`` / ``reset``
(engine/mod.rs:168-211); options mirror ``NlmeansAlgorithm`` and ``Nl4dOptions``; errors mirror
``Error`` (error.rs:3-20).  Planes are device buffers of whole u32 words; torch tensors on
``cuda`` are accepted and only their memory is used.
"""
from __future__ import annotations

import ctypes

from . import synth  # noqa: F401  (seeded clips)
from . import y4m  # noqa: F401  (YUV4MPEG2 streams)
from ._lib import (CollabPass, ContextAfterPush, DevicePlane, EngineSpec, Error, Geometry, Gpu, InvalidGeometry,
                   InvalidOptions, MotionPass, MotionSearch, NeedsReset, Nl4dOptions, NlmeansOptions, NlmWindow,
                   NothingToEmit, NullArgument, OutputsPending, PlaneMismatch, Poisoned, QueueFull, SampleFormat,
                   WindowSpan, check, load)

__all__ = [
    "Engine", "build_engine", "EngineSpec", "Geometry", "SampleFormat", "MotionSearch", "NlmeansOptions",
    "Nl4dOptions", "nlmeans_options", "nl4d_options", "nlmeans_options_for_preset", "nl4d_options_for_preset",
    "plane_bytes", "chroma_dims", "Error", "InvalidOptions", "InvalidGeometry", "PlaneMismatch", "OutputsPending",
    "NothingToEmit", "ContextAfterPush", "NeedsReset", "Gpu", "NullArgument", "QueueFull", "Poisoned", "load",
    "PRESETS", "HostDenoiser", "engine_spec", "y4m", "set_tuning", "tuning_entries",
]

PRESETS = {"veryfast": 0, "fast": 1, "base": 2, "slow": 3, "veryslow": 4}
CHROMA = {"gray": 0, "420": 1, "422": 2, "444": 3}


def _ptr(buf) -> int:
    if buf is None:
        return None
    if isinstance(buf, int):
        return buf
    return buf.data_ptr()


def _nbytes(buf) -> int:
    return buf.numel() * buf.element_size()


def motion_search(**overrides) -> MotionSearch:
    m = MotionSearch()
    load().avd_motion_search_default(ctypes.byref(m))
    for k, v in overrides.items():
        setattr(m, k, v)
    return m


def nlmeans_options(**overrides) -> NlmeansOptions:
    """NlmeansOptions defaults with overrides; setting an Option field sets its ``has_`` flag."""
    o = NlmeansOptions()
    load().avd_nlmeans_options_default(ctypes.byref(o))
    _apply(o, overrides)
    return o


def nlmeans_options_for_preset(preset: str, **overrides) -> NlmeansOptions:
    o = NlmeansOptions()
    check(load().avd_nlmeans_options_for_preset(PRESETS[preset], ctypes.byref(o)))
    _apply(o, overrides)
    return o


def nl4d_options(**overrides) -> Nl4dOptions:
    o = Nl4dOptions()
    load().avd_nl4d_options_default(ctypes.byref(o))
    _apply(o, overrides)
    return o


def nl4d_options_for_preset(preset: str, **overrides) -> Nl4dOptions:
    o = Nl4dOptions()
    check(load().avd_nl4d_options_for_preset(PRESETS[preset], ctypes.byref(o)))
    _apply(o, overrides)
    return o


def _apply(o, overrides):
    for k, v in overrides.items():
        if k == "motion" and isinstance(v, dict):
            for mk, mv in v.items():
                setattr(o.motion, mk, mv)
            continue
        setattr(o, k, v)
        if hasattr(o, "has_" + k):
            setattr(o, "has_" + k, 1)


def plane_bytes(fmt: SampleFormat, width: int, height: int) -> int:
    return int(load().avd_plane_bytes(fmt, width, height))


def chroma_dims(chroma: int, width: int, height: int):
    """ChromaLayout::chroma_dims (engine/frame.rs:30-40)."""
    if chroma == 0:
        return None
    if chroma == 1:
        return (width + 1) // 2, (height + 1) // 2
    if chroma == 2:
        return (width + 1) // 2, height
    return width, height


def plane_dims(geometry: Geometry):
    dims = [(geometry.width, geometry.height)]
    cd = chroma_dims(geometry.chroma, geometry.width, geometry.height)
    if cd:
        dims += [cd, cd]
    return dims


class Engine:
    """A stateful denoiser over device planes (the reference's ``Engine`` trait)."""

    def __init__(self, spec: EngineSpec):
        self._lib = load()
        self._h = ctypes.c_void_p()
        self.spec = spec
        check(self._lib.avd_engine_create(ctypes.byref(spec), ctypes.byref(self._h)))
        self.dims = plane_dims(spec.geometry)

    def _planes(self, planes):
        if len(planes) != len(self.dims):
            raise PlaneMismatch(f"expected {len(self.dims)} planes, got {len(planes)}")
        arr = (DevicePlane * len(planes))()
        for i, (buf, (w, h)) in enumerate(zip(planes, self.dims)):
            arr[i] = DevicePlane(_ptr(buf), _nbytes(buf), w, h)
        return arr, len(planes)

    def push(self, planes) -> int:
        arr, n = self._planes(planes)
        ready = ctypes.c_uint32()
        check(self._lib.avd_engine_push(self._h, arr, n, ctypes.byref(ready)))
        return ready.value

    def push_context(self, planes) -> None:
        arr, n = self._planes(planes)
        check(self._lib.avd_engine_push_context(self._h, arr, n))

    def emit_into(self, planes) -> None:
        arr, n = self._planes(planes)
        check(self._lib.avd_engine_emit_into(self._h, arr, n))

    def finish(self) -> int:
        tail = ctypes.c_uint32()
        check(self._lib.avd_engine_finish(self._h, ctypes.byref(tail)))
        return tail.value

    def reset(self) -> None:
        check(self._lib.avd_engine_reset(self._h))

    def synchronize(self) -> None:
        check(self._lib.avd_engine_synchronize(self._h))

    def window_span(self) -> WindowSpan:
        w = WindowSpan()
        check(self._lib.avd_engine_window_span(self._h, ctypes.byref(w)))
        return w

    def max_held_frames(self) -> int:
        v = ctypes.c_uint64()
        check(self._lib.avd_engine_max_held_frames(self._h, ctypes.byref(v)))
        return v.value

    def sample_device_time(self, every: int) -> None:
        """Samples device time of every frame (1) or of none (0) (Engine::sample_device_time)."""
        check(self._lib.avd_engine_sample_device_time(self._h, every))

    NLM_FORMS = {"plain": 0, "batched": 1, "default": 2}

    def set_nlm_form(self, form: str) -> None:
        """NLMeans engines: one launch per window ("plain"), every window of a frame and the finish
        in one launch ("batched"), or the default (spatial passes batched, temporal ones plain);
        bit-identical outputs."""
        check(self._lib.avd_engine_set_nlm_form(self._h, self.NLM_FORMS[form]))

    def timing(self) -> dict:
        """Per-stage device milliseconds, launches and work since sampling began (Engine::timing)."""
        from ._lib import Timing

        t = Timing()
        check(self._lib.avd_engine_timing(self._h, ctypes.byref(t)))
        return t.as_dict()

    def fail_for_test(self) -> None:
        check(self._lib.avd_engine_fail_for_test(self._h))

    def set_tile_cap_for_test(self, cap: int) -> None:
        check(self._lib.avd_engine_set_tile_cap_for_test(self._h, cap))

    def tile_overflows_for_test(self) -> int:
        out = ctypes.c_uint64()
        check(self._lib.avd_engine_tile_overflows_for_test(self._h, ctypes.byref(out)))
        return int(out.value)

    def close(self) -> None:
        if self._h:
            self._lib.avd_engine_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def engine_spec(kind: str, geometry: Geometry, luma=None, chroma=None, stream=None) -> EngineSpec:
    """EngineSpec (engine/spec.rs:146-157): ``kind`` is "nlmeans" or "nl4d"; a plane group left
    ``None`` passes through unchanged."""
    spec = EngineSpec()
    spec.kind = 0 if kind == "nlmeans" else 1
    spec.geometry = geometry
    spec.clean_luma = 1 if luma is not None else 0
    spec.clean_chroma = 1 if chroma is not None else 0
    if kind == "nlmeans":
        spec.nlm_luma = luma if luma is not None else nlmeans_options()
        spec.nlm_chroma = chroma if chroma is not None else nlmeans_options()
    else:
        spec.nl4d_luma = luma if luma is not None else nl4d_options()
        spec.nl4d_chroma = chroma if chroma is not None else nl4d_options()
    spec.stream = stream
    return spec


def build_engine(kind: str, geometry: Geometry, luma=None, chroma=None, stream=None) -> Engine:
    """build_engine (tune/setup.rs:79-92): ``kind`` is "nlmeans" or "nl4d"; a plane group left
    ``None`` passes through unchanged."""
    return Engine(engine_spec(kind, geometry, luma, chroma, stream))


def set_tuning(enabled: bool) -> None:
    """Turns the launch-site tuner on (the default) or off (every site runs its baseline)."""
    load().avd_tuning_set_enabled(1 if enabled else 0)


def tuning_entries() -> list:
    """The tuner's choices so far: site, shape key, chosen candidate and each candidate's time."""
    from ._lib import TuneEntry

    n = ctypes.c_uint32()
    check(load().avd_tuning_entries(None, 0, ctypes.byref(n)))
    arr = (TuneEntry * max(n.value, 1))()
    check(load().avd_tuning_entries(arr, n.value, ctypes.byref(n)))
    return [{"site": e.site.decode(), "key": e.key.decode(), "device": e.device, "choice": e.choice,
             "ms": [round(e.ms[i], 4) for i in range(e.candidates)]} for e in arr[:n.value]]


class HostDenoiser:
    """Frames in and out as host bytes (the reference's HostDenoiser, av-denoise/src/host/mod.rs):
    ``push`` takes one bytes-like buffer per plane (Y, then U and V), ``recv`` returns the oldest
    finished frame (or None), ``flush`` drains the rest and the temporal tail.  At most
    ``MAX_PENDING`` frames wait to be collected; a push past that raises QueueFull."""

    MAX_PENDING = 2

    def __init__(self, kind: str, geometry: Geometry, luma=None, chroma=None):
        self._lib = load()
        self._h = ctypes.c_void_p()
        check(self._lib.avd_host_denoiser_create(ctypes.byref(engine_spec(kind, geometry, luma, chroma)),
                                                 ctypes.byref(self._h)))
        n = ctypes.c_uint32()
        ins, outs = (ctypes.c_uint64 * 3)(), (ctypes.c_uint64 * 3)()
        check(self._lib.avd_host_denoiser_plane_lengths(self._h, ins, outs, 3, ctypes.byref(n)))
        self.input_lengths = [int(ins[i]) for i in range(n.value)]
        self.output_lengths = [int(outs[i]) for i in range(n.value)]

    @staticmethod
    def _pointers(buffers):
        import numpy as np

        arrays = [np.ascontiguousarray(np.frombuffer(b, np.uint8) if isinstance(b, (bytes, bytearray)) else b)
                  for b in buffers]
        ptrs = (ctypes.c_void_p * len(arrays))(*[a.ctypes.data for a in arrays])
        lens = (ctypes.c_uint64 * len(arrays))(*[a.nbytes for a in arrays])
        return arrays, ptrs, lens

    def push(self, planes) -> None:
        keep, ptrs, lens = self._pointers(planes)
        check(self._lib.avd_host_denoiser_push(self._h, ptrs, lens, len(keep)))

    def push_priming(self, planes) -> None:
        keep, ptrs, lens = self._pointers(planes)
        check(self._lib.avd_host_denoiser_push_priming(self._h, ptrs, lens, len(keep)))

    def recv(self):
        """The oldest frame in flight as one bytes object per plane, or None when none is."""
        bufs = [ctypes.create_string_buffer(n) for n in self.output_lengths]
        ptrs = (ctypes.c_void_p * len(bufs))(*[ctypes.addressof(b) for b in bufs])
        caps = (ctypes.c_uint64 * len(bufs))(*self.output_lengths)
        got = ctypes.c_uint32()
        check(self._lib.avd_host_denoiser_recv(self._h, ptrs, caps, len(bufs), ctypes.byref(got)))
        return [b.raw for b in bufs] if got.value else None

    def flush(self):
        """Every remaining frame, in order; the denoiser then takes a fresh stream."""
        from ._lib import FRAME_SINK

        frames = []

        def sink(_user, planes, lengths, count):
            frames.append([ctypes.string_at(planes[i], lengths[i]) for i in range(count)])

        cb = FRAME_SINK(sink)
        check(self._lib.avd_host_denoiser_flush(self._h, cb, None))
        return frames

    def reset_stream(self) -> None:
        check(self._lib.avd_host_denoiser_reset_stream(self._h))

    def pending(self) -> int:
        n = ctypes.c_uint32()
        check(self._lib.avd_host_denoiser_pending(self._h, ctypes.byref(n)))
        return n.value

    def close(self) -> None:
        if self._h:
            self._lib.avd_host_denoiser_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
