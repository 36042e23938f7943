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
from ._lib import (CollabPass, ContextAfterPush, DevicePlane, EngineSpec, Error, Geometry, Gpu, InvalidGeometry,
                   InvalidOptions, MotionPass, MotionSearch, NeedsReset, Nl4dOptions, NlmeansOptions, NlmWindow,
                   NothingToEmit, NullArgument, OutputsPending, PlaneMismatch, SampleFormat, WindowSpan, check, load)

__all__ = [
    "Engine", "build_engine", "EngineSpec", "Geometry", "SampleFormat", "MotionSearch", "NlmeansOptions",
    "Nl4dOptions", "nlmeans_options", "nl4d_options", "nlmeans_options_for_preset", "nl4d_options_for_preset",
    "plane_bytes", "chroma_dims", "Error", "InvalidOptions", "InvalidGeometry", "PlaneMismatch", "OutputsPending",
    "NothingToEmit", "ContextAfterPush", "NeedsReset", "Gpu", "NullArgument", "load", "PRESETS",
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


def build_engine(kind: str, geometry: Geometry, luma=None, chroma=None, stream=None) -> Engine:
    """build_engine (tune/setup.rs:79-92): ``kind`` is "nlmeans" or "nl4d"; a plane group left
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
    return Engine(spec)
