"""ctypes bindings of ``include/avd/avd.h``: the structs, the error classes and the loader.

The loader fails loudly when ``libavd.so`` is missing: there is no CPU fallback.
"""
from __future__ import annotations

import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libavd.so")

u32 = ctypes.c_uint32
f32 = ctypes.c_float


class SampleFormat(ctypes.Structure):
    """SampleFormat (engine/plane.rs:8-52): kind 0 U8, 1 U16 with ``depth`` bits."""
    _fields_ = [("kind", u32), ("depth", u32)]

    @classmethod
    def u8(cls):
        return cls(0, 8)

    @classmethod
    def u16(cls, depth: int):
        return cls(1, depth)

    @classmethod
    def for_depth(cls, depth: int):
        return cls.u8() if depth == 8 else cls.u16(depth)


class Geometry(ctypes.Structure):
    """Geometry (engine/frame.rs:73-97).  chroma: 0 Gray, 1 4:2:0, 2 4:2:2, 3 4:4:4."""
    _fields_ = [("width", u32), ("height", u32), ("chroma", u32), ("input", SampleFormat), ("output", SampleFormat),
                ("chroma_siting", u32)]


class DevicePlane(ctypes.Structure):
    _fields_ = [("data", ctypes.c_void_p), ("bytes", ctypes.c_uint64), ("width", u32), ("height", u32)]


class MotionSearch(ctypes.Structure):
    """MotionSearch (motion/search.rs:5-40).  estimation: 0 Auto, 1 Direct, 2 Chained."""
    _fields_ = [("blksize", u32), ("overlap", u32), ("search_radius", u32), ("max_search_escalation", u32),
                ("pyramid_levels", u32), ("estimation", u32), ("refine_radius", u32)]


class NlmeansOptions(ctypes.Structure):
    """NlmeansAlgorithm with NlmeansOptions and HqParams (nlmeans/options.rs:6-106)."""
    _fields_ = [("hq", u32), ("temporal", u32), ("temporal_radius", u32),
                ("has_search_radius", u32), ("search_radius", u32),
                ("has_patch_radius", u32), ("patch_radius", u32),
                ("has_strength", u32), ("strength", f32),
                ("has_self_weight", u32), ("self_weight", f32),
                ("motion_compensation", u32), ("motion", MotionSearch),
                ("auto_strength", u32), ("noise_floor", u32), ("has_sigma", u32), ("sigma", f32),
                ("temporal_confidence", u32), ("thsad_scale", f32), ("sigma_scale", f32)]


class Nl4dOptions(ctypes.Structure):
    """Nl4dOptions (nl4d/options.rs:11-94) plus the grid stride and group size."""
    _fields_ = [("motion", MotionSearch), ("has_sigma", u32), ("sigma", f32), ("sigma_scale", f32),
                ("thsad_scale", f32), ("refine", u32), ("spatial_radius", u32), ("has_lambda_ht", u32),
                ("lambda_ht", f32), ("lambda_ht_scale", f32), ("c_min", f32), ("kaiser_beta", f32),
                ("field_lambda", f32), ("noise_map", u32), ("pooled_threshold", u32), ("grain_data", u32),
                ("temporal_radius", u32), ("step", u32), ("group_size", u32), ("transform_form", u32)]


class EngineSpec(ctypes.Structure):
    """EngineSpec (engine/spec.rs:146-157).  kind 0 NLMeans, 1 NL4D."""
    _fields_ = [("kind", u32), ("geometry", Geometry), ("clean_luma", u32), ("clean_chroma", u32),
                ("nlm_luma", NlmeansOptions), ("nlm_chroma", NlmeansOptions),
                ("nl4d_luma", Nl4dOptions), ("nl4d_chroma", Nl4dOptions), ("stream", ctypes.c_void_p)]


class WindowSpan(ctypes.Structure):
    _fields_ = [("behind", ctypes.c_uint64), ("ahead", ctypes.c_uint64), ("edges", u32)]


STAGES = ["ingest", "motion", "collab_search", "collab_filter", "normalise", "egress", "nlm_pair", "nlm_single",
          "nlm_finish", "warp", "collab_centre", "nlm_window"]


class Timing(ctypes.Structure):
    """avd_timing: device milliseconds, launches and algorithmic work per stage."""
    _fields_ = [("milliseconds", ctypes.c_double * len(STAGES)), ("launches", ctypes.c_uint64 * len(STAGES)),
                ("work", ctypes.c_uint64 * len(STAGES))]

    def as_dict(self):
        return {name: {"ms": self.milliseconds[i], "launches": int(self.launches[i]), "work": int(self.work[i])}
                for i, name in enumerate(STAGES)}


class TuneEntry(ctypes.Structure):
    """avd_tune_entry: one launch site's tuned choice."""
    _fields_ = [("site", ctypes.c_char * 32), ("key", ctypes.c_char * 96), ("device", u32), ("choice", u32),
                ("candidates", u32), ("ms", ctypes.c_double * 8)]


class CollabPass(ctypes.Structure):
    _fields_ = [("width", u32), ("height", u32), ("channels", u32), ("step", u32), ("spatial_radius", u32),
                ("temporal_radius", u32), ("refine", u32), ("group_size", u32), ("blk_step", u32),
                ("blksize", u32), ("blocks_x", u32), ("blocks_y", u32), ("centre_slot", u32), ("slots", u32),
                ("c_min", f32), ("lam", f32 * 2), ("sigma", f32 * 2), ("weight_scale", f32),
                ("accum_scale", f32), ("kaiser", f32 * 8), ("profile", f32 * 8),
                ("code_bits", u32), ("pool_ratio", f32)]


class MotionPass(ctypes.Structure):
    _fields_ = [("width", u32), ("height", u32), ("channels", u32), ("blksize", u32), ("overlap", u32),
                ("search_radius", u32), ("pyramid_levels", u32), ("centre_slot", u32), ("neighbours", u32),
                ("max_value", f32), ("sad_floor", f32), ("thsad", f32)]


class FieldLayout(ctypes.Structure):
    """FieldLayout (motion/chroma_map.rs:18-53): a motion field's block grid and i32 field stride."""
    _fields_ = [("blocks_x", u32), ("blocks_y", u32), ("step", u32), ("blksize", u32), ("stride", u32)]


class NlmWindow(ctypes.Structure):
    _fields_ = [("width", u32), ("height", u32), ("channels", u32), ("patch_radius", u32), ("search_radius", u32),
                ("h2_inv_norm", f32), ("noise_offset", f32), ("step", u32), ("blocks_x", u32), ("blocks_y", u32)]


# ---- errors (error.rs:3-20) --------------------------------------------------------------------

class Error(RuntimeError):
    status = -1


class InvalidOptions(Error):
    status = 1


class InvalidGeometry(Error):
    status = 2


class PlaneMismatch(Error):
    status = 3


class OutputsPending(Error):
    status = 4


class NothingToEmit(Error):
    status = 5


class ContextAfterPush(Error):
    status = 6


class NeedsReset(Error):
    status = 7


class Gpu(Error):
    status = 8


class NullArgument(Error):
    status = 9


class QueueFull(Error):
    """DenoiserError::QueueFull (av-denoise/src/host/mod.rs:40-44)."""
    status = 10


class Poisoned(Error):
    """DenoiserError::Poisoned (av-denoise/src/host/mod.rs:45-49)."""
    status = 11


_BY_STATUS = {c.status: c for c in (InvalidOptions, InvalidGeometry, PlaneMismatch, OutputsPending, NothingToEmit,
                                     ContextAfterPush, NeedsReset, Gpu, NullArgument, QueueFull, Poisoned)}

# Every exported symbol and its ctypes signature (restype, argtypes).
P = ctypes.c_void_p
# avd_frame_sink: (user, planes, lengths, count).
FRAME_SINK = ctypes.CFUNCTYPE(None, ctypes.c_void_p, ctypes.POINTER(ctypes.c_void_p), ctypes.POINTER(ctypes.c_uint64),
                              ctypes.c_uint32)

SIGNATURES = {
    "avd_abi_version": (u32, []),
    "avd_last_error_message": (ctypes.c_char_p, []),
    "avd_motion_search_default": (None, [ctypes.POINTER(MotionSearch)]),
    "avd_nlmeans_options_default": (None, [ctypes.POINTER(NlmeansOptions)]),
    "avd_nlmeans_options_for_preset": (ctypes.c_int, [u32, ctypes.POINTER(NlmeansOptions)]),
    "avd_nl4d_options_default": (None, [ctypes.POINTER(Nl4dOptions)]),
    "avd_nl4d_options_for_preset": (ctypes.c_int, [u32, ctypes.POINTER(Nl4dOptions)]),
    "avd_plane_bytes": (ctypes.c_uint64, [SampleFormat, u32, u32]),
    "avd_engine_create": (ctypes.c_int, [ctypes.POINTER(EngineSpec), ctypes.POINTER(P)]),
    "avd_engine_destroy": (None, [P]),
    "avd_engine_push": (ctypes.c_int, [P, ctypes.POINTER(DevicePlane), u32, ctypes.POINTER(u32)]),
    "avd_engine_push_context": (ctypes.c_int, [P, ctypes.POINTER(DevicePlane), u32]),
    "avd_engine_emit_into": (ctypes.c_int, [P, ctypes.POINTER(DevicePlane), u32]),
    "avd_engine_finish": (ctypes.c_int, [P, ctypes.POINTER(u32)]),
    "avd_engine_reset": (ctypes.c_int, [P]),
    "avd_engine_window_span": (ctypes.c_int, [P, ctypes.POINTER(WindowSpan)]),
    "avd_engine_max_held_frames": (ctypes.c_int, [P, ctypes.POINTER(ctypes.c_uint64)]),
    "avd_engine_synchronize": (ctypes.c_int, [P]),
    "avd_engine_fail_for_test": (ctypes.c_int, [P]),
    "avd_engine_set_tile_cap_for_test": (ctypes.c_int, [P, u32]),
    "avd_engine_tile_overflows_for_test": (ctypes.c_int, [P, ctypes.POINTER(ctypes.c_uint64)]),
    "avd_engine_sample_device_time": (ctypes.c_int, [P, u32]),
    "avd_engine_set_nlm_form": (ctypes.c_int, [P, u32]),
    "avd_engine_timing": (ctypes.c_int, [P, ctypes.POINTER(Timing)]),
    "avd_site_ingest": (ctypes.c_int, [P, P, P, P, u32, u32, SampleFormat, P]),
    "avd_site_egress": (ctypes.c_int, [P, P, P, u32, u32, SampleFormat, P]),
    "avd_site_collab_search": (ctypes.c_int, [ctypes.POINTER(CollabPass), P, P, P, P, P, P, P]),
    "avd_site_collab_filter": (ctypes.c_int, [ctypes.POINTER(CollabPass), P, P, P, P, P, P, P]),
    "avd_site_collab_normalise": (ctypes.c_int, [P, P, P, u32, u32, P]),
    "avd_site_motion": (ctypes.c_int, [ctypes.POINTER(MotionPass), P, P, P, P, P]),
    "avd_site_motion_refine": (ctypes.c_int, [ctypes.POINTER(MotionPass), P, P, P, u32, P, P, P]),
    "avd_site_map_luma_fields": (ctypes.c_int, [ctypes.POINTER(FieldLayout), P, ctypes.POINTER(FieldLayout), P, u32,
                                                u32, u32, P]),
    "avd_site_nlm_single": (ctypes.c_int, [ctypes.POINTER(NlmWindow), P, P, P, P, P, P]),
    "avd_site_nlm_pair": (ctypes.c_int, [ctypes.POINTER(NlmWindow), P, P, P, P, P, P, P, P, P]),
    "avd_site_nlm_finish": (ctypes.c_int, [P, P, P, P, f32, u32, u32, P, P]),
    "avd_site_noise_spatial": (ctypes.c_int, [P, u32, u32, u32, u32, P, P, u32, P]),
    "avd_site_noise_temporal_stats": (ctypes.c_int, [P, u32, u32, u32, u32, u32, P, u32, u32, u32, u32, u32, P, P]),
    "avd_site_noise_temporal_zero": (ctypes.c_int, [P, u32, P]),
    "avd_noise_temporal_stats_len": (u32, [u32, u32, u32]),
    "avd_site_grain_measure": (ctypes.c_int, [P, P, u32, u32, u32, u32, u32, u32, P, P, u32, u32, u32, u32, u32,
                                              P, u32, u32, u32, P, P, P, P]),
    "avd_grain_partials_len": (u32, [u32, u32]),
    "avd_site_grain_reduce": (ctypes.c_int, [P, u32, P, u32, P]),
    "avd_site_grain_fit": (ctypes.c_int, [P, u32, P, P, P, P, P, P]),
    "avd_grain_bucket_edges": (None, [P]),
    "avd_grain_hist_median": (ctypes.c_double, [P, P, P]),
    "avd_grain_finest_scaling_shift": (u32, [ctypes.c_double]),
    "avd_grain_scaling_points": (ctypes.c_int, [P, P, u32, ctypes.c_double, P, P]),
    "avd_grain_quantise_ar": (None, [P, P, P]),
    "avd_grain_pooled_record": (ctypes.c_int32, [P, P, P, P]),
    "avd_grain_take_out_fine_noise": (None, [P, ctypes.c_double, ctypes.c_double]),
    "avd_grain_floor_fraction": (ctypes.c_double, [P]),
    "avd_site_grain_save_vectors": (ctypes.c_int, [P, P, P, P, u32, u32, u32, u32, P]),
    "avd_site_grain_project": (ctypes.c_int, [P, P, P, P, u32, u32, u32, u32, u32, u32, u32, P]),
    "avd_grain_persist_partials_len": (u32, [u32, u32]),
    "avd_site_grain_persist": (ctypes.c_int, [P, P, P, u32, u32, u32, u32, P]),
    "avd_host_denoiser_create": (ctypes.c_int, [ctypes.POINTER(EngineSpec), ctypes.POINTER(P)]),
    "avd_host_denoiser_destroy": (None, [P]),
    "avd_host_denoiser_plane_lengths": (ctypes.c_int, [P, ctypes.POINTER(ctypes.c_uint64),
                                                       ctypes.POINTER(ctypes.c_uint64), u32,
                                                       ctypes.POINTER(ctypes.c_uint32)]),
    "avd_host_denoiser_push": (ctypes.c_int, [P, ctypes.POINTER(P), ctypes.POINTER(ctypes.c_uint64), u32]),
    "avd_host_denoiser_push_priming": (ctypes.c_int, [P, ctypes.POINTER(P), ctypes.POINTER(ctypes.c_uint64), u32]),
    "avd_host_denoiser_recv": (ctypes.c_int, [P, ctypes.POINTER(P), ctypes.POINTER(ctypes.c_uint64), u32,
                                              ctypes.POINTER(ctypes.c_uint32)]),
    "avd_host_denoiser_flush": (ctypes.c_int, [P, FRAME_SINK, P]),
    "avd_host_denoiser_reset_stream": (ctypes.c_int, [P]),
    "avd_host_denoiser_pending": (ctypes.c_int, [P, ctypes.POINTER(ctypes.c_uint32)]),
    "avd_host_denoiser_window_span": (ctypes.c_int, [P, ctypes.POINTER(WindowSpan)]),
    "avd_tuning_set_enabled": (None, [u32]),
    "avd_tuning_entries": (ctypes.c_int, [ctypes.POINTER(TuneEntry), u32, ctypes.POINTER(ctypes.c_uint32)]),
    "avd_site_noise_spatial_noalloc": (ctypes.c_int, [P, u32, u32, u32, u32, P, u32, P]),
    "avd_noise_partials_len": (u32, [u32, u32]),
    "avd_noise_sigma_from_abs_sum": (f32, [f32, u32, u32]),
}

_LIB = None


def load() -> ctypes.CDLL:
    """Loads ``libavd.so``; raises when it has not been built."""
    global _LIB
    if _LIB is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is missing: build it with `python av-denoise_amd/avd/build.py` "
                               "(or __graft_entry__.build()); there is no CPU fallback")
        lib = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(lib, name)
            fn.restype = res
            fn.argtypes = args
        _LIB = lib
    return _LIB


def check(status: int) -> None:
    if status != 0:
        msg = load().avd_last_error_message().decode("utf-8", "replace")
        raise _BY_STATUS.get(status, Error)(msg)
