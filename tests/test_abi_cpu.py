"""The C ABI without a GPU: the library loads, exports every declared symbol, and validates
options, geometry and presets before touching the device."""
import re
import os

import pytest

import avd
from avd import _lib as L

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    text = open(os.path.join(ROOT, "include", "avd", "avd.h")).read()
    return sorted(set(re.findall(r"\b(avd_[a-z0-9_]+)\s*\(", text)))


def test_every_declared_symbol_is_exported_and_bound():
    lib = avd.load()
    names = declared_symbols()
    assert len(names) >= 30
    for name in names:
        assert hasattr(lib, name), name
        assert name in L.SIGNATURES, f"{name} has no ctypes signature"


def test_abi_version_and_defaults():
    lib = avd.load()
    assert lib.avd_abi_version() == 1
    m = avd.motion_search()
    assert (m.blksize, m.overlap, m.search_radius, m.pyramid_levels) == (16, 8, 4, 2)
    o = avd.nl4d_options()
    assert (o.refine, o.spatial_radius, o.temporal_radius, o.step, o.group_size) == (2, 9, 2, 4, 8)
    assert abs(o.c_min - 0.05) < 1e-7 and abs(o.kaiser_beta - 2.0) < 1e-7
    assert o.noise_map == 1 and o.pooled_threshold == 1 and o.has_sigma == 0
    n = avd.nlmeans_options()
    assert n.hq == 0 and n.auto_strength == 1 and n.temporal_confidence == 1


@pytest.mark.parametrize("preset,radius,search,hq", [("veryfast", 0, 2, 0), ("fast", 1, 2, 1), ("base", 2, 2, 1),
                                                      ("slow", 4, 4, 1), ("veryslow", 8, 4, 1)])
def test_nlmeans_preset_ladder(preset, radius, search, hq):
    o = avd.nlmeans_options_for_preset(preset)
    assert (o.temporal_radius, o.search_radius, o.hq) == (radius, search, hq)
    assert o.temporal == (1 if radius else 0)
    assert o.motion.max_search_escalation == (8 if preset in ("slow", "veryslow") else 4)


def test_nl4d_presets():
    assert avd.nl4d_options_for_preset("base").motion.max_search_escalation == 4
    assert avd.nl4d_options_for_preset("slow").motion.max_search_escalation == 8
    for bad in ("veryfast", "fast", "veryslow"):
        with pytest.raises(avd.InvalidOptions):
            avd.nl4d_options_for_preset(bad)


def test_plane_bytes():
    assert avd.plane_bytes(avd.SampleFormat.u8(), 5, 3) == 16
    assert avd.plane_bytes(avd.SampleFormat.u16(10), 5, 3) == 32
    assert avd.plane_bytes(avd.SampleFormat.u16(10), 3840, 2160) == 3840 * 2160 * 2


def geom(chroma=0, fmt_in=None, fmt_out=None, w=64, h=48):
    f = fmt_in or avd.SampleFormat.u16(10)
    return avd.Geometry(w, h, chroma, f, fmt_out or f, 0)


@pytest.mark.parametrize("kind,luma,chroma,g,err", [
    ("nl4d", dict(temporal_radius=0, sigma=0.02), None, geom(), avd.InvalidOptions),
    ("nl4d", dict(refine=5, sigma=0.02), None, geom(), avd.InvalidOptions),
    ("nl4d", dict(spatial_radius=17, sigma=0.02), None, geom(), avd.InvalidOptions),
    ("nl4d", dict(c_min=1.0, sigma=0.02), None, geom(), avd.InvalidOptions),
    ("nl4d", dict(kaiser_beta=9.0, sigma=0.02), None, geom(), avd.InvalidOptions),
    ("nl4d", dict(lambda_ht_scale=20.0, sigma=0.02), None, geom(), avd.InvalidOptions),
    ("nl4d", dict(pooled_threshold=0, noise_map=0), None, geom(), avd.InvalidOptions),  # no sigma
    ("nl4d", dict(sigma=0.02, pooled_threshold=0), dict(sigma=0.02), geom(0), avd.InvalidGeometry),  # gray chroma
    ("nl4d", dict(sigma=0.02, pooled_threshold=0), None, geom(0, avd.SampleFormat.u16(16)), avd.InvalidGeometry),
    ("nl4d", dict(sigma=0.02, pooled_threshold=0), None, geom(0, w=7, h=7), avd.InvalidGeometry),
    ("nl4d", dict(sigma=0.02, pooled_threshold=0, transform_form=2), None, geom(), avd.InvalidOptions),
    ("nl4d", dict(sigma=0.02, pooled_threshold=1, transform_form=1), None, geom(), avd.InvalidOptions),  # mixed: hard only
    # Unbuilt reference behaviour is refused, not replaced: the truemotion regulariser
    # (field_lambda > 0, nl4d/denoiser.rs:235) and Auto estimation at temporal radius >= 3, which
    # the reference resolves to Chained (motion/search.rs:133-138).
    ("nl4d", dict(sigma=0.02, pooled_threshold=0, field_lambda=1.0, motion={"estimation": 1}), None, geom(),
     avd.InvalidOptions),
    ("nl4d", dict(sigma=0.02, pooled_threshold=0, field_lambda=0.0, temporal_radius=3, motion={"estimation": 0}),
     None, geom(), avd.InvalidOptions),
    ("nl4d", dict(sigma=0.02, pooled_threshold=0, field_lambda=0.0, motion={"estimation": 2}), None, geom(),
     avd.InvalidOptions),
    ("nlmeans", dict(hq=1, sigma=0.02, temporal=1, temporal_radius=3, motion_compensation=1,
                     motion={"estimation": 0}), None, geom(), avd.InvalidOptions),
    # The chroma pass takes its motion from the luma pass (engine/spec.rs:320-341): both search the
    # same way at the same temporal radius, or neither searches.
    ("nl4d", dict(sigma=0.02, pooled_threshold=0),
     dict(sigma=0.02, pooled_threshold=0, motion={"estimation": 1, "blksize": 8, "overlap": 4}), geom(1),
     avd.InvalidOptions),
    ("nl4d", dict(sigma=0.02, pooled_threshold=0), dict(sigma=0.02, pooled_threshold=0, temporal_radius=1), geom(1),
     avd.InvalidOptions),
    ("nlmeans", dict(temporal=1, temporal_radius=1, motion_compensation=1, motion={"estimation": 1}),
     dict(temporal=1, temporal_radius=1, motion_compensation=0), geom(1), avd.InvalidOptions),
    ("nlmeans", dict(temporal=1, temporal_radius=1, motion_compensation=1, motion={"estimation": 1}),
     dict(temporal=1, temporal_radius=1, motion_compensation=1, motion={"estimation": 1, "search_radius": 2}),
     geom(1), avd.InvalidOptions),
    # MotionCompensationMode::validate (motion/search.rs:226-290): block size 4..=32 and even, search
    # radius 1..=8, escalation between the radius and 8, 1..=3 pyramid levels (3 is not built here).
    *[("nl4d", dict(sigma=0.02, pooled_threshold=0, motion={"estimation": 1, **m}), None, geom(), avd.InvalidOptions)
      for m in ({"blksize": 2, "overlap": 0}, {"blksize": 34, "overlap": 8}, {"blksize": 9, "overlap": 4},
                {"search_radius": 0, "max_search_escalation": 0}, {"max_search_escalation": 2},
                {"max_search_escalation": 9}, {"pyramid_levels": 0}, {"pyramid_levels": 3})],
    ("nlmeans", dict(hq=1, sigma=0.02, temporal=1, temporal_radius=1, motion_compensation=1,
                     motion={"estimation": 1, "search_radius": 9, "max_search_escalation": 9}), None, geom(),
     avd.InvalidOptions),
    ("nlmeans", dict(patch_radius=17), None, geom(), avd.InvalidOptions),
    ("nlmeans", dict(search_radius=9), None, geom(), avd.InvalidOptions),
    ("nlmeans", dict(hq=1), None, geom(), avd.InvalidOptions),  # HQ needs a fixed sigma here
    ("nlmeans", dict(), None, geom(0, avd.SampleFormat(2, 0)), avd.InvalidGeometry),  # F32 planes
    ("nlmeans", dict(), None, geom(1, avd.SampleFormat.u8(), avd.SampleFormat.u16(10)), avd.InvalidGeometry),
])
def test_engine_validation_errors(kind, luma, chroma, g, err):
    if kind == "nl4d":
        # Options the build runs, unless the case sets them: no truemotion, Direct estimation.
        luma = {"field_lambda": 0.0, "motion": {"estimation": 1}, **luma}
        chroma = None if chroma is None else {"field_lambda": 0.0, "motion": {"estimation": 1}, **chroma}
    mk = avd.nl4d_options if kind == "nl4d" else avd.nlmeans_options
    with pytest.raises(err):
        avd.build_engine(kind, g, luma=mk(**luma), chroma=mk(**chroma) if chroma is not None else None)


def test_the_defaults_that_resolve_to_built_behaviour_are_accepted():
    # Auto below temporal radius 3 is Direct (built); field_lambda 0 runs no truemotion.  Creating
    # an engine allocates device memory, so only the option check is reached here: without a GPU
    # the failure, if any, is a GPU error, never InvalidOptions.
    for opts in (dict(sigma=0.02, pooled_threshold=0, field_lambda=0.0, temporal_radius=2, motion={"estimation": 0}),
                 dict(sigma=0.02, pooled_threshold=0, field_lambda=0.0, temporal_radius=3, motion={"estimation": 1})):
        try:
            avd.build_engine("nl4d", geom(), luma=avd.nl4d_options(**opts))
        except avd.Gpu:
            pass


def test_a_spec_must_clean_something():
    with pytest.raises(avd.InvalidOptions):
        avd.build_engine("nlmeans", geom(1), luma=None, chroma=None)


def test_noise_site_rejects_bad_arguments_before_any_launch():
    import ctypes

    from avd import _lib as L

    lib = L.load()
    buf = ctypes.c_void_p(1)  # never dereferenced: validation fails first
    assert lib.avd_site_noise_spatial(None, 64, 48, 1, 0, buf, buf, 0, None) == avd.NullArgument.status
    assert lib.avd_site_noise_spatial(buf, 64, 48, 3, 0, buf, buf, 0, None) == avd.InvalidGeometry.status
    assert lib.avd_site_noise_spatial(buf, 0, 48, 1, 0, buf, buf, 0, None) == avd.InvalidGeometry.status


def test_noalloc_noise_entry_refuses_frames_past_its_static_scratch():
    import ctypes

    from avd import _lib as L

    lib = L.load()
    w, h = 8192, 8200  # 256 x 1025 tiles: 524800 partials, above the 2^19 floats of the array
    assert lib.avd_noise_partials_len(w, h) > (1 << 19)
    assert lib.avd_noise_partials_len(8192, 8192) == (1 << 19)
    buf = ctypes.c_void_p(1)  # never dereferenced: the checks fail before any launch
    assert lib.avd_site_noise_spatial_noalloc(buf, w, h, 1, 0, buf, 0, None) == avd.InvalidGeometry.status
    assert lib.avd_site_noise_spatial_noalloc(None, 64, 48, 1, 0, buf, 0, None) == avd.NullArgument.status
    assert lib.avd_site_noise_spatial_noalloc(buf, 64, 48, 3, 0, buf, 0, None) == avd.InvalidGeometry.status
