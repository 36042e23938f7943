// The extern "C" boundary: status codes, thread-local error messages, defaults and presets, the
// engine entry points and the per-stage launch sites.  No exception crosses this file.
#include <cmath>
#include <cstring>
#include <limits>
#include <new>
#include <string>

#include "engine.h"
#include "host.h"
#include "tune.h"

struct avd_engine {
    avd::Engine impl;
    explicit avd_engine(const avd_engine_spec &s) : impl(s) {}
};

struct avd_host_denoiser {
    avd::HostDenoiser impl;
    explicit avd_host_denoiser(const avd_engine_spec &s) : impl(s) {}
};

namespace {

thread_local std::string g_error;

avd_status fail(avd_status s, const std::string &m)
{
    g_error = m;
    return s;
}

template <typename F>
avd_status guard(F &&f)
{
    try {
        g_error.clear();
        f();
        return AVD_OK;
    } catch (const avd::Failure &e) {
        return fail(e.status, e.message);
    } catch (const std::bad_alloc &) {
        return fail(AVD_ERR_GPU, "host allocation failed");
    } catch (const std::exception &e) {
        return fail(AVD_ERR_GPU, e.what());
    }
}

avd_status hip_status(hipError_t e, const char *what)
{
    if (e == hipSuccess) return AVD_OK;
    if (e == hipErrorInvalidValue) return fail(AVD_ERR_INVALID_OPTIONS, std::string(what) + ": unsupported parameters");
    return fail(AVD_ERR_GPU, std::string(what) + ": " + hipGetErrorString(e));
}

avd_status check_format(avd_sample_format f)
{
    return guard([&] { avd::validate_format(f); });
}

}  // namespace

extern "C" {

uint32_t avd_abi_version(void) { return AVD_ABI_VERSION; }

const char *avd_last_error_message(void) { return g_error.c_str(); }

// MotionSearch::default (motion/search.rs:29-40).
void avd_motion_search_default(avd_motion_search *out)
{
    if (!out) return;
    std::memset(out, 0, sizeof(*out));
    out->blksize = 16;
    out->overlap = 8;
    out->search_radius = 4;
    out->max_search_escalation = 4;
    out->pyramid_levels = 2;
    out->estimation = 0;  // Auto
    out->refine_radius = 2;
}

// NlmeansOptions / HqParams defaults (nlmeans/options.rs:6-27, nlmeans/params.rs:139-152).
void avd_nlmeans_options_default(avd_nlmeans_options *out)
{
    if (!out) return;
    std::memset(out, 0, sizeof(*out));
    avd_motion_search_default(&out->motion);
    out->auto_strength = 1;
    out->noise_floor = 1;
    out->temporal_confidence = 1;
    out->thsad_scale = 1.0f;
    out->sigma_scale = 1.0f;
}

// nlmeans_variant_for / nlmeans_temporal_radius_for / nlmeans_search_radius_for
// (nlmeans/options.rs:58-82) and search_escalation_for (options.rs:32-38).
avd_status avd_nlmeans_options_for_preset(uint32_t preset, avd_nlmeans_options *out)
{
    if (!out) return fail(AVD_ERR_NULL_ARGUMENT, "out is null");
    if (preset > 4) return fail(AVD_ERR_INVALID_OPTIONS, "unknown preset");
    static const uint32_t radius[5] = {0, 1, 2, 4, 8};
    static const uint32_t search[5] = {2, 2, 2, 4, 4};
    avd_nlmeans_options_default(out);
    out->hq = preset == 0 ? 0 : 1;
    out->temporal_radius = radius[preset];
    out->temporal = radius[preset] > 0 ? 1 : 0;
    out->has_search_radius = 1;
    out->search_radius = search[preset];
    if (preset >= 3) out->motion.max_search_escalation = 8;
    return AVD_OK;
}

// Nl4dOptions::default (nl4d/options.rs:68-94, nl4d/params.rs:107-129).
void avd_nl4d_options_default(avd_nl4d_options *out)
{
    if (!out) return;
    std::memset(out, 0, sizeof(*out));
    avd_motion_search_default(&out->motion);
    out->sigma_scale = 1.0f;
    out->thsad_scale = 1.0f;
    out->refine = 2;
    out->spatial_radius = 9;
    out->lambda_ht_scale = 1.0f;
    out->c_min = 0.05f;
    out->kaiser_beta = 2.0f;
    out->field_lambda = 1.0f;
    out->noise_map = 1;
    out->pooled_threshold = 1;
    out->grain_data = 0;
    out->temporal_radius = 2;
    out->step = 4;
    out->group_size = 8;
}

// nl4d runs the base and slow presets only; slow raises the search escalation ceiling to 8
// (av-denoise/src/bin/cli/nl4d.rs:231-275).
avd_status avd_nl4d_options_for_preset(uint32_t preset, avd_nl4d_options *out)
{
    if (!out) return fail(AVD_ERR_NULL_ARGUMENT, "out is null");
    if (preset != 2 && preset != 3) return fail(AVD_ERR_INVALID_OPTIONS, "nl4d only runs the base and slow presets");
    avd_nl4d_options_default(out);
    if (preset == 3) out->motion.max_search_escalation = 8;
    return AVD_OK;
}

uint64_t avd_plane_bytes(avd_sample_format format, uint32_t width, uint32_t height)
{
    return avd::plane_bytes(format, (uint64_t)width * height);
}

avd_status avd_engine_create(const avd_engine_spec *spec, avd_engine **out)
{
    if (!spec || !out) return fail(AVD_ERR_NULL_ARGUMENT, "spec or out is null");
    *out = nullptr;
    return guard([&] { *out = new avd_engine(*spec); });
}

void avd_engine_destroy(avd_engine *engine) { delete engine; }

avd_status avd_engine_push(avd_engine *e, const avd_device_plane *planes, uint32_t count, uint32_t *ready)
{
    if (!e || !planes || !ready) return fail(AVD_ERR_NULL_ARGUMENT, "engine, planes or ready is null");
    return guard([&] { *ready = e->impl.push(planes, count); });
}

avd_status avd_engine_push_context(avd_engine *e, const avd_device_plane *planes, uint32_t count)
{
    if (!e || !planes) return fail(AVD_ERR_NULL_ARGUMENT, "engine or planes is null");
    return guard([&] { e->impl.push_context(planes, count); });
}

avd_status avd_engine_emit_into(avd_engine *e, const avd_device_plane *planes, uint32_t count)
{
    if (!e || !planes) return fail(AVD_ERR_NULL_ARGUMENT, "engine or planes is null");
    return guard([&] { e->impl.emit(planes, count); });
}

avd_status avd_engine_finish(avd_engine *e, uint32_t *tail)
{
    if (!e || !tail) return fail(AVD_ERR_NULL_ARGUMENT, "engine or tail is null");
    return guard([&] { *tail = e->impl.finish(); });
}

avd_status avd_engine_reset(avd_engine *e)
{
    if (!e) return fail(AVD_ERR_NULL_ARGUMENT, "engine is null");
    return guard([&] { e->impl.reset(); });
}

avd_status avd_engine_window_span(const avd_engine *e, avd_window_span *out)
{
    if (!e || !out) return fail(AVD_ERR_NULL_ARGUMENT, "engine or out is null");
    return guard([&] { *out = e->impl.window_span(); });
}

avd_status avd_engine_max_held_frames(const avd_engine *e, uint64_t *out)
{
    if (!e || !out) return fail(AVD_ERR_NULL_ARGUMENT, "engine or out is null");
    return guard([&] { *out = e->impl.max_held_frames(); });
}

avd_status avd_engine_synchronize(avd_engine *e)
{
    if (!e) return fail(AVD_ERR_NULL_ARGUMENT, "engine is null");
    return guard([&] { e->impl.synchronize(); });
}

avd_status avd_engine_fail_for_test(avd_engine *e)
{
    if (!e) return fail(AVD_ERR_NULL_ARGUMENT, "engine is null");
    return guard([&] { e->impl.fail_for_test(); });
}

avd_status avd_engine_set_tile_cap_for_test(avd_engine *e, uint32_t cap)
{
    if (!e) return fail(AVD_ERR_NULL_ARGUMENT, "engine is null");
    return guard([&] { e->impl.set_tile_cap_for_test(cap); });
}

avd_status avd_engine_tile_overflows_for_test(const avd_engine *e, uint64_t *pairs)
{
    if (!e || !pairs) return fail(AVD_ERR_NULL_ARGUMENT, "engine or pairs is null");
    return guard([&] { *pairs = e->impl.tile_overflows_for_test(); });
}

avd_status avd_engine_sample_device_time(avd_engine *e, uint32_t every)
{
    if (!e) return fail(AVD_ERR_NULL_ARGUMENT, "engine is null");
    return guard([&] { e->impl.sample_device_time(every); });
}

avd_status avd_engine_set_nlm_form(avd_engine *e, uint32_t form)
{
    if (!e) return fail(AVD_ERR_NULL_ARGUMENT, "engine is null");
    return guard([&] { e->impl.set_nlm_form(form); });
}

avd_status avd_engine_timing(avd_engine *e, avd_timing *out)
{
    if (!e || !out) return fail(AVD_ERR_NULL_ARGUMENT, "engine or out is null");
    return guard([&] { *out = e->impl.timing(); });
}

// ---- launch sites ----------------------------------------------------------------------------

avd_status avd_host_denoiser_create(const avd_engine_spec *spec, avd_host_denoiser **out)
{
    if (!spec || !out) return fail(AVD_ERR_NULL_ARGUMENT, "spec or out is null");
    *out = nullptr;
    return guard([&] { *out = new avd_host_denoiser(*spec); });
}

void avd_host_denoiser_destroy(avd_host_denoiser *d) { delete d; }

avd_status avd_host_denoiser_plane_lengths(const avd_host_denoiser *d, uint64_t *input, uint64_t *output,
                                           uint32_t count, uint32_t *planes)
{
    if (!d || !planes) return fail(AVD_ERR_NULL_ARGUMENT, "denoiser or planes is null");
    *planes = d->impl.planes();
    for (uint32_t i = 0; i < count && i < d->impl.planes(); ++i) {
        if (input) input[i] = d->impl.input_length(i);
        if (output) output[i] = d->impl.output_length(i);
    }
    return AVD_OK;
}

avd_status avd_host_denoiser_push(avd_host_denoiser *d, const void *const *planes, const uint64_t *lengths,
                                  uint32_t count)
{
    if (!d || !planes || !lengths) return fail(AVD_ERR_NULL_ARGUMENT, "denoiser, planes or lengths is null");
    return guard([&] { d->impl.push(planes, lengths, count); });
}

avd_status avd_host_denoiser_push_priming(avd_host_denoiser *d, const void *const *planes, const uint64_t *lengths,
                                          uint32_t count)
{
    if (!d || !planes || !lengths) return fail(AVD_ERR_NULL_ARGUMENT, "denoiser, planes or lengths is null");
    return guard([&] { d->impl.push_priming(planes, lengths, count); });
}

avd_status avd_host_denoiser_recv(avd_host_denoiser *d, void *const *planes, const uint64_t *capacities,
                                  uint32_t count, uint32_t *received)
{
    if (!d || !planes || !capacities || !received)
        return fail(AVD_ERR_NULL_ARGUMENT, "denoiser, planes, capacities or received is null");
    *received = 0;
    return guard([&] { *received = d->impl.recv(planes, capacities, count) ? 1u : 0u; });
}

avd_status avd_host_denoiser_flush(avd_host_denoiser *d, avd_frame_sink sink, void *user)
{
    if (!d || !sink) return fail(AVD_ERR_NULL_ARGUMENT, "denoiser or sink is null");
    return guard([&] {
        d->impl.flush([&](const void *const *planes, const uint64_t *lengths, uint32_t count) {
            sink(user, planes, lengths, count);
        });
    });
}

avd_status avd_host_denoiser_reset_stream(avd_host_denoiser *d)
{
    if (!d) return fail(AVD_ERR_NULL_ARGUMENT, "denoiser is null");
    return guard([&] { d->impl.reset_stream(); });
}

avd_status avd_host_denoiser_pending(const avd_host_denoiser *d, uint32_t *pending)
{
    if (!d || !pending) return fail(AVD_ERR_NULL_ARGUMENT, "denoiser or pending is null");
    *pending = d->impl.pending();
    return AVD_OK;
}

avd_status avd_host_denoiser_window_span(const avd_host_denoiser *d, avd_window_span *out)
{
    if (!d || !out) return fail(AVD_ERR_NULL_ARGUMENT, "denoiser or out is null");
    return guard([&] { *out = d->impl.window_span(); });
}

void avd_tuning_set_enabled(uint32_t enabled) { avd::Tuner::instance().set_enabled(enabled != 0); }

avd_status avd_tuning_entries(avd_tune_entry *out, uint32_t capacity, uint32_t *count)
{
    if (!count || (capacity && !out)) return fail(AVD_ERR_NULL_ARGUMENT, "out or count is null");
    return guard([&] {
        const std::vector<avd::TuneEntry> all = avd::Tuner::instance().entries();
        *count = (uint32_t)all.size();
        for (uint32_t i = 0; i < capacity && i < all.size(); ++i) {
            avd_tune_entry &e = out[i];
            std::memset(&e, 0, sizeof(e));
            std::strncpy(e.site, all[i].site.c_str(), sizeof(e.site) - 1);
            std::strncpy(e.key, all[i].key.c_str(), sizeof(e.key) - 1);
            e.device = (uint32_t)all[i].device;
            e.choice = (uint32_t)all[i].choice;
            e.candidates = (uint32_t)all[i].candidates;
            for (size_t k = 0; k < all[i].ms.size() && k < 8; ++k) e.ms[k] = all[i].ms[k];
        }
    });
}

avd_status avd_site_ingest(const void *plane0, const void *plane1, float *slot, uint16_t *codes, uint32_t pixels,
                           uint32_t channels, avd_sample_format format, void *stream)
{
    if (!plane0 || !slot || (channels == 2 && !plane1)) return fail(AVD_ERR_NULL_ARGUMENT, "null plane or slot");
    if (channels < 1 || channels > 2) return fail(AVD_ERR_INVALID_GEOMETRY, "channels must be 1 or 2");
    avd_status st = check_format(format);
    if (st != AVD_OK) return st;
    return hip_status(avd::launch_ingest(static_cast<const uint32_t *>(plane0),
                                         static_cast<const uint32_t *>(plane1 ? plane1 : plane0), slot, codes, pixels,
                                         (int)channels, (int)avd::samples_per_word(format), avd::max_value(format),
                                         static_cast<hipStream_t>(stream)),
                      "ingest");
}

avd_status avd_site_egress(const float *frame, void *plane0, void *plane1, uint32_t pixels, uint32_t channels,
                           avd_sample_format format, void *stream)
{
    if (!frame || !plane0 || (channels == 2 && !plane1)) return fail(AVD_ERR_NULL_ARGUMENT, "null frame or plane");
    if (channels < 1 || channels > 2) return fail(AVD_ERR_INVALID_GEOMETRY, "channels must be 1 or 2");
    avd_status st = check_format(format);
    if (st != AVD_OK) return st;
    return hip_status(avd::launch_egress(frame, static_cast<uint32_t *>(plane0),
                                         static_cast<uint32_t *>(plane1 ? plane1 : plane0), pixels, (int)channels,
                                         (int)avd::samples_per_word(format), avd::max_value(format),
                                         static_cast<hipStream_t>(stream)),
                      "egress");
}

avd_status avd_site_collab_search(const avd_collab_pass *pass, const uint16_t *ring_codes, const float *ring_values,
                                  const uint32_t *neighbour_slots, const int32_t *mv_field, const float *confidence,
                                  uint32_t *members, void *stream)
{
    if (!pass || !members || (!ring_codes && !ring_values)) return fail(AVD_ERR_NULL_ARGUMENT, "null argument");
    if (pass->temporal_radius > 0 && (!neighbour_slots || !mv_field || !confidence))
        return fail(AVD_ERR_NULL_ARGUMENT, "a temporal pass needs neighbour slots, motion and confidence");
    const hipStream_t s = static_cast<hipStream_t>(stream);
    if (!ring_codes || !avd::centre_on_matrix_cores(*pass))
        return hip_status(avd::launch_collab_search(*pass, ring_codes, ring_values, neighbour_slots, mv_field,
                                                    confidence, members, s),
                          "collab_search");
    // The matrix-core centre search's scratch lives for this call, in stream order.
    void *scratch = nullptr;
    hipError_t e = hipMallocAsync(&scratch, avd::centre_scratch_bytes(pass->width, pass->height, pass->channels), s);
    if (e != hipSuccess) return hip_status(e, "collab_search scratch");
    const avd::CentreScratch cs = avd::centre_scratch_at(scratch, pass->width, pass->height, pass->channels);
    e = avd::launch_collab_search(*pass, ring_codes, ring_values, neighbour_slots, mv_field, confidence, members, s,
                                  nullptr, &cs);
    const hipError_t f = hipFreeAsync(scratch, s);
    return hip_status(e != hipSuccess ? e : f, "collab_search");
}

avd_status avd_site_collab_filter(const avd_collab_pass *pass, const float *ring_values,
                                  const uint32_t *neighbour_slots, const uint32_t *members, uint64_t *aggregate,
                                  int32_t *accum_second, float *group_weight, void *stream)
{
    if (!pass || !ring_values || !members || !aggregate || !group_weight || (pass->channels == 2 && !accum_second))
        return fail(AVD_ERR_NULL_ARGUMENT, "null argument");
    return hip_status(avd::launch_collab_filter(*pass, ring_values, neighbour_slots, members,
                                                reinterpret_cast<unsigned long long *>(aggregate), accum_second,
                                                group_weight, static_cast<hipStream_t>(stream)),
                      "collab_filter");
}

avd_status avd_site_collab_normalise(uint64_t *aggregate_region, int32_t *accum_second_region, float *out,
                                     uint32_t pixels, uint32_t channels, void *stream)
{
    if (!aggregate_region || !out || (channels == 2 && !accum_second_region))
        return fail(AVD_ERR_NULL_ARGUMENT, "null argument");
    if (channels < 1 || channels > 2) return fail(AVD_ERR_INVALID_GEOMETRY, "channels must be 1 or 2");
    return hip_status(avd::launch_collab_normalise(reinterpret_cast<unsigned long long *>(aggregate_region),
                                                   accum_second_region, out, pixels, (int)channels,
                                                   static_cast<hipStream_t>(stream)),
                      "collab_normalise");
}

avd_status avd_site_noise_spatial(const float *ring_values, uint32_t width, uint32_t height, uint32_t channels,
                                  uint32_t frame, float *partials, float *results, uint32_t slot, void *stream)
{
    if (!ring_values || !partials || !results) return fail(AVD_ERR_NULL_ARGUMENT, "null argument");
    if (width == 0 || height == 0 || (channels != 1 && channels != 2))
        return fail(AVD_ERR_INVALID_GEOMETRY, "noise measurement needs a non-empty frame of 1 or 2 channels");
    return hip_status(avd::launch_noise_spatial(ring_values, width, height, channels, frame, partials, results, slot,
                                                static_cast<hipStream_t>(stream)),
                      "noise_spatial");
}

avd_status avd_site_noise_spatial_noalloc(const float *ring_values, uint32_t width, uint32_t height,
                                          uint32_t channels, uint32_t frame, float *results, uint32_t slot,
                                          void *stream)
{
    if (!ring_values || !results) return fail(AVD_ERR_NULL_ARGUMENT, "null argument");
    if (width == 0 || height == 0 || (channels != 1 && channels != 2))
        return fail(AVD_ERR_INVALID_GEOMETRY, "noise measurement needs a non-empty frame of 1 or 2 channels");
    if (avd::noise_partials_len(width, height) > avd::kNoiseScratchFloats)
        return fail(AVD_ERR_INVALID_GEOMETRY, "frame of " + std::to_string(width) + "x" + std::to_string(height) +
                                                  " needs more partials than the static scratch holds (" +
                                                  std::to_string(avd::kNoiseScratchFloats) + " floats)");
    float *partials = nullptr;
    const hipError_t e = avd::noise_scratch_address(&partials);
    if (e != hipSuccess) return hip_status(e, "noise_spatial_noalloc scratch");
    return hip_status(avd::launch_noise_spatial(ring_values, width, height, channels, frame, partials, results, slot,
                                                static_cast<hipStream_t>(stream)),
                      "noise_spatial_noalloc");
}

uint32_t avd_noise_partials_len(uint32_t width, uint32_t height) { return avd::noise_partials_len(width, height); }

avd_status avd_site_noise_temporal_stats(const float *ring_values, uint32_t width, uint32_t height, uint32_t channels,
                                         uint32_t slot_new, uint32_t slot_prev, const int32_t *mv_field,
                                         uint32_t mv_offset, uint32_t mv_step, uint32_t mv_blocks_x,
                                         uint32_t mv_blocks_y, uint32_t luma_fields, float *stats, void *stream)
{
    if (!ring_values || !stats) return fail(AVD_ERR_NULL_ARGUMENT, "null argument");
    if (width == 0 || height == 0 || width > (1u << 20) || height > (1u << 20) || (channels != 1 && channels != 2))
        return fail(AVD_ERR_INVALID_GEOMETRY, "temporal statistics need a non-empty frame of 1 or 2 channels");
    if ((uint64_t)((width + 15) / 16) * ((height + 15) / 16) * (3 * channels + 26) > 0xFFFFFFFFull)
        return fail(AVD_ERR_INVALID_GEOMETRY, "the frame's temporal records exceed 32-bit offsets");
    if (mv_field && (mv_step == 0 || mv_blocks_x == 0 || mv_blocks_y == 0))
        return fail(AVD_ERR_INVALID_OPTIONS, "a motion field needs a non-zero step and block grid");
    return hip_status(avd::launch_temporal_stats(ring_values, width, height, channels, slot_new, slot_prev, mv_field,
                                                 mv_offset, mv_step, mv_blocks_x, mv_blocks_y, luma_fields != 0, stats,
                                                 static_cast<hipStream_t>(stream)),
                      "noise_temporal_stats");
}

avd_status avd_site_noise_temporal_zero(float *stats, uint32_t len, void *stream)
{
    if (!stats) return fail(AVD_ERR_NULL_ARGUMENT, "null argument");
    return hip_status(avd::launch_temporal_zero(stats, len, static_cast<hipStream_t>(stream)), "noise_temporal_zero");
}

uint32_t avd_noise_temporal_stats_len(uint32_t width, uint32_t height, uint32_t channels)
{
    return avd::temporal_stats_len(width, height, channels);
}

avd_status avd_site_grain_measure(const float *ring_values, const float *denoised, uint32_t width, uint32_t height,
                                  uint32_t channels, uint32_t channel, uint32_t slot_t, uint32_t slot_next,
                                  const int32_t *saved_mv, const float *saved_conf, uint32_t blocks_x,
                                  uint32_t blocks_y, uint32_t step, uint32_t source_entry, uint32_t has_source,
                                  const float *luma_means, uint32_t has_luma_means, uint32_t means_blocks_x,
                                  uint32_t luma_band, const float *edges, int32_t *hist, float *partials,
                                  void *stream)
{
    if (!ring_values || !denoised || !saved_mv || !saved_conf || !edges || !hist || !partials ||
        (means_blocks_x > 0 && !luma_means))
        return fail(AVD_ERR_NULL_ARGUMENT, "null argument");
    if (width == 0 || height == 0 || width > (1u << 16) || height > (1u << 16) || channels < 1 || channels > 4 ||
        channel >= channels)
        return fail(AVD_ERR_INVALID_GEOMETRY, "grain measurement needs a frame of 1 to 4 channels and a channel inside it");
    if (blocks_x == 0 || blocks_y == 0 || step == 0)
        return fail(AVD_ERR_INVALID_OPTIONS, "grain measurement needs a non-empty motion block grid and step");
    avd::GrainMeasure m{ring_values, denoised, saved_mv, saved_conf, edges, luma_means, hist, partials,
                        width, height, channels, channel, slot_t, slot_next, blocks_x, blocks_y, step,
                        source_entry, means_blocks_x, has_source != 0, has_luma_means != 0, luma_band != 0};
    return hip_status(avd::launch_grain_measure(m, static_cast<hipStream_t>(stream)), "grain_measure");
}

uint32_t avd_grain_partials_len(uint32_t width, uint32_t height) { return avd::grain_partials_len(width, height); }

avd_status avd_site_grain_save_vectors(const int32_t *mv_field, const float *confidence, int32_t *saved_mv,
                                       float *saved_conf, uint32_t mv_offset, uint32_t conf_offset, uint32_t entry,
                                       uint32_t blocks, void *stream)
{
    if (!mv_field || !confidence || !saved_mv || !saved_conf) return fail(AVD_ERR_NULL_ARGUMENT, "null argument");
    return hip_status(avd::launch_grain_save_vectors(mv_field, confidence, saved_mv, saved_conf, mv_offset,
                                                     conf_offset, entry, blocks, static_cast<hipStream_t>(stream)),
                      "grain_save_vectors");
}

avd_status avd_site_grain_project(const float *ring_values, const float *denoised, const uint32_t *patterns,
                                  float *history, uint32_t width, uint32_t height, uint32_t channels,
                                  uint32_t channel, uint32_t frame_slot, uint32_t history_slot, uint32_t luma_band,
                                  void *stream)
{
    if (!ring_values || !denoised || !patterns || !history) return fail(AVD_ERR_NULL_ARGUMENT, "null argument");
    if (width == 0 || height == 0 || width > (1u << 16) || height > (1u << 16) || channels < 1 || channels > 4 ||
        channel >= channels)
        return fail(AVD_ERR_INVALID_GEOMETRY, "grain projection needs a frame of 1 to 4 channels and a channel inside it");
    if (history_slot >= 9) return fail(AVD_ERR_INVALID_OPTIONS, "the grain history holds 9 slots");
    return hip_status(avd::launch_grain_project(ring_values, denoised, patterns, history, width, height, channels,
                                                channel, frame_slot, history_slot, luma_band != 0,
                                                static_cast<hipStream_t>(stream)),
                      "grain_project");
}

uint32_t avd_grain_persist_partials_len(uint32_t cells, uint32_t lags) { return avd::grain_persist_blocks(cells) * lags * 3; }

avd_status avd_site_grain_persist(const float *history, float *partials, float *record, uint32_t record_offset,
                                  uint32_t now_slot, uint32_t cells, uint32_t lags, void *stream)
{
    if (!history || !partials || !record) return fail(AVD_ERR_NULL_ARGUMENT, "null argument");
    if (now_slot >= 9 || lags > 8) return fail(AVD_ERR_INVALID_OPTIONS, "the grain history holds 9 slots, 8 lags at most");
    if ((uint64_t)cells * 9 * 9 > 0xFFFFFFFFull) return fail(AVD_ERR_INVALID_GEOMETRY, "too many grain cells");
    return hip_status(avd::launch_grain_persist(history, partials, record, record_offset, now_slot, cells, lags,
                                                static_cast<hipStream_t>(stream)),
                      "grain_persist");
}

avd_status avd_site_grain_fit(const double *records, uint32_t count, double *corrected, double *weights,
                              int32_t *quantised, uint32_t *shifts, int32_t *solved, void *stream)
{
    if (!records || !corrected || !weights || !quantised || !shifts || !solved)
        return fail(AVD_ERR_NULL_ARGUMENT, "null argument");
    return hip_status(avd::launch_grain_fit(records, count, corrected, weights, quantised, shifts, solved,
                                            static_cast<hipStream_t>(stream)),
                      "grain_fit");
}

avd_status avd_site_grain_reduce(const float *partials, uint32_t cells, float *chunk, uint32_t record_offset,
                                 void *stream)
{
    if (!partials || !chunk) return fail(AVD_ERR_NULL_ARGUMENT, "null argument");
    if ((uint64_t)cells * 48 > 0xFFFFFFFFull)
        return fail(AVD_ERR_INVALID_GEOMETRY, "more grain cells than 32-bit partial offsets reach");
    return hip_status(avd::launch_grain_reduce(partials, cells, chunk, record_offset, static_cast<hipStream_t>(stream)),
                      "grain_reduce");
}

float avd_noise_sigma_from_abs_sum(float abs_sum, uint32_t width, uint32_t height)
{
    // The 3x3 mask has no interior response below 3 samples on an axis (the reference's version
    // underflows there); NaN says that no estimate exists.
    if (width < 3 || height < 3) return std::numeric_limits<float>::quiet_NaN();
    const float interior = (float)(width - 2) * (float)(height - 2);
    return std::sqrt(1.57079632679489661923f) * abs_sum / (6.0f * interior);
}

avd_status avd_site_motion(const avd_motion_pass *p, const uint16_t *ring_codes, const uint32_t *neighbour_slots,
                           int32_t *mv_field, float *confidence, void *stream)
{
    if (!p || !ring_codes || !neighbour_slots || !mv_field) return fail(AVD_ERR_NULL_ARGUMENT, "null argument");
    if (p->overlap >= p->blksize || p->width == 0 || p->height == 0 || p->channels < 1 || p->channels > 2)
        return fail(AVD_ERR_INVALID_OPTIONS, "invalid motion geometry");
    hipStream_t s = static_cast<hipStream_t>(stream);
    avd::MotionGeom g{(int)p->width, (int)p->height, (int)p->channels, (int)p->blksize, (int)(p->blksize - p->overlap),
                      (int)p->search_radius, (int)p->pyramid_levels, 0, 0};
    g.blocks_x = std::max((int)avd::div_up(p->width, g.step), 1);
    g.blocks_y = std::max((int)avd::div_up(p->height, g.step), 1);
    // Level-1 sums of every slot the pass reads, in scratch owned by this call.
    const size_t l1 = (size_t)std::max(p->width / 2, 1u) * std::max(p->height / 2, 1u);
    uint32_t slots = p->centre_slot + 1;
    std::vector<uint32_t> host(p->neighbours);
    hipError_t e = hipMemcpyAsync(host.data(), neighbour_slots, p->neighbours * sizeof(uint32_t),
                                  hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (e != hipSuccess) return hip_status(e, "read neighbour slots");
    for (uint32_t n : host) slots = std::max(slots, n + 1);
    int32_t *pyr = nullptr;
    if (p->pyramid_levels >= 2) {
        e = hipMalloc(&pyr, slots * l1 * sizeof(int32_t));
        if (e != hipSuccess) return hip_status(e, "hipMalloc");
        const size_t slot_len = (size_t)p->width * p->height * p->channels;
        for (uint32_t sl = 0; sl < slots && e == hipSuccess; ++sl)
            e = avd::launch_downscale(ring_codes + sl * slot_len, p->width, p->height, p->channels, pyr + sl * l1, s);
    }
    if (e == hipSuccess)
        e = avd::launch_motion(g, ring_codes, pyr, (int)p->centre_slot, neighbour_slots, (int)p->neighbours, mv_field,
                               confidence, p->max_value, p->sad_floor, p->thsad, s);
    if (pyr) {
        hipError_t e2 = hipStreamSynchronize(s);
        (void)hipFree(pyr);
        if (e == hipSuccess) e = e2;
    }
    return hip_status(e, "motion");
}

avd_status avd_site_motion_refine(const avd_motion_pass *p, const uint16_t *ring_codes,
                                  const uint32_t *neighbour_slots, const int32_t *seeds, uint32_t refine_radius,
                                  int32_t *mv_field, float *confidence, void *stream)
{
    if (!p || !ring_codes || !neighbour_slots || !seeds || !mv_field) return fail(AVD_ERR_NULL_ARGUMENT, "null argument");
    if (p->overlap >= p->blksize || p->width == 0 || p->height == 0 || p->channels < 1 || p->channels > 2 ||
        refine_radius > 8)
        return fail(AVD_ERR_INVALID_OPTIONS, "invalid motion geometry");
    avd::MotionGeom g{(int)p->width, (int)p->height, (int)p->channels, (int)p->blksize, (int)(p->blksize - p->overlap),
                      (int)refine_radius, 1, 0, 0};
    g.blocks_x = std::max((int)avd::div_up(p->width, g.step), 1);
    g.blocks_y = std::max((int)avd::div_up(p->height, g.step), 1);
    return hip_status(avd::launch_motion_refine(g, ring_codes, (int)p->centre_slot, neighbour_slots, (int)p->neighbours,
                                                seeds, mv_field, confidence, p->max_value, p->sad_floor, p->thsad,
                                                (int)refine_radius, static_cast<hipStream_t>(stream)),
                      "motion refine");
}

avd_status avd_site_map_luma_fields(const avd_field_layout *luma, const int32_t *luma_fields,
                                    const avd_field_layout *chroma, int32_t *chroma_fields, uint32_t shift_x,
                                    uint32_t shift_y, uint32_t fields, void *stream)
{
    if (!luma || !chroma || !luma_fields || !chroma_fields) return fail(AVD_ERR_NULL_ARGUMENT, "null argument");
    auto bad = [](const avd_field_layout *l) {
        return l->blocks_x == 0 || l->blocks_y == 0 || l->step == 0 || l->blksize == 0 || l->blocks_x > 65535 ||
               l->blocks_y > 65535 || (uint64_t)l->stride < (uint64_t)l->blocks_x * l->blocks_y * 2 ||
               l->stride > (1u << 30);
    };
    if (bad(luma) || bad(chroma) || shift_x > 1 || shift_y > 1)
        return fail(AVD_ERR_INVALID_OPTIONS, "invalid field layout");
    const avd::FieldMap m{(int)luma->blocks_x,   (int)luma->blocks_y,   (int)luma->step,    (int)luma->blksize,
                          (int)luma->stride,     (int)chroma->blocks_x, (int)chroma->blocks_y, (int)chroma->step,
                          (int)chroma->blksize,  (int)chroma->stride,   (int)shift_x,       (int)shift_y};
    return hip_status(avd::launch_map_luma_fields(m, luma_fields, chroma_fields, (int)fields,
                                                  static_cast<hipStream_t>(stream)),
                      "map luma fields");
}

avd_status avd_site_nlm_single(const avd_nlm_window *w, const float *frame, const float *lut, float *accum,
                               float *wsum, float *maxw, void *stream)
{
    if (!w || !frame || !lut || !accum || !wsum || !maxw) return fail(AVD_ERR_NULL_ARGUMENT, "null argument");
    avd::NlmWindow k{(int)w->width, (int)w->height, (int)w->channels, (int)w->patch_radius, (int)w->search_radius,
                     w->h2_inv_norm * 1.4426950408889634f, w->noise_offset, 1, 1, 1};
    return hip_status(avd::launch_nlm_single(k, frame, lut, accum, wsum, maxw, static_cast<hipStream_t>(stream)),
                      "nlm_single");
}

avd_status avd_site_nlm_pair(const avd_nlm_window *w, const float *centre, const float *fwd, const float *bwd,
                             const float *conf_fwd, const float *conf_bwd, float *accum, float *wsum, float *maxw,
                             void *stream)
{
    if (!w || !centre || !fwd || !bwd || !accum || !wsum || !maxw) return fail(AVD_ERR_NULL_ARGUMENT, "null argument");
    avd::NlmWindow k{(int)w->width, (int)w->height, (int)w->channels, (int)w->patch_radius, (int)w->search_radius,
                     w->h2_inv_norm * 1.4426950408889634f, w->noise_offset, (int)w->step, (int)w->blocks_x,
                     (int)w->blocks_y};
    return hip_status(avd::launch_nlm_pair(k, centre, fwd, bwd, conf_fwd, conf_bwd, accum, wsum, maxw,
                                           static_cast<hipStream_t>(stream)),
                      "nlm_pair");
}

avd_status avd_site_nlm_finish(const float *frame, const float *accum, const float *wsum, const float *maxw,
                               float self_weight, uint32_t pixels, uint32_t channels, float *out, void *stream)
{
    if (!frame || !accum || !wsum || !maxw || !out) return fail(AVD_ERR_NULL_ARGUMENT, "null argument");
    return hip_status(avd::launch_nlm_finish(frame, accum, wsum, maxw, self_weight, pixels, (int)channels, out,
                                             static_cast<hipStream_t>(stream)),
                      "nlm_finish");
}

}  // extern "C"
