/*
 * avd.h -- C ABI of the MI355X-native av-denoise denoising core.
 *
 * Every entry point is plain C: pointers, sizes and POD structs, no C++ or torch types.  A
 * device plane is a pointer into HIP device memory; the caller owns it and lends it for the call.
 *
 * Reference interfaces each declaration replaces (paths relative to the reference checkout):
 *   avd_engine_*        Engine trait          av-denoise-core/src/engine/mod.rs:168-211
 *   avd_engine_create   build_engine          av-denoise-core/src/tune/setup.rs:79-92
 *   avd_engine_spec     EngineSpec            av-denoise-core/src/engine/spec.rs:146-157
 *   avd_geometry        Geometry              av-denoise-core/src/engine/frame.rs:73-97
 *   avd_sample_format   SampleFormat          av-denoise-core/src/engine/plane.rs:8-52
 *   avd_device_plane    DevicePlane           av-denoise-core/src/engine/plane.rs:54-86
 *   avd_nlmeans_options NlmeansAlgorithm      av-denoise-core/src/nlmeans/options.rs:6-106
 *   avd_nl4d_options    Nl4dOptions           av-denoise-core/src/nl4d/options.rs:11-94
 *   avd_motion_search   MotionSearch          av-denoise-core/src/motion/search.rs:5-40
 *   avd_status          Error                 av-denoise-core/src/error.rs:3-20
 *   avd_preset_*        Preset ladder         av-denoise-core/src/options.rs:16-38,
 *                                             av-denoise-core/src/nlmeans/options.rs:58-82
 *   avd_site_*          launch sites          av-denoise-core/src/tune/site.rs:56-157 (the
 *                       kernel-level boundary: collab_search / collab_filter of the SplitSearch
 *                       profile, fused/split.rs:34-281, and the motion snapshot of
 *                       nl4d/snapshot.rs), exposed so a caller can run one pass of one stage.
 *
 * Threading: an engine is driven by one thread at a time (the reference's Engine is Send, not
 * Sync).  Engines on different threads are independent.  Work runs in order on the engine's HIP
 * stream.  Any HIP failure poisons the engine until avd_engine_reset.
 */
#ifndef AVD_AVD_H
#define AVD_AVD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define AVD_ABI_VERSION 1u

/* Error classes, one per reference Error variant (error.rs:3-20), plus HIP failures. */
typedef enum avd_status {
    AVD_OK = 0,
    AVD_ERR_INVALID_OPTIONS = 1,
    AVD_ERR_INVALID_GEOMETRY = 2,
    AVD_ERR_PLANE_MISMATCH = 3,
    AVD_ERR_OUTPUTS_PENDING = 4,
    AVD_ERR_NOTHING_TO_EMIT = 5,
    AVD_ERR_CONTEXT_AFTER_PUSH = 6,
    AVD_ERR_NEEDS_RESET = 7,
    AVD_ERR_GPU = 8,
    AVD_ERR_NULL_ARGUMENT = 9,
    /* The host wire API's own errors (DenoiserError, av-denoise/src/host/mod.rs:36-50). */
    AVD_ERR_QUEUE_FULL = 10,
    AVD_ERR_POISONED = 11
} avd_status;

/* SampleFormat: U8, or U16 holding `depth` significant bits (9..16). */
typedef enum avd_format_kind { AVD_FORMAT_U8 = 0, AVD_FORMAT_U16 = 1, AVD_FORMAT_F32 = 2 } avd_format_kind;

typedef struct avd_sample_format {
    uint32_t kind;  /* avd_format_kind */
    uint32_t depth; /* 8 for U8, 9..16 for U16, ignored for F32 */
} avd_sample_format;

typedef enum avd_chroma_layout {
    AVD_CHROMA_GRAY = 0,
    AVD_CHROMA_420 = 1,
    AVD_CHROMA_422 = 2,
    AVD_CHROMA_444 = 3
} avd_chroma_layout;

typedef struct avd_geometry {
    uint32_t width;  /* luma plane */
    uint32_t height;
    uint32_t chroma; /* avd_chroma_layout */
    avd_sample_format input;
    avd_sample_format output;
    uint32_t chroma_siting; /* carried for callers, never changes the output */
} avd_geometry;

/* A tightly packed plane of whole u32 words in device memory (DevicePlane). */
typedef struct avd_device_plane {
    void *data;      /* hipDeviceptr_t */
    uint64_t bytes;  /* capacity of `data`; must be >= the plane's word-rounded size */
    uint32_t width;
    uint32_t height;
} avd_device_plane;

/* MotionSearch / MotionCompensationMode::Mvtools.  estimation: 0 Auto, 1 Direct, 2 Chained. */
typedef struct avd_motion_search {
    uint32_t blksize;
    uint32_t overlap;
    uint32_t search_radius;
    uint32_t max_search_escalation;
    uint32_t pyramid_levels;
    uint32_t estimation;
    uint32_t refine_radius; /* for Chained */
} avd_motion_search;

/* NlmeansAlgorithm: Fast or Hq, with NlmeansOptions (+ HqParams).  `has_*` flags stand in for
 * the reference's Option<> fields. */
typedef struct avd_nlmeans_options {
    uint32_t hq;                 /* 0 Fast, 1 Hq */
    uint32_t temporal;           /* DenoisingMode: 0 Spacial, 1 Temporal */
    uint32_t temporal_radius;
    uint32_t has_search_radius, search_radius;
    uint32_t has_patch_radius, patch_radius;
    uint32_t has_strength;
    float strength;
    uint32_t has_self_weight;
    float self_weight;
    uint32_t motion_compensation; /* 0 None, 1 Mvtools */
    avd_motion_search motion;
    /* HqParams */
    uint32_t auto_strength;
    uint32_t noise_floor;
    uint32_t has_sigma;
    float sigma;                 /* normalised, 0..1 */
    uint32_t temporal_confidence;
    float thsad_scale;
    float sigma_scale;
} avd_nlmeans_options;

/* Nl4dOptions.  `step` and `group_size` are this library's reference-grid stride and group size;
 * the reference fixes them at 4 and 8 (collab/mod.rs:16-20).  This build runs any stride in 1..8
 * and groups of 8 or 16 members (16: volume grids of 4 frames x 4 volumes, or 2 x 8 at temporal
 * radius 1; stacks of up to 16 otherwise). */
typedef struct avd_nl4d_options {
    avd_motion_search motion;
    uint32_t has_sigma;
    float sigma;
    float sigma_scale;
    float thsad_scale;
    uint32_t refine;
    uint32_t spatial_radius;
    uint32_t has_lambda_ht;
    float lambda_ht;
    float lambda_ht_scale;
    float c_min;
    float kaiser_beta;
    float field_lambda;
    uint32_t noise_map;
    uint32_t pooled_threshold;
    uint32_t grain_data;
    uint32_t temporal_radius;
    uint32_t step;
    uint32_t group_size;
    /* Filter transform form (no counterpart in the reference): 0, the FP32 form, bit-exact with
     * the reference's f32 arithmetic (default); 1, the mixed form, the 2-D DCT, the group Haar and
     * their inverses as f16 matrix products with f32 accumulation (hard threshold only, so
     * pooled_threshold must be 0; output within 0.05 dB PSNR of the FP32 form). */
    uint32_t transform_form;
} avd_nl4d_options;

typedef enum avd_engine_kind { AVD_ENGINE_NLMEANS = 0, AVD_ENGINE_NL4D = 1 } avd_engine_kind;

/* EngineSpec: which engine, the geometry, and the options of each plane group it cleans. */
typedef struct avd_engine_spec {
    uint32_t kind; /* avd_engine_kind */
    avd_geometry geometry;
    uint32_t clean_luma;
    uint32_t clean_chroma;
    avd_nlmeans_options nlm_luma;
    avd_nlmeans_options nlm_chroma;
    avd_nl4d_options nl4d_luma;
    avd_nl4d_options nl4d_chroma;
    void *stream; /* hipStream_t to run on; NULL runs on the default stream */
} avd_engine_spec;

typedef struct avd_engine avd_engine;

/* Window span (WindowSpan, engine/mod.rs:41-60).  edges: 0 Repeat, 1 Shifted. */
typedef struct avd_window_span {
    uint64_t behind;
    uint64_t ahead;
    uint32_t edges;
} avd_window_span;

/* ---- library ---------------------------------------------------------------------------- */

uint32_t avd_abi_version(void);
/* The message of the last error on this thread (empty when none). */
const char *avd_last_error_message(void);

/* ---- option defaults and presets (options.rs:16-38, nlmeans/options.rs:58-82,
 *      nl4d/options.rs:68-94) -------------------------------------------------------------- */

/* preset: 0 veryfast, 1 fast, 2 base, 3 slow, 4 veryslow. */
void avd_motion_search_default(avd_motion_search *out);
void avd_nlmeans_options_default(avd_nlmeans_options *out);
avd_status avd_nlmeans_options_for_preset(uint32_t preset, avd_nlmeans_options *out);
void avd_nl4d_options_default(avd_nl4d_options *out);
avd_status avd_nl4d_options_for_preset(uint32_t preset, avd_nl4d_options *out);
/* Bytes a plane of width x height samples occupies in `format` (SampleFormat::plane_bytes). */
uint64_t avd_plane_bytes(avd_sample_format format, uint32_t width, uint32_t height);

/* ---- engine (Engine trait) ------------------------------------------------------------- */

avd_status avd_engine_create(const avd_engine_spec *spec, avd_engine **out);
void avd_engine_destroy(avd_engine *engine);
/* Ingests one frame (all planes, luma first); *ready = frames ready to emit. */
avd_status avd_engine_push(avd_engine *engine, const avd_device_plane *planes, uint32_t count, uint32_t *ready);
/* Ingests one frame as context only (Engine::push_context). */
avd_status avd_engine_push_context(avd_engine *engine, const avd_device_plane *planes, uint32_t count);
/* Writes the oldest ready frame into planes (Engine::emit_into). */
avd_status avd_engine_emit_into(avd_engine *engine, const avd_device_plane *planes, uint32_t count);
/* Ends the stream; *tail = tail frames to emit (Engine::finish). */
avd_status avd_engine_finish(avd_engine *engine, uint32_t *tail);
/* Abandons the current stream and clears a poisoned engine (Engine::reset). */
avd_status avd_engine_reset(avd_engine *engine);
avd_status avd_engine_window_span(const avd_engine *engine, avd_window_span *out);
avd_status avd_engine_max_held_frames(const avd_engine *engine, uint64_t *out);
/* Waits for every queued launch of the engine. */
avd_status avd_engine_synchronize(avd_engine *engine);
/* Poisons the engine as a GPU failure would (fail_through_guard_for_test, nlmeans/engine.rs:195). */
avd_status avd_engine_fail_for_test(avd_engine *engine);
/* Test hooks with no reference counterpart, for the binned aggregation of NL4D engines: shrink
 * every tile list to `cap` entries (256 up to the engine's own, 768 for groups of 8 and 2048 for
 * 16) so that small frames take the overflow path, and read how many (member, tile) pairs
 * overflowed since (each pass then waits for its filter to read the count). */
avd_status avd_engine_set_tile_cap_for_test(avd_engine *engine, uint32_t cap);
avd_status avd_engine_tile_overflows_for_test(const avd_engine *engine, uint64_t *pairs);

/* Device time per stage (EngineTiming / DeviceTimer, engine/timing.rs:13-200): with sampling on,
 * every launch of a stage is bracketed by HIP events on the engine's stream.  NL4D's grouping
 * search is two stages: AVD_STAGE_COLLAB_CENTRE (the centre search of every reference) and
 * AVD_STAGE_COLLAB_SEARCH (the trajectory search and the grid decision).  `work` counts the
 * stage's algorithmic operations: both search stages 192 per candidate patch scored per
 * channel (64 samples x subtract, multiply, add); AVD_STAGE_MOTION 3 per pixel of every scored
 * block offset (subtract, absolute value, add) on both pyramid levels; AVD_STAGE_COLLAB_FILTER per
 * reference 8 members x (C x (4096 + 576 + 256) + 64 x (C + 1)) (forward and inverse 2-D DCT as
 * 8x8x8 products, three Haar levels of 3 operations, 4 per threshold, one fixed-point add per
 * accumulator); AVD_STAGE_NLM_PAIR, AVD_STAGE_NLM_SINGLE and AVD_STAGE_NLM_WINDOW (5C + 2(2P + 1)
 * + 7) per candidate pixel (squared differences, channel scale, separable box sum, weight and
 * accumulation); pixels for the rest. */
enum {
    AVD_STAGE_INGEST = 0,
    AVD_STAGE_MOTION = 1,
    AVD_STAGE_COLLAB_SEARCH = 2,
    AVD_STAGE_COLLAB_FILTER = 3,
    AVD_STAGE_NORMALISE = 4,
    AVD_STAGE_EGRESS = 5,
    AVD_STAGE_NLM_PAIR = 6,
    AVD_STAGE_NLM_SINGLE = 7,
    AVD_STAGE_NLM_FINISH = 8,
    AVD_STAGE_WARP = 9,
    AVD_STAGE_COLLAB_CENTRE = 10,
    AVD_STAGE_NLM_WINDOW = 11,  /* the batched form's one launch per frame: every window and the finish */
    AVD_STAGE_COUNT = 12
};

typedef struct avd_timing {
    double milliseconds[AVD_STAGE_COUNT];
    uint64_t launches[AVD_STAGE_COUNT];
    uint64_t work[AVD_STAGE_COUNT];
} avd_timing;

/* Samples device time of every frame (every = 1) or of none (every = 0) (sample_device_time). */
avd_status avd_engine_sample_device_time(avd_engine *engine, uint32_t every);
/* Waits for the sampled launches and returns the totals since sampling began (Engine::timing). */
avd_status avd_engine_timing(avd_engine *engine, avd_timing *out);

/* No counterpart in the reference, whose NLMeans launches once per temporal pair, once for the
 * single window and once for the finish (nlmeans/dispatch/mod.rs:665-679): an NLMeans engine
 * runs each frame's windows that way (AVD_NLM_FORM_PLAIN) or batched (AVD_NLM_FORM_BATCHED:
 * every pair, the single window and the finish of a plane group in one launch, the window sums
 * kept on chip).  The outputs are bit-identical.  AVD_NLM_FORM_DEFAULT, the form an engine starts
 * in, batches spatial passes and keeps temporal ones plain (the faster of the two on an MI355X at
 * 640x360 spatial and 1920x1080 temporal).  NL4D engines have one form and ignore the call. */
enum { AVD_NLM_FORM_PLAIN = 0, AVD_NLM_FORM_BATCHED = 1, AVD_NLM_FORM_DEFAULT = 2 };
avd_status avd_engine_set_nlm_form(avd_engine *engine, uint32_t form);

/* ---- host wire API (HostDenoiser, av-denoise/src/host/mod.rs:27-345) ------------------------ */

/* A denoiser over frames held as host bytes: one buffer per plane (Y, then U and V for a chroma
 * layout), each plane `width * height` samples of 1 byte (U8) or 2 bytes little-endian (U16).  It
 * drives an engine of `spec` (whose `stream` is ignored: the denoiser runs on a stream of its own)
 * and keeps at most AVD_HOST_MAX_PENDING readbacks in flight (MAX_PENDING, host/mod.rs:27). */
#define AVD_HOST_MAX_PENDING 2u
typedef struct avd_host_denoiser avd_host_denoiser;
/* Receives one output frame: `count` planes of `lengths[i]` bytes, valid for the call only. */
typedef void (*avd_frame_sink)(void *user, const void *const *planes, const uint64_t *lengths, uint32_t count);

avd_status avd_host_denoiser_create(const avd_engine_spec *spec, avd_host_denoiser **out);
void avd_host_denoiser_destroy(avd_host_denoiser *denoiser);
/* Planes per frame and the wire bytes of each input and output plane (`count` entries each). */
avd_status avd_host_denoiser_plane_lengths(const avd_host_denoiser *denoiser, uint64_t *input, uint64_t *output,
                                           uint32_t count, uint32_t *planes);
/* HostDenoiser::push: uploads one frame and starts denoising the frames it completes.
 * AVD_ERR_QUEUE_FULL once AVD_HOST_MAX_PENDING frames wait to be collected (no state changes; recv,
 * then push the same frame again); AVD_ERR_PLANE_MISMATCH for a wrong plane count or length (no
 * state changes); any other failure poisons the stream (AVD_ERR_POISONED until reset_stream). */
avd_status avd_host_denoiser_push(avd_host_denoiser *denoiser, const void *const *planes, const uint64_t *lengths,
                                  uint32_t count);
/* HostDenoiser::push_priming: one frame of context only (Engine::push_context). */
avd_status avd_host_denoiser_push_priming(avd_host_denoiser *denoiser, const void *const *planes,
                                          const uint64_t *lengths, uint32_t count);
/* HostDenoiser::recv: waits for the oldest frame in flight and copies it into `planes`
 * (`capacities[i]` bytes each); *received = 0 when nothing is in flight. */
avd_status avd_host_denoiser_recv(avd_host_denoiser *denoiser, void *const *planes, const uint64_t *capacities,
                                  uint32_t count, uint32_t *received);
/* HostDenoiser::flush: hands the frames in flight and the temporal tail to `sink` in order; the
 * denoiser then takes a fresh stream. */
avd_status avd_host_denoiser_flush(avd_host_denoiser *denoiser, avd_frame_sink sink, void *user);
/* HostDenoiser::reset_stream: drops the frames in flight and the stream, and clears the poison. */
avd_status avd_host_denoiser_reset_stream(avd_host_denoiser *denoiser);
/* Frames in flight. */
avd_status avd_host_denoiser_pending(const avd_host_denoiser *denoiser, uint32_t *pending);
avd_status avd_host_denoiser_window_span(const avd_host_denoiser *denoiser, avd_window_span *out);

/* ---- launch-site tuner (KernelSelector, tune/selector.rs:109; VariantSite, tune/site.rs:56-157) -- */

/* An engine's tunable launch sites (the NLM windows' rows per wave; the NL4D trajectory search's
 * references per wave) offer candidate launches that compute bit-identical outputs.  With tuning on
 * (the default) the first engine of a shape runs every candidate on its own buffers, drops any whose
 * outputs differ from the baseline's (candidate 0) by a bit, times the rest with HIP events and
 * keeps the fastest for that device, site and shape for the rest of the process.  Off, every site
 * runs its baseline. */
void avd_tuning_set_enabled(uint32_t enabled);
typedef struct avd_tune_entry {
    char site[32];
    char key[96];
    uint32_t device;
    uint32_t choice;
    uint32_t candidates;
    double ms[8]; /* best time of each candidate; negative when its outputs differed */
} avd_tune_entry;
/* Copies up to `capacity` of the tuner's choices so far; *count = how many exist. */
avd_status avd_tuning_entries(avd_tune_entry *out, uint32_t capacity, uint32_t *count);

/* ---- launch sites: one stage of one pass over caller-owned device buffers ----------------- */

/* Plane ingest (engine/kernels.rs:25-63): u32 words of `channels` planes into an interleaved f32
 * slot and its u16 code mirror (codes may be NULL). */
avd_status avd_site_ingest(const void *plane0, const void *plane1, float *slot, uint16_t *codes, uint32_t pixels,
                           uint32_t channels, avd_sample_format format, void *stream);
/* Plane egress (engine/kernels.rs:103-158). */
avd_status avd_site_egress(const float *frame, void *plane0, void *plane1, uint32_t pixels, uint32_t channels,
                           avd_sample_format format, void *stream);

/* Parameters of one NL4D grouping pass (collab_fused's bindings, collab/kernels/fused/mod.rs:170-211). */
typedef struct avd_collab_pass {
    uint32_t width, height, channels;
    uint32_t step;             /* reference-grid stride */
    uint32_t spatial_radius;
    uint32_t temporal_radius;  /* 0 for a single-frame ring */
    uint32_t refine;
    uint32_t group_size;       /* k_max */
    uint32_t blk_step, blksize, blocks_x, blocks_y;
    uint32_t centre_slot;
    uint32_t slots;            /* frames in the ring */
    float c_min;
    float lambda[2];
    float sigma[2];
    float weight_scale;
    float accum_scale;
    float kaiser[8];
    float profile[8];
    uint32_t code_bits;        /* every u16 code is below 2^code_bits; 0 when unknown (up to 16) */
    float pool_ratio;          /* > 0: pooled threshold at lambda * pool_ratio (fused/pooled.rs);
                                  0: the per-coefficient hard threshold */
} avd_collab_pass;

/* search_group over every reference (collab_search, fused/split.rs:34-170): writes 9 words per
 * reference (8 pack_pos_t members and the group word).  `ring_codes` (u16 codes, exact SSD) or,
 * when NULL, `ring_values` (f32, the reference's F32Search distance) is searched.  mv_field holds
 * 2*radius slices of blocks_x*blocks_y (dx,dy) pairs, confidence the matching floats. */
avd_status avd_site_collab_search(const avd_collab_pass *pass, const uint16_t *ring_codes, const float *ring_values,
                                  const uint32_t *neighbour_slots, const int32_t *mv_field, const float *confidence,
                                  uint32_t *members, void *stream);
/* filter_group over every reference (collab_filter, fused/split.rs:181-281): transforms, hard
 * threshold and the fixed-point scatter (aggregate.rs:147-211).  The reference keeps an i32 sum per
 * pixel and channel and an i32 weight sum per pixel; here `aggregate` (slots*px u64) holds, per
 * pixel, weight_sum * 2^32 + first_channel_sum (the sum sign-extended), so one 64-bit atomic add
 * scatters both; both sums stay within i32 by the accumulation scale, so the low word is the sum
 * and the high word plus (sum < 0) the weight sum.  `accum_second` (slots*px i32) holds the second
 * channel's sum when channels == 2 and may be NULL otherwise. */
avd_status avd_site_collab_filter(const avd_collab_pass *pass, const float *ring_values,
                                  const uint32_t *neighbour_slots, const uint32_t *members, uint64_t *aggregate,
                                  int32_t *accum_second, float *group_weight, void *stream);
/* collab_normalise (aggregate.rs:252-290) of one region (aggregate and, for two channels,
 * accum_second), which it leaves zeroed. */
avd_status avd_site_collab_normalise(uint64_t *aggregate_region, int32_t *accum_second_region, float *out,
                                     uint32_t pixels, uint32_t channels, void *stream);

/* Motion of every neighbour slot against the centre (mc_block_match_coarse + mc_block_match_fine
 * with confidence, motion/analyse.rs:175-267) on the lane-0 codes of an interleaved u16 ring.
 * mv_field/confidence as in avd_site_collab_search.  pyramid_levels 1 or 2. */
typedef struct avd_motion_pass {
    uint32_t width, height, channels;
    uint32_t blksize, overlap, search_radius, pyramid_levels;
    uint32_t centre_slot;
    uint32_t neighbours;
    float max_value;
    float sad_floor;
    float thsad;
} avd_motion_pass;

avd_status avd_site_motion(const avd_motion_pass *pass, const uint16_t *ring_codes, const uint32_t *neighbour_slots,
                           int32_t *mv_field, float *confidence, void *stream);

/* run_seeded_refine (motion/analyse.rs:321-367) at the single-level refine geometry
 * (MotionCtx::single_level, motion/context.rs:108-116): the fine block match on level 0 only,
 * refine_radius around each seed, with confidence.  `seeds` holds one (x, y) per block per
 * neighbour, laid out as mv_field, and may be mv_field itself.  pass->search_radius and
 * pyramid_levels are not read. */
avd_status avd_site_motion_refine(const avd_motion_pass *pass, const uint16_t *ring_codes,
                                  const uint32_t *neighbour_slots, const int32_t *seeds, uint32_t refine_radius,
                                  int32_t *mv_field, float *confidence, void *stream);

/* FieldLayout (motion/chroma_map.rs:18-53): a motion field's block grid and the i32 stride
 * between its fields (at least blocks_x * blocks_y * 2). */
typedef struct avd_field_layout {
    uint32_t blocks_x, blocks_y, step, blksize, stride;
} avd_field_layout;

/* map_luma_fields (motion/chroma_map.rs:56-98): `fields` luma motion fields onto the chroma block
 * grid, each chroma vector the per-component lower median of the luma vectors its area covers,
 * divided by the subsampling (shift 0 or 1 per axis) rounding half away from zero. */
avd_status avd_site_map_luma_fields(const avd_field_layout *luma, const int32_t *luma_fields,
                                    const avd_field_layout *chroma, int32_t *chroma_fields, uint32_t shift_x,
                                    uint32_t shift_y, uint32_t fields, void *stream);

/* Parameters of one NLM window launch (nlm_fused_single_window / nlm_fused_pair_accumulate_window,
 * nlmeans/kernels/fused.rs:138-530). */
typedef struct avd_nlm_window {
    uint32_t width, height, channels;
    uint32_t patch_radius, search_radius;
    float h2_inv_norm;
    float noise_offset;        /* pair: flat offset; single: offsets from `lut` */
    uint32_t step, blocks_x, blocks_y; /* confidence block grid */
} avd_nlm_window;

avd_status avd_site_nlm_single(const avd_nlm_window *w, const float *frame, const float *lut, float *accum,
                               float *wsum, float *maxw, void *stream);
avd_status avd_site_nlm_pair(const avd_nlm_window *w, const float *centre, const float *fwd, const float *bwd,
                             const float *conf_fwd, const float *conf_bwd, float *accum, float *wsum, float *maxw,
                             void *stream);
avd_status avd_site_nlm_finish(const float *frame, const float *accum, const float *wsum, const float *maxw,
                               float self_weight, uint32_t pixels, uint32_t channels, float *out, void *stream);

/* Spatial noise measurement: both stages of the Immerkaer estimate of frame `frame` of a ring of
 * f32 [frame][height][width][channels] samples (run_noise_estimate, noise/spatial.rs:93-125, with
 * NoiseCtx's fields; kernels nlm_noise_partial / nlm_noise_reduce, noise/kernels/stats.rs:24-140).
 * `partials` holds avd_noise_partials_len(width, height) floats (two per 32x8 tile); the two
 * per-channel sums of absolute mask responses land in results[2 slot], results[2 slot + 1] (lane 1
 * is 0 for one channel), bit-identical to the reference's f32 sums. */
avd_status avd_site_noise_spatial(const float *ring_values, uint32_t width, uint32_t height, uint32_t channels,
                                  uint32_t frame, float *partials, float *results, uint32_t slot, void *stream);
/* No counterpart in the reference: avd_site_noise_spatial for callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).  It
 * takes no partials workspace and allocates nothing: the partials live in a static device array of
 * AVD_NOISE_SCRATCH_FLOATS floats (2 MiB), one per GPU the library runs on, held until the process
 * ends; frames whose avd_noise_partials_len exceeds it (above about 8192x8192) return
 * AVD_ERR_INVALID_GEOMETRY.  Runs on the current device; `stream` must belong to it.  Calls on one
 * GPU share the array, so calls that overlap on one GPU race.  Results match avd_site_noise_spatial
 * bit for bit. */
#define AVD_NOISE_SCRATCH_FLOATS (1u << 19)
avd_status avd_site_noise_spatial_noalloc(const float *ring_values, uint32_t width, uint32_t height,
                                          uint32_t channels, uint32_t frame, float *results, uint32_t slot,
                                          void *stream);
/* partials_len (noise/spatial.rs:31-34): floats of one frame's partials. */
uint32_t avd_noise_partials_len(uint32_t width, uint32_t height);
/* sigma_from_abs_sum (noise/spatial.rs:134-138): sqrt(pi / 2) sum / (6 (width - 2)(height - 2)),
 * the Immerkaer sigma of a frame-wide sum (in the samples' normalised units).  NaN when width or
 * height is below 3, where the mask has no interior sample. */
float avd_noise_sigma_from_abs_sum(float abs_sum, uint32_t width, uint32_t height);

/* Temporal block statistics: one record per 16x16 block of frame `slot_new` of a ring of f32
 * [frame][height][width][channels] samples, its residual taken against frame `slot_prev`
 * (run_temporal_noise_stats, noise/temporal.rs:206-257; kernel nlm_temporal_noise_stats,
 * noise/kernels/stats.rs:142-604).  `stats` holds avd_noise_temporal_stats_len floats: the blocks
 * row-major, each record 3 channels + 26 floats laid out as temporal_stats_record_len and
 * texture_lane say (noise/temporal.rs:113-131): every channel's sum_d, every sum_d2, the lag-1 sum,
 * channel 0's texture, the moved flag, four quarter records of six fields (sum_d, sum_d2, luma sum,
 * flatness, luma min, luma max; zero when `luma_fields` is 0), then every further channel's
 * texture.  `mv_field` NULL measures the zero vector; otherwise each block reads the whole-pixel
 * vector (x, y) of the motion block covering its origin, from the `mv_step` grid of mv_blocks_x by
 * mv_blocks_y blocks at mv_field[mv_offset..], and follows it when its gain check passes.  The
 * records are bit-identical to the reference's f32 records. */
avd_status avd_site_noise_temporal_stats(const float *ring_values, uint32_t width, uint32_t height, uint32_t channels,
                                         uint32_t slot_new, uint32_t slot_prev, const int32_t *mv_field,
                                         uint32_t mv_offset, uint32_t mv_step, uint32_t mv_blocks_x,
                                         uint32_t mv_blocks_y, uint32_t luma_fields, float *stats, void *stream);
/* zero_temporal_stats_slot (noise/temporal.rs:260-287; kernel nlm_temporal_stats_zero,
 * noise/kernels/stats.rs:662-677): `len` zeroes from `stats`, by a kernel (no memset node). */
avd_status avd_site_noise_temporal_zero(float *stats, uint32_t len, void *stream);
/* temporal_stats_slot_len (noise/temporal.rs:141-144): floats of one frame's records. */
uint32_t avd_noise_temporal_stats_len(uint32_t width, uint32_t height, uint32_t channels);

/* Film-grain measurement of one completed frame (grain_measure, nl4d/kernels/grain.rs:243-513;
 * launched per plane by the grain stage, nl4d/grain/chunk.rs).  Every 8x8 cell of channel `channel`
 * of ring slots `slot_t` / `slot_next` (f32 [slot][height][width][channels]) is judged against the
 * denoised frame `denoised` (f32 [height][width][channels]): source grain is the difference of the
 * two slots along the cell's saved vector (`saved_mv`, `saved_conf`: entry `source_entry` of
 * blocks_x x blocks_y blocks on a `step` grid) over sqrt 2, one-frame grain the cell's high-pass
 * std.  Accepted cells add one count each into `hist` (the frame's 2 x 16 x 64 i32 histograms,
 * source grain then one-frame grain, by brightness bin and std bucket of the 65 `edges`); a
 * one-frame cell bins by the colour grid `luma_means` (3 floats per 16x16 block, means_blocks_x
 * wide) when means_blocks_x > 0 and `has_luma_means` is set.  `partials` receives
 * avd_grain_partials_len floats: per cell the 46 lag sums of its mean-free source grain (LAGS
 * order, nl4d/grain/consts.rs:27), its pixel count and strength group, zeros when rejected.
 * Histogram counts and partials are bit-identical to the reference's. */
avd_status avd_site_grain_measure(const float *ring_values, const float *denoised, uint32_t width, uint32_t height,
                                  uint32_t channels, uint32_t channel, uint32_t slot_t, uint32_t slot_next,
                                  const int32_t *saved_mv, const float *saved_conf, uint32_t blocks_x,
                                  uint32_t blocks_y, uint32_t step, uint32_t source_entry, uint32_t has_source,
                                  const float *luma_means, uint32_t has_luma_means, uint32_t means_blocks_x,
                                  uint32_t luma_band, const float *edges, int32_t *hist, float *partials,
                                  void *stream);
/* Floats of one frame's grain partials: 48 per whole 8x8 cell. */
uint32_t avd_grain_partials_len(uint32_t width, uint32_t height);
/* grain_reduce_partials (nl4d/kernels/grain.rs:515-560; launched per frame by the grain stage,
 * nl4d/grain/export.rs:389-400): each of `cells` partials (48 floats from avd_site_grain_measure)
 * is added into its strength group's record of `chunk` (16 records of 47 floats from
 * `record_offset`: the 46 lag sums and the pixel count), keeping the records' earlier sums.  The
 * sums are bit-identical to the reference's. */
avd_status avd_site_grain_reduce(const float *partials, uint32_t cells, float *chunk, uint32_t record_offset,
                                 void *stream);
/* grain_save_vectors (nl4d/kernels/grain.rs:50-75; launched by nl4d/grain/export.rs:264-290): one
 * neighbour's `blocks` motion vectors (mv_field[mv_offset..], x then y) and confidences
 * (confidence[conf_offset..]) into saved entry `entry` of saved_mv / saved_conf. */
avd_status avd_site_grain_save_vectors(const int32_t *mv_field, const float *confidence, int32_t *saved_mv,
                                       float *saved_conf, uint32_t mv_offset, uint32_t conf_offset, uint32_t entry,
                                       uint32_t blocks, void *stream);
/* grain_project (nl4d/kernels/grain_persistence.rs:54-153): per whole 8x8 cell of channel `channel`
 * of ring slot `frame_slot`, the 8 projections onto `patterns` (PATTERNS, nl4d/grain/consts.rs:76-85,
 * as two u32 words each, the low word first) and the flat flag of the denoised cell, into history
 * slot `history_slot` (of 9): 9 floats per cell at (history_slot * cells + cell) * 9. */
avd_status avd_site_grain_project(const float *ring_values, const float *denoised, const uint32_t *patterns,
                                  float *history, uint32_t width, uint32_t height, uint32_t channels,
                                  uint32_t channel, uint32_t frame_slot, uint32_t history_slot, uint32_t luma_band,
                                  void *stream);
/* Floats of grain_persist's partials for `cells` cells at `lags` lags: 3 per block of 1024 cells per lag. */
uint32_t avd_grain_persist_partials_len(uint32_t cells, uint32_t lags);
/* grain_persist then grain_persist_reduce (nl4d/kernels/grain_persistence.rs:155-279; launched by
 * nl4d/grain/persistence.rs:245-300): history slot `now_slot` against each of the `lags` (at most 8)
 * slots before it, per lag the cell count, auto sum and cross sum of the cells flat in both, added
 * into record[record_offset + 3 lag ..]; `partials` holds avd_grain_persist_partials_len floats.
 * Bit-identical to the reference's f32 sums. */
avd_status avd_site_grain_persist(const float *history, float *partials, float *record, uint32_t record_offset,
                                  uint32_t now_slot, uint32_t cells, uint32_t lags, void *stream);
/* Host side of the grain table (nl4d/grain/fit.rs), f64 arithmetic in the reference's order.
 * avd_grain_bucket_edges: the 65 std bucket edges, log-spaced from 0.05/255 to 32/255
 * (bucket_edges, fit.rs:28-38).  avd_grain_hist_median: the median std of a 64-bucket histogram,
 * interpolated geometrically inside its bucket; *found = 0 for an empty histogram (hist_median,
 * fit.rs:51-74).  avd_grain_finest_scaling_shift: the finest shift 11..9 keeping `peak` within a
 * byte, else 8 (fit.rs:335-346).  avd_grain_scaling_points: each point's AV1 scaling value
 * sigma * 255 / sigma_template at the finest shift for the largest, rounded and capped at 255
 * (scaling_points, fit.rs:310-360).  avd_grain_quantise_ar: the 24 weights' AV1 integers at the
 * finest AR shift 9..7 that fits a signed byte, else 6, clamped (quantise_ar, fit.rs:272-305). */
void avd_grain_bucket_edges(float *edges);
double avd_grain_hist_median(const uint32_t *counts, const float *edges, int32_t *found);
uint32_t avd_grain_finest_scaling_shift(double peak);
avd_status avd_grain_scaling_points(const uint8_t *luma, const double *sigma, uint32_t count, double sigma_template,
                                    uint8_t *scaled, uint32_t *shift);
void avd_grain_quantise_ar(const double *weights, int32_t *quantised, uint32_t *shift);
/* pooled_record (nl4d/grain/segment.rs:170-180, 286-311): the strength groups whose std range
 * (edges[4 g] .. edges[4 g + 4]) meets half to twice the median of the source histogram (16 x 64
 * counts, brightness rows merged) summed into one 47-value record (46 lag sums, pixel count) from
 * `group_records` (16 records of 47); returns 1, or 0 for an empty histogram or fewer than 50000
 * pixels.  take_out_fine_noise (segment.rs:313-318): min(fine_noise, 0.95) * floor of the zero lag
 * taken off a corrected record. */
int32_t avd_grain_pooled_record(const uint32_t *source_hist, const double *group_records, const float *edges,
                                double *record);
void avd_grain_take_out_fine_noise(double *corrected, double floor, double fine_noise);
/* floor_fraction (nl4d/grain/fit.rs:218-270): the share of the AR filter's power spectrum (24
 * weights in AR_OFFSETS order, 64 x 64 frequencies) that a flat white floor accounts for, its lowest
 * point over its mean, clamped to 0..1 (0 when not finite). */
double avd_grain_floor_fraction(const double *weights);
/* The AR(3) grain fit in f64, on the GPU (undo_mean_removal, yule_walker, solve and quantise_ar,
 * nl4d/grain/fit.rs:120-300; called per pooled record by nl4d/grain/segment.rs:262-265 and
 * nl4d/grain/chroma.rs:159-160).  Each of `count` records holds 47 f64 values (the 46 lag sums in
 * LAGS order and the pixel count); per record it writes the record with the cell-mean bias undone
 * (corrected[47]), the 24 Yule-Walker weights in AR_OFFSETS order (weights[24]), their AV1
 * integers (quantised[24], at the finest shift 6..9 that fits a signed byte, clamped) and the shift
 * (shifts[1]), and solved[1] = 1, or 0 with zero weights where the reference returns None (an
 * empty record, a pivot below 1e-18 or a non-finite solution).  Bit-identical to the reference's
 * f64 arithmetic. */
avd_status avd_site_grain_fit(const double *records, uint32_t count, double *corrected, double *weights,
                              int32_t *quantised, uint32_t *shifts, int32_t *solved, void *stream);

#ifdef __cplusplus
}
#endif

#endif /* AVD_AVD_H */
