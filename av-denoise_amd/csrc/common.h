// Shared declarations of the HIP kernels' host launchers and the small device helpers they use.
// Every launcher enqueues on `stream` and returns the launch's hipError_t.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "avd/avd.h"

namespace avd {

constexpr int kPatch = 8;

inline uint32_t div_up(uint32_t a, uint32_t b) { return (a + b - 1) / b; }

// ---- io.hip -----------------------------------------------------------------------------
hipError_t launch_ingest(const uint32_t *p0, const uint32_t *p1, float *slot, uint16_t *codes, uint32_t pixels,
                         int channels, int samples_per_word, float max_value, hipStream_t s);
hipError_t launch_egress(const float *frame, uint32_t *p0, uint32_t *p1, uint32_t pixels, int channels,
                         int samples_per_word, float max_value, hipStream_t s);
hipError_t launch_copy_words(const uint32_t *src, uint32_t *dst, uint64_t words, hipStream_t s);
hipError_t launch_fill_u32(void *dst, uint32_t value, uint64_t words, hipStream_t s);

// ---- nlm.hip ----------------------------------------------------------------------------
struct NlmWindow {
    int W, H, C, P, S;
    float h2_log2e;      // h2_inv_norm * log2(e): the weight is exp2(-max(d - off, 0) * h2_log2e)
    float offset;        // pair: the flat noise offset
    int step, blocks_x, blocks_y;
};
// `variant` picks the rows each wave walks (0: 16, 1: 8, 2: 4); every variant gives bit-identical
// sums, so the tuner may choose any (kNlmVariants of them).
constexpr int kNlmVariants = 3;
int nlm_variant_rows(int variant);
hipError_t launch_nlm_single(const NlmWindow &w, const float *frame, const float *lut, float *accum, float *wsum,
                             float *maxw, hipStream_t s, int variant = 0);
hipError_t launch_nlm_pair(const NlmWindow &w, const float *centre, const float *fwd, const float *bwd,
                           const float *conf_f, const float *conf_b, float *accum, float *wsum, float *maxw,
                           hipStream_t s, int variant = 0);
hipError_t launch_nlm_finish(const float *frame, const float *accum, const float *wsum, const float *maxw,
                             float self_weight, uint32_t pixels, int C, float *out, hipStream_t s);
bool nlm_supported(int C, int P, int S);
// The temporal pairs of one window, oldest first (q = -r .. -1): the forward frame at -q (the
// older neighbour), the backward one at +q, and their block confidences (null without).
constexpr int kNlmMaxPairs = 8;
struct NlmFrames {
    int pairs;
    const float *fwd[kNlmMaxPairs], *bwd[kNlmMaxPairs];
    const float *conf_f[kNlmMaxPairs], *conf_b[kNlmMaxPairs];
};
// The batched form: every pair, the single window and the finish of one frame in one launch, the
// window sums kept on chip; bit-identical to launch_nlm_pair (each pair) + launch_nlm_single +
// launch_nlm_finish on zeroed accumulators.  `variant` picks 8 (0) or 4 (1) rows per wave.
constexpr int kNlmBatchedVariants = 2;
hipError_t launch_nlm_window(const NlmWindow &w, const NlmFrames &f, const float *centre, const float *lut,
                             float self_weight, float *out, hipStream_t s, int variant = 0);

// ---- motion.hip -------------------------------------------------------------------------
struct MotionGeom {
    int W, H, C;
    int blksize, step, radius, levels;
    int blocks_x, blocks_y;
};
// Level-1 sums of lane 0 of one u16 ring slot.
hipError_t launch_downscale(const uint16_t *codes_slot, int W, int H, int C, int32_t *level1, hipStream_t s);
// Coarse + fine search of `neighbours` slots against `centre`; pyr1 holds one level-1 plane per slot.
hipError_t launch_motion(const MotionGeom &g, const uint16_t *codes, const int32_t *pyr1, int centre_slot,
                         const uint32_t *nslots_dev, int neighbours, int32_t *mv, float *conf, float max_value,
                         float sad_floor, float thsad, hipStream_t s);
hipError_t launch_warp(const float *src, float *dst, const int32_t *mv, int W, int H, int C, int step,
                       int blocks_x, int blocks_y, hipStream_t s);
// The fine pass alone at `radius` around per-block seeds (`seeds`, laid out as `mv`), on level 0:
// the chroma pass's refine of the luma pass's mapped vectors.  `seeds` may be `mv`.
hipError_t launch_motion_refine(const MotionGeom &g, const uint16_t *codes, int centre_slot, const uint32_t *nslots_dev,
                                int neighbours, const int32_t *seeds, int32_t *mv, float *conf, float max_value,
                                float sad_floor, float thsad, int radius, hipStream_t s);
// Two motion block grids, luma and chroma, with the i32 stride between fields of each and the
// chroma subsampling as shifts per axis.
struct FieldMap {
    int luma_bx, luma_by, luma_step, luma_blk, luma_stride;
    int chroma_bx, chroma_by, chroma_step, chroma_blk, chroma_stride;
    int shift_x, shift_y;
};
// `fields` luma motion fields onto the chroma block grid.
hipError_t launch_map_luma_fields(const FieldMap &m, const int32_t *luma, int32_t *chroma, int fields, hipStream_t s);

// ---- collab.hip -------------------------------------------------------------------------
// `scored_count`, when non-null, points at kScoreSlots counters kScoreStride apart; their sum gains
// the candidate patches the launch scores times the channels.
constexpr int kScoreSlots = 64, kScoreStride = 16;
// Scratch of the matrix-core centre search: the centre frame as f16 patterns of 1024 + code
// (`plane`, W*H*C u16; `plane_odd`, the same shifted by one code, W*H u16, one channel only) and
// the 8x8 patch energy of every position (`energy`, W*H u32), and the references searched again
// with exact keys (`redo`: a count, then the references).
struct CentreScratch {
    uint16_t *plane, *plane_odd;
    uint32_t *energy, *redo;
};
size_t centre_scratch_bytes(uint32_t W, uint32_t H, uint32_t C);
CentreScratch centre_scratch_at(void *base, uint32_t W, uint32_t H, uint32_t C);
// Whether a pass takes the matrix-core centre search (integer codes below 1024 and a spatial
// radius whose two grid rows of candidates fit 32 rows); it then needs a CentreScratch.
bool centre_on_matrix_cores(const avd_collab_pass &p);
// The search as two launches (the engine times them apart): the centre search of every
// reference, then the trajectory search and grid decision over integer codes.
hipError_t launch_collab_centre(const avd_collab_pass &p, const uint16_t *codes, uint32_t *members, hipStream_t s,
                                const CentreScratch *centre = nullptr);
hipError_t launch_collab_trajectory(const avd_collab_pass &p, const uint16_t *codes, const uint32_t *nslots,
                                    const int32_t *mv, const float *conf, uint32_t *members, hipStream_t s,
                                    unsigned long long *scored_count = nullptr);
hipError_t launch_collab_search(const avd_collab_pass &p, const uint16_t *codes, const float *values,
                                const uint32_t *nslots, const int32_t *mv, const float *conf, uint32_t *members,
                                hipStream_t s, unsigned long long *scored_count = nullptr,
                                const CentreScratch *centre = nullptr);
// Accumulators of the fixed-point scatter: `aggregate` holds per pixel (weight sum << 32) plus the
// first channel's sum, one 64-bit atomic add per pixel; `accum_second` the second channel's sum.
hipError_t launch_collab_filter(const avd_collab_pass &p, const float *values, const uint32_t *nslots,
                                const uint32_t *members, unsigned long long *aggregate, int32_t *accum_second,
                                float *group_weight, hipStream_t s);
// Spatial noise measurement (noise.hip): the Immerkaer partials of frame `frame` of a ring of f32
// [frame][H][W][C] samples, then their per-channel totals into row `slot` of `results`.
uint32_t noise_partials_len(uint32_t width, uint32_t height);
hipError_t launch_noise_spatial(const float *ring, uint32_t width, uint32_t height, uint32_t channels,
                                uint32_t frame, float *partials, float *results, uint32_t slot, hipStream_t s);
// The current device's copy of the static partials array of the allocation-free entry point
// (kNoiseScratchFloats floats), resolved on the first call per device.
constexpr uint32_t kNoiseScratchFloats = AVD_NOISE_SCRATCH_FLOATS;
hipError_t noise_scratch_address(float **out);
// Temporal block statistics (noise.hip): one record of 3 C + 26 floats per 16x16 block of frame
// slot_new against slot_prev of the same ring; mv (may be null) holds whole-pixel vectors of an
// mv_step grid of mv_bx x mv_by blocks from mv[mv_offset].
uint32_t temporal_stats_len(uint32_t width, uint32_t height, uint32_t channels);
hipError_t launch_temporal_stats(const float *ring, uint32_t width, uint32_t height, uint32_t channels,
                                 uint32_t slot_new, uint32_t slot_prev, const int32_t *mv, uint32_t mv_offset,
                                 uint32_t mv_step, uint32_t mv_bx, uint32_t mv_by, bool luma_fields, float *stats,
                                 hipStream_t s);
hipError_t launch_temporal_zero(float *stats, uint32_t len, hipStream_t s);
// Film-grain measurement (grain.hip): every 8x8 cell of channel `channel` of ring slots slot_t /
// slot_next (f32 [slot][H][W][C]) against the denoised frame (f32 [H][W][C]); counts into the
// frame's 2 x 16 x 64 i32 histograms, one 48-float partial per cell.
struct GrainMeasure {
    const float *ring, *denoised;
    const int32_t *saved_mv;
    const float *saved_conf, *edges, *luma_means;
    int32_t *hist;
    float *partials;
    uint32_t width, height, channels, channel, slot_t, slot_next;
    uint32_t blocks_x, blocks_y, step, source_entry, means_blocks_x;
    bool has_source, has_luma_means, luma_band;
};
uint32_t grain_partials_len(uint32_t width, uint32_t height);
hipError_t launch_grain_measure(const GrainMeasure &m, hipStream_t s);
// Every cell's partial added into its strength group's 47-float record of `chunk` (16 records from
// record_offset, earlier sums kept).
hipError_t launch_grain_reduce(const float *partials, uint32_t cells, float *chunk, uint32_t record_offset,
                               hipStream_t s);
// The f64 AR(3) fit of `count` 47-value records: corrected records, 24 weights, their AV1 integers
// and shift, and whether each system solved.
// Grain persistence: the vector save, the per-cell projections into a history slot, and the
// lag-wise persistence sums of the newest slot against the earlier ones, added into `record`.
hipError_t launch_grain_save_vectors(const int32_t *mv_field, const float *confidence, int32_t *saved_mv,
                                    float *saved_conf, uint32_t mv_offset, uint32_t conf_offset, uint32_t entry,
                                    uint32_t blocks, hipStream_t s);
hipError_t launch_grain_project(const float *ring, const float *out, const uint32_t *patterns, float *history,
                                uint32_t width, uint32_t height, uint32_t channels, uint32_t channel,
                                uint32_t frame_slot, uint32_t history_slot, bool luma_band, hipStream_t s);
uint32_t grain_persist_blocks(uint32_t cells);
hipError_t launch_grain_persist(const float *history, float *partials, float *record, uint32_t record_offset,
                                uint32_t now_slot, uint32_t cells, uint32_t lags, hipStream_t s);
hipError_t launch_grain_fit(const double *records, uint32_t count, double *corrected, double *weights,
                            int32_t *quantised, uint32_t *shifts, int32_t *solved, hipStream_t s);
hipError_t launch_collab_normalise(unsigned long long *aggregate, int32_t *accum_second, float *out, uint32_t pixels,
                                   int C, hipStream_t s);

// Binned aggregation: the filter stages each member's fixed-point values (`stage`, [ref][member][c]
// [64] i32) and enters the member in the list of every 32x32 tile of its ring slot that its patch
// overlaps (`count` / `list`: `cap` entries per tile; pairs past that go to `overflow`, up to
// refs * members * 4 pairs); each tile then sums its members in LDS and adds to the accumulators
// once.  Member ids are (reference << log2 members) | member.
constexpr int kTileShift = 5;
// Entries per tile list: groups of 16 put twice the members of groups of 8 in each tile.
inline uint32_t tile_cap_for(uint32_t group_size) { return group_size == 16 ? 2048u : 768u; }
constexpr uint32_t kTileCapMin = 256;  // the gather reads a list's first 256 entries with its count
struct TileBins {
    int32_t *stage;
    uint32_t *count, *list, *overflow, *overflow_count;
    uint32_t tiles_x, per_slot;  // tiles per row, tiles per ring slot
    uint32_t cap;                // entries per tile list (at least kTileCapMin)
};
hipError_t launch_collab_filter_binned(const avd_collab_pass &p, const float *values, const uint32_t *nslots,
                                       const uint32_t *members, const TileBins &tiles,
                                       unsigned long long *aggregate, int32_t *accum_second, float *group_weight,
                                       hipStream_t s, bool mixed = false, const uint16_t *codes = nullptr);

}  // namespace avd
