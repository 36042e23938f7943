// Host-side engine types (see engine.cpp).
#pragma once

#include <memory>
#include <string>
#include <vector>

#include "common.h"

namespace avd {

// An error carrying the C ABI status it maps to.
struct Failure {
    avd_status status;
    std::string message;
    Failure(avd_status s, std::string m) : status(s), message(std::move(m)) {}
};

void check_hip(hipError_t e, const char *what);

// Owned device memory.
struct DeviceBuffer {
    void *ptr = nullptr;
    size_t bytes = 0;
    DeviceBuffer() = default;
    DeviceBuffer(const DeviceBuffer &) = delete;
    DeviceBuffer &operator=(const DeviceBuffer &) = delete;
    ~DeviceBuffer();
    void allocate(size_t n);
    void release();
    template <typename T>
    T *as() const
    {
        return static_cast<T *>(ptr);
    }
};

class LumaMotionShare;

uint32_t samples_per_word(avd_sample_format f);
float max_value(avd_sample_format f);
uint64_t plane_bytes(avd_sample_format f, uint64_t pixels);
void validate_format(avd_sample_format f);

float h2_inv_norm(float strength, uint32_t patch_radius);
float noise_offset(float sigma, int channels, uint32_t patch_radius);
float sad_noise_floor(uint32_t blksize, float sigma);
float thsad(uint32_t blksize, float scale);
void kaiser_window(float beta, float out[8]);
float cross_frame_accum_scale(uint32_t spatial_radius, uint32_t temporal_radius, uint32_t step,
                              uint32_t group_size);
float weight_scale(float sigma, float g_max);

// The 2r+1 ingested frames of one plane group: f32 values, their u16 codes and, for motion, the
// level-1 sums of lane 0.
struct Ring {
    uint32_t W = 0, H = 0, N = 0;
    int C = 1;
    size_t px = 0, l1_px = 0;
    DeviceBuffer values, codes, pyr1;
    std::vector<uint32_t> slot_frames;
    uint32_t head = 0, loaded = 0, real = 0;

    void init(uint32_t W, uint32_t H, int C, uint32_t slots, bool pyramid);
    void reset();
    uint32_t phys(int logical) const;
    bool ready() const { return loaded >= N; }
    void write(const avd_device_plane *planes, avd_sample_format f, hipStream_t s);
    void duplicate_last(hipStream_t s);
    float *value_slot(uint32_t s) const { return values.as<float>() + (size_t)s * px * C; }
    uint16_t *code_slot(uint32_t s) const { return codes.as<uint16_t>() + (size_t)s * px * C; }
    int32_t *l1_slot(uint32_t s) const { return pyr1.as<int32_t>() + (size_t)s * l1_px; }
    // Device table of neighbour slots for every (ring start, centre) pair: entry
    // (start * N + centre) holds the N-1 physical slots of the other logical positions, oldest
    // first, so a pass binds its neighbours without a host-to-device copy.
    DeviceBuffer neighbour_table;
    const uint32_t *neighbours_of(uint32_t centre_logical) const
    {
        return neighbour_table.as<uint32_t>() + ((size_t)(head % N) * N + centre_logical) * (N > 1 ? N - 1 : 1);
    }
    const float *value_ptr() const { return values.as<float>(); }
    const uint16_t *code_ptr() const { return codes.as<uint16_t>(); }
    const int32_t *l1_ptr() const { return pyr1.as<int32_t>(); }
};

// Device time per stage, bracketed by HIP events on the engine's stream while sampling is on
// (the reference's DeviceTimer, engine/timing.rs:13-200).  Passes share their engine's timer.
class StageTimer {
  public:
    ~StageTimer();
    void set_sampling(bool on, hipStream_t s);
    bool on() const { return on_; }
    // Records the start of a launch of `stage` and returns its handle (or -1 when not sampling).
    int begin(int stage, hipStream_t s);
    void end(int handle, uint64_t work, hipStream_t s);
    void add_work(int stage, uint64_t work) { work_[stage] += work; }
    // Waits for every pending launch and folds it into the totals.
    void harvest();
    avd_timing totals() const;
    unsigned long long *scored_counter() const { return on_ ? counter_.as<unsigned long long>() : nullptr; }
    void fold_counter(hipStream_t s);

  private:
    struct Pending {
        int stage;
        hipEvent_t start, stop;
        uint64_t work;
    };
    bool on_ = false;
    std::vector<Pending> pending_;
    double ms_[AVD_STAGE_COUNT] = {};
    uint64_t launches_[AVD_STAGE_COUNT] = {};
    uint64_t work_[AVD_STAGE_COUNT] = {};
    DeviceBuffer counter_;
};

// One plane group's denoiser (the reference's PassEngine + FramePass).
class Pass {
  public:
    virtual ~Pass() = default;
    virtual uint32_t push(const avd_device_plane *planes) = 0;
    virtual void push_context(const avd_device_plane *planes) = 0;
    virtual void emit(const avd_device_plane *planes) = 0;
    virtual uint32_t finish() = 0;
    virtual void reset_stream() = 0;
    virtual size_t max_held() const = 0;
    virtual bool has_pending() const = 0;
    uint32_t emitted_frame() const { return emitted_frame_; }
    void set_timer(StageTimer *t) { timer_ = t; }
    // The motion search the pass runs (none for a pass that searches no motion), and its block grid.
    virtual bool searches_motion() const = 0;
    virtual uint32_t temporal_radius() const = 0;
    virtual const avd_motion_search &motion_search() const = 0;
    virtual const MotionGeom &motion_geom() const = 0;
    // The luma pass writes its motion into `share`; the chroma pass (`reads`) refines it instead of
    // searching its own, so it drops its coarse pyramid level.
    virtual void set_motion_share(LumaMotionShare *share, bool reads)
    {
        share_ = share;
        reads_motion_ = reads;
    }
    // The stream the pass enqueues on (the engine moves the chroma pass between its own stream and
    // the caller's; the engine's fork and join order the switch).
    void set_stream(hipStream_t s) { stream_ = s; }

  protected:
    explicit Pass(hipStream_t s) : stream_(s) {}
    // The window's motion of `neighbours` slots against `centre` into mv/conf: the pass's own
    // search, written into the share when the pass writes one (and `shares`), or, on the pass that
    // reads it, the refine of the luma pass's mapped vectors.
    void search_motion(const MotionGeom &g, const Ring &ring, uint32_t centre, const uint32_t *nslots, int neighbours,
                       int32_t *mv, float *conf, float mx, float sad_floor, float th, bool shares);
    hipStream_t stream_;
    // Runs `launch` (returning hipError_t) as one timed launch of `stage`.
    template <typename F>
    void timed(int stage, uint64_t work, hipStream_t s, const char *what, F &&launch)
    {
        const int h = timer_ ? timer_->begin(stage, s) : -1;
        check_hip(launch(), what);
        if (timer_) timer_->end(h, work, s);
    }
    uint32_t emitted_frame_ = 0;
    StageTimer *timer_ = nullptr;
    LumaMotionShare *share_ = nullptr;
    bool reads_motion_ = false;
};

// The luma pass's motion, mapped onto the chroma block grid for the chroma pass of the same centre
// frame (the reference's LumaMotionRing, motion/luma_ring.rs:10-150).  One slot per window frame,
// chosen and tagged by the centre's stream frame: luma never runs more passes ahead of chroma than
// that, since the engine runs a call's luma passes, then its chroma passes, and each call's end
// orders the chroma stream before the next call's luma work.  The luma stream maps into a slot and
// records its event; the chroma stream waits for that event before it reads the slot.
class LumaMotionShare {
  public:
    LumaMotionShare(const MotionGeom &luma, const MotionGeom &chroma, int shift_x, int shift_y, uint32_t radius);
    ~LumaMotionShare();
    LumaMotionShare(const LumaMotionShare &) = delete;
    LumaMotionShare &operator=(const LumaMotionShare &) = delete;
    // Maps the 2r window fields `luma_mv` of centre stream frame `frame` into its slot, on `s`.
    hipError_t write(uint32_t frame, const int32_t *luma_mv, hipStream_t s);
    // The mapped fields of centre stream frame `frame`, once `s` has waited for them; fails when
    // the slot holds another frame.
    const int32_t *read(uint32_t frame, hipStream_t s);
    void clear();

  private:
    FieldMap map_{};
    uint32_t slots_ = 0, fields_ = 0;
    DeviceBuffer buf_;
    std::vector<hipEvent_t> ready_;
    std::vector<int64_t> tags_;
};

// How far the chroma pass searches around each mapped luma vector (MAPPED_REFINE_RADIUS,
// nlmeans/dispatch/motion.rs:42-43).
constexpr int kMappedRefineRadius = 1;

// Whether two passes run the same motion search (the reference compares MotionCompensationMode and
// the temporal radius, engine/spec.rs:320-341).
bool same_motion_search(const avd_motion_search &a, const avd_motion_search &b);

class NlmPass : public Pass {
  public:
    NlmPass(const avd_nlmeans_options &o, uint32_t W, uint32_t H, int C, avd_sample_format in, avd_sample_format out,
            hipStream_t s);
    uint32_t push(const avd_device_plane *planes) override;
    void push_context(const avd_device_plane *planes) override;
    void emit(const avd_device_plane *planes) override;
    uint32_t finish() override;
    void reset_stream() override;
    size_t max_held() const override;
    bool has_pending() const override;
    bool searches_motion() const override { return motion_; }
    uint32_t temporal_radius() const override { return r_; }
    // The batched form (all windows of a frame and the finish in one launch) or the plain one
    // (one launch per pair, the single window and the finish); the outputs are bit-identical.
    // By default a spatial pass (one window) is batched and a temporal one plain.
    void set_batched(bool on) { batched_ = on; }
    const avd_motion_search &motion_search() const override { return motion_opts_; }
    const MotionGeom &motion_geom() const override { return mg_; }
    void set_motion_share(LumaMotionShare *share, bool reads) override;

  private:
    void ingest(const avd_device_plane *planes);
    void denoise();
    void window_motion(uint32_t centre, const uint32_t *nslots, int neighbours, float mx);
    NlmWindow window() const;
    void tune_windows();
    void egress(const float *frame, const avd_device_plane *planes);

    uint32_t W_, H_;
    int C_;
    avd_sample_format in_, out_;
    uint32_t r_ = 0, S_ = 2, P_ = 4;
    bool hq_ = false, has_sigma_ = false, motion_ = false, confidence_ = false, conf_only_ = false, pushed_ = false;
    float sigma_ = 0.0f, self_weight_ = 1.0f, h2_ = 0.0f, offset_ = 0.0f, thsad_scale_ = 1.0f;
    avd_motion_search motion_opts_{};
    MotionGeom mg_{};
    Ring ring_;
    DeviceBuffer accum_, wsum_, maxw_, lut_, comp_, mv_, conf_;
    DeviceBuffer outputs_[2];
    bool ready_ = false;
    int ready_output_ = 0, next_output_ = 0;
    int variant_single_ = 0, variant_pair_ = 0;  // the tuner's rows per wave of the window sites
    int variant_window_ = 0;                     // and of the batched form's window
    bool batched_ = false;                       // every window of a frame in one launch
    uint32_t tail_remaining_ = 0;
};

class Nl4dPass : public Pass {
  public:
    Nl4dPass(const avd_nl4d_options &o, uint32_t W, uint32_t H, int C, avd_sample_format in, avd_sample_format out,
             hipStream_t s);
    uint32_t push(const avd_device_plane *planes) override;
    void push_context(const avd_device_plane *planes) override;
    void emit(const avd_device_plane *planes) override;
    uint32_t finish() override;
    void reset_stream() override;
    size_t max_held() const override;
    bool has_pending() const override;
    bool searches_motion() const override { return true; }
    uint32_t temporal_radius() const override { return r_; }
    const avd_motion_search &motion_search() const override { return motion_opts_; }
    const MotionGeom &motion_geom() const override { return mg_; }
    void set_motion_share(LumaMotionShare *share, bool reads) override;
    // Test hooks: shrink the tile lists so the overflow path runs, and count its pairs.
    void set_tile_cap_for_test(uint32_t cap);
    uint64_t overflow_pairs_seen() const { return overflow_seen_; }

  private:
    enum class Clear { Whole, Newest, None };
    uint32_t submit();
    void run_pass(uint32_t centre_logical, Clear clear);
    void zero_region(uint32_t slot);

    uint32_t W_, H_;
    int C_;
    avd_sample_format in_, out_;
    uint32_t r_ = 1, N_ = 3;
    float sigma_ = 0.0f, thsad_scale_ = 1.0f;
    avd_motion_search motion_opts_{};
    MotionGeom mg_{};
    avd_collab_pass pass_{};
    size_t refs_ = 0;
    Ring ring_;
    // Per ring slot: aggregate_ holds (weight sum << 32) + first-channel sum per pixel, accum_second_
    // the second channel's sum (chroma passes only).
    DeviceBuffer aggregate_, accum_second_, members_, group_weight_, output_, mv_, conf_;
    // Binned aggregation: staged member values, per-tile counts / offsets / cursors, member lists.
    DeviceBuffer stage_, tile_count_, tile_list_, tile_overflow_;
    TileBins bins_{};
    // The matrix-core centre search's f16 centre frame and patch energies (when the pass takes it).
    DeviceBuffer centre_buf_;
    CentreScratch centre_{};
    uint64_t centre_ops_ = 0;  // algorithmic operations of one centre search launch
    bool mixed_ = false;       // the mixed-precision (f16 matrix) transform form
    std::vector<uint32_t> pending_, tail_;
    std::vector<bool> zeroed_;
    uint32_t passes_ = 0;
    bool continuation_ = false, tail_emitting_ = false;
    bool watch_overflow_ = false;
    uint64_t overflow_seen_ = 0;
};

class Engine {
  public:
    explicit Engine(const avd_engine_spec &spec);
    ~Engine();
    uint32_t push(const avd_device_plane *planes, uint32_t count);
    void push_context(const avd_device_plane *planes, uint32_t count);
    void emit(const avd_device_plane *planes, uint32_t count);
    uint32_t finish();
    void reset();
    void synchronize();
    size_t max_held_frames() const;
    avd_window_span window_span() const;
    void fail_for_test();
    void set_tile_cap_for_test(uint32_t cap);
    uint64_t tile_overflows_for_test() const;
    void sample_device_time(uint32_t every);
    avd_timing timing();
    // NLMeans engines: the plain (0), batched (1) or default (2) launch form (NL4D engines have one
    // form).
    void set_nlm_form(uint32_t form);

  private:
    Pass &lead();
    const Pass &lead() const;
    void check_usable() const;
    void check_planes_for(const avd_device_plane *planes, uint32_t count, avd_sample_format f) const;
    void store_uncleaned(const avd_device_plane *planes, uint32_t frame);
    void load_uncleaned(const avd_device_plane *planes, uint32_t frame);
    void restart();
    template <typename F>
    auto guarded(F &&f) -> decltype(f());
    // With both plane groups cleaned, the chroma pass runs on its own stream (side_): each call
    // forks it from the caller's stream at its start and joins it back at its end, so the caller
    // sees the engine's work in order on one stream while the two passes overlap.
    void fork();
    void join();

    avd_engine_spec spec_;
    hipStream_t stream_ = nullptr, side_ = nullptr;
    hipEvent_t fork_ev_ = nullptr, join_ev_ = nullptr;
    StageTimer timer_;
    std::unique_ptr<LumaMotionShare> share_;  // outlives the passes that point at it
    std::unique_ptr<Pass> luma_, chroma_;
    uint32_t planes_ = 1, chroma_w_ = 0, chroma_h_ = 0;
    DeviceBuffer passthrough_;
    uint32_t pt_slots_ = 0, pt_planes_ = 0;
    uint64_t pt_words_ = 0;
    uint32_t next_frame_ = 0, tail_left_ = 0;
    bool pushed_ = false, poisoned_ = false;
};

}  // namespace avd
