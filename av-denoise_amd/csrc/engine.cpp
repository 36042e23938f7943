// Host engine: option resolution, plane-group passes and the Engine push/emit/finish protocol.
//
// Mirrors av-denoise-core's host side: EngineSpec validation (engine/spec.rs:291-344), the planar
// engine that runs a luma pass and a chroma pass in lockstep with a passthrough ring for an
// uncleaned group (engine/planar.rs:120-560), the NLMeans pass with repeated edges
// (nlmeans/engine.rs:201-262, nlmeans/denoiser/{mod,ring,edges}.rs, nlmeans/dispatch/mod.rs:627-679)
// and the NL4D pass with shifted edges and head/middle/tail passes (nl4d/engine.rs:264-360,
// nl4d/denoiser.rs:370-600, nl4d/collab_pass.rs:50-200).  Parameter closed forms follow
// nlmeans/params.rs:229-268, nlmeans/options.rs:106-137, nl4d/options.rs:118-185,
// collab/kernels/aggregate.rs:36-125 and motion/confidence.rs:17-29.
#include "engine.h"
#include "tune.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <stdexcept>

namespace avd {

// ---------------------------------------------------------------------------------------------
// Errors and buffers
// ---------------------------------------------------------------------------------------------

void check_hip(hipError_t e, const char *what)
{
    if (e != hipSuccess) throw Failure(AVD_ERR_GPU, std::string(what) + ": " + hipGetErrorString(e));
}

DeviceBuffer::~DeviceBuffer() { release(); }

void DeviceBuffer::release()
{
    if (ptr) (void)hipFree(ptr);
    ptr = nullptr;
    bytes = 0;
}

void DeviceBuffer::allocate(size_t n)
{
    release();
    if (n == 0) n = 4;
    check_hip(hipMalloc(&ptr, n), "hipMalloc");
    bytes = n;
}

// ---------------------------------------------------------------------------------------------
// Formats
// ---------------------------------------------------------------------------------------------

uint32_t samples_per_word(avd_sample_format f) { return f.kind == AVD_FORMAT_U8 ? 4u : 2u; }

float max_value(avd_sample_format f)
{
    const uint32_t depth = f.kind == AVD_FORMAT_U8 ? 8u : f.depth;
    return (float)((1u << depth) - 1u);
}

uint64_t plane_bytes(avd_sample_format f, uint64_t pixels)
{
    if (f.kind == AVD_FORMAT_F32) return pixels * 4;
    const uint64_t spw = samples_per_word(f);
    return (pixels + spw - 1) / spw * 4;
}

void validate_format(avd_sample_format f)
{
    if (f.kind == AVD_FORMAT_U8) return;
    if (f.kind == AVD_FORMAT_U16) {
        if (f.depth < 9 || f.depth > 16)
            throw Failure(AVD_ERR_INVALID_GEOMETRY, "U16 depth must be between 9 and 16, got " + std::to_string(f.depth));
        return;
    }
    throw Failure(AVD_ERR_INVALID_GEOMETRY, "F32 planes are not supported by this library; use U8 or U16 planes");
}

static void check_planes(const avd_device_plane *planes, uint32_t count, uint32_t expected, uint32_t W, uint32_t H,
                         avd_sample_format f)
{
    if (count != expected)
        throw Failure(AVD_ERR_PLANE_MISMATCH,
                      "expected " + std::to_string(expected) + " planes, got " + std::to_string(count));
    const uint64_t need = plane_bytes(f, (uint64_t)W * H);
    for (uint32_t i = 0; i < count; ++i) {
        if (planes[i].width != W || planes[i].height != H)
            throw Failure(AVD_ERR_PLANE_MISMATCH, "plane " + std::to_string(i) + " is " + std::to_string(planes[i].width) +
                                                      "x" + std::to_string(planes[i].height) + ", expected " +
                                                      std::to_string(W) + "x" + std::to_string(H));
        if (planes[i].bytes < need || planes[i].data == nullptr)
            throw Failure(AVD_ERR_PLANE_MISMATCH, "plane " + std::to_string(i) + " holds " +
                                                      std::to_string(planes[i].bytes) + " bytes, needs " +
                                                      std::to_string(need));
    }
}

// ---------------------------------------------------------------------------------------------
// Closed forms
// ---------------------------------------------------------------------------------------------

static const float kHqStrengthLuma[9] = {0.45f, 0.45f, 0.42f, 0.42f, 0.35f, 0.35f, 0.35f, 0.30f, 0.30f};
static const float kHqStrengthChroma[9] = {1.00f, 0.85f, 0.70f, 0.70f, 0.70f, 0.70f, 0.70f, 0.70f, 0.70f};

float h2_inv_norm(float strength, uint32_t patch_radius)
{
    const float area = (float)((2 * patch_radius + 1) * (2 * patch_radius + 1));
    return (255.0f * 255.0f) / (3.0f * strength * strength * area);
}

float noise_offset(float sigma, int channels, uint32_t patch_radius)
{
    const float area = (float)((2 * patch_radius + 1) * (2 * patch_radius + 1));
    float sum_sq = 0.0f;
    for (int c = 0; c < channels; ++c) sum_sq += sigma * sigma;
    return 2.0f * (channels == 1 ? 3.0f : 1.5f) * sum_sq * area;
}

float sad_noise_floor(uint32_t blksize, float sigma)
{
    const float area = (float)(blksize * blksize);
    return area * 2.0f * sigma / (float)std::sqrt(M_PI);
}

float thsad(uint32_t blksize, float scale) { return scale * (float)(blksize * blksize) * 0.02f; }

static double bessel_i0(double x)
{
    double half = x / 2.0, term = 1.0, sum = 1.0;
    for (int k = 1; k < 32; ++k) {
        term *= half / k;
        sum += term * term;
    }
    return sum;
}

void kaiser_window(float beta, float out[8])
{
    const double denom = bessel_i0((double)beta);
    for (int i = 0; i < 8; ++i) {
        const double pos = 2.0 * i / 7.0 - 1.0;
        out[i] = (float)(bessel_i0((double)beta * std::sqrt(1.0 - pos * pos)) / denom);
    }
}

// Algorithmic operations of one motion launch: 3 per pixel of every scored block offset (subtract,
// absolute value, add), the coarse level on half-size blocks of the 2x2 level.
uint64_t motion_ops(const MotionGeom &g, int neighbours)
{
    const uint64_t side2 = (uint64_t)(2 * g.radius + 1) * (2 * g.radius + 1);
    uint64_t ops = (uint64_t)g.blocks_x * g.blocks_y * side2 * (uint64_t)g.blksize * g.blksize * 3;
    if (g.levels == 2) {
        const int lw = std::max((int)g.W / 2, 1), lh = std::max((int)g.H / 2, 1);
        const int cblk = std::max(g.blksize / 2, 2), cstep = std::max(g.step / 2, 1);
        const uint64_t cb = (uint64_t)std::max((lw + cstep - 1) / cstep, 1) * std::max((lh + cstep - 1) / cstep, 1);
        ops += cb * side2 * (uint64_t)cblk * cblk * 3;
    }
    return ops * (uint64_t)neighbours;
}

// Per-candidate NLM operations (5C + 2(2P + 1) + 7).
uint64_t nlm_candidate_ops(int C, int P) { return (uint64_t)(5 * C + 2 * (2 * P + 1) + 7); }

// cross_frame_accum_scale (aggregate.rs:36-55) with the reference's STEP and MAX_K as the grid
// stride and the group size, so the stride-3, 16-member grid gets the bound of its own worst case.
float cross_frame_accum_scale(uint32_t spatial_radius, uint32_t temporal_radius, uint32_t step, uint32_t group_size)
{
    const uint32_t per_axis = ((8 - 1) + 2 * spatial_radius) / step + 1;
    const double contribs = (double)per_axis * (double)per_axis * (double)group_size;
    const double passes = (double)(4 * temporal_radius + 1);
    const double max_raw = contribs * passes * 5.0;
    const double budget = (double)INT32_MAX / 2.0;
    return (float)std::pow(2.0, std::floor(std::log2(budget / max_raw)));
}

float weight_scale(float sigma, float g_max)
{
    const float norm = sigma * sigma * g_max * g_max;
    return (std::isfinite(norm) && norm > 1e-12f) ? norm : 1e-12f;
}

// ---------------------------------------------------------------------------------------------
// Stage timing
// ---------------------------------------------------------------------------------------------

StageTimer::~StageTimer()
{
    for (auto &p : pending_) {
        (void)hipEventDestroy(p.start);
        (void)hipEventDestroy(p.stop);
    }
}

void StageTimer::set_sampling(bool on, hipStream_t s)
{
    harvest();
    on_ = on;
    for (int i = 0; i < AVD_STAGE_COUNT; ++i) {
        ms_[i] = 0.0;
        launches_[i] = 0;
        work_[i] = 0;
    }
    if (on) {
        if (!counter_.ptr) counter_.allocate(sizeof(unsigned long long) * kScoreSlots * kScoreStride);
        check_hip(launch_fill_u32(counter_.ptr, 0, 2 * kScoreSlots * kScoreStride, s), "zero counter");
    }
}

int StageTimer::begin(int stage, hipStream_t s)
{
    if (!on_) return -1;
    Pending p{stage, nullptr, nullptr, 0};
    check_hip(hipEventCreate(&p.start), "hipEventCreate");
    check_hip(hipEventCreate(&p.stop), "hipEventCreate");
    check_hip(hipEventRecord(p.start, s), "hipEventRecord");
    pending_.push_back(p);
    return (int)pending_.size() - 1;
}

void StageTimer::end(int handle, uint64_t work, hipStream_t s)
{
    if (handle < 0) return;
    Pending &p = pending_[(size_t)handle];
    p.work = work;
    check_hip(hipEventRecord(p.stop, s), "hipEventRecord");
}

void StageTimer::harvest()
{
    for (auto &p : pending_) {
        check_hip(hipEventSynchronize(p.stop), "hipEventSynchronize");
        float ms = 0.0f;
        check_hip(hipEventElapsedTime(&ms, p.start, p.stop), "hipEventElapsedTime");
        ms_[p.stage] += ms;
        launches_[p.stage] += 1;
        work_[p.stage] += p.work;
        (void)hipEventDestroy(p.start);
        (void)hipEventDestroy(p.stop);
    }
    pending_.clear();
}

void StageTimer::fold_counter(hipStream_t s)
{
    if (!on_ || !counter_.ptr) return;
    std::vector<unsigned long long> v((size_t)kScoreSlots * kScoreStride);
    check_hip(hipMemcpyAsync(v.data(), counter_.ptr, v.size() * sizeof(v[0]), hipMemcpyDeviceToHost, s),
              "read counter");
    check_hip(hipStreamSynchronize(s), "read counter");
    // 192 operations per candidate patch and channel: 64 samples x (subtract, multiply, add).
    for (int i = 0; i < kScoreSlots; ++i) work_[AVD_STAGE_COLLAB_SEARCH] += 192 * v[(size_t)i * kScoreStride];
    check_hip(launch_fill_u32(counter_.ptr, 0, 2 * kScoreSlots * kScoreStride, s), "zero counter");
}

avd_timing StageTimer::totals() const
{
    avd_timing t;
    for (int i = 0; i < AVD_STAGE_COUNT; ++i) {
        t.milliseconds[i] = ms_[i];
        t.launches[i] = launches_[i];
        t.work[i] = work_[i];
    }
    return t;
}

// ---------------------------------------------------------------------------------------------
// Frame ring shared by both passes (nlmeans/denoiser/ring.rs)
// ---------------------------------------------------------------------------------------------

void Ring::init(uint32_t W_, uint32_t H_, int C_, uint32_t slots_, bool pyramid)
{
    W = W_;
    H = H_;
    C = C_;
    N = slots_;
    px = (size_t)W * H;
    values.allocate(N * px * C * sizeof(float));
    codes.allocate(N * px * C * sizeof(uint16_t));
    if (pyramid) {
        l1_px = (size_t)std::max(W / 2, 1u) * std::max(H / 2, 1u);
        pyr1.allocate(N * l1_px * sizeof(int32_t));
    }
    slot_frames.assign(N, 0);
    const uint32_t per = N > 1 ? N - 1 : 1;
    std::vector<uint32_t> table((size_t)N * N * per, 0);
    for (uint32_t start = 0; start < N; ++start)
        for (uint32_t c = 0; c < N; ++c) {
            uint32_t *e = table.data() + ((size_t)start * N + c) * per;
            for (uint32_t l = 0; l < N; ++l)
                if (l != c) *e++ = (start + l) % N;
        }
    neighbour_table.allocate(table.size() * sizeof(uint32_t));
    check_hip(hipMemcpy(neighbour_table.ptr, table.data(), table.size() * sizeof(uint32_t), hipMemcpyHostToDevice),
              "upload neighbour table");
    reset();
}

void Ring::reset()
{
    head = 0;
    loaded = 0;
    real = 0;
}

uint32_t Ring::phys(int logical) const
{
    const int n = (int)N;
    const int wrapped = ((logical % n) + n) % n;
    return (uint32_t)(((int)(head % N) + wrapped) % n);
}

void Ring::write(const avd_device_plane *planes, avd_sample_format f, hipStream_t s)
{
    const uint32_t slot = head % N;
    const uint32_t *p0 = static_cast<const uint32_t *>(planes[0].data);
    const uint32_t *p1 = C > 1 ? static_cast<const uint32_t *>(planes[1].data) : p0;
    check_hip(launch_ingest(p0, p1, value_slot(slot), code_slot(slot), (uint32_t)px, C, (int)samples_per_word(f),
                            max_value(f), s),
              "ingest");
    if (pyr1.ptr) check_hip(launch_downscale(code_slot(slot), W, H, C, l1_slot(slot), s), "downscale");
    slot_frames[slot] = real;
    ++head;
    if (loaded < N) ++loaded;
    ++real;
}

void Ring::duplicate_last(hipStream_t s)
{
    const uint32_t last = (head + N - 1) % N, next = head % N;
    check_hip(hipMemcpyAsync(value_slot(next), value_slot(last), px * C * sizeof(float), hipMemcpyDeviceToDevice, s),
              "copy slot");
    check_hip(hipMemcpyAsync(code_slot(next), code_slot(last), px * C * sizeof(uint16_t), hipMemcpyDeviceToDevice, s),
              "copy slot");
    if (pyr1.ptr)
        check_hip(hipMemcpyAsync(l1_slot(next), l1_slot(last), l1_px * sizeof(int32_t), hipMemcpyDeviceToDevice, s),
                  "copy slot");
    slot_frames[next] = slot_frames[last];
    ++head;
}

// MotionEstimation::resolve (motion/search.rs:118-140): Auto is Direct below temporal radius 3
// and Chained from it on.  Only Direct is built, so a request that resolves to Chained is refused
// rather than run as something else.
static void check_estimation(const avd_motion_search &m, uint32_t temporal_radius)
{
    if (m.estimation > 2) throw Failure(AVD_ERR_INVALID_OPTIONS, "estimation must be 0 (Auto), 1 (Direct) or 2 (Chained)");
    if (m.estimation == 2)
        throw Failure(AVD_ERR_INVALID_OPTIONS, "Chained motion estimation is not built yet; use Direct");
    if (m.estimation == 0 && temporal_radius >= 3)
        throw Failure(AVD_ERR_INVALID_OPTIONS,
                      "Auto motion estimation resolves to Chained at temporal radius 3 and above, which is not built "
                      "yet; use Direct");
}

// MotionCompensationMode::validate (motion/search.rs:226-290, MotionEstimation::validate :143-155),
// then what this library runs of it: two pyramid levels at most, Direct estimation, no escalation.
static void validate_motion_search(const avd_motion_search &m, uint32_t temporal_radius)
{
    auto invalid = [](const std::string &what) { throw Failure(AVD_ERR_INVALID_OPTIONS, "motion-compensation " + what); };
    if (m.blksize < 4) invalid("blksize=" + std::to_string(m.blksize) + " is too small, the minimum is 4 pixels per side");
    if (m.blksize > 32) invalid("blksize=" + std::to_string(m.blksize) + " exceeds the supported maximum of 32");
    if (m.blksize % 2 != 0)
        invalid("blksize=" + std::to_string(m.blksize) + " must be even so the /2 coarse level is well-defined");
    if (m.overlap >= m.blksize) invalid("overlap must be strictly less than blksize");
    if (m.search_radius == 0 || m.search_radius > 8) invalid("search_radius must be in 1..=8");
    if (m.max_search_escalation < m.search_radius || m.max_search_escalation > 8)
        invalid("max_search_escalation must be between search_radius and 8");
    if (m.pyramid_levels == 0 || m.pyramid_levels > 3) invalid("pyramid_levels must be in 1..=3");
    if (m.estimation == 2 && (m.refine_radius == 0 || m.refine_radius > 8))
        throw Failure(AVD_ERR_INVALID_OPTIONS, "motion-estimation refine_radius must be in 1..=8");
    if (m.pyramid_levels > 2) throw Failure(AVD_ERR_INVALID_OPTIONS, "three-level motion pyramids are not built yet");
    check_estimation(m, temporal_radius);
    if (m.max_search_escalation > m.search_radius)
        throw Failure(AVD_ERR_INVALID_OPTIONS, "search escalation is not built yet");
}

// ---------------------------------------------------------------------------------------------
// Luma motion shared with the chroma pass (motion/luma_ring.rs, motion/chroma_map.rs)
// ---------------------------------------------------------------------------------------------

bool same_motion_search(const avd_motion_search &a, const avd_motion_search &b)
{
    return a.blksize == b.blksize && a.overlap == b.overlap && a.search_radius == b.search_radius &&
           a.max_search_escalation == b.max_search_escalation && a.pyramid_levels == b.pyramid_levels &&
           a.estimation == b.estimation && a.refine_radius == b.refine_radius;
}

void Pass::search_motion(const MotionGeom &g, const Ring &ring, uint32_t centre, const uint32_t *nslots,
                         int neighbours, int32_t *mv, float *conf, float mx, float sad_floor, float th, bool shares)
{
    const uint32_t frame = ring.slot_frames[centre];
    if (share_ && reads_motion_ && shares) {
        const int32_t *seeds = share_->read(frame, stream_);
        MotionGeom rg = g;
        rg.radius = kMappedRefineRadius;
        rg.levels = 1;
        timed(AVD_STAGE_MOTION, motion_ops(rg, neighbours), stream_, "motion refine", [&] {
            return launch_motion_refine(g, ring.code_ptr(), (int)centre, nslots, neighbours, seeds, mv, conf, mx,
                                        sad_floor, th, kMappedRefineRadius, stream_);
        });
        return;
    }
    timed(AVD_STAGE_MOTION, motion_ops(g, neighbours), stream_, "motion", [&] {
        return launch_motion(g, ring.code_ptr(), ring.l1_ptr(), (int)centre, nslots, neighbours, mv, conf, mx,
                             sad_floor, th, stream_);
    });
    if (share_ && shares)
        timed(AVD_STAGE_MOTION, 0, stream_, "share motion", [&] { return share_->write(frame, mv, stream_); });
}

LumaMotionShare::LumaMotionShare(const MotionGeom &luma, const MotionGeom &chroma, int shift_x, int shift_y,
                                 uint32_t radius)
{
    map_ = FieldMap{luma.blocks_x,  luma.blocks_y,  luma.step,   luma.blksize, luma.blocks_x * luma.blocks_y * 2,
                    chroma.blocks_x, chroma.blocks_y, chroma.step, chroma.blksize,
                    chroma.blocks_x * chroma.blocks_y * 2, shift_x, shift_y};
    slots_ = 2 * radius + 1;
    fields_ = 2 * radius;
    buf_.allocate((size_t)slots_ * fields_ * map_.chroma_stride * sizeof(int32_t));
    ready_.assign(slots_, nullptr);
    for (auto &e : ready_) check_hip(hipEventCreateWithFlags(&e, hipEventDisableTiming), "create motion share event");
    tags_.assign(slots_, -1);
}

LumaMotionShare::~LumaMotionShare()
{
    for (hipEvent_t e : ready_)
        if (e) (void)hipEventDestroy(e);
}

hipError_t LumaMotionShare::write(uint32_t frame, const int32_t *luma_mv, hipStream_t s)
{
    const uint32_t slot = frame % slots_;
    int32_t *dst = buf_.as<int32_t>() + (size_t)slot * fields_ * map_.chroma_stride;
    hipError_t e = launch_map_luma_fields(map_, luma_mv, dst, (int)fields_, s);
    if (e == hipSuccess) e = hipEventRecord(ready_[slot], s);
    if (e == hipSuccess) tags_[slot] = frame;
    return e;
}

const int32_t *LumaMotionShare::read(uint32_t frame, hipStream_t s)
{
    const uint32_t slot = frame % slots_;
    if (tags_[slot] != (int64_t)frame)
        throw Failure(AVD_ERR_GPU, "the luma pass has not shared the motion of stream frame " + std::to_string(frame));
    check_hip(hipStreamWaitEvent(s, ready_[slot], 0), "wait for luma motion");
    return buf_.as<int32_t>() + (size_t)slot * fields_ * map_.chroma_stride;
}

void LumaMotionShare::clear() { std::fill(tags_.begin(), tags_.end(), -1); }

// ---------------------------------------------------------------------------------------------
// NLMeans pass
// ---------------------------------------------------------------------------------------------

NlmPass::NlmPass(const avd_nlmeans_options &o, uint32_t W, uint32_t H, int C, avd_sample_format in,
                 avd_sample_format out, hipStream_t s)
    : Pass(s), W_(W), H_(H), C_(C), in_(in), out_(out)
{
    // resolve_params (nlmeans/options.rs:106-137)
    r_ = o.temporal ? o.temporal_radius : 0;
    hq_ = o.hq != 0;
    S_ = o.has_search_radius ? o.search_radius : 2;
    P_ = o.has_patch_radius ? o.patch_radius : 4;
    const bool auto_strength = hq_ && o.auto_strength;
    float strength = 1.2f;
    if (auto_strength) strength = (C == 1 ? kHqStrengthLuma : kHqStrengthChroma)[std::min(r_, 8u)];
    if (o.has_strength) strength = o.strength;
    self_weight_ = o.has_self_weight ? o.self_weight : 1.0f;
    has_sigma_ = hq_ && o.has_sigma;
    sigma_ = o.sigma;
    motion_ = o.motion_compensation != 0 && r_ > 0;
    confidence_ = hq_ && o.temporal_confidence && r_ > 0;
    thsad_scale_ = hq_ ? o.thsad_scale : 1.0f;

    // NlmParams::validate (nlmeans/params.rs:276-340) plus this library's kernel limits.
    if (P_ > 16) throw Failure(AVD_ERR_INVALID_OPTIONS, "patch_radius=" + std::to_string(P_) + " exceeds 16");
    if (S_ > 8) throw Failure(AVD_ERR_INVALID_OPTIONS, "search_radius=" + std::to_string(S_) + " exceeds 8");
    if (r_ > 8) throw Failure(AVD_ERR_INVALID_OPTIONS, "temporal_radius=" + std::to_string(r_) + " exceeds 8");
    if (!(strength > 0.0f) || !std::isfinite(strength))
        throw Failure(AVD_ERR_INVALID_OPTIONS, "strength must be finite and above 0");
    if (!nlm_supported(C, (int)P_, (int)S_))
        throw Failure(AVD_ERR_INVALID_OPTIONS, "patch_radius above 8 (the separable path) is not supported by this library");
    if (hq_ && !has_sigma_)
        throw Failure(AVD_ERR_INVALID_OPTIONS,
                      "HQ noise estimation is not built yet; pass a fixed sigma (HqParams::sigma_override)");
    // Confidence with motion compensation off scores each block where it stands
    // (MotionCtx::confidence_only, motion/context.rs:182-200: 16x16 blocks at step 8, one level,
    // radius 0) and leaves the neighbours unwarped.
    conf_only_ = confidence_ && !motion_;
    if (conf_only_) {
        if (in.kind == AVD_FORMAT_U16 && in.depth > 12)
            throw Failure(AVD_ERR_INVALID_GEOMETRY, "motion search on more than 12-bit samples is not supported");
        mg_ = MotionGeom{(int)W, (int)H, C, 16, 8, 0, 1, 0, 0};
        mg_.blocks_x = std::max((int)div_up(W, mg_.step), 1);
        mg_.blocks_y = std::max((int)div_up(H, mg_.step), 1);
    }
    if (motion_) {
        const avd_motion_search &m = o.motion;
        validate_motion_search(m, r_);
        if (in.kind == AVD_FORMAT_U16 && in.depth > 12)
            throw Failure(AVD_ERR_INVALID_GEOMETRY, "motion search on more than 12-bit samples is not supported");
        motion_opts_ = m;
        mg_ = MotionGeom{(int)W, (int)H, C, (int)m.blksize, (int)(m.blksize - m.overlap), (int)m.search_radius,
                         (int)m.pyramid_levels, 0, 0};
        mg_.blocks_x = std::max((int)div_up(W, mg_.step), 1);
        mg_.blocks_y = std::max((int)div_up(H, mg_.step), 1);
    }

    float h2_strength = strength;
    if (auto_strength && has_sigma_) h2_strength = strength * sigma_ * 255.0f;
    h2_ = h2_inv_norm(h2_strength, P_);
    offset_ = (hq_ && o.noise_floor && has_sigma_) ? noise_offset(sigma_, C, P_) : 0.0f;

    const uint32_t N = 2 * r_ + 1;
    ring_.init(W, H, C, N, motion_);
    const size_t px = (size_t)W * H;
    accum_.allocate(px * C * sizeof(float));
    wsum_.allocate(px * sizeof(float));
    maxw_.allocate(px * sizeof(float));
    for (auto &b : outputs_) b.allocate(px * C * sizeof(float));
    const uint32_t side = 2 * S_ + 1;
    std::vector<float> lut(side * side, offset_);
    lut[S_ * side + S_] = 0.0f;  // build_spatial_offset_lut at rho = 0 (noise/correlation.rs:75-88)
    lut_.allocate(lut.size() * sizeof(float));
    check_hip(hipMemcpy(lut_.ptr, lut.data(), lut.size() * sizeof(float), hipMemcpyHostToDevice), "upload lut");
    if (motion_ || conf_only_) {
        const size_t blocks = (size_t)mg_.blocks_x * mg_.blocks_y;
        if (motion_) comp_.allocate(N * px * C * sizeof(float));
        mv_.allocate(2 * r_ * blocks * 2 * sizeof(int32_t));
        conf_.allocate(2 * r_ * blocks * sizeof(float));
    }
    tune_windows();
}

void NlmPass::set_motion_share(LumaMotionShare *share, bool reads)
{
    Pass::set_motion_share(share, reads);
    if (reads) ring_.pyr1.release();
}

// The window's motion: the pass's own search (run_match, nlmeans/dispatch/motion.rs:180-224), shared
// when the pass writes luma motion (share_window_motion, :226-252), or the refine of the luma
// pass's mapped vectors (refine_luma_motion, :254-313).
void NlmPass::window_motion(uint32_t centre, const uint32_t *nslots, int neighbours, float mx)
{
    search_motion(mg_, ring_, centre, nslots, neighbours, mv_.as<int32_t>(), conf_.as<float>(), mx,
                        sad_noise_floor(mg_.blksize, sigma_), thsad(mg_.blksize, thsad_scale_), motion_);
}

NlmWindow NlmPass::window() const
{
    NlmWindow w{(int)W_, (int)H_, C_, (int)P_, (int)S_, h2_ * 1.4426950408889634f, offset_, 1, 1, 1};
    if (motion_ || conf_only_) {
        w.step = mg_.step;
        w.blocks_x = mg_.blocks_x;
        w.blocks_y = mg_.blocks_y;
    }
    return w;
}

// The window launch sites' rows per wave, picked by the tuner on this pass's own buffers: ring
// slots of flat mid-grey (the window walks every offset whatever the content, so the time is the
// content's) and, with confidence, blocks of confidence 1.  The stream then starts from scratch.
void NlmPass::tune_windows()
{
    const size_t px = (size_t)W_ * H_;
    check_hip(launch_fill_u32(ring_.values.ptr, 0x3f000000u, px * C_ * ring_.N, stream_), "tuning frames");
    if (confidence_)
        check_hip(launch_fill_u32(conf_.ptr, 0x3f800000u, conf_.bytes / sizeof(float), stream_), "tuning confidence");
    const NlmWindow w = window();
    auto clear = [&]() -> hipError_t {
        hipError_t e = launch_fill_u32(accum_.ptr, 0, px * C_, stream_);
        if (e == hipSuccess) e = launch_fill_u32(wsum_.ptr, 0, px, stream_);
        if (e == hipSuccess) e = launch_fill_u32(maxw_.ptr, 0, px, stream_);
        return e;
    };
    auto read = [&]() {
        std::vector<uint8_t> out(px * (C_ + 2) * sizeof(float));
        check_hip(hipMemcpy(out.data(), accum_.ptr, px * C_ * sizeof(float), hipMemcpyDeviceToHost), "tuning read");
        check_hip(hipMemcpy(out.data() + px * C_ * 4, wsum_.ptr, px * sizeof(float), hipMemcpyDeviceToHost),
                  "tuning read");
        check_hip(hipMemcpy(out.data() + px * (C_ + 1) * 4, maxw_.ptr, px * sizeof(float), hipMemcpyDeviceToHost),
                  "tuning read");
        return out;
    };
    const std::string key = "C" + std::to_string(C_) + " P" + std::to_string(P_) + " S" + std::to_string(S_) + " " +
                            std::to_string(W_) + "x" + std::to_string(H_);
    const float *f = ring_.value_slot(0);
    variant_single_ = Tuner::instance().choose(
        "nlm_single", key, kNlmVariants, stream_,
        [&](int v) { return launch_nlm_single(w, f, lut_.as<float>(), accum_.as<float>(), wsum_.as<float>(),
                                              maxw_.as<float>(), stream_, v); },
        clear, read);
    {
        NlmFrames fr{};
        fr.pairs = (int)r_;
        for (uint32_t k = 0; k < r_; ++k) {
            fr.fwd[k] = fr.bwd[k] = f;
            fr.conf_f[k] = fr.conf_b[k] = confidence_ ? conf_.as<float>() : nullptr;
        }
        float *out = outputs_[0].as<float>();
        variant_window_ = Tuner::instance().choose(
            "nlm_window", key + " r" + std::to_string(r_) + (confidence_ ? " confidence" : ""), kNlmBatchedVariants,
            stream_, [&](int v) { return launch_nlm_window(w, fr, f, lut_.as<float>(), self_weight_, out, stream_, v); },
            [&]() -> hipError_t { return launch_fill_u32(out, 0, px * C_, stream_); },
            [&]() {
                std::vector<uint8_t> o(px * C_ * sizeof(float));
                check_hip(hipMemcpy(o.data(), out, o.size(), hipMemcpyDeviceToHost), "tuning read");
                return o;
            });
    }
    if (r_ > 0) {
        const float *c = confidence_ ? conf_.as<float>() : nullptr;
        variant_pair_ = Tuner::instance().choose(
            "nlm_pair", key + (c ? " confidence" : ""), kNlmVariants, stream_,
            [&](int v) { return launch_nlm_pair(w, f, f, f, c, c, accum_.as<float>(), wsum_.as<float>(),
                                                maxw_.as<float>(), stream_, v); },
            clear, read);
    }
}

size_t NlmPass::max_held() const { return r_; }

void NlmPass::reset_stream()
{
    ring_.reset();
    ready_ = false;
    tail_remaining_ = 0;
    next_output_ = 0;
    pushed_ = false;
}

void NlmPass::ingest(const avd_device_plane *planes)
{
    timed(AVD_STAGE_INGEST, ring_.px, stream_, "ingest", [&] {
        ring_.write(planes, in_, stream_);
        return hipSuccess;
    });
    if (ring_.real == 1 && r_ > 0) {  // prime_leading_edge_if_first (ring.rs:106-125)
        for (uint32_t i = 0; i < r_; ++i) {
            ring_.duplicate_last(stream_);
            ++ring_.loaded;
        }
    }
}

uint32_t NlmPass::push(const avd_device_plane *planes)
{
    ingest(planes);
    pushed_ = true;
    if (ring_.ready()) {
        denoise();
        ready_ = true;
        return 1;
    }
    return 0;
}

void NlmPass::push_context(const avd_device_plane *planes) { ingest(planes); }

// run_windowed_kernels (nlmeans/dispatch/mod.rs:665-679): zero, pairs q_k = -r..-1, single, finish.
void NlmPass::denoise()
{
    const size_t px = (size_t)W_ * H_;
    const int C = C_;
    const uint32_t centre = ring_.phys((int)r_);
    emitted_frame_ = ring_.slot_frames[centre];
    const float *cf = ring_.value_slot(centre);
    auto frame_of = [&](int k) -> const float * { return ring_.value_slot(ring_.phys((int)r_ + k)); };
    std::vector<const float *> nb(2 * r_ + 1, nullptr);
    for (int k = -(int)r_; k <= (int)r_; ++k) nb[k + r_] = frame_of(k);
    if (motion_ || conf_only_) {
        // run_motion_compensation (nlmeans/dispatch/motion.rs:427-520), or run_confidence_pass
        // (:554-612) when only confidence is on: neighbours oldest first.
        std::vector<uint32_t> slots;
        for (int k = -(int)r_; k <= (int)r_; ++k)
            if (k != 0) slots.push_back(ring_.phys((int)r_ + k));
        const float mx = max_value(in_);
        const size_t blocks = (size_t)mg_.blocks_x * mg_.blocks_y;
        window_motion(centre, ring_.neighbours_of(r_), (int)slots.size(), mx);
        for (size_t i = 0; i < slots.size() && motion_; ++i) {
            const int k = (int)i < (int)r_ ? (int)i - (int)r_ : (int)i - (int)r_ + 1;
            float *dst = comp_.as<float>() + (size_t)slots[i] * px * C;
            timed(AVD_STAGE_WARP, px, stream_, "warp", [&] {
                return launch_warp(ring_.value_slot(slots[i]), dst, mv_.as<int32_t>() + i * blocks * 2, W_, H_, C,
                                   mg_.step, mg_.blocks_x, mg_.blocks_y, stream_);
            });
            nb[k + r_] = dst;
        }
    }
    const NlmWindow w = window();
    const size_t blocks = (motion_ || conf_only_) ? (size_t)mg_.blocks_x * mg_.blocks_y : 0;
    auto conf_of = [&](int k) -> const float * {
        if (!confidence_) return nullptr;
        const int idx = k < 0 ? k + (int)r_ : k + (int)r_ - 1;
        return conf_.as<float>() + (size_t)idx * blocks;
    };
    const uint64_t window = (uint64_t)(2 * S_ + 1) * (2 * S_ + 1);
    if (batched_) {
        // Pairs q = -r .. -1 (the forward frame the older neighbour), the single window and the
        // finish in one launch.
        NlmFrames fr{};
        fr.pairs = (int)r_;
        for (int qk = -(int)r_; qk < 0; ++qk) {
            const int k = qk + (int)r_;
            fr.fwd[k] = nb[qk + r_];
            fr.bwd[k] = nb[-qk + r_];
            fr.conf_f[k] = conf_of(qk);
            fr.conf_b[k] = conf_of(-qk);
        }
        float *out = outputs_[next_output_].as<float>();
        const uint64_t candidates = px * (window * 2 * r_ + window - 1);
        timed(AVD_STAGE_NLM_WINDOW, candidates * nlm_candidate_ops(C, P_), stream_, "nlm window", [&] {
            return launch_nlm_window(w, fr, cf, lut_.as<float>(), self_weight_, out, stream_, variant_window_);
        });
        ready_output_ = next_output_;
        next_output_ = (next_output_ + 1) % 2;
        return;
    }
    check_hip(launch_fill_u32(accum_.ptr, 0, px * C, stream_), "zero");
    check_hip(launch_fill_u32(wsum_.ptr, 0, px, stream_), "zero");
    check_hip(launch_fill_u32(maxw_.ptr, 0, px, stream_), "zero");
    for (int qk = -(int)r_; qk < 0; ++qk)
        timed(AVD_STAGE_NLM_PAIR, px * window * 2 * nlm_candidate_ops(C, P_), stream_, "nlm pair", [&] {
            return launch_nlm_pair(w, cf, nb[qk + r_], nb[-qk + r_], conf_of(qk), conf_of(-qk), accum_.as<float>(),
                                   wsum_.as<float>(), maxw_.as<float>(), stream_, variant_pair_);
        });
    timed(AVD_STAGE_NLM_SINGLE, px * (window - 1) * nlm_candidate_ops(C, P_), stream_, "nlm single", [&] {
        return launch_nlm_single(w, cf, lut_.as<float>(), accum_.as<float>(), wsum_.as<float>(), maxw_.as<float>(),
                                 stream_, variant_single_);
    });
    float *out = outputs_[next_output_].as<float>();
    timed(AVD_STAGE_NLM_FINISH, px, stream_, "nlm finish", [&] {
        return launch_nlm_finish(cf, accum_.as<float>(), wsum_.as<float>(), maxw_.as<float>(), self_weight_,
                                 (uint32_t)px, C, out, stream_);
    });
    ready_output_ = next_output_;
    next_output_ = (next_output_ + 1) % 2;
}

void NlmPass::egress(const float *frame, const avd_device_plane *planes)
{
    uint32_t *p0 = static_cast<uint32_t *>(planes[0].data);
    uint32_t *p1 = C_ > 1 ? static_cast<uint32_t *>(planes[1].data) : p0;
    timed(AVD_STAGE_EGRESS, (uint64_t)W_ * H_, stream_, "egress", [&] {
        return launch_egress(frame, p0, p1, W_ * H_, C_, (int)samples_per_word(out_), max_value(out_), stream_);
    });
}

void NlmPass::emit(const avd_device_plane *planes)
{
    if (ready_) {
        ready_ = false;
        egress(outputs_[ready_output_].as<float>(), planes);
        return;
    }
    if (tail_remaining_ == 0) throw Failure(AVD_ERR_NOTHING_TO_EMIT, "no frame is ready to emit");
    // next_tail_frame (nlmeans/engine.rs:153-160): duplicate forward until the window is full.
    for (;;) {
        ring_.duplicate_last(stream_);
        if (ring_.loaded < ring_.N) ++ring_.loaded;
        if (ring_.ready()) break;
    }
    denoise();
    egress(outputs_[ready_output_].as<float>(), planes);
    if (--tail_remaining_ == 0) reset_stream();
}

uint32_t NlmPass::finish()
{
    tail_remaining_ = std::min<uint32_t>(ring_.real, r_);
    if (tail_remaining_ == 0) reset_stream();
    return tail_remaining_;
}

bool NlmPass::has_pending() const { return ready_; }

// ---------------------------------------------------------------------------------------------
// NL4D pass
// ---------------------------------------------------------------------------------------------

Nl4dPass::Nl4dPass(const avd_nl4d_options &o, uint32_t W, uint32_t H, int C, avd_sample_format in,
                   avd_sample_format out, hipStream_t s)
    : Pass(s), W_(W), H_(H), C_(C), in_(in), out_(out)
{
    r_ = o.temporal_radius;
    if (r_ == 0) throw Failure(AVD_ERR_INVALID_OPTIONS, "nl4d needs a temporal radius of at least 1");
    if (!(o.lambda_ht_scale >= 0.1f && o.lambda_ht_scale <= 10.0f))
        throw Failure(AVD_ERR_INVALID_OPTIONS, "lambda_ht_scale must be finite and in 0.1..=10.0");
    const float lam = (o.has_lambda_ht ? o.lambda_ht : (C == 1 ? 4.158f : 3.234f)) * o.lambda_ht_scale;
    // Nl4dParams::validate (nl4d/params.rs:131-198)
    if (r_ > 8) throw Failure(AVD_ERR_INVALID_OPTIONS, "temporal_radius must be in 1..=8");
    if (o.refine < 1 || o.refine > 4) throw Failure(AVD_ERR_INVALID_OPTIONS, "refine must be in 1..=4");
    if (o.spatial_radius < 1 || o.spatial_radius > 16)
        throw Failure(AVD_ERR_INVALID_OPTIONS, "spatial_radius must be in 1..=16");
    if (!(lam > 0.0f) || !std::isfinite(lam)) throw Failure(AVD_ERR_INVALID_OPTIONS, "lambda_ht must be above 0");
    if (!(o.c_min >= 0.0f && o.c_min < 1.0f)) throw Failure(AVD_ERR_INVALID_OPTIONS, "c_min must be in 0..1");
    if (!(o.kaiser_beta >= 0.0f && o.kaiser_beta <= 8.0f))
        throw Failure(AVD_ERR_INVALID_OPTIONS, "kaiser_beta must be in 0..=8");
    const avd_motion_search &m = o.motion;
    validate_motion_search(m, r_);
    const uint32_t step = m.blksize - m.overlap;
    if ((m.blksize + step - 1) / step > 4)
        throw Failure(AVD_ERR_INVALID_OPTIONS, "more than 4 motion blocks would cover a patch on each axis");
    // The truemotion regulariser (nl4d/denoiser.rs:235, truemotion/stage.rs:61) runs whenever
    // field_lambda is above 0; it is not built, so only field_lambda = 0 keeps the reference's output.
    if (o.field_lambda != 0.0f)
        throw Failure(AVD_ERR_INVALID_OPTIONS, "the truemotion regulariser is not built yet; set field_lambda = 0");
    if (!o.has_sigma)
        throw Failure(AVD_ERR_INVALID_OPTIONS, "noise estimation is not built yet; pass a fixed sigma");
    if (o.grain_data) throw Failure(AVD_ERR_INVALID_OPTIONS, "grain measurement is not built yet");
    if (o.transform_form > 1) throw Failure(AVD_ERR_INVALID_OPTIONS, "transform_form must be 0 (FP32) or 1 (mixed)");
    if (o.transform_form == 1 && o.pooled_threshold)
        throw Failure(AVD_ERR_INVALID_OPTIONS, "the mixed transform form thresholds hard: set pooled_threshold = 0");
    mixed_ = o.transform_form == 1;
    // The reference's grid is stride 4 with groups of 8 (collab/mod.rs:16-20); this library also runs
    // any stride up to 8 and groups of 16 (volume grids of 4 frames x 4 volumes, 2 x 8).
    if (o.step < 1 || o.step > 8) throw Failure(AVD_ERR_INVALID_OPTIONS, "step must be in 1..=8");
    if (o.group_size != 8 && o.group_size != 16)
        throw Failure(AVD_ERR_INVALID_OPTIONS, "group_size must be 8 (the reference's) or 16");
    if (W < 8 || H < 8) throw Failure(AVD_ERR_INVALID_GEOMETRY, "frames must be at least 8x8 for the patch grid");
    if (W >= 8192 || H >= 8192) throw Failure(AVD_ERR_INVALID_GEOMETRY, "frames must be below 8192 on each axis");
    if (in.kind == AVD_FORMAT_U16 && in.depth > 12)
        throw Failure(AVD_ERR_INVALID_GEOMETRY, "grouping on more than 12-bit samples is not supported by this library");

    sigma_ = o.sigma;
    thsad_scale_ = o.thsad_scale;
    motion_opts_ = m;
    mg_ = MotionGeom{(int)W, (int)H, C, (int)m.blksize, (int)step, (int)m.search_radius, (int)m.pyramid_levels, 0, 0};
    mg_.blocks_x = std::max((int)div_up(W, step), 1);
    mg_.blocks_y = std::max((int)div_up(H, step), 1);

    N_ = 2 * r_ + 1;
    std::memset(&pass_, 0, sizeof(pass_));
    pass_.width = W;
    pass_.height = H;
    pass_.channels = (uint32_t)C;
    pass_.step = o.step;
    pass_.spatial_radius = o.spatial_radius;
    pass_.temporal_radius = r_;
    pass_.refine = o.refine;
    pass_.group_size = o.group_size;
    pass_.blk_step = step;
    pass_.blksize = m.blksize;
    pass_.blocks_x = (uint32_t)mg_.blocks_x;
    pass_.blocks_y = (uint32_t)mg_.blocks_y;
    pass_.slots = N_;
    pass_.c_min = o.c_min;
    pass_.lambda[0] = pass_.lambda[1] = lam;
    pass_.sigma[0] = pass_.sigma[1] = sigma_;
    for (int i = 0; i < 8; ++i) pass_.profile[i] = 1.0f;  // dct_noise_profile(0)
    pass_.weight_scale = weight_scale(sigma_, 1.0f);
    pass_.accum_scale = cross_frame_accum_scale(o.spatial_radius, r_, o.step, o.group_size);
    kaiser_window(o.kaiser_beta, pass_.kaiser);
    pass_.code_bits = in.kind == AVD_FORMAT_U8 ? 8u : in.depth;
    // nl4d_pool_ratio (nl4d/options.rs:109-117): the calibrated 2.42 over the plane's default lambda.
    pass_.pool_ratio = o.pooled_threshold ? 2.42f / (C == 1 ? 4.158f : 3.234f) : 0.0f;

    ring_.init(W, H, C, N_, true);
    const size_t px = (size_t)W * H;
    refs_ = (size_t)((W - 8 + o.step - 1) / o.step + 1) * ((H - 8 + o.step - 1) / o.step + 1);
    aggregate_.allocate(N_ * px * sizeof(unsigned long long));
    if (C == 2) accum_second_.allocate(N_ * px * sizeof(int32_t));
    const size_t K = o.group_size;
    members_.allocate(refs_ * (K + 1) * sizeof(uint32_t));  // K member words, then the group word
    group_weight_.allocate(refs_ * sizeof(float));
    {
        const uint32_t side = 1u << kTileShift;
        const uint32_t tiles_x = (W + side - 1) / side, per_slot = tiles_x * ((H + side - 1) / side);
        const size_t n_tiles = (size_t)per_slot * N_;
        const uint32_t cap = tile_cap_for(o.group_size);
        stage_.allocate(refs_ * K * C * 64 * sizeof(int32_t));
        tile_count_.allocate((n_tiles + 1) * sizeof(uint32_t));  // the last word counts overflow pairs
        tile_list_.allocate(n_tiles * cap * sizeof(uint32_t));
        tile_overflow_.allocate(refs_ * K * 4 * 2 * sizeof(uint32_t));  // up to four tiles per member
        uint32_t *counts = tile_count_.as<uint32_t>();
        bins_ = TileBins{stage_.as<int32_t>(), counts, tile_list_.as<uint32_t>(), tile_overflow_.as<uint32_t>(),
                         counts + n_tiles, tiles_x, per_slot, cap};
    }
    // The centre search scores every position of each reference's clipped rectangle but its own.
    centre_ops_ = 0;
    {
        const int st = (int)o.step, S = (int)o.spatial_radius;
        const int rx_n = ((int)W - 8 + st - 1) / st + 1, ry_n = ((int)H - 8 + st - 1) / st + 1;
        uint64_t rows = 0, cols = 0;
        for (int i = 0; i < rx_n; ++i) {
            const int x = std::min(i * st, (int)W - 8);
            cols += (uint64_t)(std::min(x + S, (int)W - 8) - std::max(x - S, 0) + 1);
        }
        for (int i = 0; i < ry_n; ++i) {
            const int y = std::min(i * st, (int)H - 8);
            rows += (uint64_t)(std::min(y + S, (int)H - 8) - std::max(y - S, 0) + 1);
        }
        centre_ops_ = (rows * cols - (uint64_t)rx_n * ry_n) * 192u * (uint64_t)C;
    }
    if (centre_on_matrix_cores(pass_)) {
        centre_buf_.allocate(centre_scratch_bytes(W, H, (uint32_t)C));
        centre_ = centre_scratch_at(centre_buf_.ptr, W, H, (uint32_t)C);
    }
    output_.allocate(px * C * sizeof(float));
    const size_t blocks = (size_t)mg_.blocks_x * mg_.blocks_y;
    mv_.allocate(2 * r_ * blocks * 2 * sizeof(int32_t));
    conf_.allocate(2 * r_ * blocks * sizeof(float));
    zeroed_.assign(N_, false);
}

size_t Nl4dPass::max_held() const { return 2 * r_; }

void Nl4dPass::set_motion_share(LumaMotionShare *share, bool reads)
{
    Pass::set_motion_share(share, reads);
    if (reads) ring_.pyr1.release();
}

void Nl4dPass::reset_stream()
{
    ring_.reset();
    pending_.clear();
    tail_.clear();
    passes_ = 0;
    continuation_ = false;
    std::fill(zeroed_.begin(), zeroed_.end(), false);
}

void Nl4dPass::zero_region(uint32_t slot)
{
    const size_t px = (size_t)W_ * H_;
    check_hip(launch_fill_u32(aggregate_.as<uint32_t>() + slot * px * 2, 0, px * 2, stream_), "zero region");
    if (C_ == 2) check_hip(launch_fill_u32(accum_second_.as<uint32_t>() + slot * px, 0, px, stream_), "zero region");
}

// run_pass + scatter_pass (nl4d/collab_pass.rs:50-200) for the pass centred on logical `centre`.
void Nl4dPass::run_pass(uint32_t centre_logical, Clear clear)
{
    const uint32_t centre = ring_.phys((int)centre_logical);
    if (clear == Clear::Whole) {
        for (uint32_t s = 0; s < N_; ++s) zero_region(s);
    } else if (clear == Clear::Newest) {
        const uint32_t newest = (centre + r_) % N_;
        if (!zeroed_[newest]) zero_region(newest);
    }
    std::fill(zeroed_.begin(), zeroed_.end(), false);
    const uint32_t *nslots = ring_.neighbours_of(centre_logical);
    search_motion(mg_, ring_, centre, nslots, (int)(N_ - 1), mv_.as<int32_t>(), conf_.as<float>(), max_value(in_),
                  sad_noise_floor(mg_.blksize, sigma_), thsad(mg_.blksize, thsad_scale_), true);
    avd_collab_pass p = pass_;
    p.centre_slot = centre;
    unsigned long long *counter = timer_ ? timer_->scored_counter() : nullptr;
    timed(AVD_STAGE_COLLAB_CENTRE, centre_ops_, stream_, "collab centre", [&] {
        return launch_collab_centre(p, ring_.code_ptr(), members_.as<uint32_t>(), stream_,
                                    centre_.plane ? &centre_ : nullptr);
    });
    timed(AVD_STAGE_COLLAB_SEARCH, 0, stream_, "collab search", [&] {
        return launch_collab_trajectory(p, ring_.code_ptr(), nslots, mv_.as<int32_t>(), conf_.as<float>(),
                                        members_.as<uint32_t>(), stream_, counter);
    });
    const uint64_t filter_ops =
        (uint64_t)refs_ * pass_.group_size * ((uint64_t)C_ * (4096 + 576 + 256) + 64 * (C_ + 1));
    timed(AVD_STAGE_COLLAB_FILTER, filter_ops, stream_, "collab filter", [&] {
        return launch_collab_filter_binned(p, ring_.value_ptr(), nslots, members_.as<uint32_t>(), bins_,
                                           aggregate_.as<unsigned long long>(), accum_second_.as<int32_t>(),
                                           group_weight_.as<float>(), stream_, mixed_, ring_.code_ptr());
    });
    if (watch_overflow_) {  // test hook: count the (member, tile) pairs past the lists
        uint32_t pairs = 0;
        check_hip(hipMemcpyAsync(&pairs, bins_.overflow_count, sizeof(pairs), hipMemcpyDeviceToHost, stream_),
                  "read overflow count");
        check_hip(hipStreamSynchronize(stream_), "read overflow count");
        overflow_seen_ += pairs;
    }
    ++passes_;
}

uint32_t Nl4dPass::submit()
{
    if (!ring_.ready()) return 0;
    Clear clear = passes_ == 0 ? Clear::Whole : Clear::Newest;
    if (passes_ == 0 && !continuation_) {
        for (uint32_t c = 0; c < r_; ++c) {
            run_pass(c, clear);
            clear = Clear::None;
        }
    }
    run_pass(r_, clear);
    if (passes_ > r_) {
        const uint32_t centre = ring_.phys((int)r_);
        pending_.push_back((centre + N_ - r_) % N_);
        return 1;
    }
    return 0;
}

uint32_t Nl4dPass::push(const avd_device_plane *planes)
{
    timed(AVD_STAGE_INGEST, ring_.px, stream_, "ingest", [&] {
        ring_.write(planes, in_, stream_);
        return hipSuccess;
    });
    submit();
    return (uint32_t)pending_.size();
}

void Nl4dPass::push_context(const avd_device_plane *planes)
{
    if (ring_.real == 0) continuation_ = true;
    ring_.write(planes, in_, stream_);
}

void Nl4dPass::emit(const avd_device_plane *planes)
{
    if (pending_.empty() && !tail_.empty()) {
        pending_.push_back(tail_.front());
        tail_.erase(tail_.begin());
        tail_emitting_ = true;
    }
    if (pending_.empty()) throw Failure(AVD_ERR_NOTHING_TO_EMIT, "no frame is ready to emit");
    const uint32_t slot = pending_.front();
    pending_.erase(pending_.begin());
    emitted_frame_ = ring_.slot_frames[slot];
    const size_t px = (size_t)W_ * H_;
    // read_region + normalise_region (nl4d/denoiser.rs:402-410, :720-760): the region is left zeroed.
    timed(AVD_STAGE_NORMALISE, px, stream_, "normalise", [&] {
        return launch_collab_normalise(aggregate_.as<unsigned long long>() + slot * px,
                                       C_ == 2 ? accum_second_.as<int32_t>() + slot * px : nullptr,
                                       output_.as<float>(), (uint32_t)px, C_, stream_);
    });
    zeroed_[slot] = true;
    uint32_t *p0 = static_cast<uint32_t *>(planes[0].data);
    uint32_t *p1 = C_ > 1 ? static_cast<uint32_t *>(planes[1].data) : p0;
    timed(AVD_STAGE_EGRESS, px, stream_, "egress", [&] {
        return launch_egress(output_.as<float>(), p0, p1, (uint32_t)px, C_, (int)samples_per_word(out_),
                             max_value(out_), stream_);
    });
    if (tail_emitting_ && tail_.empty() && pending_.empty()) {
        tail_emitting_ = false;
        reset_stream();
    }
}

// finish_passes (nl4d/denoiser.rs:534-585).
uint32_t Nl4dPass::finish()
{
    const uint32_t real = ring_.real;
    const uint32_t emit = std::min(real, 2 * r_);
    if (emit == 0) {
        reset_stream();
        return 0;
    }
    const bool short_stream = real < N_;
    if (short_stream) {
        while (!ring_.ready()) {
            ring_.duplicate_last(stream_);
            ++ring_.loaded;
        }
    }
    std::vector<uint32_t> centres;
    uint32_t last_real;
    if (short_stream) {
        for (uint32_t c = 0; c < emit; ++c) centres.push_back(c);
        last_real = emit - 1;
    } else {
        for (uint32_t c = r_ + 1; c <= 2 * r_; ++c) centres.push_back(c);
        last_real = 2 * r_;
    }
    Clear clear = (short_stream || passes_ == 0) ? Clear::Whole : Clear::None;
    for (uint32_t c : centres) {
        run_pass(c, clear);
        clear = Clear::None;
    }
    tail_.clear();
    for (uint32_t l = last_real + 1 - emit; l <= last_real; ++l) tail_.push_back(ring_.phys((int)l));
    return emit;
}

bool Nl4dPass::has_pending() const { return !pending_.empty(); }

// ---------------------------------------------------------------------------------------------
// Planar engine
// ---------------------------------------------------------------------------------------------

// EngineSpec::validate (engine/spec.rs:320-341) for a spec that cleans both plane groups: the
// passes read the same window span and, since the chroma pass takes its motion from the luma pass,
// either neither searches motion or both run the same search.
static void check_same_motion(const avd_engine_spec &spec)
{
    struct Motion {
        bool active;
        uint32_t span;
        avd_motion_search m;
    };
    auto nlm = [](const avd_nlmeans_options &o) {
        const uint32_t r = o.temporal ? o.temporal_radius : 0;
        return Motion{o.motion_compensation != 0 && r > 0, r, o.motion};
    };
    auto nl4d = [](const avd_nl4d_options &o) { return Motion{true, 2 * o.temporal_radius, o.motion}; };
    const bool is_nl4d = spec.kind == AVD_ENGINE_NL4D;
    const Motion l = is_nl4d ? nl4d(spec.nl4d_luma) : nlm(spec.nlm_luma);
    const Motion c = is_nl4d ? nl4d(spec.nl4d_chroma) : nlm(spec.nlm_chroma);
    if (l.span != c.span)
        throw Failure(AVD_ERR_INVALID_OPTIONS, "the luma pass holds " + std::to_string(l.span) +
                                                   " frames back but the chroma pass holds " + std::to_string(c.span));
    if (l.active != c.active || (l.active && !same_motion_search(l.m, c.m)))
        throw Failure(AVD_ERR_INVALID_OPTIONS, "the chroma pass takes its motion from the luma pass, so both need "
                                               "the same motion search and temporal radius");
}

Engine::Engine(const avd_engine_spec &spec) : spec_(spec)
{
    stream_ = static_cast<hipStream_t>(spec.stream);
    const avd_geometry &g = spec.geometry;
    validate_format(g.input);
    validate_format(g.output);
    if (g.width == 0 || g.height == 0) throw Failure(AVD_ERR_INVALID_GEOMETRY, "frame dimensions must be non-zero");
    if (g.chroma > AVD_CHROMA_444) throw Failure(AVD_ERR_INVALID_GEOMETRY, "unknown chroma layout");
    if (spec.kind > AVD_ENGINE_NL4D) throw Failure(AVD_ERR_INVALID_OPTIONS, "unknown engine kind");
    const bool gray = g.chroma == AVD_CHROMA_GRAY;
    if (!spec.clean_luma && !spec.clean_chroma)
        throw Failure(AVD_ERR_INVALID_OPTIONS, "a spec must clean luma, chroma or both");
    if (spec.clean_chroma && gray) throw Failure(AVD_ERR_INVALID_GEOMETRY, "a gray frame has no chroma planes to clean");
    const bool passes_through = !gray && (spec.clean_luma != 0) != (spec.clean_chroma != 0);
    if (passes_through && (g.input.kind != g.output.kind || g.input.depth != g.output.depth))
        throw Failure(AVD_ERR_INVALID_GEOMETRY, "an uncleaned plane needs the same input and output format");
    if (spec.clean_luma && spec.clean_chroma && !gray) check_same_motion(spec);
    planes_ = gray ? 1 : 3;
    chroma_w_ = g.width;
    chroma_h_ = g.height;
    if (g.chroma == AVD_CHROMA_420) {
        chroma_w_ = (g.width + 1) / 2;
        chroma_h_ = (g.height + 1) / 2;
    } else if (g.chroma == AVD_CHROMA_422) {
        chroma_w_ = (g.width + 1) / 2;
    }
    hipStream_t chroma_stream = stream_;
    if (spec.clean_luma && spec.clean_chroma && !gray) {
        check_hip(hipStreamCreateWithFlags(&side_, hipStreamNonBlocking), "create chroma stream");
        check_hip(hipEventCreateWithFlags(&fork_ev_, hipEventDisableTiming), "create event");
        check_hip(hipEventCreateWithFlags(&join_ev_, hipEventDisableTiming), "create event");
        chroma_stream = side_;
    }
    if (spec.kind == AVD_ENGINE_NLMEANS) {
        if (spec.clean_luma) luma_.reset(new NlmPass(spec.nlm_luma, g.width, g.height, 1, g.input, g.output, stream_));
        if (spec.clean_chroma)
            chroma_.reset(new NlmPass(spec.nlm_chroma, chroma_w_, chroma_h_, 2, g.input, g.output, chroma_stream));
    } else {
        if (spec.clean_luma) luma_.reset(new Nl4dPass(spec.nl4d_luma, g.width, g.height, 1, g.input, g.output, stream_));
        if (spec.clean_chroma)
            chroma_.reset(new Nl4dPass(spec.nl4d_chroma, chroma_w_, chroma_h_, 2, g.input, g.output, chroma_stream));
    }
    if (luma_) luma_->set_timer(&timer_);
    if (chroma_) chroma_->set_timer(&timer_);
    set_nlm_form(AVD_NLM_FORM_DEFAULT);
    if (luma_ && chroma_ && luma_->max_held() != chroma_->max_held())
        throw Failure(AVD_ERR_INVALID_OPTIONS, "the luma and chroma passes read different window spans");
    // luma_motion_ring (engine/spec.rs:350-365): when both passes search motion, the chroma pass
    // refines the luma pass's instead of searching its own.
    if (luma_ && chroma_ && luma_->searches_motion() && chroma_->searches_motion()) {
        const int shift_x = g.chroma == AVD_CHROMA_444 ? 0 : 1, shift_y = g.chroma == AVD_CHROMA_420 ? 1 : 0;
        share_.reset(new LumaMotionShare(luma_->motion_geom(), chroma_->motion_geom(), shift_x, shift_y,
                                         luma_->temporal_radius()));
        luma_->set_motion_share(share_.get(), false);
        chroma_->set_motion_share(share_.get(), true);
    }
    if (passes_through) {
        pt_slots_ = lead().max_held() + 1;
        const bool luma_uncleaned = !spec.clean_luma;
        pt_planes_ = luma_uncleaned ? 1 : 2;
        const uint32_t w = luma_uncleaned ? g.width : chroma_w_, h = luma_uncleaned ? g.height : chroma_h_;
        pt_words_ = plane_bytes(g.input, (uint64_t)w * h) / 4;
        passthrough_.allocate(pt_slots_ * pt_planes_ * pt_words_ * 4);
    }
}

Engine::~Engine()
{
    if (side_) {
        // Work already enqueued on the chroma stream finishes before the passes free their buffers.
        hipStreamSynchronize(side_);
        luma_.reset();
        chroma_.reset();
        hipStreamDestroy(side_);
    }
    if (fork_ev_) hipEventDestroy(fork_ev_);
    if (join_ev_) hipEventDestroy(join_ev_);
}

void Engine::fork()
{
    if (!side_) return;
    check_hip(hipEventRecord(fork_ev_, stream_), "fork");
    check_hip(hipStreamWaitEvent(side_, fork_ev_, 0), "fork");
}

void Engine::join()
{
    if (!side_) return;
    check_hip(hipEventRecord(join_ev_, side_), "join");
    check_hip(hipStreamWaitEvent(stream_, join_ev_, 0), "join");
}

Pass &Engine::lead() { return luma_ ? *luma_ : *chroma_; }
const Pass &Engine::lead() const { return luma_ ? *luma_ : *chroma_; }

void Engine::check_usable() const
{
    if (poisoned_) throw Failure(AVD_ERR_NEEDS_RESET, "an earlier call failed, reset the engine before using it again");
}

void Engine::check_planes_for(const avd_device_plane *planes, uint32_t count, avd_sample_format f) const
{
    if (count != planes_)
        throw Failure(AVD_ERR_PLANE_MISMATCH, "expected " + std::to_string(planes_) + " planes, got " +
                                                  std::to_string(count));
    check_planes(planes, 1, 1, spec_.geometry.width, spec_.geometry.height, f);
    if (planes_ == 3) check_planes(planes + 1, 2, 2, chroma_w_, chroma_h_, f);
}

void Engine::store_uncleaned(const avd_device_plane *planes, uint32_t frame)
{
    if (!passthrough_.ptr) return;
    const avd_device_plane *src = spec_.clean_luma ? planes + 1 : planes;
    const uint32_t slot = frame % pt_slots_;
    for (uint32_t i = 0; i < pt_planes_; ++i)
        check_hip(launch_copy_words(static_cast<const uint32_t *>(src[i].data),
                                    passthrough_.as<uint32_t>() + ((size_t)slot * pt_planes_ + i) * pt_words_, pt_words_,
                                    stream_),
                  "store passthrough");
}

void Engine::load_uncleaned(const avd_device_plane *planes, uint32_t frame)
{
    if (!passthrough_.ptr) return;
    const avd_device_plane *dst = spec_.clean_luma ? planes + 1 : planes;
    const uint32_t slot = frame % pt_slots_;
    for (uint32_t i = 0; i < pt_planes_; ++i)
        check_hip(launch_copy_words(passthrough_.as<uint32_t>() + ((size_t)slot * pt_planes_ + i) * pt_words_,
                                    static_cast<uint32_t *>(dst[i].data), pt_words_, stream_),
                  "load passthrough");
}

template <typename F>
auto Engine::guarded(F &&f) -> decltype(f())
{
    check_usable();
    try {
        return f();
    } catch (const Failure &e) {
        if (e.status == AVD_ERR_GPU) poisoned_ = true;
        throw;
    }
}

uint32_t Engine::push(const avd_device_plane *planes, uint32_t count)
{
    return guarded([&] {
        check_planes_for(planes, count, spec_.geometry.input);
        if (lead().has_pending()) throw Failure(AVD_ERR_OUTPUTS_PENDING, "every ready frame must be emitted before the next push");
        const uint32_t frame = next_frame_++;
        pushed_ = true;
        uint32_t ready = 0;
        fork();
        if (luma_) ready = luma_->push(planes);
        if (chroma_) {
            const uint32_t rc = chroma_->push(planes + 1);
            if (luma_ && rc != ready) {
                poisoned_ = true;
                throw Failure(AVD_ERR_GPU, "the luma and chroma passes disagree on ready frames");
            }
            ready = rc;
        }
        join();
        store_uncleaned(planes, frame);
        return ready;
    });
}

void Engine::push_context(const avd_device_plane *planes, uint32_t count)
{
    guarded([&] {
        check_planes_for(planes, count, spec_.geometry.input);
        if (lead().has_pending()) throw Failure(AVD_ERR_OUTPUTS_PENDING, "every ready frame must be emitted before the next push");
        if (pushed_) throw Failure(AVD_ERR_CONTEXT_AFTER_PUSH, "context frames must come before the first push of a stream");
        const uint32_t frame = next_frame_++;
        fork();
        if (luma_) luma_->push_context(planes);
        if (chroma_) chroma_->push_context(planes + 1);
        join();
        store_uncleaned(planes, frame);
        return 0;
    });
}

void Engine::emit(const avd_device_plane *planes, uint32_t count)
{
    guarded([&] {
        check_planes_for(planes, count, spec_.geometry.output);
        fork();
        if (luma_) luma_->emit(planes);
        if (chroma_) chroma_->emit(planes + 1);
        join();
        load_uncleaned(planes, lead().emitted_frame());
        if (tail_left_ > 0 && --tail_left_ == 0) restart();
        return 0;
    });
}

uint32_t Engine::finish()
{
    return guarded([&] {
        if (lead().has_pending()) throw Failure(AVD_ERR_OUTPUTS_PENDING, "every ready frame must be emitted before the next push");
        uint32_t n = 0;
        if (luma_) n = luma_->finish();
        if (chroma_) {
            const uint32_t nc = chroma_->finish();
            if (luma_ && nc != n) {
                poisoned_ = true;
                throw Failure(AVD_ERR_GPU, "the luma and chroma passes disagree on tail frames");
            }
            n = nc;
        }
        tail_left_ = n;
        if (n == 0) restart();
        return n;
    });
}

void Engine::set_nlm_form(uint32_t form)
{
    if (form > AVD_NLM_FORM_DEFAULT) throw Failure(AVD_ERR_INVALID_OPTIONS, "unknown NLMeans launch form");
    if (spec_.kind != AVD_ENGINE_NLMEANS) return;
    for (Pass *p : {luma_.get(), chroma_.get()}) {
        if (!p) continue;
        NlmPass &n = static_cast<NlmPass &>(*p);
        n.set_batched(form == AVD_NLM_FORM_DEFAULT ? n.temporal_radius() == 0 : form == AVD_NLM_FORM_BATCHED);
    }
}

void Engine::restart()
{
    if (share_) share_->clear();
    next_frame_ = 0;
    pushed_ = false;
    tail_left_ = 0;
}

void Engine::reset()
{
    if (luma_) luma_->reset_stream();
    if (chroma_) chroma_->reset_stream();
    restart();
    poisoned_ = false;
}

void Engine::synchronize()
{
    guarded([&] {
        check_hip(hipStreamSynchronize(stream_), "synchronize");
        return 0;
    });
}

size_t Engine::max_held_frames() const { return lead().max_held(); }

avd_window_span Engine::window_span() const
{
    avd_window_span w;
    w.behind = w.ahead = max_held_frames();
    w.edges = spec_.kind == AVD_ENGINE_NL4D ? 1u : 0u;
    return w;
}

void Engine::sample_device_time(uint32_t every)
{
    guarded([&] {
        timer_.set_sampling(every != 0, stream_);
        // While device time is sampled the chroma pass runs on the caller's stream, so each stage's
        // launches are timed alone rather than beside the other plane group's.
        if (side_ && chroma_) chroma_->set_stream(every != 0 ? stream_ : side_);
        return 0;
    });
}

avd_timing Engine::timing()
{
    return guarded([&] {
        timer_.fold_counter(stream_);
        timer_.harvest();
        return timer_.totals();
    });
}

void Nl4dPass::set_tile_cap_for_test(uint32_t cap)
{
    if (cap < kTileCapMin || cap > tile_cap_for(pass_.group_size))
        throw Failure(AVD_ERR_INVALID_OPTIONS, "tile list capacity out of range");
    bins_.cap = cap;
    watch_overflow_ = true;
}

void Engine::set_tile_cap_for_test(uint32_t cap)
{
    bool any = false;
    for (Pass *pass : {luma_.get(), chroma_.get()})
        if (auto *n = dynamic_cast<Nl4dPass *>(pass)) {
            n->set_tile_cap_for_test(cap);
            any = true;
        }
    if (!any) throw Failure(AVD_ERR_INVALID_OPTIONS, "only NL4D engines bin their aggregation");
}

uint64_t Engine::tile_overflows_for_test() const
{
    uint64_t total = 0;
    for (const Pass *pass : {luma_.get(), chroma_.get()})
        if (auto *n = dynamic_cast<const Nl4dPass *>(pass)) total += n->overflow_pairs_seen();
    return total;
}

void Engine::fail_for_test()
{
    try {
        guarded([&]() -> int { throw Failure(AVD_ERR_GPU, "forced failure"); });
    } catch (const Failure &) {
    }
}

}  // namespace avd
