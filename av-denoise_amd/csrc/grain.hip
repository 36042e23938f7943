// Film-grain measurement (grain_measure, nl4d/kernels/grain.rs:243-513; constants
// nl4d/grain/consts.rs:1-45).
//
// One wave per 8x8 cell of channel `channel` at a time, lane l on pixel (l % 8, l / 8); a workgroup
// of four waves walks 64 consecutive cells, the four waves on neighbouring cells.  Source grain is
// the motion-compensated difference of ring slots slot_next and slot_t over sqrt 2; one-frame grain
// is the high-pass std of slot_t over the cell.  A cell that
// passes the flatness, band, clipping, interior, confidence and std gates adds one count to its
// brightness bin and std bucket of the source histogram (i32 atomics); a usable cell adds one to
// the one-frame histogram.  Every cell writes its partial: the 46 lag sums of its mean-free source
// grain over the cell and a halo of 6 columns each side and 3 rows below (zeros when rejected), its
// pixel count and its strength group.  Every f32 sum follows the reference's order: the cell sums
// are 64-slot halving trees (wave shuffles pair the same slots), each lag sum adds the 64 pixels row
// by row; so the histograms and partials are bit-identical to the reference's.

#include "common.h"

namespace avd {
namespace {

constexpr int kCell = 8, kTileW = 20, kTileH = 11, kHalo = 6, kLags = 46, kPartial = 48;
constexpr int kBuckets = 64, kBins = 16, kHist = kBins * kBuckets;
constexpr int kWaves = 4, kCellsPerGroup = 64;  // 16 cells per wave

__device__ inline float tree64(float v)
{
#pragma unroll
    for (int stride = 32; stride >= 1; stride >>= 1) v += __shfl_down(v, stride);
    return __shfl(v, 0);
}

__device__ inline float min64(float v)
{
#pragma unroll
    for (int stride = 32; stride >= 1; stride >>= 1) v = fminf(v, __shfl_xor(v, stride));
    return v;
}

__device__ inline float max64(float v)
{
#pragma unroll
    for (int stride = 32; stride >= 1; stride >>= 1) v = fmaxf(v, __shfl_xor(v, stride));
    return v;
}

// Rust's saturating float-to-u32 cast, for the brightness bins.
__device__ inline uint32_t sat_u32(float v) { return !(v >= 0.0f) ? 0u : v >= 4294967296.0f ? 0xFFFFFFFFu : (uint32_t)v; }

__device__ inline int shift_clamp(int v, int s, int limit) { return min(max(v + s, 0), limit - 1); }

__device__ inline uint32_t block_for(int cx, int cy, uint32_t bx, uint32_t by, uint32_t step)
{
    return min((uint32_t)(cx * kCell) / step, bx - 1) + min((uint32_t)(cy * kCell) / step, by - 1) * bx;
}

// The highest bucket whose lower edge is at or below s (bucket_for, grain.rs:193-205).
__device__ inline uint32_t bucket_for(float s, const float *edges)
{
    uint32_t b = 0;
    for (uint32_t i = 1; i < kBuckets; ++i) b = s >= edges[i] ? i : b;
    return b;
}

__device__ inline uint32_t luma_bin(float mean) { return min(sat_u32(mean * (float)kBins), (uint32_t)kBins - 1); }

__global__ __launch_bounds__(64 * kWaves) void k_grain_measure(
    const float *__restrict__ ring, const float *__restrict__ out_t, const int32_t *__restrict__ saved_mv,
    const float *__restrict__ saved_conf, const float *__restrict__ edges, const float *__restrict__ luma_means,
    int *__restrict__ hist, float *__restrict__ partials, int W, int H, int C, int ch, uint32_t slot_t,
    uint32_t slot_next, uint32_t bx, uint32_t by, uint32_t step, uint32_t entry, int has_source, int has_luma_means,
    uint32_t means_bx, int luma_band)
{
    // The workgroup's counts gather in LDS and reach `hist` as one atomic per non-empty bin: a flat
    // region sends every cell to the same bin, where per-cell global atomics would serialise.
    __shared__ int s_hist[2 * kHist];
    __shared__ float s_tile[kWaves][kTileW * kTileH];
    const int t = (int)threadIdx.x, lane = t & 63, wave = t >> 6;
    for (int i = t; i < 2 * kHist; i += 64 * kWaves) s_hist[i] = 0;
    __syncthreads();

    const int cells_x = W / kCell, cells_y = H / kCell, cells = cells_x * cells_y;
    const int lx = lane & 7, ly = lane >> 3;
    const size_t plane = (size_t)W * H;
    const float *ft = ring + (size_t)slot_t * plane * C + ch, *fnx = ring + (size_t)slot_next * plane * C + ch;
    const uint32_t blocks = bx * by, base = entry * blocks * 2;
    constexpr float kR2 = 0.70710677f;
    float *tile = s_tile[wave];

    // Wave w takes cells first + w, first + w + 4, ...: the four waves work on neighbouring cells.
    const int first = (int)blockIdx.x * kCellsPerGroup;
    for (int k = 0; k < kCellsPerGroup / kWaves; ++k) {
        const int cell = first + wave + k * kWaves;
        if (cell >= cells) break;  // uniform per wave; only the final barrier follows the loop
        // The previous cell's lag sums have read the tile.
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        const int cx = cell % cells_x, cy = cell / cells_x, ox = cx * kCell, oy = cy * kCell;
        const int x = ox + lx, y = oy + ly;
        const uint32_t blk = block_for(cx, cy, bx, by, step);
        const int dx = saved_mv[base + blk * 2], dy = saved_mv[base + blk * 2 + 1];

        const float cur = ft[((size_t)y * W + x) * C];
        const float nxt = fnx[((size_t)shift_clamp(y, dy, H) * W + shift_clamp(x, dx, W)) * C];
        const float g = (nxt - cur) * kR2;
        const float clean = out_t[((size_t)y * W + x) * C + ch];
        const float origin = ft[((size_t)oy * W + ox) * C];
        tile[ly * kTileW + lx + kHalo] = g;

        // One-frame high pass (neighbour_squares, grain.rs:251-269): squared steps to the right and
        // lower neighbours inside the cell, relative to the origin sample.
        const float fv = cur - origin;
        const float right = __shfl_down(fv, 1), down = __shfl_down(fv, 8);
        float hp = 0.0f;
        if (lx + 1 < kCell) hp += (right - fv) * (right - fv);
        if (ly + 1 < kCell) hp += (down - fv) * (down - fv);

        const float sum = tree64(g), sum_sq = tree64(g * g), clean_sum = tree64(clean), hp_sum = tree64(hp);
        const float clean_lo = min64(clean), clean_hi = max64(clean);
        const float src_lo = min64(fminf(cur, nxt)), src_hi = max64(fmaxf(cur, nxt));

        const float count = 64.0f;
        const float var = (sum_sq - sum * sum / count) / (count - 1.0f);
        const float grain_std = sqrtf(fmaxf(var, 0.0f));
        const float mean = clean_sum / count;
        const float frame_std = sqrtf(hp_sum / 224.0f);
        const float range = clean_hi - clean_lo;
        const bool interior = cx >= 1 && cx + 2 <= cells_x && cy + 2 <= cells_y;
        const float conf = saved_conf[entry * blocks + blk];
        const bool band = !luma_band || (mean > 20.0f / 255.0f && mean < 235.0f / 255.0f);
        const bool usable = has_source && range < 5.0f / 255.0f && band && src_lo >= 4.0f / 255.0f &&
                            src_hi <= 251.0f / 255.0f;
        const bool passes = usable && interior && conf >= 0.8f && grain_std > 0.05f / 255.0f;
        const bool frame_passes = usable && (means_bx == 0 || has_luma_means);
        uint32_t group = 0;
        if (passes) {
            const uint32_t b = bucket_for(grain_std, edges);
            group = b / 4;
            if (lane == 0) atomicAdd(s_hist + luma_bin(mean) * kBuckets + b, 1);
        }
        if (frame_passes && lane == 0) {
            const uint32_t b = bucket_for(frame_std, edges);
            uint32_t bin = luma_bin(mean);
            if (means_bx > 0) {
                const uint32_t mb = (uint32_t)(cy * kCell / 16) * means_bx + (uint32_t)(cx * kCell / 16);
                bin = min(sat_u32(luma_means[mb * 3] / 16.0f), (uint32_t)kBins - 1);
            }
            atomicAdd(s_hist + kHist + bin * kBuckets + b, 1);
        }

        // The accepted cell's halo of source grain, each pixel with its own cell's vector, clamped
        // into the frame (grain.rs:441-475).  Four unrolled steps: every vector load goes out, then
        // every sample load.
        if (passes) {
            constexpr int kSteps = (kTileW * kTileH + 63) / 64;
            int hx[kSteps], hy[kSteps], ex[kSteps], ey[kSteps];
            bool halo[kSteps];
#pragma unroll
            for (int j = 0; j < kSteps; ++j) {
                const int f = lane + 64 * j, tx = f % kTileW, ty = f / kTileW;
                halo[j] = f < kTileW * kTileH && !(ty < kCell && tx >= kHalo && tx < kHalo + kCell);
                hx[j] = shift_clamp(0, ox + tx - kHalo, W);
                hy[j] = shift_clamp(oy + min(ty, kTileH - 1), 0, H);
                const uint32_t b2 = block_for(hx[j] / kCell, hy[j] / kCell, bx, by, step);
                ex[j] = saved_mv[base + b2 * 2];
                ey[j] = saved_mv[base + b2 * 2 + 1];
            }
            float vn[kSteps], vt[kSteps];
#pragma unroll
            for (int j = 0; j < kSteps; ++j) {
                vn[j] = fnx[((size_t)shift_clamp(hy[j], ey[j], H) * W + shift_clamp(hx[j], ex[j], W)) * C];
                vt[j] = ft[((size_t)hy[j] * W + hx[j]) * C];
            }
#pragma unroll
            for (int j = 0; j < kSteps; ++j)
                if (halo[j]) tile[lane + 64 * j] = (vn[j] - vt[j]) * kR2;
        }
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");

        float *part = partials + (size_t)cell * kPartial;
        if (lane < kLags) {
            // Lane l sums lag l (dy 0: dx 0..6; dy 1..3: dx -6..6) over the cell's pixels in row order.
            const int row = lane > kHalo ? 1 + (lane - kHalo - 1) / 13 : 0;
            const int col = lane > kHalo ? (lane - kHalo - 1) % 13 : lane + kHalo;
            const float gmean = sum / 64.0f;
            float acc = 0.0f;
            if (passes) {
#pragma unroll 16
                for (int p = 0; p < 64; ++p) {
                    const int px = p & 7, py = p >> 3;
                    const float c = tile[py * kTileW + px + kHalo] - gmean;
                    const float n = tile[(py + row) * kTileW + px + col] - gmean;
                    acc += c * n;
                }
            }
            part[lane] = acc;
        } else if (lane == kLags) {
            part[kLags] = passes ? 64.0f : 0.0f;
        } else if (lane == kLags + 1) {
            part[kLags + 1] = (float)group;
        }
    }

    __syncthreads();
    for (int i = t; i < 2 * kHist; i += 64 * kWaves) {
        const int n = s_hist[i];
        if (n) atomicAdd(hist + i, n);
    }
}

// grain_reduce_partials (grain.rs:515-560): workgroup `lane` of 47 takes one record lane; thread t
// of 128 adds lane `lane` of cells t, t + 128, ... in turn into its strength group's column (the
// group clamped to 15), then each group's 128 columns fold by a halving tree (the 64-stride step
// through LDS, the rest as wave shuffles, the same slot pairs) and group g's total is added into
// chunk[record_offset + 47 g + lane].  The 16 columns of a thread stay in registers: each cell adds
// its value to its own group's and an exact +0.0 to the others (a column starts at +0.0 and so
// never holds -0.0, which an added +0.0 would change), keeping the reference's order without an LDS
// read-modify-write on the chain.
constexpr int kReduceThreads = 128, kGroups = 16, kRecordLen = kLags + 1;

__global__ __launch_bounds__(kReduceThreads) void k_grain_reduce(const float *__restrict__ partials,
                                                                 float *__restrict__ chunk, uint32_t record_offset,
                                                                 uint32_t cells)
{
    __shared__ float s_col[kGroups][kReduceThreads];
    const uint32_t lane = blockIdx.x, t = threadIdx.x;
    float acc[kGroups];
#pragma unroll
    for (int g = 0; g < kGroups; ++g) acc[g] = 0.0f;
    auto add = [&](float v, float gv) {
        const uint32_t group = min(sat_u32(gv), (uint32_t)kGroups - 1);
#pragma unroll
        for (int g = 0; g < kGroups; ++g) acc[g] += group == (uint32_t)g ? v : 0.0f;
    };
    constexpr uint32_t kBatch = 8;
    uint32_t c = t;
    for (; c + (kBatch - 1) * kReduceThreads < cells; c += kBatch * kReduceThreads) {
        float v[kBatch], gv[kBatch];
#pragma unroll
        for (uint32_t j = 0; j < kBatch; ++j) {
            const size_t at = (size_t)(c + j * kReduceThreads) * kPartial;
            v[j] = partials[at + lane];
            gv[j] = partials[at + kLags + 1];
        }
#pragma unroll
        for (uint32_t j = 0; j < kBatch; ++j) add(v[j], gv[j]);
    }
    for (; c < cells; c += kReduceThreads) {
        const size_t at = (size_t)c * kPartial;
        add(partials[at + lane], partials[at + kLags + 1]);
    }
#pragma unroll
    for (int g = 0; g < kGroups; ++g) s_col[g][t] = acc[g];
    __syncthreads();
    if (t < 64) {
        float total[kGroups];
#pragma unroll
        for (int g = 0; g < kGroups; ++g) {
            float x = s_col[g][t] + s_col[g][t + 64];
#pragma unroll
            for (int stride = 32; stride >= 1; stride >>= 1) x += __shfl_down(x, stride);
            total[g] = x;
        }
        if (t == 0) {
#pragma unroll
            for (int g = 0; g < kGroups; ++g) chunk[record_offset + g * kRecordLen + lane] += total[g];
        }
    }
}

// The AR(3) fit in f64 (undo_mean_removal, yule_walker, solve and quantise_ar, nl4d/grain/fit.rs:
// 79-300): one wave per 47-value record.  Each of the 20 mean-removal rounds has lane l < 46
// compute lag l's cell-mean term over the 15 x 15 cell offsets in the reference's order and lane 46
// the zero lag's; elimination gives row r to lane r (the factors
// first, then each row's update), the pivot is the last of equal maxima (as Iterator::max_by
// keeps), back substitution and quantisation run in lane 0.  Every f64 operation is the
// reference's, in its order, so the results are bit-identical.
__device__ inline int fit_lane(int dy, int dx)
{
    if (dy < 0 || (dy == 0 && dx < 0)) {
        dy = -dy;
        dx = -dx;
    }
    if (dy > 3 || dx > 6 || dx < -6) return -1;
    return dy == 0 ? dx : 7 + (dy - 1) * 13 + dx + 6;
}

__device__ inline double cell_term(const double *rho, int dy, int dx)
{
    double total = 0.0;
    for (int oy = -7; oy <= 7; ++oy)
        for (int ox = -7; ox <= 7; ++ox) {
            const int lane = fit_lane(dy + oy, dx + ox);
            if (lane < 0) continue;
            const double pairs = (double)((8 - abs(oy)) * (8 - abs(ox)));
            total += pairs / 4096.0 * rho[lane];
        }
    return total;
}

__global__ __launch_bounds__(64) void k_grain_fit(const double *__restrict__ records, double *__restrict__ corrected,
                                                  double *__restrict__ weights, int32_t *__restrict__ quantised,
                                                  uint32_t *__restrict__ shifts, int32_t *__restrict__ solved)
{
    __shared__ double s_rho[kLags], s_next[kLags], s_m[24][25], s_factor[24], s_sol[24];
    __shared__ int s_oy[24], s_ox[24], s_ok;
    __shared__ double s_zero_term;
    const int lane = (int)threadIdx.x;
    const size_t r = blockIdx.x;
    const double *rec = records + r * kRecordLen;
    double *cor = corrected + r * kRecordLen;
    auto sync = [] {
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
        __builtin_amdgcn_s_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
    };
    if (lane < 24) {  // AR_OFFSETS: dy -3..0, dx -3..3, the causal half
        const int k = lane < 21 ? lane : lane - 21, dy = lane < 21 ? -3 + k / 7 : 0;
        s_oy[lane] = dy;
        s_ox[lane] = lane < 21 ? k % 7 - 3 : -3 + k;
    }
    const double zero_lag = rec[0];
    const double measured = lane < kLags ? rec[lane] / zero_lag : 0.0;
    if (lane < kLags) s_rho[lane] = lane == 0 ? 1.0 : measured;
    sync();
    if (zero_lag > 0.0) {
        const int dy = lane < 7 ? 0 : lane < 20 ? 1 : lane < 33 ? 2 : 3;
        const int dx = lane < 7 ? lane : (lane - 7) % 13 - 6;
        for (int round = 0; round < 20; ++round) {
            // Lane 46 computes the zero lag's term once for every lane.
            const double term = lane < kLags ? cell_term(s_rho, dy, dx) : lane == kLags ? cell_term(s_rho, 0, 0) : 0.0;
            if (lane == kLags) s_zero_term = term;
            sync();
            if (lane < kLags) s_next[lane] = measured * (1.0 - s_zero_term) + term;
            sync();
            if (lane < kLags) s_rho[lane] = s_next[lane];
            sync();
        }
        if (lane < kLags) cor[lane] = s_rho[lane] * zero_lag;
        if (lane == kLags) cor[kLags] = rec[kLags];
    } else if (lane < kRecordLen) {
        cor[lane] = rec[lane];
    }
    // The corrected record, read back as the solve's input.
    __shared__ double s_cor[kRecordLen];
    if (lane < kLags) s_cor[lane] = zero_lag > 0.0 ? s_rho[lane] * zero_lag : rec[lane];
    if (lane == kLags) s_cor[kLags] = rec[kLags];
    sync();
    const double pixels = s_cor[kLags];
    if (lane == 0) s_ok = pixels > 0.0;
    if (lane < 24 && pixels > 0.0) {
        for (int j = 0; j < 24; ++j) s_m[lane][j] = s_cor[fit_lane(s_oy[lane] - s_oy[j], s_ox[lane] - s_ox[j])] / pixels;
        s_m[lane][24] = s_cor[fit_lane(s_oy[lane], s_ox[lane])] / pixels;
    }
    sync();
    for (int col = 0; col < 24 && s_ok; ++col) {
        if (lane == 0) {
            int pivot = col;
            for (int i = col + 1; i < 24; ++i)
                if (fabs(s_m[i][col]) >= fabs(s_m[pivot][col])) pivot = i;
            if (fabs(s_m[pivot][col]) < 1e-18) s_ok = 0;
            else if (pivot != col)
                for (int j = 0; j < 25; ++j) {
                    const double t = s_m[col][j];
                    s_m[col][j] = s_m[pivot][j];
                    s_m[pivot][j] = t;
                }
        }
        sync();
        if (!s_ok) break;
        if (lane > col && lane < 24) s_factor[lane] = s_m[lane][col] / s_m[col][col];
        sync();
        if (lane > col && lane < 24)
            for (int j = col; j < 25; ++j) s_m[lane][j] -= s_factor[lane] * s_m[col][j];
        sync();
    }
    double *w = weights + r * 24;
    int32_t *q = quantised + r * 24;
    if (lane == 0) {
        bool ok = s_ok != 0;
        if (ok) {
            for (int i = 23; i >= 0; --i) {
                double v = s_m[i][24];
                for (int j = i + 1; j < 24; ++j) v -= s_m[i][j] * s_sol[j];
                s_sol[i] = v / s_m[i][i];
            }
            for (int i = 0; i < 24; ++i) ok = ok && isfinite(s_sol[i]);
        }
        uint32_t shift = 6;
        if (ok) {
            for (uint32_t cand = 9; cand >= 7; --cand) {
                bool fits = true;
                for (int i = 0; i < 24; ++i) {
                    const double v = round(s_sol[i] * (double)(1u << cand));
                    fits = fits && v >= -128.0 && v <= 127.0;
                }
                if (fits) {
                    shift = cand;
                    break;
                }
            }
        }
        for (int i = 0; i < 24; ++i) {
            const double v = ok ? fmin(fmax(round(s_sol[i] * (double)(1u << shift)), -128.0), 127.0) : 0.0;
            w[i] = ok ? s_sol[i] : 0.0;
            q[i] = (int32_t)v;
        }
        shifts[r] = ok ? shift : 0u;
        solved[r] = ok ? 1 : 0;
    }
}

// grain_save_vectors (grain.rs:50-75): one neighbour's vectors and confidences into saved entry
// `entry`, a grid-stride copy.
__global__ void k_grain_save_vectors(const int32_t *__restrict__ mv_field, const float *__restrict__ confidence,
                                     int32_t *__restrict__ saved_mv, float *__restrict__ saved_conf,
                                     uint32_t mv_offset, uint32_t conf_offset, uint32_t entry, uint32_t blocks)
{
    for (uint32_t b = blockIdx.x * blockDim.x + threadIdx.x; b < blocks; b += gridDim.x * blockDim.x) {
        const uint32_t s = entry * blocks + b;
        saved_mv[s * 2] = mv_field[mv_offset + b * 2];
        saved_mv[s * 2 + 1] = mv_field[mv_offset + b * 2 + 1];
        saved_conf[s] = confidence[conf_offset + b];
    }
}

// grain_project (grain_persistence.rs:54-153): one wave per 8x8 cell, four cells per workgroup.  The
// eight pattern projections (pixel y * 8 + x added where the pattern bit is set, taken away where
// clear) and the denoised mean are 64-slot shuffle trees, the extremes butterflies; lane 0 writes
// the cell's history entry (8 projections and the flat flag).
constexpr int kProjections = 8, kEntry = kProjections + 1;

__global__ __launch_bounds__(256) void k_grain_project(const float *__restrict__ ring, const float *__restrict__ out,
                                                       const uint32_t *__restrict__ patterns,
                                                       float *__restrict__ history, int W, int H, int C, int ch,
                                                       uint32_t frame_slot, uint32_t history_slot, int luma_band)
{
    const int lane = (int)threadIdx.x & 63, cells_x = W / kCell, cells = cells_x * (H / kCell);
    const int cell = (int)blockIdx.x * 4 + ((int)threadIdx.x >> 6);
    if (cell >= cells) return;  // a whole wave; the kernel has no barrier
    const int x = cell % cells_x * kCell + (lane & 7), y = cell / cells_x * kCell + (lane >> 3);
    const float sample = ring[(((size_t)frame_slot * H + y) * W + x) * C + ch];
    const float clean = out[((size_t)y * W + x) * C + ch];
    float proj[kProjections];
#pragma unroll
    for (int p = 0; p < kProjections; ++p) {
        const uint32_t bit = (patterns[p * 2 + (lane >> 5)] >> (lane & 31)) & 1u;
        proj[p] = tree64(((float)bit * 2.0f - 1.0f) * sample);
    }
    const float clean_sum = tree64(clean);
    const float clean_lo = min64(clean), clean_hi = max64(clean), src_lo = min64(sample), src_hi = max64(sample);
    if (lane == 0) {
        float *e = history + ((size_t)history_slot * cells + cell) * kEntry;
#pragma unroll
        for (int p = 0; p < kProjections; ++p) e[p] = proj[p];
        const float mean = clean_sum / 64.0f, range = clean_hi - clean_lo;
        const bool band = !luma_band || (mean > 20.0f / 255.0f && mean < 235.0f / 255.0f);
        const bool flat = range < 5.0f / 255.0f && band && src_lo >= 4.0f / 255.0f && src_hi <= 251.0f / 255.0f;
        e[kProjections] = flat ? 1.0f : 0.0f;
    }
}

// grain_persist (grain_persistence.rs:155-230): workgroup (block, n) of 256 threads covers 1024
// cells at lag n + 1; thread t walks cells t, t + 256, t + 512, t + 768 of the block in turn, then
// each lane (count, auto, cross) folds by a 256-slot halving tree (the 128- and 64-stride steps
// through LDS, the rest as shuffles).
constexpr int kPersistThreads = 256, kPersistCells = 1024, kHistoryFrames = 9;

__global__ __launch_bounds__(kPersistThreads) void k_grain_persist(const float *__restrict__ history,
                                                                   float *__restrict__ partials, uint32_t now_slot,
                                                                   uint32_t cells, uint32_t blocks)
{
    __shared__ float s[3][kPersistThreads];
    const uint32_t t = threadIdx.x, block = blockIdx.x, lag = blockIdx.y;
    const uint32_t earlier = (now_slot + kHistoryFrames - 1 - lag) % kHistoryFrames;
    float count = 0.0f, auto_sum = 0.0f, cross_sum = 0.0f;
#pragma unroll
    for (uint32_t step = 0; step < kPersistCells / kPersistThreads; ++step) {
        const uint32_t cell = block * kPersistCells + t + step * kPersistThreads;
        if (cell < cells) {
            const float *now = history + ((size_t)now_slot * cells + cell) * kEntry;
            const float *ear = history + ((size_t)earlier * cells + cell) * kEntry;
            if (now[kProjections] * ear[kProjections] > 0.5f) {
                float sq = 0.0f, pr = 0.0f;
#pragma unroll
                for (int p = 0; p < kProjections; ++p) {
                    sq += now[p] * now[p] + ear[p] * ear[p];
                    pr += now[p] * ear[p];
                }
                count += 1.0f;
                auto_sum += sq * (1.0f / 1024.0f);
                cross_sum += pr * (1.0f / 512.0f);
            }
        }
    }
    s[0][t] = count;
    s[1][t] = auto_sum;
    s[2][t] = cross_sum;
    __syncthreads();
    if (t < 64) {
        float v[3];
#pragma unroll
        for (int l = 0; l < 3; ++l) {
            v[l] = (s[l][t] + s[l][t + 128]) + (s[l][t + 64] + s[l][t + 192]);
#pragma unroll
            for (int stride = 32; stride >= 1; stride >>= 1) v[l] += __shfl_down(v[l], stride);
        }
        if (t == 0) {
#pragma unroll
            for (int l = 0; l < 3; ++l) partials[((size_t)lag * blocks + block) * 3 + l] = v[l];
        }
    }
}

// grain_persist_reduce (grain_persistence.rs:232-279): workgroup n of 128 threads; thread t adds
// blocks t, t + 128, ... in turn per lane, then a 128-slot tree (the 64-stride step through LDS, the
// rest as shuffles), added into record[record_offset + 3 n + lane].
__global__ __launch_bounds__(kReduceThreads) void k_grain_persist_reduce(const float *__restrict__ partials,
                                                                         float *__restrict__ record,
                                                                         uint32_t record_offset, uint32_t blocks)
{
    __shared__ float s[3][kReduceThreads];
    const uint32_t t = threadIdx.x, lag = blockIdx.x;
    float acc[3] = {0.0f, 0.0f, 0.0f};
    for (uint32_t b = t; b < blocks; b += kReduceThreads)
#pragma unroll
        for (int l = 0; l < 3; ++l) acc[l] += partials[((size_t)lag * blocks + b) * 3 + l];
#pragma unroll
    for (int l = 0; l < 3; ++l) s[l][t] = acc[l];
    __syncthreads();
    if (t < 64) {
        float v[3];
#pragma unroll
        for (int l = 0; l < 3; ++l) {
            v[l] = s[l][t] + s[l][t + 64];
#pragma unroll
            for (int stride = 32; stride >= 1; stride >>= 1) v[l] += __shfl_down(v[l], stride);
        }
        if (t == 0) {
#pragma unroll
            for (int l = 0; l < 3; ++l) record[record_offset + lag * 3 + l] += v[l];
        }
    }
}

}  // namespace

uint32_t grain_partials_len(uint32_t width, uint32_t height)
{
    return (width / kCell) * (height / kCell) * kPartial;
}

hipError_t launch_grain_measure(const GrainMeasure &m, hipStream_t s)
{
    const uint32_t cells = (m.width / kCell) * (m.height / kCell);
    if (cells == 0) return hipSuccess;
    const uint32_t groups = (cells + kCellsPerGroup - 1) / kCellsPerGroup;
    k_grain_measure<<<groups, 64 * kWaves, 0, s>>>(
        m.ring, m.denoised, m.saved_mv, m.saved_conf, m.edges, m.luma_means, m.hist, m.partials, (int)m.width,
        (int)m.height, (int)m.channels, (int)m.channel, m.slot_t, m.slot_next, m.blocks_x, m.blocks_y, m.step,
        m.source_entry, m.has_source ? 1 : 0, m.has_luma_means ? 1 : 0, m.means_blocks_x, m.luma_band ? 1 : 0);
    return hipGetLastError();
}

hipError_t launch_grain_reduce(const float *partials, uint32_t cells, float *chunk, uint32_t record_offset,
                               hipStream_t s)
{
    if (!partials || !chunk) return hipErrorInvalidValue;
    k_grain_reduce<<<kRecordLen, kReduceThreads, 0, s>>>(partials, chunk, record_offset, cells);
    return hipGetLastError();
}

hipError_t launch_grain_fit(const double *records, uint32_t count, double *corrected, double *weights,
                            int32_t *quantised, uint32_t *shifts, int32_t *solved, hipStream_t s)
{
    if (!records || !corrected || !weights || !quantised || !shifts || !solved) return hipErrorInvalidValue;
    if (count == 0) return hipSuccess;
    k_grain_fit<<<count, 64, 0, s>>>(records, corrected, weights, quantised, shifts, solved);
    return hipGetLastError();
}

hipError_t launch_grain_save_vectors(const int32_t *mv_field, const float *confidence, int32_t *saved_mv,
                                    float *saved_conf, uint32_t mv_offset, uint32_t conf_offset, uint32_t entry,
                                    uint32_t blocks, hipStream_t s)
{
    if (!mv_field || !confidence || !saved_mv || !saved_conf) return hipErrorInvalidValue;
    if (blocks == 0) return hipSuccess;
    k_grain_save_vectors<<<std::min<uint32_t>((blocks + 255) / 256, 1024), 256, 0, s>>>(
        mv_field, confidence, saved_mv, saved_conf, mv_offset, conf_offset, entry, blocks);
    return hipGetLastError();
}

hipError_t launch_grain_project(const float *ring, const float *out, const uint32_t *patterns, float *history,
                                uint32_t width, uint32_t height, uint32_t channels, uint32_t channel,
                                uint32_t frame_slot, uint32_t history_slot, bool luma_band, hipStream_t s)
{
    const uint32_t cells = (width / kCell) * (height / kCell);
    if (cells == 0) return hipSuccess;
    k_grain_project<<<(cells + 3) / 4, 256, 0, s>>>(ring, out, patterns, history, (int)width, (int)height,
                                                   (int)channels, (int)channel, frame_slot, history_slot,
                                                   luma_band ? 1 : 0);
    return hipGetLastError();
}

uint32_t grain_persist_blocks(uint32_t cells) { return (cells + kPersistCells - 1) / kPersistCells; }

hipError_t launch_grain_persist(const float *history, float *partials, float *record, uint32_t record_offset,
                                uint32_t now_slot, uint32_t cells, uint32_t lags, hipStream_t s)
{
    const uint32_t blocks = grain_persist_blocks(cells);
    if (blocks == 0 || lags == 0) return hipSuccess;
    k_grain_persist<<<dim3(blocks, lags), kPersistThreads, 0, s>>>(history, partials, now_slot, cells, blocks);
    k_grain_persist_reduce<<<lags, kReduceThreads, 0, s>>>(partials, record, record_offset, blocks);
    return hipGetLastError();
}

}  // namespace avd
