// Spatial noise measurement: the two stages of the Immerkaer estimate
// (nlm_noise_partial / nlm_noise_reduce, noise/kernels/stats.rs:24-140; driven by
// noise/spatial.rs:93-125).
//
// Stage 1: per 32x8 tile of the frame (eight tiles per workgroup).  Every interior pixel applies the 3x3 mask
//   [1 -2 1; -2 4 -2; 1 -2 1], which cancels affine content and leaves mostly noise, and the tile
//   sums the absolute responses of each channel into its own lane of the tile's two partials.
//   Border pixels, pixels past the frame and unused lanes contribute zero.
// Stage 2: one workgroup of 256 threads sums every partial into the two per-channel totals of row
//   `slot` of the results.
// Both sums are f32 in the reference's order (thread l of stage 2 adds partials l, l + 256, ... in
// turn; each total then adds its 256 thread sums in thread order, as does each tile in stage 1),
// so the totals are bit-identical to the reference's.  The work is one read of the frame (HBM).

#include <mutex>

#include "common.h"

namespace avd {

// The partials of the entry point that allocates nothing: a file-scope device array, one copy per
// GPU at its own address, resolved once per device by hipGetSymbolAddress.  Host code only ever
// uses that resolved address.  Every partial the reduce reads is written by the partial kernel of
// the same call first, so the array needs no clearing; calls that overlap on one GPU race on it.
__device__ __align__(256) float scratch[kNoiseScratchFloats];

namespace {

constexpr int kImmBlockX = 32, kImmBlockY = 8;  // the tile (noise/mod.rs:117-119)
constexpr int kImmLanes = 2;                    // partials per tile and totals per slot
constexpr int kReduceBlock = 256;               // stage 2's threads (noise/mod.rs:122)

// A workgroup takes kTilesPerBlock tiles stacked vertically: it computes their responses one tile
// at a time (thread t on pixel t of the tile), then lane (tile, channel) of the first wave adds its
// tile's 256 responses in thread order, so 16 of the ordered sums run side by side.
constexpr int kTilesPerBlock = 8;
constexpr int kTileStride = kImmBlockX * kImmBlockY * kImmLanes + 2;  // floats; 2r + lane picks the bank

template <int C>
__global__ __launch_bounds__(kImmBlockX * kImmBlockY) void k_noise_partial(const float *__restrict__ ring,
                                                                           float *__restrict__ partials,
                                                                           uint32_t frame, int W, int H)
{
    constexpr int kThreads = kImmBlockX * kImmBlockY;
    __shared__ float s_abs[kTilesPerBlock * kTileStride];
    const int t = (int)threadIdx.x;  // = row * 32 + column of a tile
    const int tiles_x = (W + kImmBlockX - 1) / kImmBlockX, tiles_y = (H + kImmBlockY - 1) / kImmBlockY;
    const int x = (int)blockIdx.x * kImmBlockX + (t & (kImmBlockX - 1));
    const float *f = ring + (size_t)frame * W * H * C;
#pragma unroll
    for (int r = 0; r < kTilesPerBlock; ++r) {
        const int y = ((int)blockIdx.y * kTilesPerBlock + r) * kImmBlockY + t / kImmBlockX;
        float response[C];
#pragma unroll
        for (int c = 0; c < C; ++c) response[c] = 0.0f;
        if (x >= 1 && x < W - 1 && y >= 1 && y < H - 1) {
            const float *row0 = f + ((size_t)(y - 1) * W + x) * C, *row1 = row0 + (size_t)W * C,
                        *row2 = row1 + (size_t)W * C;
#pragma unroll
            for (int c = 0; c < C; ++c) {
                const float centre = row1[c], left = row1[c - C], right = row1[c + C];
                const float up = row0[c], down = row2[c];
                const float up_left = row0[c - C], up_right = row0[c + C], down_left = row2[c - C],
                            down_right = row2[c + C];
                // The reference's expression order: centre 4 - (left + right + up + down) 2 + the corners.
                response[c] =
                    centre * 4.0f - (left + right + up + down) * 2.0f + (up_left + up_right + down_left + down_right);
            }
        }
#pragma unroll
        for (int lane = 0; lane < kImmLanes; ++lane)
            s_abs[r * kTileStride + t * kImmLanes + lane] = lane < C ? fabsf(response[lane < C ? lane : 0]) : 0.0f;
    }
    __syncthreads();
    if (t < kTilesPerBlock * kImmLanes) {
        const int r = t / kImmLanes, lane = t % kImmLanes;
        const int ty = (int)blockIdx.y * kTilesPerBlock + r;
        if (ty < tiles_y) {
            float sum = 0.0f;
#pragma unroll 16
            for (int k = 0; k < kThreads; ++k) sum += s_abs[r * kTileStride + k * kImmLanes + lane];
            partials[((size_t)ty * tiles_x + blockIdx.x) * kImmLanes + lane] = sum;
        }
    }
}

// Thread t adds partials t, t + 256, ... of each lane in turn, 32 loads in flight at a time.
__global__ __launch_bounds__(kReduceBlock) void k_noise_reduce(const float *__restrict__ partials,
                                                                float *__restrict__ results, uint32_t slot,
                                                                uint32_t count)
{
    constexpr uint32_t kUnroll = 32;
    __shared__ float s_sum[kReduceBlock * kImmLanes];
    const uint32_t t = threadIdx.x;
    float sum0 = 0.0f, sum1 = 0.0f;
    for (uint32_t i = t; i < count; i += kReduceBlock * kUnroll) {
        float a[kUnroll], b[kUnroll];
#pragma unroll
        for (uint32_t u = 0; u < kUnroll; ++u) {
            const uint32_t k = i + u * kReduceBlock;
            a[u] = k < count ? partials[(size_t)k * kImmLanes] : 0.0f;
            b[u] = k < count ? partials[(size_t)k * kImmLanes + 1] : 0.0f;
        }
#pragma unroll
        for (uint32_t u = 0; u < kUnroll; ++u) {
            if (i + u * kReduceBlock < count) {
                sum0 += a[u];
                sum1 += b[u];
            }
        }
    }
    s_sum[t * kImmLanes] = sum0;
    s_sum[t * kImmLanes + 1] = sum1;
    __syncthreads();
    if (t < (uint32_t)kImmLanes) {
        float total = 0.0f;
        for (int k = 0; k < kReduceBlock; ++k) total += s_sum[k * kImmLanes + t];
        results[(size_t)slot * kImmLanes + t] = total;
    }
}

}  // namespace

uint32_t noise_partials_len(uint32_t width, uint32_t height)
{
    return ((width + kImmBlockX - 1) / kImmBlockX) * ((height + kImmBlockY - 1) / kImmBlockY) * kImmLanes;
}

hipError_t launch_noise_spatial(const float *ring, uint32_t width, uint32_t height, uint32_t channels,
                                uint32_t frame, float *partials, float *results, uint32_t slot, hipStream_t s)
{
    if (!ring || !partials || !results || width == 0 || height == 0 || (channels != 1 && channels != 2) ||
        (uint64_t)width * height * channels * (frame + 1ull) > (1ull << 40))
        return hipErrorInvalidValue;
    const uint32_t tiles_x = (width + kImmBlockX - 1) / kImmBlockX, tiles_y = (height + kImmBlockY - 1) / kImmBlockY;
    const dim3 grid(tiles_x, (tiles_y + kTilesPerBlock - 1) / kTilesPerBlock);
    if (channels == 1)
        k_noise_partial<1><<<grid, kImmBlockX * kImmBlockY, 0, s>>>(ring, partials, frame, (int)width, (int)height);
    else
        k_noise_partial<2><<<grid, kImmBlockX * kImmBlockY, 0, s>>>(ring, partials, frame, (int)width, (int)height);
    k_noise_reduce<<<1, kReduceBlock, 0, s>>>(partials, results, slot, tiles_x * tiles_y);
    return hipGetLastError();
}

hipError_t noise_scratch_address(float **out)
{
    constexpr int kMaxDevices = 64;
    static std::mutex mu;
    static float *resolved[kMaxDevices] = {};
    int dev = -1;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return e;
    if (dev < 0 || dev >= kMaxDevices) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lock(mu);
    if (!resolved[dev]) {
        void *ptr = nullptr;
        e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return e;
        resolved[dev] = static_cast<float *>(ptr);
    }
    *out = resolved[dev];
    return hipSuccess;
}

// ---------------------------------------------------------------------------------------------
// Temporal block statistics (nlm_temporal_noise_stats, noise/kernels/stats.rs:142-604; record layout
// noise/temporal.rs:18-131; driven by run_temporal_noise_stats, noise/temporal.rs:206-257).
//
// One workgroup per 16x16 block of the frame, thread t on pixel (t % 16, t / 16).  The residual is
// frame slot_new minus frame slot_prev of the ring.  A record holds per channel sum_d and sum_d2, the
// lag-1 product sum of channel 0's row neighbours, channel 0's cell texture, the moved flag, four
// quarter records of channel 0 (sum_d, sum_d2, luma sum, flatness, luma min, luma max) and every
// further channel's texture.  Each f32 sum follows the reference's order: a lane total adds its 256
// pixel values in pixel order, the gain check and the quarter sums are halving trees (slot s + stride
// into slot s, stride n/2 .. 1; the first two steps of the 256-long trees go through LDS, the rest
// are wave shuffles, which pair the same slots), cell means add their pixels row by row.  So the
// records are bit-identical to the reference's f32 records.  After one barrier the four waves split
// the reductions: wave 0 the lane totals, wave 1 the cell textures, wave 2 the flatness grid, wave 3
// the quarter sums and extremes.  The work is one read of both frames: a few ms at most at 4K even
// with the sequential totals, which bound the kernel (one dependent add per pixel and lane).
// ---------------------------------------------------------------------------------------------

namespace {

constexpr int kTB = 16, kTThreads = kTB * kTB;
constexpr uint32_t kXcds = 8;
constexpr int kTLaneStride = kTThreads + 1;  // lane-major totals: lanes land in different banks
constexpr float kUnmeasuredFlatness = 3.0e38f;
constexpr float kCompensationGainMax = 0.5f;

__device__ inline float tree_wave(float v)
{
#pragma unroll
    for (int stride = 32; stride >= 1; stride >>= 1) v += __shfl_down(v, stride);
    return v;
}

// The mean squared step between adjacent 4x4-cell means fully inside a bw x bh block, or 0 with no
// such pair; `cells` holds the 16 means row-major (cell_texture, stats.rs:626-656).
__device__ inline float cell_texture(const float *cells, int bw, int bh)
{
    const int cols = bw >> 2, rows = bh >> 2;
    float energy = 0.0f;
    int pairs = 0;
    for (int cy = 0; cy < rows; ++cy)
        for (int cx = 0; cx < cols; ++cx) {
            const float here = cells[cy * 4 + cx];
            if (cx + 1 < cols) {
                const float step = here - cells[cy * 4 + cx + 1];
                energy += step * step;
                pairs++;
            }
            if (cy + 1 < rows) {
                const float step = here - cells[(cy + 1) * 4 + cx];
                energy += step * step;
                pairs++;
            }
        }
    return pairs ? energy / (float)pairs : 0.0f;
}

template <int C, bool LUMA, bool COMP>
__global__ __launch_bounds__(kTThreads) void k_temporal_stats(const float *__restrict__ ring, float *__restrict__ stats,
                                                              const int32_t *__restrict__ mv, uint32_t slot_new,
                                                              uint32_t slot_prev, uint32_t mv_offset, uint32_t mv_step,
                                                              uint32_t mv_bx, uint32_t mv_by, int W, int H)
{
    constexpr int kLanes = 2 * C + 1, kRec = 3 * C + 26;
    __shared__ float s_lane[kLanes * kTLaneStride];
    __shared__ float s_val[C * kTThreads];                    // the new frame's values per channel
    __shared__ float s_quarter[LUMA ? 5 * kTThreads : 1];     // quarter-major: sum, d, d2, min, max
    __shared__ float s_mean[LUMA ? kTThreads : 1];            // the temporal mean of channel 0
    __shared__ float s_cells[C * 16];
    __shared__ float s_gain[COMP ? 4 * kTThreads : 1];
    __shared__ int s_shift[3];

    const int t = (int)threadIdx.x, lane = t & 63, wave = t >> 6, lx = t & (kTB - 1), ly = t >> 4;
    // Workgroups reach the eight XCDs round robin; XCD x takes a contiguous run of the row-major
    // blocks, so the two blocks sharing a 128-byte row segment meet in one L2.
    const uint32_t nbx = ((uint32_t)W + kTB - 1) / kTB, blocks = nbx * (((uint32_t)H + kTB - 1) / kTB);
    const uint32_t xcd = blockIdx.x % kXcds, per = blocks / kXcds, extra = blocks % kXcds;
    const uint32_t b = xcd * per + min(xcd, extra) + blockIdx.x / kXcds;
    const int ox = (int)(b % nbx) * kTB, oy = (int)(b / nbx) * kTB;
    const int bw = min(kTB, W - ox), bh = min(kTB, H - oy);
    const int x = ox + lx, y = oy + ly;
    const bool in_frame = x < W && y < H;
    const size_t plane = (size_t)W * H * C;
    const float *fn = ring + (size_t)slot_new * plane, *fp = ring + (size_t)slot_prev * plane;
    float *out = stats + (size_t)b * kRec;

    float nv[C], pv[C];
#pragma unroll
    for (int c = 0; c < C; ++c) nv[c] = pv[c] = 0.0f;
    if (in_frame) {
        const size_t at = ((size_t)y * W + x) * C;
#pragma unroll
        for (int c = 0; c < C; ++c) {
            nv[c] = fn[at + c];
            pv[c] = fp[at + c];
        }
    }

    int sx = 0, sy = 0;
    bool matched = true;
    float moved = 0.0f;
    if constexpr (COMP) {
        // The gain check (stats.rs:257-341): the block follows its vector when channel 0's residual
        // variance along it, over the pixels whose match stays in the frame, is at most half the
        // zero-vector variance, or when no match stays in the frame; a followed block whose shifted
        // extent leaves the frame is unmatched and writes zeros.
        const uint32_t col = min((uint32_t)ox / mv_step, mv_bx - 1), row = min((uint32_t)oy / mv_step, mv_by - 1);
        const int32_t *v = mv + mv_offset + (size_t)(row * mv_bx + col) * 2;
        const int vx = v[0], vy = v[1];
        float zr = 0.0f, vr = 0.0f;
        if (in_frame) {
            zr = nv[0] - pv[0];
            const long mx = (long)x + vx, my = (long)y + vy;
            if (mx >= 0 && mx < W && my >= 0 && my < H) vr = nv[0] - fp[((size_t)my * W + (size_t)mx) * C];
        }
        s_gain[t] = zr;
        s_gain[kTThreads + t] = zr * zr;
        s_gain[2 * kTThreads + t] = vr;
        s_gain[3 * kTThreads + t] = vr * vr;
        __syncthreads();
        if (wave == 0) {
            float sums[4];
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                const float *r = s_gain + k * kTThreads;
                sums[k] = tree_wave((r[lane] + r[lane + 128]) + (r[lane + 64] + r[lane + 192]));
            }
            if (lane == 0) {
                const long left = (long)ox + vx, top = (long)oy + vy, right = left + bw, bottom = top + bh;
                const long span_w = max(min(right, (long)W) - max(left, 0l), 0l);
                const long span_h = max(min(bottom, (long)H) - max(top, 0l), 0l);
                const float matched_px = (float)(span_w * span_h), block_px = (float)(bw * bh);
                const float zmean = sums[0] / block_px;
                const float zvar = sums[1] / block_px - zmean * zmean;
                const float safe = fmaxf(matched_px, 1.0f);
                const float vmean = sums[2] / safe;
                const float vvar = sums[3] / safe - vmean * vmean;
                const bool follows = (vx != 0 || vy != 0) && (matched_px == 0.0f || vvar <= kCompensationGainMax * zvar);
                const bool fits = left >= 0 && right <= W && top >= 0 && bottom <= H;
                s_shift[0] = follows ? vx : 0;
                s_shift[1] = follows ? vy : 0;
                s_shift[2] = follows && !fits ? 0 : 1;
            }
        }
        __syncthreads();
        sx = s_shift[0];
        sy = s_shift[1];
        matched = s_shift[2] != 0;
        moved = (sx != 0 || sy != 0) && matched ? 1.0f : 0.0f;
        if ((sx != 0 || sy != 0) && matched && in_frame) {
            const size_t at = ((size_t)(y + sy) * W + (x + sx)) * C;
#pragma unroll
            for (int c = 0; c < C; ++c) pv[c] = fp[at + c];
        }
    }

    const bool valid = in_frame && matched;
    float res[C];
#pragma unroll
    for (int c = 0; c < C; ++c) {
        res[c] = valid ? nv[c] - pv[c] : 0.0f;
        s_lane[c * kTLaneStride + t] = res[c];
        s_lane[(C + c) * kTLaneStride + t] = res[c] * res[c];
        s_val[c * kTThreads + t] = valid ? nv[c] : 0.0f;
    }
    // Pixel t + 1 is the row neighbour in the same wave whenever the pair lies inside the block.
    const float next = __shfl_down(res[0], 1);
    s_lane[2 * C * kTLaneStride + t] = valid && lx + 1 < bw ? res[0] * next : 0.0f;
    if constexpr (LUMA) {
        const int slot = ((ly >> 3) * 2 + (lx >> 3)) * 64 + (ly & 7) * 8 + (lx & 7);
        s_quarter[slot] = valid ? nv[0] : 0.0f;
        s_quarter[kTThreads + slot] = res[0];
        s_quarter[2 * kTThreads + slot] = res[0] * res[0];
        s_quarter[3 * kTThreads + slot] = valid ? nv[0] : 1.0e30f;
        s_quarter[4 * kTThreads + slot] = valid ? nv[0] : -1.0e30f;
        s_mean[t] = valid ? 0.5f * (nv[0] + pv[0]) : 0.0f;
    }
    __syncthreads();

    if (wave == 0) {
        // Lane totals (stats.rs:400-410): one lane per record lane adds the 256 pixels in order.
        if (lane < kLanes) {
            const float *v = s_lane + lane * kTLaneStride;
            float total = 0.0f;
#pragma unroll 32
            for (int k = 0; k < kTThreads; ++k) total += v[k];
            out[lane] = total;
        }
    } else if (wave == 1) {
        // Cell means of each channel (write_cell_means, stats.rs:606-624), then the textures.
        if (lane < 16 * C) {
            const int c = lane >> 4, cell = lane & 15, cx = cell & 3, cy = cell >> 2;
            const float *v = s_val + c * kTThreads;
            float sum = 0.0f;
#pragma unroll
            for (int dy = 0; dy < 4; ++dy)
#pragma unroll
                for (int dx = 0; dx < 4; ++dx) sum += v[(4 * cy + dy) * kTB + 4 * cx + dx];
            s_cells[lane] = sum / 16.0f;
        }
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
        if (lane < C) {
            const float texture = cell_texture(s_cells + 16 * lane, bw, bh);
            out[lane == 0 ? 2 * C + 1 : 2 * C + 26 + lane] = texture;
        }
        if (lane == 0) out[2 * C + 2] = moved;
    } else if (wave == 2) {
        float *flat = out + 2 * C + 3 + 3;
        if constexpr (LUMA) {
            // Flatness (stats.rs:520-585): the 8x8 grid of 2x2 means of the temporal mean, each
            // cell's squared steps to its right and lower neighbour inside its quarter, then per
            // quarter a halving tree over its 16 cells (stored quarter-major) over 24 pairs.
            const int gx = lane & 7, gy = lane >> 3;
            float cell = 0.0f;
#pragma unroll
            for (int dy = 0; dy < 2; ++dy)
#pragma unroll
                for (int dx = 0; dx < 2; ++dx) cell += s_mean[(2 * gy + dy) * kTB + 2 * gx + dx];
            cell = cell / 4.0f;
            const float right = __shfl_down(cell, 1), below = __shfl_down(cell, 8);
            float e = 0.0f;
            if ((gx & 3) != 3) e += (cell - right) * (cell - right);
            if ((gy & 3) != 3) e += (cell - below) * (cell - below);
            // Reorder the scores quarter-major: slot q * 16 + (gy % 4) * 4 + gx % 4.
            const int dest = ((gy >> 2) * 2 + (gx >> 2)) * 16 + (gy & 3) * 4 + (gx & 3);
            float q = __builtin_bit_cast(float, __builtin_amdgcn_ds_permute(dest * 4, __builtin_bit_cast(int, e)));
#pragma unroll
            for (int stride = 8; stride >= 1; stride >>= 1) q += __shfl_down(q, stride);
            if ((lane & 15) == 0) {
                const int quarter = lane >> 4;
                const bool full = bw >= ((quarter & 1) + 1) * 8 && bh >= ((quarter >> 1) + 1) * 8;
                flat[quarter * 6] = !matched ? 0.0f : full ? q / 24.0f : kUnmeasuredFlatness;
            }
        } else {
            if (lane < 4) flat[lane * 6] = 0.0f;
        }
    } else {
        float *quarters = out + 2 * C + 3;
        if constexpr (LUMA) {
#pragma unroll
            for (int quarter = 0; quarter < 4; ++quarter) {
                const float *qv = s_quarter + quarter * 64 + lane;
                const float luma = tree_wave(qv[0]);
                const float d = tree_wave(qv[kTThreads]);
                const float d2 = tree_wave(qv[2 * kTThreads]);
                float lo = qv[3 * kTThreads], hi = qv[4 * kTThreads];
#pragma unroll
                for (int stride = 32; stride >= 1; stride >>= 1) {
                    lo = fminf(lo, __shfl_xor(lo, stride));
                    hi = fmaxf(hi, __shfl_xor(hi, stride));
                }
                if (lane == 0) {
                    float *f = quarters + quarter * 6;
                    f[0] = matched ? d : 0.0f;
                    f[1] = matched ? d2 : 0.0f;
                    f[2] = matched ? luma : 0.0f;
                    f[4] = matched ? lo : 0.0f;
                    f[5] = matched ? hi : 0.0f;
                }
            }
        } else {
            if (lane < 24 && lane % 6 != 3) quarters[lane] = 0.0f;
        }
    }
}

__global__ void k_temporal_zero(float *__restrict__ dst, uint32_t len)
{
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < len; i += gridDim.x * blockDim.x) dst[i] = 0.0f;
}

template <int C, bool LUMA>
void launch_temporal_typed(dim3 grid, hipStream_t s, const float *ring, float *stats, const int32_t *mv,
                           uint32_t slot_new, uint32_t slot_prev, uint32_t mv_offset, uint32_t mv_step, uint32_t mv_bx,
                           uint32_t mv_by, int W, int H)
{
    if (mv)
        k_temporal_stats<C, LUMA, true><<<grid, kTThreads, 0, s>>>(ring, stats, mv, slot_new, slot_prev, mv_offset,
                                                                   mv_step, mv_bx, mv_by, W, H);
    else
        k_temporal_stats<C, LUMA, false><<<grid, kTThreads, 0, s>>>(ring, stats, mv, slot_new, slot_prev, 0, 1, 1,
                                                                    1, W, H);
}

}  // namespace

uint32_t temporal_stats_len(uint32_t width, uint32_t height, uint32_t channels)
{
    return ((width + kTB - 1) / kTB) * ((height + kTB - 1) / kTB) * (3 * channels + 26);
}

hipError_t launch_temporal_stats(const float *ring, uint32_t width, uint32_t height, uint32_t channels,
                                 uint32_t slot_new, uint32_t slot_prev, const int32_t *mv, uint32_t mv_offset,
                                 uint32_t mv_step, uint32_t mv_bx, uint32_t mv_by, bool luma_fields, float *stats,
                                 hipStream_t s)
{
    if (!ring || !stats || width == 0 || height == 0 || width > (1u << 20) || height > (1u << 20) ||
        (channels != 1 && channels != 2) || (mv && (mv_step == 0 || mv_bx == 0 || mv_by == 0)) ||
        (uint64_t)((width + kTB - 1) / kTB) * ((height + kTB - 1) / kTB) * (3 * channels + 26) > 0xFFFFFFFFull)
        return hipErrorInvalidValue;
    const dim3 grid(((width + kTB - 1) / kTB) * ((height + kTB - 1) / kTB));
    const int W = (int)width, H = (int)height;
    if (channels == 1) {
        if (luma_fields) launch_temporal_typed<1, true>(grid, s, ring, stats, mv, slot_new, slot_prev, mv_offset, mv_step, mv_bx, mv_by, W, H);
        else launch_temporal_typed<1, false>(grid, s, ring, stats, mv, slot_new, slot_prev, mv_offset, mv_step, mv_bx, mv_by, W, H);
    } else {
        if (luma_fields) launch_temporal_typed<2, true>(grid, s, ring, stats, mv, slot_new, slot_prev, mv_offset, mv_step, mv_bx, mv_by, W, H);
        else launch_temporal_typed<2, false>(grid, s, ring, stats, mv, slot_new, slot_prev, mv_offset, mv_step, mv_bx, mv_by, W, H);
    }
    return hipGetLastError();
}

hipError_t launch_temporal_zero(float *stats, uint32_t len, hipStream_t s)
{
    if (!stats) return hipErrorInvalidValue;
    if (len == 0) return hipSuccess;
    const uint32_t blocks = std::min<uint32_t>((len + 255) / 256, 4096);
    k_temporal_zero<<<blocks, 256, 0, s>>>(stats, len);
    return hipGetLastError();
}

}  // namespace avd
