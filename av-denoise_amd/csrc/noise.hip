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

}  // namespace avd
