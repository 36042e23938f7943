// Plane ingest/egress and the small copy/fill kernels.
//
// ingest follows engine/kernels.rs:25-63: each sample is `(word >> shift) & mask` (8- or 16-bit
// mask, not the depth) divided by the format's largest code, written to an interleaved f32 ring
// slot.  This build also writes the masked code to a u16 mirror of the slot, which the integer
// block matches read.  egress follows engine/kernels.rs:103-158: clamp to 0..1, `u32(v*max+0.5)`,
// whole words per thread, tail lanes zero.
#include "common.h"

namespace avd {

namespace {

constexpr int kThreads = 256;

// One thread per input word: every word is read once and its samples land in consecutive slots.
template <int SPW, int C>
__global__ __launch_bounds__(kThreads) void k_ingest(const uint32_t *__restrict__ p0,
                                                     const uint32_t *__restrict__ p1, float *__restrict__ slot,
                                                     uint16_t *__restrict__ codes, uint32_t pixels,
                                                     float max_value)
{
    constexpr int kBits = 32 / SPW;
    constexpr uint32_t kMask = (1u << kBits) - 1u;
    const uint32_t words = (pixels + SPW - 1) / SPW;
    for (uint32_t w = blockIdx.x * kThreads + threadIdx.x; w < words; w += gridDim.x * kThreads) {
        const uint32_t a = p0[w];
        const uint32_t b = (C > 1) ? p1[w] : 0u;
#pragma unroll
        for (int l = 0; l < SPW; ++l) {
            const uint32_t p = w * SPW + l;
            if (p >= pixels) break;
            const uint32_t ca = (a >> (l * kBits)) & kMask;
            slot[(size_t)p * C] = (float)ca / max_value;
            if (codes) codes[(size_t)p * C] = (uint16_t)ca;
            if (C > 1) {
                const uint32_t cb = (b >> (l * kBits)) & kMask;
                slot[(size_t)p * C + 1] = (float)cb / max_value;
                if (codes) codes[(size_t)p * C + 1] = (uint16_t)cb;
            }
        }
    }
}

template <int SPW, int C>
__global__ __launch_bounds__(kThreads) void k_egress(const float *__restrict__ frame, uint32_t *__restrict__ p0,
                                                     uint32_t *__restrict__ p1, uint32_t pixels, float max_value)
{
    constexpr int kBits = 32 / SPW;
    const uint32_t words = (pixels + SPW - 1) / SPW;
    for (uint32_t w = blockIdx.x * kThreads + threadIdx.x; w < words; w += gridDim.x * kThreads) {
        uint32_t a = 0, b = 0;
#pragma unroll
        for (int l = 0; l < SPW; ++l) {
            const uint32_t p = w * SPW + l;
            if (p >= pixels) break;
            const float v = fminf(fmaxf(frame[(size_t)p * C], 0.0f), 1.0f);
            a |= (uint32_t)(v * max_value + 0.5f) << (l * kBits);
            if (C > 1) {
                const float u = fminf(fmaxf(frame[(size_t)p * C + 1], 0.0f), 1.0f);
                b |= (uint32_t)(u * max_value + 0.5f) << (l * kBits);
            }
        }
        p0[w] = a;
        if (C > 1) p1[w] = b;
    }
}

__global__ __launch_bounds__(kThreads) void k_copy_words(const uint32_t *__restrict__ src, uint32_t *__restrict__ dst,
                                                         uint64_t words)
{
    for (uint64_t i = (uint64_t)blockIdx.x * kThreads + threadIdx.x; i < words; i += (uint64_t)gridDim.x * kThreads)
        dst[i] = src[i];
}

__global__ __launch_bounds__(kThreads) void k_fill_u32(uint32_t *__restrict__ dst, uint32_t value, uint64_t words)
{
    for (uint64_t i = (uint64_t)blockIdx.x * kThreads + threadIdx.x; i < words; i += (uint64_t)gridDim.x * kThreads)
        dst[i] = value;
}

// A grid-stride launch sized to keep every CU busy without an oversized grid.
inline unsigned grid_for(uint64_t items)
{
    uint64_t blocks = (items + kThreads - 1) / kThreads;
    if (blocks < 1) blocks = 1;
    if (blocks > 4096) blocks = 4096;
    return (unsigned)blocks;
}

}  // namespace

hipError_t launch_ingest(const uint32_t *p0, const uint32_t *p1, float *slot, uint16_t *codes, uint32_t pixels,
                         int channels, int spw, float max_value, hipStream_t s)
{
    const unsigned g = grid_for((pixels + spw - 1) / spw);
    if (spw == 4 && channels == 1) k_ingest<4, 1><<<g, kThreads, 0, s>>>(p0, p1, slot, codes, pixels, max_value);
    else if (spw == 4 && channels == 2) k_ingest<4, 2><<<g, kThreads, 0, s>>>(p0, p1, slot, codes, pixels, max_value);
    else if (spw == 2 && channels == 1) k_ingest<2, 1><<<g, kThreads, 0, s>>>(p0, p1, slot, codes, pixels, max_value);
    else if (spw == 2 && channels == 2) k_ingest<2, 2><<<g, kThreads, 0, s>>>(p0, p1, slot, codes, pixels, max_value);
    else return hipErrorInvalidValue;
    return hipGetLastError();
}

hipError_t launch_egress(const float *frame, uint32_t *p0, uint32_t *p1, uint32_t pixels, int channels, int spw,
                         float max_value, hipStream_t s)
{
    const unsigned g = grid_for((pixels + spw - 1) / spw);
    if (spw == 4 && channels == 1) k_egress<4, 1><<<g, kThreads, 0, s>>>(frame, p0, p1, pixels, max_value);
    else if (spw == 4 && channels == 2) k_egress<4, 2><<<g, kThreads, 0, s>>>(frame, p0, p1, pixels, max_value);
    else if (spw == 2 && channels == 1) k_egress<2, 1><<<g, kThreads, 0, s>>>(frame, p0, p1, pixels, max_value);
    else if (spw == 2 && channels == 2) k_egress<2, 2><<<g, kThreads, 0, s>>>(frame, p0, p1, pixels, max_value);
    else return hipErrorInvalidValue;
    return hipGetLastError();
}

hipError_t launch_copy_words(const uint32_t *src, uint32_t *dst, uint64_t words, hipStream_t s)
{
    if (words == 0) return hipSuccess;
    k_copy_words<<<grid_for(words), kThreads, 0, s>>>(src, dst, words);
    return hipGetLastError();
}

hipError_t launch_fill_u32(void *dst, uint32_t value, uint64_t words, hipStream_t s)
{
    if (words == 0) return hipSuccess;
    k_fill_u32<<<grid_for(words), kThreads, 0, s>>>(static_cast<uint32_t *>(dst), value, words);
    return hipGetLastError();
}

}  // namespace avd
