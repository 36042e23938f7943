// Block-match motion search, confidence and warp on integer codes.
//
// Follows motion/kernels/block_match.rs: the coarse pass (:117-211) searches a (2R+1)^2 window
// around zero on pyramid level 1 and writes the winner times 2 to every fine block it covers; the
// fine pass (:306-369, :489-535) searches around that seed on level 0 and writes the vector and
// the confidence max(0, (T^2 - e^2) / (T^2 + e^2)), e = max(SAD - floor, 0).  A window's winner
// is the lowest SAD, then the lowest row-major index, and the window centre (zero or the seed)
// wins a tie with that pick (:700-716, :940-1000).  Level 0 is lane 0 of the ring's u16 codes and
// level 1 the 2x2 sum of those codes (pyramid.rs:11-44 averages the normalised samples; the sum
// is that mean times 4*max), so every SAD here is an exact integer and ranks candidates as the
// exact SAD does.  The confidence reads the level-0 SAD divided by max, one f32 division.
#include "common.h"

#include <algorithm>

namespace avd {

namespace {

constexpr int kMaxR = 8;
constexpr int kMaxBlk = 32;

__device__ __forceinline__ int clampi(int v, int hi) { return v < 0 ? 0 : (v > hi ? hi : v); }

__global__ __launch_bounds__(256) void k_downscale(const uint16_t *__restrict__ codes, int W, int H, int C,
                                                   int32_t *__restrict__ l1)
{
    const int dw = max(W / 2, 1), dh = max(H / 2, 1);
    const int n = dw * dh;
    for (int i = blockIdx.x * 256 + threadIdx.x; i < n; i += gridDim.x * 256) {
        const int y = i / dw, x = i - y * dw;
        const int sx = 2 * x, sy = 2 * y;
        const int sx1 = sx + 1 < W ? sx + 1 : sx, sy1 = sy + 1 < H ? sy + 1 : sy;
        l1[i] = (int32_t)codes[((size_t)sy * W + sx) * C] + codes[((size_t)sy * W + sx1) * C] +
                codes[((size_t)sy1 * W + sx) * C] + codes[((size_t)sy1 * W + sx1) * C];
    }
}

// Reads a clamped level sample: level 0 is lane 0 of the interleaved codes, level 1 a flat plane.
struct Level {
    const uint16_t *codes;  // level 0 when non-null
    const int32_t *plane;   // level 1 otherwise
    int w, h, C;
    __device__ __forceinline__ uint32_t at(int x, int y) const
    {
        x = clampi(x, w - 1);
        y = clampi(y, h - 1);
        return codes ? (uint32_t)codes[((size_t)y * w + x) * C] : (uint32_t)plane[(size_t)y * w + x];
    }
};

// |a - b| + acc on unsigned words in one instruction.
__device__ __forceinline__ uint32_t sad_u32(uint32_t a, uint32_t b, uint32_t acc)
{
    uint32_t r;
    asm("v_sad_u32 %0, %1, %2, %3" : "=v"(r) : "v"(a), "v"(b), "v"(acc));
    return r;
}

constexpr int kBmThreads = 512;
constexpr int kBmLdsBudget = 64 * 1024;  // window bytes staged per batch of neighbours

// Block-match of one block against every neighbour frame: one workgroup per block.  The block is
// staged in LDS once; the neighbours go in batches whose (blk + 2R)^2 windows, centred on each
// neighbour's seed, fit the LDS budget.  Thread i scores candidate (neighbour i / side^2, offset
// i % side^2) over the blk^2 pixels with v_sad_u32 (the block row is a broadcast read, the window
// row the thread's own).  Each neighbour's minimum (SAD << 9 | offset) falls out of an LDS
// atomic min, so the lowest SAD and then the lowest row-major offset wins; the window centre then
// wins a tie with that pick.  Writes through `finish(n, dx, dy, sad)` (thread n of the block).
// `nb` is the neighbours' level with its base at slot 0; neighbour n sits nslots[n] levels on.
template <typename Seed, typename Finish>
__device__ __forceinline__ void match_block(const Level &cen, const Level &nb, size_t level_len,
                                            const uint32_t *__restrict__ nslots, int neighbours, int ox, int oy,
                                            int blk, int R, Seed &&seed, Finish &&finish, uint32_t *lds)
{
    const int side = 2 * R + 1, side2 = side * side, ww = blk + 2 * R;
    const int win = ww * ww;
    int batch = kBmLdsBudget / (win * 4);
    batch = max(1, min(batch, min(neighbours, kBmThreads / side2 > 0 ? kBmThreads / side2 : 1)));
    batch = max(1, min(batch, 16));
    uint32_t *s_blk = lds;                       // blk * blk
    uint32_t *s_key = s_blk + blk * blk;          // 16
    uint32_t *s_csad = s_key + 16;                // 16
    int *s_seed = (int *)(s_csad + 16);           // 32
    uint32_t *s_slot = (uint32_t *)(s_seed + 32); // 16
    uint32_t *s_win = s_slot + 16;                // batch * win
    for (int i = threadIdx.x; i < blk * blk; i += kBmThreads) {
        const int py = i / blk, px = i - py * blk;
        s_blk[i] = cen.at(ox + px, oy + py);
    }
    for (int n0 = 0; n0 < neighbours; n0 += batch) {
        const int nb_n = min(batch, neighbours - n0);
        if ((int)threadIdx.x < nb_n) {
            int sdx, sdy;
            seed(n0 + threadIdx.x, sdx, sdy);
            s_seed[2 * threadIdx.x] = sdx;
            s_seed[2 * threadIdx.x + 1] = sdy;
            s_slot[threadIdx.x] = nslots[n0 + threadIdx.x];
            s_key[threadIdx.x] = ~0u;
        }
        __syncthreads();
        for (int i = threadIdx.x; i < nb_n * win; i += kBmThreads) {
            const int nl = i / win, j = i - nl * win;
            const int wy = j / ww, wx = j - wy * ww;
            Level l = nb;
            if (l.codes) l.codes += (size_t)s_slot[nl] * level_len;
            else l.plane += (size_t)s_slot[nl] * level_len;
            s_win[i] = l.at(ox + s_seed[2 * nl] - R + wx, oy + s_seed[2 * nl + 1] - R + wy);
        }
        __syncthreads();
        for (int task = threadIdx.x; task < nb_n * side2; task += kBmThreads) {
            const int nl = task / side2, c = task - nl * side2;
            const int dy = c / side, dx = c - dy * side;
            const uint32_t *w = s_win + nl * win + dy * ww + dx;
            uint32_t sad = 0;
            for (int py = 0; py < blk; ++py) {
                const uint32_t *brow = s_blk + py * blk;
                const uint32_t *wrow = w + py * ww;
                if (blk == 16) {
#pragma unroll
                    for (int px = 0; px < 16; ++px) sad = sad_u32(brow[px], wrow[px], sad);
                } else if (blk == 8) {
#pragma unroll
                    for (int px = 0; px < 8; ++px) sad = sad_u32(brow[px], wrow[px], sad);
                } else {
                    for (int px = 0; px < blk; ++px) sad = sad_u32(brow[px], wrow[px], sad);
                }
            }
            atomicMin(&s_key[nl], (sad << 9) | (uint32_t)c);
            if (c == R * side + R) s_csad[nl] = sad;
        }
        __syncthreads();
        if ((int)threadIdx.x < nb_n) {
            const int nl = threadIdx.x;
            const uint32_t key = s_key[nl];
            const uint32_t best = key >> 9;
            const int bi = (int)(key & 511u);
            const int sdx = s_seed[2 * nl], sdy = s_seed[2 * nl + 1];
            if (s_csad[nl] <= best) finish(n0 + nl, sdx, sdy, s_csad[nl]);
            else finish(n0 + nl, sdx + (bi % side) - R, sdy + (bi / side) - R, best);
        }
        __syncthreads();
    }
}

constexpr int kRowsLdsBudget = 48 * 1024;

// Window rows of the rows form are u16 samples (codes of at most 12 bits, or level-1 sums of four
// of them), each row padded to a multiple of 8 samples so rows start on 16-byte boundaries.
__host__ __device__ constexpr int rows_stride(int ww) { return (ww + 7) & ~7; }

__host__ __device__ inline int match_rows_batch(int blk, int R, int neighbours)
{
    const int ww = blk + 2 * R;
    int batch = kRowsLdsBudget / (ww * rows_stride(ww) * 2);
    batch = batch < 1 ? 1 : batch;
    batch = batch < neighbours ? batch : neighbours;
    return batch < 16 ? batch : 16;
}

size_t match_rows_lds_bytes(int blk, int R, int neighbours)
{
    const int ww = blk + 2 * R;
    return (size_t)(blk * blk / 2 + 16 + 16 + 32 + 16) * 4 +
           (size_t)match_rows_batch(blk, R, neighbours) * ww * rows_stride(ww) * 2;
}

size_t match_lds_bytes(int blk, int R, int neighbours)
{
    const int side2 = (2 * R + 1) * (2 * R + 1), ww = blk + 2 * R, win = ww * ww;
    int batch = kBmLdsBudget / (win * 4);
    batch = std::max(1, std::min(batch, std::min(neighbours, kBmThreads / side2 > 0 ? kBmThreads / side2 : 1)));
    batch = std::max(1, std::min(batch, 16));
    return (size_t)(blk * blk + 16 + 16 + 32 + 16) * 4 + (size_t)batch * win * 4;
}

// Copies the `cols` x `rows` rectangle at (x0, y0) of a level into dst (row stride `stride`
// words), clamping to the level: lanes take whole rows; a row inside the level is read with plain
// offset loads, a row crossing its left or right edge sample by sample.
template <int COLS>
__device__ __forceinline__ void stage_rect(const Level &l, int x0, int y0, int rows, uint32_t *dst, int stride,
                                           int lane, int lanes)
{
    const bool inside = x0 >= 0 && x0 + COLS <= l.w;
    for (int row = lane; row < rows; row += lanes) {
        const int y = clampi(y0 + row, l.h - 1);
        uint32_t *d = dst + row * stride;
        if (inside) {
            if (l.codes) {
                const uint16_t *src = l.codes + ((size_t)y * l.w + x0) * l.C;
#pragma unroll
                for (int x = 0; x < COLS; ++x) d[x] = src[x * l.C];
            } else {
                const int32_t *src = l.plane + (size_t)y * l.w + x0;
#pragma unroll
                for (int x = 0; x < COLS; ++x) d[x] = (uint32_t)src[x];
            }
        } else {
#pragma unroll
            for (int x = 0; x < COLS; ++x) d[x] = l.at(x0 + x, y);
        }
    }
}

// One lane's share of the COLS x ROWS rectangle at (x0, y0) of a level: sample lane + 64 k in
// register k, the wave's lanes sweeping consecutive samples (so a load instruction covers a few
// cache lines, not one per lane).  Every load of the rectangle is issued before any is used; a
// rectangle inside the level reads without clamps.
template <int COLS, int ROWS>
struct TileRegs {
    static constexpr int N = COLS * ROWS, PER = (N + 63) / 64;
    uint32_t v[PER];
};

template <int COLS, int ROWS>
__device__ __forceinline__ void load_tile16(const Level &l, int x0, int y0, int lane, TileRegs<COLS, ROWS> &t)
{
    using T = TileRegs<COLS, ROWS>;
    const bool inside = x0 >= 0 && y0 >= 0 && x0 + COLS <= l.w && y0 + ROWS <= l.h;
    // Sample k's (column, row) offset from (x0, y0), clamped to the level when the rectangle passes
    // an edge; 32-bit sample offsets from the level's start.
    auto offset = [&](int k) -> uint32_t {
        const int i = min(lane + 64 * k, T::N - 1), r = i / COLS, c = i - r * COLS;
        const int x = inside ? x0 + c : clampi(x0 + c, l.w - 1), y = inside ? y0 + r : clampi(y0 + r, l.h - 1);
        return (uint32_t)(y * l.w + x);
    };
    if (l.codes) {
#pragma unroll
        for (int k = 0; k < T::PER; ++k) t.v[k] = l.codes[(size_t)offset(k) * l.C];
    } else {
#pragma unroll
        for (int k = 0; k < T::PER; ++k) t.v[k] = (uint32_t)l.plane[offset(k)];
    }
}

// The COLS x ROWS rectangle at (x0, y0) of a level into u16 rows `stride` samples apart, one
// sample per lane at a time (for rectangles too large to hold in registers).
template <int COLS, int ROWS>
__device__ __forceinline__ void stage_tile16(const Level &l, int x0, int y0, uint16_t *dst, int stride, int lane)
{
    const bool inside = x0 >= 0 && y0 >= 0 && x0 + COLS <= l.w && y0 + ROWS <= l.h;
    for (int i = lane; i < COLS * ROWS; i += 64) {
        const int r = i / COLS, c = i - r * COLS;
        uint32_t v;
        if (inside)
            v = l.codes ? (uint32_t)l.codes[((size_t)(y0 + r) * l.w + x0 + c) * l.C]
                        : (uint32_t)l.plane[(size_t)(y0 + r) * l.w + x0 + c];
        else
            v = l.at(x0 + c, y0 + r);
        dst[r * stride + c] = (uint16_t)v;
    }
}

// The rectangle loaded by load_tile16 into u16 rows `stride` samples apart.
template <int COLS, int ROWS>
__device__ __forceinline__ void store_tile16(const TileRegs<COLS, ROWS> &t, uint16_t *dst, int stride, int lane)
{
    using T = TileRegs<COLS, ROWS>;
#pragma unroll
    for (int k = 0; k < T::PER; ++k) {
        const int i = lane + 64 * k, r = i / COLS, c = i - r * COLS;
        if (i < T::N) dst[r * stride + c] = (uint16_t)t.v[k];
    }
}

// This workgroup's block (column, row) of the 2-D grid, renumbered so that each XCD (workgroups are
// dealt round-robin over eight) takes a contiguous run of rows: neighbouring blocks' overlapping
// windows then meet in one L2.
__device__ __forceinline__ void xcd_block_2d(int &bc, int &br)
{
    const int n = (int)(gridDim.x * gridDim.y), b = (int)(blockIdx.y * gridDim.x + blockIdx.x);
    const int xcd = b & 7, i = b >> 3, q = n >> 3, rem = n & 7;
    const int v = xcd * q + min(xcd, rem) + i;
    br = v / (int)gridDim.x;
    bc = v - br * (int)gridDim.x;
}

// |a.lo16 - b.lo16| + |a.hi16 - b.hi16| + acc: two packed samples per instruction (gfx950's
// v_sad_u16 sums both halves into the 32-bit accumulator).
__device__ __forceinline__ uint32_t sad2_u16(uint32_t a, uint32_t b, uint32_t acc)
{
    uint32_t r;
    asm("v_sad_u16 %0, %1, %2, %3" : "=v"(r) : "v"(a), "v"(b), "v"(acc));
    return r;
}

// The same search specialised for a compile-time block size and radius: one wave per block, lane
// (neighbour, dy) scores the whole row of offsets (dx = 0..2R) for its dy.  Per block row it holds
// the block row and the window row (BLK + 2R samples) in registers, so each window row is read from
// LDS once for 2R + 1 candidates instead of once per candidate.
template <int BLK, int R, typename Seed, typename Finish>
__device__ __forceinline__ void match_rows(const Level &cen, const Level &nb, size_t level_len,
                                           const uint32_t *__restrict__ nslots, int neighbours, int ox, int oy,
                                           Seed &&seed, Finish &&finish, uint32_t *lds)
{
    constexpr int SIDE = 2 * R + 1, WW = BLK + 2 * R, WS = rows_stride(WW), WIN = WW * WS;
    const int lane = threadIdx.x;
    const int batch = match_rows_batch(BLK, R, neighbours);
    uint16_t *s_blk = reinterpret_cast<uint16_t *>(lds);  // BLK * BLK u16
    uint32_t *s_key = lds + BLK * BLK / 2;                   // 16
    uint32_t *s_csad = s_key + 16;                           // 16
    int *s_seed = (int *)(s_csad + 16);                      // 32
    uint32_t *s_slot = (uint32_t *)(s_seed + 32);            // 16
    // batch windows of WW rows of WS u16 samples, rows 16-byte aligned
    uint16_t *s_win = reinterpret_cast<uint16_t *>(s_slot + 16);
    {
        TileRegs<BLK, BLK> tb;
        load_tile16<BLK, BLK>(cen, ox, oy, lane, tb);
        store_tile16<BLK, BLK>(tb, s_blk, BLK, lane);
    }
    for (int n0 = 0; n0 < neighbours; n0 += batch) {
        const int nb_n = min(batch, neighbours - n0);
        if (lane < nb_n) {
            int sdx, sdy;
            seed(n0 + lane, sdx, sdy);
            s_seed[2 * lane] = sdx;
            s_seed[2 * lane + 1] = sdy;
            s_slot[lane] = nslots[n0 + lane];
            s_key[lane] = ~0u;
        }
        __syncthreads();
        // The whole wave stages each neighbour's window in turn, the next window's loads issued
        // before this one's stores.
        auto load_win = [&](int nl, TileRegs<WW, WW> &t) {
            Level l = nb;
            if (l.codes) l.codes += (size_t)s_slot[nl] * level_len;
            else l.plane += (size_t)s_slot[nl] * level_len;
            load_tile16<WW, WW>(l, ox + s_seed[2 * nl] - R, oy + s_seed[2 * nl + 1] - R, lane, t);
        };
        if constexpr (TileRegs<WW, WW>::PER <= 16) {
            TileRegs<WW, WW> ta, tb;
            load_win(0, ta);
            for (int nl = 0; nl < nb_n; nl += 2) {
                if (nl + 1 < nb_n) load_win(nl + 1, tb);
                store_tile16<WW, WW>(ta, s_win + nl * WIN, WS, lane);
                if (nl + 1 < nb_n) {
                    if (nl + 2 < nb_n) load_win(nl + 2, ta);
                    store_tile16<WW, WW>(tb, s_win + (nl + 1) * WIN, WS, lane);
                }
            }
        } else {  // windows past 1024 samples: sample by sample, to spare registers
            for (int nl = 0; nl < nb_n; ++nl) {
                Level l = nb;
                if (l.codes) l.codes += (size_t)s_slot[nl] * level_len;
                else l.plane += (size_t)s_slot[nl] * level_len;
                stage_tile16<WW, WW>(l, ox + s_seed[2 * nl] - R, oy + s_seed[2 * nl + 1] - R, s_win + nl * WIN, WS,
                                     lane);
            }
        }
        __syncthreads();
        for (int task = lane; task < nb_n * SIDE; task += 64) {
            const int nl = task / SIDE, dy = task - nl * SIDE;
            uint32_t sad[SIDE];
#pragma unroll
            for (int j = 0; j < SIDE; ++j) sad[j] = 0;
            const uint16_t *wbase = s_win + nl * WIN + dy * WS;
#pragma unroll 2
            for (int py = 0; py < BLK; ++py) {
                // Rows start on 16-byte boundaries: BLK is a multiple of 8 samples, WS of 8, and the
                // window area starts (BLK^2 / 2 + 80) words into the 16-byte aligned LDS block.
                const uint32_t *brow = (const uint32_t *)__builtin_assume_aligned(s_blk + py * BLK, 16);
                const uint32_t *wrow = (const uint32_t *)__builtin_assume_aligned(wbase + py * WS, 16);
                // Packed pairs: the block's (2k, 2k + 1), the window's from an even sample (wp) and
                // from an odd one (wo[k] = samples 2k + 1, 2k + 2).
                uint32_t bp[BLK / 2], wp[(WW + 1) / 2], wo[WW / 2];
#pragma unroll
                for (int k = 0; k < BLK / 2; ++k) bp[k] = brow[k];
#pragma unroll
                for (int k = 0; k < (WW + 1) / 2; ++k) wp[k] = wrow[k];
#pragma unroll
                for (int k = 0; k + 1 < (WW + 1) / 2; ++k) wo[k] = __builtin_amdgcn_alignbyte(wp[k + 1], wp[k], 2);
#pragma unroll
                for (int j = 0; j < SIDE; ++j)
#pragma unroll
                    for (int k = 0; k < BLK / 2; ++k)
                        sad[j] = sad2_u16(bp[k], (j & 1) ? wo[(j >> 1) + k] : wp[(j >> 1) + k], sad[j]);
            }
            uint32_t best = ~0u;
#pragma unroll
            for (int j = 0; j < SIDE; ++j) best = min(best, (sad[j] << 9) | (uint32_t)(dy * SIDE + j));
            atomicMin(&s_key[nl], best);
            if (dy == R) s_csad[nl] = sad[R];
        }
        __syncthreads();
        if (lane < nb_n) {
            const uint32_t key = s_key[lane];
            const uint32_t best = key >> 9;
            const int bi = (int)(key & 511u);
            const int sdx = s_seed[2 * lane], sdy = s_seed[2 * lane + 1];
            if (s_csad[lane] <= best) finish(n0 + lane, sdx, sdy, s_csad[lane]);
            else finish(n0 + lane, sdx + (bi % SIDE) - R, sdy + (bi / SIDE) - R, best);
        }
        __syncthreads();
    }
}

// Coarse pass (block_match.rs:117-211) on level 1 around a zero seed; the winner times 2 seeds every
// fine block this coarse block covers.
__global__ __launch_bounds__(kBmThreads) void k_bm_coarse(const int32_t *__restrict__ pyr1, int lw, int lh,
                                                          int blk, int step, int R, int centre_slot,
                                                          const uint32_t *__restrict__ nslots, int neighbours,
                                                          int32_t *__restrict__ mv, int fine_bx, int fine_by,
                                                          int fine_step)
{
    extern __shared__ uint32_t lds[];
    int bc, br;
    xcd_block_2d(bc, br);
    const size_t lp = (size_t)lw * lh;
    const Level cen{nullptr, pyr1 + (size_t)centre_slot * lp, lw, lh, 1};
    const Level nb{nullptr, pyr1, lw, lh, 1};
    const int c0 = min(bc * step * 2 / fine_step, fine_bx), r0 = min(br * step * 2 / fine_step, fine_by);
    const int c1 = (bc == (int)gridDim.x - 1) ? fine_bx : min((bc + 1) * step * 2 / fine_step, fine_bx);
    const int r1 = (br == (int)gridDim.y - 1) ? fine_by : min((br + 1) * step * 2 / fine_step, fine_by);
    match_block(
        cen, nb, lp, nslots, neighbours, bc * step, br * step, blk, R,
        [](int, int &sdx, int &sdy) {
            sdx = 0;
            sdy = 0;
        },
        [&](int n, int dx, int dy, uint32_t) {
            int32_t *m = mv + (size_t)n * fine_bx * fine_by * 2;
            for (int fr = r0; fr < r1; ++fr)
                for (int fc = c0; fc < c1; ++fc) {
                    m[((size_t)fr * fine_bx + fc) * 2] = dx * 2;
                    m[((size_t)fr * fine_bx + fc) * 2 + 1] = dy * 2;
                }
        },
        lds);
}

// Fine pass (block_match.rs:306-369) on level 0 around the seed, then the confidence
// (:489-535): the level-0 SAD over max, one f32 division.
__global__ __launch_bounds__(kBmThreads) void k_bm_fine(const uint16_t *__restrict__ codes, int W, int H, int C,
                                                        int blk, int step, int R, const int32_t *seeds,
                                                        int centre_slot, const uint32_t *__restrict__ nslots,
                                                        int neighbours, int32_t *mv, float *__restrict__ conf,
                                                        float max_value, float sad_floor, float thsad)
{
    extern __shared__ uint32_t lds[];
    int bc, br;
    xcd_block_2d(bc, br);
    const int bx = gridDim.x, by = gridDim.y;
    const size_t sp = (size_t)W * H * C;
    const Level cen{codes + (size_t)centre_slot * sp, nullptr, W, H, C};
    const Level nb{codes, nullptr, W, H, C};
    const size_t blkid = (size_t)br * bx + bc;
    const size_t per_n = (size_t)bx * by;
    match_block(
        cen, nb, sp, nslots, neighbours, bc * step, br * step, blk, R,
        [&](int n, int &sdx, int &sdy) {
            const size_t b = (size_t)n * per_n + blkid;
            sdx = seeds ? seeds[b * 2] : 0;
            sdy = seeds ? seeds[b * 2 + 1] : 0;
        },
        [&](int n, int dx, int dy, uint32_t sad) {
            const size_t b = (size_t)n * per_n + blkid;
            mv[b * 2] = dx;
            mv[b * 2 + 1] = dy;
            if (conf) {
                float excess = (float)sad / max_value - sad_floor;
                if (excess < 0.0f) excess = 0.0f;
                const float t2 = thsad * thsad, e2 = excess * excess;
                const float c = (t2 - e2) / (t2 + e2);
                conf[b] = c < 0.0f ? 0.0f : c;
            }
        },
        lds);
}

template <int BLK, int R>
__global__ __launch_bounds__(64) void k_bm_coarse_rows(const int32_t *__restrict__ pyr1, int lw, int lh, int step,
                                                       int centre_slot, const uint32_t *__restrict__ nslots,
                                                       int neighbours, int32_t *__restrict__ mv, int fine_bx,
                                                       int fine_by, int fine_step)
{
    extern __shared__ __align__(16) uint32_t lds_rows[];
    int bc, br;
    xcd_block_2d(bc, br);
    const size_t lp = (size_t)lw * lh;
    const Level cen{nullptr, pyr1 + (size_t)centre_slot * lp, lw, lh, 1};
    const Level nb{nullptr, pyr1, lw, lh, 1};
    const int c0 = min(bc * step * 2 / fine_step, fine_bx), r0 = min(br * step * 2 / fine_step, fine_by);
    const int c1 = (bc == (int)gridDim.x - 1) ? fine_bx : min((bc + 1) * step * 2 / fine_step, fine_bx);
    const int r1 = (br == (int)gridDim.y - 1) ? fine_by : min((br + 1) * step * 2 / fine_step, fine_by);
    match_rows<BLK, R>(
        cen, nb, lp, nslots, neighbours, bc * step, br * step,
        [](int, int &sdx, int &sdy) {
            sdx = 0;
            sdy = 0;
        },
        [&](int n, int dx, int dy, uint32_t) {
            int32_t *m = mv + (size_t)n * fine_bx * fine_by * 2;
            for (int fr = r0; fr < r1; ++fr)
                for (int fc = c0; fc < c1; ++fc) {
                    m[((size_t)fr * fine_bx + fc) * 2] = dx * 2;
                    m[((size_t)fr * fine_bx + fc) * 2 + 1] = dy * 2;
                }
        },
        lds_rows);
}

template <int BLK, int R>
__global__ __launch_bounds__(64) void k_bm_fine_rows(const uint16_t *__restrict__ codes, int W, int H, int C, int step,
                                                     const int32_t *seeds, int centre_slot,
                                                     const uint32_t *__restrict__ nslots, int neighbours,
                                                     int32_t *mv, float *__restrict__ conf,
                                                     float max_value, float sad_floor, float thsad)
{
    extern __shared__ __align__(16) uint32_t lds_rows[];
    int bc, br;
    xcd_block_2d(bc, br);
    const int bx = gridDim.x, by = gridDim.y;
    const size_t sp = (size_t)W * H * C;
    const Level cen{codes + (size_t)centre_slot * sp, nullptr, W, H, C};
    const Level nb{codes, nullptr, W, H, C};
    const size_t blkid = (size_t)br * bx + bc;
    const size_t per_n = (size_t)bx * by;
    match_rows<BLK, R>(
        cen, nb, sp, nslots, neighbours, bc * step, br * step,
        [&](int n, int &sdx, int &sdy) {
            const size_t b = (size_t)n * per_n + blkid;
            sdx = seeds ? seeds[b * 2] : 0;
            sdy = seeds ? seeds[b * 2 + 1] : 0;
        },
        [&](int n, int dx, int dy, uint32_t sad) {
            const size_t b = (size_t)n * per_n + blkid;
            mv[b * 2] = dx;
            mv[b * 2 + 1] = dy;
            if (conf) {
                float excess = (float)sad / max_value - sad_floor;
                if (excess < 0.0f) excess = 0.0f;
                const float t2 = thsad * thsad, e2 = excess * excess;
                const float c = (t2 - e2) / (t2 + e2);
                conf[b] = c < 0.0f ? 0.0f : c;
            }
        },
        lds_rows);
}

// Specialised forms exist for these block sizes and radii (radius 1 is the chroma refine's).
inline bool rows_form(int blk, int R)
{
    return (blk == 8 || blk == 16 || blk == 32) && (R == 1 || R == 2 || R == 4 || R == 8);
}

#define AVD_ROWS_RADII(B, RV, CALL)                                                                         \
    do {                                                                                                    \
        if ((RV) == 1) CALL(B, 1); else if ((RV) == 2) CALL(B, 2); else if ((RV) == 4) CALL(B, 4);          \
        else CALL(B, 8);                                                                                    \
    } while (0)
#define AVD_ROWS_DISPATCH(BLKV, RV, CALL)                                                                   \
    do {                                                                                                    \
        if ((BLKV) == 8) AVD_ROWS_RADII(8, RV, CALL);                                                       \
        else if ((BLKV) == 16) AVD_ROWS_RADII(16, RV, CALL);                                                \
        else AVD_ROWS_RADII(32, RV, CALL);                                                                  \
    } while (0)

__global__ __launch_bounds__(256) void k_warp(const float *__restrict__ src, float *__restrict__ dst,
                                              const int32_t *__restrict__ mv, int W, int H, int C, int step,
                                              int blocks_x, int blocks_y)
{
    const int n = W * H;
    for (int i = blockIdx.x * 256 + threadIdx.x; i < n; i += gridDim.x * 256) {
        const int y = i / W, x = i - y * W;
        const int bcol = min(x / step, blocks_x - 1), brow = min(y / step, blocks_y - 1);
        const int32_t *v = mv + ((size_t)brow * blocks_x + bcol) * 2;
        const int sx = clampi(x + v[0], W - 1), sy = clampi(y + v[1], H - 1);
        for (int c = 0; c < C; ++c) dst[(size_t)i * C + c] = src[((size_t)sy * W + sx) * C + c];
    }
}

inline unsigned grid_for(uint64_t items)
{
    uint64_t g = (items + 255) / 256;
    return (unsigned)(g < 1 ? 1 : (g > 4096 ? 4096 : g));
}

// The fine pass on level 0 at radius R around `seeds` (zero when null), into mv and conf.
hipError_t launch_fine(const MotionGeom &g, int R, const uint16_t *codes, int centre_slot, const uint32_t *nslots_dev,
                       int neighbours, const int32_t *seeds, int32_t *mv, float *conf, float max_value,
                       float sad_floor, float thsad, hipStream_t s)
{
    const dim3 grid(g.blocks_x, g.blocks_y);
    if (rows_form(g.blksize, R)) {
#define AVD_FINE(B, RR)                                                                                      \
    k_bm_fine_rows<B, RR><<<grid, 64, match_rows_lds_bytes(B, RR, neighbours), s>>>(                         \
        codes, g.W, g.H, g.C, g.step, seeds, centre_slot, nslots_dev, neighbours, mv, conf, max_value,       \
        sad_floor, thsad)
        AVD_ROWS_DISPATCH(g.blksize, R, AVD_FINE);
#undef AVD_FINE
    } else {
        k_bm_fine<<<grid, kBmThreads, match_lds_bytes(g.blksize, R, neighbours), s>>>(
            codes, g.W, g.H, g.C, g.blksize, g.step, R, seeds, centre_slot, nslots_dev, neighbours, mv, conf,
            max_value, sad_floor, thsad);
    }
    return hipGetLastError();
}

// The lower median of `n` values `stride` words apart: the value with at most (n - 1) / 2 of the
// others below it and more than that at or below it (chroma_map.rs:22-44 ranks the same way).
__device__ __forceinline__ int lower_median(const int32_t *v, int cols, int rows, int row_pitch)
{
    const int n = cols * rows, mid = (n - 1) / 2;
    for (int i = 0; i < n; ++i) {
        const int vi = v[((i / cols) * row_pitch + i % cols) * 2];
        int below = 0, equal = 0;
        for (int j = 0; j < n; ++j) {
            const int vj = v[((j / cols) * row_pitch + j % cols) * 2];
            below += vj < vi;
            equal += vj == vi;
        }
        if (below <= mid && mid < below + equal) return vi;
    }
    return v[0];
}

// Divides by 1 << shift, rounding half away from zero (chroma_map.rs:6-16).
__device__ __forceinline__ int scale_down(int v, int shift)
{
    if (shift == 0) return v;
    const int m = v < 0 ? -v : v, r = (m + (1 << (shift - 1))) >> shift;
    return v < 0 ? -r : r;
}

// One thread per (field, chroma block): the chroma block covers (chroma_blk << shift) luma samples
// a side from (col * chroma_step) << shift, so covers = (area - luma_blk) / luma_step + 1 luma blocks
// a side from start = origin / luma_step; those past the luma grid drop out (the first one never
// does).  The per-component lower median of the rest, scaled down by the subsampling, is the
// chroma block's vector.  The candidates are re-read from the (cached) luma field, so any cover
// size runs without per-thread arrays.
__global__ __launch_bounds__(256) void k_map_luma_fields(FieldMap m, const int32_t *__restrict__ luma,
                                                         int32_t *__restrict__ chroma, int fields)
{
    const int area_x = m.chroma_blk << m.shift_x, area_y = m.chroma_blk << m.shift_y;
    const int covers_x = max(area_x - m.luma_blk, 0) / m.luma_step + 1;
    const int covers_y = max(area_y - m.luma_blk, 0) / m.luma_step + 1;
    const int blocks = m.chroma_bx * m.chroma_by;
    const int total = blocks * fields;
    for (int i = blockIdx.x * 256 + threadIdx.x; i < total; i += gridDim.x * 256) {
        const int f = i / blocks, b = i - f * blocks;
        const int row = b / m.chroma_bx, col = b - row * m.chroma_bx;
        const int sc = ((col * m.chroma_step) << m.shift_x) / m.luma_step;
        const int sr = ((row * m.chroma_step) << m.shift_y) / m.luma_step;
        const int last_c = min(sc + covers_x - 1, m.luma_bx - 1), last_r = min(sr + covers_y - 1, m.luma_by - 1);
        const int first_c = min(sc, last_c), first_r = min(sr, last_r);
        const int32_t *src = luma + (size_t)f * m.luma_stride + ((size_t)first_r * m.luma_bx + first_c) * 2;
        const int cols = last_c - first_c + 1, rows = last_r - first_r + 1;
        int32_t *dst = chroma + (size_t)f * m.chroma_stride + (size_t)b * 2;
        dst[0] = scale_down(lower_median(src, cols, rows, m.luma_bx), m.shift_x);
        dst[1] = scale_down(lower_median(src + 1, cols, rows, m.luma_bx), m.shift_y);
    }
}

}  // namespace

hipError_t launch_downscale(const uint16_t *codes_slot, int W, int H, int C, int32_t *level1, hipStream_t s)
{
    const int n = max(W / 2, 1) * max(H / 2, 1);
    k_downscale<<<grid_for(n), 256, 0, s>>>(codes_slot, W, H, C, level1);
    return hipGetLastError();
}

hipError_t launch_motion(const MotionGeom &g, const uint16_t *codes, const int32_t *pyr1, int centre_slot,
                         const uint32_t *nslots_dev, int neighbours, int32_t *mv, float *conf, float max_value,
                         float sad_floor, float thsad, hipStream_t s)
{
    if (neighbours <= 0) return hipSuccess;
    if (g.radius > kMaxR || g.blksize > kMaxBlk || g.blksize < 2 || g.levels < 1 || g.levels > 2 || neighbours > 16)
        return hipErrorInvalidValue;
    int use_seed = 0;
    if (g.levels == 2) {
        const int lw = max(g.W / 2, 1), lh = max(g.H / 2, 1);
        const int cblk = max(g.blksize / 2, 2), cstep = max(g.step / 2, 1);
        const dim3 grid(max((int)div_up(lw, cstep), 1), max((int)div_up(lh, cstep), 1));
        if (rows_form(cblk, g.radius)) {
#define AVD_COARSE(B, RR)                                                                                    \
    k_bm_coarse_rows<B, RR><<<grid, 64, match_rows_lds_bytes(B, RR, neighbours), s>>>(                       \
        pyr1, lw, lh, cstep, centre_slot, nslots_dev, neighbours, mv, g.blocks_x, g.blocks_y, g.step)
            AVD_ROWS_DISPATCH(cblk, g.radius, AVD_COARSE);
#undef AVD_COARSE
        } else {
            k_bm_coarse<<<grid, kBmThreads, match_lds_bytes(cblk, g.radius, neighbours), s>>>(
                pyr1, lw, lh, cblk, cstep, g.radius, centre_slot, nslots_dev, neighbours, mv, g.blocks_x,
                g.blocks_y, g.step);
        }
        use_seed = 1;
        hipError_t e = hipGetLastError();
        if (e != hipSuccess) return e;
    }
    return launch_fine(g, g.radius, codes, centre_slot, nslots_dev, neighbours, use_seed ? mv : nullptr, mv, conf,
                       max_value, sad_floor, thsad, s);
}

// run_seeded_refine (motion/analyse.rs:321-367) at the single-level geometry
// (MotionCtx::single_level, motion/context.rs:108-116): no coarse pass, the fine pass `radius`
// around each seed.
hipError_t launch_motion_refine(const MotionGeom &g, const uint16_t *codes, int centre_slot, const uint32_t *nslots_dev,
                                int neighbours, const int32_t *seeds, int32_t *mv, float *conf, float max_value,
                                float sad_floor, float thsad, int radius, hipStream_t s)
{
    if (neighbours <= 0) return hipSuccess;
    if (radius < 0 || radius > kMaxR || g.blksize > kMaxBlk || g.blksize < 2 || neighbours > 16 || !seeds)
        return hipErrorInvalidValue;
    return launch_fine(g, radius, codes, centre_slot, nslots_dev, neighbours, seeds, mv, conf, max_value, sad_floor,
                       thsad, s);
}

// map_luma_fields (motion/chroma_map.rs:56-98) over mc_map_luma_fields (kernels/chroma_map.rs:57-123).
hipError_t launch_map_luma_fields(const FieldMap &m, const int32_t *luma, int32_t *chroma, int fields, hipStream_t s)
{
    if (fields <= 0) return hipSuccess;
    if (m.luma_bx < 1 || m.luma_by < 1 || m.chroma_bx < 1 || m.chroma_by < 1 || m.luma_step < 1 || m.chroma_step < 1 ||
        m.shift_x < 0 || m.shift_x > 1 || m.shift_y < 0 || m.shift_y > 1 ||
        m.luma_stride < m.luma_bx * m.luma_by * 2 || m.chroma_stride < m.chroma_bx * m.chroma_by * 2)
        return hipErrorInvalidValue;
    const uint64_t items = (uint64_t)m.chroma_bx * m.chroma_by * fields;
    k_map_luma_fields<<<grid_for(items), 256, 0, s>>>(m, luma, chroma, fields);
    return hipGetLastError();
}

hipError_t launch_warp(const float *src, float *dst, const int32_t *mv, int W, int H, int C, int step, int blocks_x,
                       int blocks_y, hipStream_t s)
{
    k_warp<<<grid_for((uint64_t)W * H), 256, 0, s>>>(src, dst, mv, W, H, C, step, blocks_x, blocks_y);
    return hipGetLastError();
}

}  // namespace avd
