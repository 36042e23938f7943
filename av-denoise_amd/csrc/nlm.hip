// NLMeans window kernels for gfx950.
//
// Semantics follow nlmeans/kernels/fused.rs: nlm_fused_single_window (:364-530) walks every
// offset q != 0 of the (2s+1)^2 window of a frame against itself; nlm_fused_pair_accumulate_window
// (:138-347) walks every q of one temporal pair, the forward frame at +q and the backward frame at
// -q.  Per offset the per-pixel distance is sum_c (x - y)^2 * scale (3 luma, 1.5 chroma) with
// clamp-to-edge reads, the patch distance is its (2P+1)^2 box sum, the weight is
// exp(-max(D - offset, 0) * h2) (times the block confidence on a pair), and each pixel gathers
// sum(w * y), sum(w) and max(w) in registers, added to the accumulators once.  nlm_finish
// (accumulate.rs:54-101) then blends in the centre with self_weight * max_w.
//
// MI355X layout.  A wavefront owns 64 adjacent distance columns: lane l holds the distance of
// image column X0 - P + l, and after the horizontal box sum lane l (l >= 2P) owns output column
// X0 - 2P + l.  The horizontal sum is a chain of 2P DPP wave_shr:1 adds (each a single v_add_f32
// with a DPP source), so it needs no LDS and no barrier.  Each lane walks ROWS output rows of its
// column: the centre values of its ROWS + 2P distance rows stay in registers for the whole
// window, and the vertical box sum adds 2P+1 of the row sums kept in registers.  ROWS (16, 8 or 4)
// trades the 2P halo rows each wave recomputes against the number of waves a small frame spreads
// over; every choice gives bit-identical sums, and the tuner picks one per window shape.  The neighbour
// frame(s) of the workgroup are staged once in LDS with their (P+s) halo, so the offset loop runs
// with no barrier at all.  On planes up to 2^20 pixels each row group's offsets are split over SPLIT
// waves whose sums then add in LDS in part order, so small frames still fill the chip.
// exp is v_exp_f32 (exp2) with log2(e) folded into h2 on the host.
#include "common.h"

namespace avd {

namespace {

constexpr int kMaxP = 8;
constexpr int kMaxS = 8;

__device__ __forceinline__ float wave_shr1(float v)
{
    // DPP wave_shr:1: lane l reads lane l-1; lane 0 reads 0.
    return __builtin_amdgcn_update_dpp(0.0f, v, 0x138, 0xf, 0xf, false);
}

__device__ __forceinline__ int clampi(int v, int hi) { return v < 0 ? 0 : (v > hi ? hi : v); }

// Loads a clamped (rows x cols) window of an interleaved frame into LDS.
template <int C, int WV>
__device__ void stage_tile(float *tile, const float *__restrict__ frame, int W, int H, int x0, int y0, int cols,
                           int rows)
{
    const int n = cols * rows;
    for (int i = threadIdx.x; i < n; i += WV * 64) {
        const int ty = i / cols, tx = i - ty * cols;
        const int sx = clampi(x0 + tx, W - 1), sy = clampi(y0 + ty, H - 1);
        const float *src = frame + ((size_t)sy * W + sx) * C;
#pragma unroll
        for (int c = 0; c < C; ++c) tile[(size_t)i * C + c] = src[c];
    }
}

// The offset loop of one window for this lane's ROWS output rows, into acc / ws / mw (zeroed
// here): every offset q (q != 0 for a single window) against the staged tile `tf` (the forward
// frame of a pair, or the frame itself) and, for a pair, `tb` (the backward frame at -q).  `cen`
// holds the lane's centre column of ROWS + 2P distance rows; `tcols` is the tiles' row pitch.
template <int C, int P, int ROWS, bool PAIR, bool CONF, int SPLIT>
__device__ __forceinline__ void window_sums(const NlmWindow &w, const float *tf, const float *tb, int tcols, int wave,
                                            int lane, int part, const float (&cen)[ROWS + 2 * P][C],
                                            const float *__restrict__ lut, const float (&conf_fv)[ROWS],
                                            const float (&conf_bv)[ROWS], float (&acc)[ROWS][C], float (&ws)[ROWS],
                                            float (&mw)[ROWS])
{
    constexpr int DR = ROWS + 2 * P;
    const int S = w.S;
    const float scale = (C == 1) ? 3.0f : 1.5f;
    const float h2l = w.h2_log2e;
#pragma unroll
    for (int j = 0; j < ROWS; ++j) {
#pragma unroll
        for (int c = 0; c < C; ++c) acc[j][c] = 0.0f;
        ws[j] = 0.0f;
        mw[j] = 0.0f;
    }
    const int side = 2 * S + 1;
    // Tile coordinates of this lane's distance position at offset 0.
    const int trow0 = wave * ROWS + S;  // + r + qy
    const int tcol0 = lane + S;          // + qx
    // Gather column of the output pixel: lane - P + S (+/- qx).  Lanes below 2P own no output and
    // read a clamped column so every LDS read stays inside the tile.
    const int gcol0 = max(lane, P) - P + S;
    // This wave's share of the offsets: the rows qy of part `part` of SPLIT consecutive runs of
    // offset rows (a test inside the loop for a finer split cost the loop its shape and doubled its
    // registers).
    const int qy_lo = -S + part * side / SPLIT, qy_hi = -S + (part + 1) * side / SPLIT;
    for (int qy = qy_lo; qy < qy_hi; ++qy) {
        for (int qx = -S; qx <= S; ++qx) {
            if (!PAIR && qx == 0 && qy == 0) continue;
            const float off = PAIR ? w.offset : lut[(qy + S) * side + (qx + S)];
            float hf[DR], hb[DR];
#pragma unroll
            for (int r = 0; r < DR; ++r) {
                const float *nf = tf + ((size_t)(trow0 + r + qy) * tcols + (tcol0 + qx)) * C;
                float df = 0.0f;
#pragma unroll
                for (int c = 0; c < C; ++c) {
                    const float e = cen[r][c] - nf[c];
                    df += e * e;
                }
                df *= scale;
                float a = df;
#pragma unroll
                for (int k = 0; k < 2 * P; ++k) a = df + wave_shr1(a);
                hf[r] = a;
                if (PAIR) {
                    const float *nb = tb + ((size_t)(trow0 + r - qy) * tcols + (tcol0 - qx)) * C;
                    float db = 0.0f;
#pragma unroll
                    for (int c = 0; c < C; ++c) {
                        const float e = cen[r][c] - nb[c];
                        db += e * e;
                    }
                    db *= scale;
                    float b = db;
#pragma unroll
                    for (int k = 0; k < 2 * P; ++k) b = db + wave_shr1(b);
                    hb[r] = b;
                }
            }
#pragma unroll
            for (int j = 0; j < ROWS; ++j) {
                float Df = 0.0f;
#pragma unroll
                for (int k = 0; k <= 2 * P; ++k) Df += hf[j + k];
                float wf = __builtin_amdgcn_exp2f(-fmaxf(Df - off, 0.0f) * h2l);
                if (CONF) wf *= conf_fv[j];
                // The gathered neighbour sits at the output pixel plus q: distance row j+P, column lane-P.
                const float *gf = tf + ((size_t)(trow0 + j + P + qy) * tcols + (gcol0 + qx)) * C;
                if (PAIR) {
                    float Db = 0.0f;
#pragma unroll
                    for (int k = 0; k <= 2 * P; ++k) Db += hb[j + k];
                    float wb = __builtin_amdgcn_exp2f(-fmaxf(Db - off, 0.0f) * h2l);
                    if (CONF) wb *= conf_bv[j];
                    const float *gb = tb + ((size_t)(trow0 + j + P - qy) * tcols + (gcol0 - qx)) * C;
#pragma unroll
                    for (int c = 0; c < C; ++c) acc[j][c] = acc[j][c] + gf[c] * wf + gb[c] * wb;
                    ws[j] += wf + wb;
                    mw[j] = fmaxf(mw[j], fmaxf(wf, wb));
                } else {
#pragma unroll
                    for (int c = 0; c < C; ++c) acc[j][c] += gf[c] * wf;
                    ws[j] += wf;
                    mw[j] = fmaxf(mw[j], wf);
                }
            }
        }
    }
}

// Adds the window sums of parts 1 .. SPLIT-1 into part 0's, in part order, through `red`: LDS of
// (SPLIT - 1) * WV * kCombineRows * (C + 2) * 64 floats, kCombineRows output rows at a time (the
// tiles are done with, so the caller passes their memory).  Every wave of the workgroup calls it;
// part 0 holds the totals afterwards.
constexpr int kCombineRows = 4;

template <int C, int ROWS, int WV, int SPLIT>
__device__ __forceinline__ void combine_parts(float *red, int wave, int lane, int part, float (&acc)[ROWS][C],
                                              float (&ws)[ROWS], float (&mw)[ROWS])
{
    if constexpr (SPLIT > 1) {
        constexpr int CR = ROWS < kCombineRows ? ROWS : kCombineRows, F = CR * (C + 2);
#pragma unroll
        for (int j0 = 0; j0 < ROWS; j0 += CR) {
            __syncthreads();  // every wave is done with the tiles, or part 0 with the last rows
            if (part > 0) {
                float *dst = red + ((size_t)((part - 1) * WV + wave) * F) * 64 + lane;
#pragma unroll
                for (int j = 0; j < CR; ++j) {
#pragma unroll
                    for (int c = 0; c < C; ++c) dst[(size_t)(j * (C + 2) + c) * 64] = acc[j0 + j][c];
                    dst[(size_t)(j * (C + 2) + C) * 64] = ws[j0 + j];
                    dst[(size_t)(j * (C + 2) + C + 1) * 64] = mw[j0 + j];
                }
            }
            __syncthreads();
            if (part == 0) {
#pragma unroll
                for (int p = 1; p < SPLIT; ++p) {
                    const float *src = red + ((size_t)((p - 1) * WV + wave) * F) * 64 + lane;
#pragma unroll
                    for (int j = 0; j < CR; ++j) {
#pragma unroll
                        for (int c = 0; c < C; ++c)
                            acc[j0 + j][c] = acc[j0 + j][c] + src[(size_t)(j * (C + 2) + c) * 64];
                        ws[j0 + j] = ws[j0 + j] + src[(size_t)(j * (C + 2) + C) * 64];
                        mw[j0 + j] = fmaxf(mw[j0 + j], src[(size_t)(j * (C + 2) + C + 1) * 64]);
                    }
                }
            }
        }
    }
}

// The centre values of this lane's distance column, one per distance row, and the confidence of
// the blocks its output rows fall in.
template <int C, int P, int ROWS>
__device__ __forceinline__ void load_centre(const NlmWindow &w, const float *__restrict__ centre, int X0, int Y0,
                                            int wave, int lane, float (&cen)[ROWS + 2 * P][C])
{
    const int xd = X0 - P + lane;
    const int ybase = Y0 + wave * ROWS - P;
#pragma unroll
    for (int r = 0; r < ROWS + 2 * P; ++r) {
        const float *src = centre + ((size_t)clampi(ybase + r, w.H - 1) * w.W + clampi(xd, w.W - 1)) * C;
#pragma unroll
        for (int c = 0; c < C; ++c) cen[r][c] = src[c];
    }
}

template <int ROWS>
__device__ __forceinline__ void load_conf(const NlmWindow &w, const float *__restrict__ conf, int ox, int Y0, int wave,
                                          float (&out)[ROWS])
{
    const int bcol = min(clampi(ox, w.W - 1) / w.step, w.blocks_x - 1);
#pragma unroll
    for (int j = 0; j < ROWS; ++j) {
        const int oy = clampi(Y0 + wave * ROWS + j, w.H - 1);
        out[j] = conf[min(oy / w.step, w.blocks_y - 1) * w.blocks_x + bcol];
    }
}

template <int C, int P, int WV, int ROWS, bool PAIR, bool CONF, int SPLIT>
__global__ __launch_bounds__(WV * SPLIT * 64) void k_nlm(NlmWindow w, const float *__restrict__ centre,
                                                             const float *__restrict__ fwd, const float *__restrict__ bwd,
                                                             const float *__restrict__ lut,
                                                             const float *__restrict__ conf_f,
                                                             const float *__restrict__ conf_b, float *__restrict__ accum,
                                                             float *__restrict__ wsum, float *__restrict__ maxw)
{
    extern __shared__ float lds[];
    const int S = w.S, W = w.W, H = w.H;
    // Wave (wave, part): row group `wave` of the workgroup's WV, offset part `part` of SPLIT.
    const int lane = threadIdx.x & 63, wave = (threadIdx.x >> 6) % WV;
    const int part = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6) / WV);  // uniform per wave
    const int X0 = blockIdx.x * (64 - 2 * P);
    const int Y0 = blockIdx.y * (WV * ROWS);

    // Neighbour tiles: columns X0-P-S .. X0-P+63+S, rows Y0-P-S .. Y0+WV*ROWS+P-1+S.
    const int tcols = 64 + 2 * S, trows = WV * ROWS + 2 * P + 2 * S;
    const int tx0 = X0 - P - S, ty0 = Y0 - P - S;
    float *tf = lds;
    float *tb = lds + (size_t)tcols * trows * C;
    stage_tile<C, WV * SPLIT>(tf, fwd, W, H, tx0, ty0, tcols, trows);
    if (PAIR) stage_tile<C, WV * SPLIT>(tb, bwd, W, H, tx0, ty0, tcols, trows);
    float cen[ROWS + 2 * P][C];
    load_centre<C, P, ROWS>(w, centre, X0, Y0, wave, lane, cen);
    __syncthreads();

    const int ox = X0 - 2 * P + lane;  // output column of this lane (valid for lane >= 2P)
    float conf_fv[ROWS], conf_bv[ROWS];
    if (CONF) {
        load_conf<ROWS>(w, conf_f, ox, Y0, wave, conf_fv);
        load_conf<ROWS>(w, conf_b, ox, Y0, wave, conf_bv);
    }
    float acc[ROWS][C], ws[ROWS], mw[ROWS];
    window_sums<C, P, ROWS, PAIR, CONF, SPLIT>(w, tf, tb, tcols, wave, lane, part, cen, lut, conf_fv, conf_bv, acc, ws,
                                               mw);
    combine_parts<C, ROWS, WV, SPLIT>(lds, wave, lane, part, acc, ws, mw);

    if (part != 0 || lane < 2 * P || ox >= W) return;
#pragma unroll
    for (int j = 0; j < ROWS; ++j) {
        const int oy = Y0 + wave * ROWS + j;
        if (oy >= H) break;
        const size_t p = (size_t)oy * W + ox;
#pragma unroll
        for (int c = 0; c < C; ++c) accum[p * C + c] += acc[j][c];
        wsum[p] += ws[j];
        maxw[p] = fmaxf(maxw[p], mw[j]);
    }
}

// The batched form: one launch per frame and plane group walks every temporal pair (oldest first),
// then the single window, then the finish, with the window sums of each tile in registers.  Each
// window's sums start from zero and join the running totals in the plain form's order (the plain
// form adds every launch's sums to accumulators that start at zero), so the output is bit-identical
// to the plain launches' while the accumulators never reach memory.  The pair tiles are re-staged
// per pair; the single window stages the centre frame into the first of them.
template <int C, int P, int WV, int ROWS, bool CONF, int SPLIT>
__global__ __launch_bounds__(WV * SPLIT * 64) void k_nlm_window(NlmWindow w, NlmFrames f,
                                                                const float *__restrict__ centre,
                                                                const float *__restrict__ lut, float self_weight,
                                                                float *__restrict__ out)
{
    extern __shared__ float lds[];
    const int S = w.S, W = w.W, H = w.H;
    const int lane = threadIdx.x & 63, wave = (threadIdx.x >> 6) % WV;
    const int part = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6) / WV);  // uniform per wave
    const int X0 = blockIdx.x * (64 - 2 * P);
    const int Y0 = blockIdx.y * (WV * ROWS);
    const int tcols = 64 + 2 * S, trows = WV * ROWS + 2 * P + 2 * S;
    const int tx0 = X0 - P - S, ty0 = Y0 - P - S;
    float *tf = lds;
    float *tb = lds + (size_t)tcols * trows * C;
    float cen[ROWS + 2 * P][C];
    load_centre<C, P, ROWS>(w, centre, X0, Y0, wave, lane, cen);
    const int ox = X0 - 2 * P + lane;

    float tacc[ROWS][C], tws[ROWS], tmw[ROWS];
#pragma unroll
    for (int j = 0; j < ROWS; ++j) {
#pragma unroll
        for (int c = 0; c < C; ++c) tacc[j][c] = 0.0f;
        tws[j] = 0.0f;
        tmw[j] = 0.0f;
    }
    float acc[ROWS][C], ws[ROWS], mw[ROWS];
    float conf_fv[ROWS], conf_bv[ROWS];
    for (int k = 0; k < f.pairs; ++k) {
        if (k > 0) __syncthreads();  // every wave is done with the previous pair's tiles
        stage_tile<C, WV * SPLIT>(tf, f.fwd[k], W, H, tx0, ty0, tcols, trows);
        stage_tile<C, WV * SPLIT>(tb, f.bwd[k], W, H, tx0, ty0, tcols, trows);
        __syncthreads();
        if (CONF) {
            load_conf<ROWS>(w, f.conf_f[k], ox, Y0, wave, conf_fv);
            load_conf<ROWS>(w, f.conf_b[k], ox, Y0, wave, conf_bv);
        }
        window_sums<C, P, ROWS, true, CONF, SPLIT>(w, tf, tb, tcols, wave, lane, part, cen, nullptr, conf_fv, conf_bv,
                                                   acc, ws, mw);
        combine_parts<C, ROWS, WV, SPLIT>(lds, wave, lane, part, acc, ws, mw);
#pragma unroll
        for (int j = 0; j < ROWS; ++j) {
#pragma unroll
            for (int c = 0; c < C; ++c) tacc[j][c] = tacc[j][c] + acc[j][c];
            tws[j] = tws[j] + ws[j];
            tmw[j] = fmaxf(tmw[j], mw[j]);
        }
    }
    if (f.pairs > 0) __syncthreads();
    stage_tile<C, WV * SPLIT>(tf, centre, W, H, tx0, ty0, tcols, trows);
    __syncthreads();
    window_sums<C, P, ROWS, false, false, SPLIT>(w, tf, tf, tcols, wave, lane, part, cen, lut, conf_fv, conf_bv, acc,
                                                 ws, mw);
    combine_parts<C, ROWS, WV, SPLIT>(lds, wave, lane, part, acc, ws, mw);

    if (part != 0 || lane < 2 * P || ox >= W) return;
    // The finish (accumulate.rs:54-101), as k_finish computes it.
#pragma unroll
    for (int j = 0; j < ROWS; ++j) {
        const int oy = Y0 + wave * ROWS + j;
        if (oy >= H) break;
        float a[C];
#pragma unroll
        for (int c = 0; c < C; ++c) a[c] = tacc[j][c] + acc[j][c];
        const float wsum = tws[j] + ws[j];
        const float maxw = fmaxf(tmw[j], mw[j]);
        const size_t p = (size_t)oy * W + ox;
        const float sw = self_weight * maxw;
        const float den = sw + wsum;
        if (den > 1e-30f) {
            const float inv = 1.0f / den;
#pragma unroll
            for (int c = 0; c < C; ++c) out[p * C + c] = (centre[p * C + c] * sw + a[c]) * inv;
        } else {
#pragma unroll
            for (int c = 0; c < C; ++c) out[p * C + c] = centre[p * C + c];
        }
    }
}

template <int C>
__global__ __launch_bounds__(256) void k_finish(const float *__restrict__ frame, const float *__restrict__ accum,
                                                const float *__restrict__ wsum, const float *__restrict__ maxw,
                                                float self_weight, uint32_t pixels, float *__restrict__ out)
{
    for (uint32_t p = blockIdx.x * 256 + threadIdx.x; p < pixels; p += gridDim.x * 256) {
        const float sw = self_weight * maxw[p];
        const float den = sw + wsum[p];
        if (den > 1e-30f) {
            const float inv = 1.0f / den;
#pragma unroll
            for (int c = 0; c < C; ++c) out[(size_t)p * C + c] = (frame[(size_t)p * C + c] * sw + accum[(size_t)p * C + c]) * inv;
        } else {
#pragma unroll
            for (int c = 0; c < C; ++c) out[(size_t)p * C + c] = frame[(size_t)p * C + c];
        }
    }
}

// Row groups per workgroup: pairs of two-channel frames stage two tiles of twice the width, so they
// take half the rows to keep the LDS footprint near that of a luma pass; and a workgroup holds at
// most eight waves (row groups x offset parts), four for two channels, whose windows with large
// patches need more than 256 VGPRs.
constexpr int waves_for(int C, bool pair, int split)
{
    const int base = (pair && C == 2) ? 2 : 4, cap = C == 2 ? 4 : 8;
    return base * split > cap ? (cap / split > 0 ? cap / split : 1) : base;
}

// Offset parts per row group: each workgroup runs SPLIT waves per row group, each on a consecutive
// run of the window's offset rows, so a frame puts SPLIT times the waves on the chip.  Planes up to
// 2^20 pixels, which would leave SIMDs idle otherwise, take four (one channel) or two (two
// channels, whose workgroups hold at most four waves); larger ones fill the chip with one (on an
// MI355X: 640x360 spatial luma 12.3k frames/s against 6.3k at one part; 1920x1080 4:2:0 at
// temporal radius 2, 643 frames/s against 596 at one part everywhere, 610 with four for chroma and
// ≈620 with two for luma).  The part sums add in part order, the same for every tile shape and both
// launch forms, which therefore stay bit-identical.
inline int offset_split(int W, int H, int C) { return (size_t)W * H <= (1u << 20) ? (C == 1 ? 4 : 2) : 1; }

// LDS of a workgroup: its tile(s), or the part sums while they combine, whichever is larger.
size_t tile_bytes(int C, int P, int S, bool pair, int rows, int split)
{
    const int wv = waves_for(C, pair, split);
    const size_t one = (size_t)(64 + 2 * S) * (wv * rows + 2 * P + 2 * S) * C * sizeof(float);
    const size_t tiles = pair ? 2 * one : one;
    const int cr = rows < kCombineRows ? rows : kCombineRows;
    const size_t parts = (size_t)(split - 1) * wv * cr * (C + 2) * 64 * sizeof(float);
    return tiles > parts ? tiles : parts;
}

int device_lds_limit()
{
    static int limit = -1;
    if (limit < 0) {
        int dev = 0, v = 0;
        if (hipGetDevice(&dev) != hipSuccess ||
            hipDeviceGetAttribute(&v, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) != hipSuccess)
            v = 64 * 1024;
        limit = v;
    }
    return limit;
}

template <int C, int P, int ROWS, int SPLIT>
hipError_t dispatch_split(const NlmWindow &w, const float *centre, const float *fwd, const float *bwd,
                          const float *lut, const float *cf, const float *cb, float *accum, float *wsum, float *maxw,
                          bool pair, bool conf, hipStream_t s)
{
    constexpr int WS = waves_for(C, false, SPLIT), WP = waves_for(C, true, SPLIT);
    const size_t lds = tile_bytes(C, P, w.S, pair, ROWS, SPLIT);
    if (!pair) {
        const dim3 grid(div_up(w.W, 64 - 2 * P), div_up(w.H, WS * ROWS));
        k_nlm<C, P, WS, ROWS, false, false, SPLIT><<<grid, WS * SPLIT * 64, lds, s>>>(w, centre, fwd, bwd, lut, cf,
                                                                                     cb, accum, wsum, maxw);
    } else {
        const dim3 grid(div_up(w.W, 64 - 2 * P), div_up(w.H, WP * ROWS));
        if (conf)
            k_nlm<C, P, WP, ROWS, true, true, SPLIT><<<grid, WP * SPLIT * 64, lds, s>>>(w, centre, fwd, bwd, lut, cf,
                                                                                       cb, accum, wsum, maxw);
        else
            k_nlm<C, P, WP, ROWS, true, false, SPLIT><<<grid, WP * SPLIT * 64, lds, s>>>(w, centre, fwd, bwd, lut, cf,
                                                                                        cb, accum, wsum, maxw);
    }
    return hipGetLastError();
}

template <int C, int P, int ROWS>
hipError_t dispatch_rows(const NlmWindow &w, const float *centre, const float *fwd, const float *bwd, const float *lut,
                         const float *cf, const float *cb, float *accum, float *wsum, float *maxw, bool pair, bool conf,
                         hipStream_t s)
{
    switch (offset_split(w.W, w.H, C)) {
    case 4: return dispatch_split<C, P, ROWS, 4>(w, centre, fwd, bwd, lut, cf, cb, accum, wsum, maxw, pair, conf, s);
    case 2: return dispatch_split<C, P, ROWS, 2>(w, centre, fwd, bwd, lut, cf, cb, accum, wsum, maxw, pair, conf, s);
    default: return dispatch_split<C, P, ROWS, 1>(w, centre, fwd, bwd, lut, cf, cb, accum, wsum, maxw, pair, conf, s);
    }
}

template <int C, int P>
hipError_t dispatch(const NlmWindow &w, const float *centre, const float *fwd, const float *bwd, const float *lut,
                    const float *cf, const float *cb, float *accum, float *wsum, float *maxw, bool pair, bool conf,
                    int variant, hipStream_t s)
{
    switch (variant) {
    case 1: return dispatch_rows<C, P, 8>(w, centre, fwd, bwd, lut, cf, cb, accum, wsum, maxw, pair, conf, s);
    case 2: return dispatch_rows<C, P, 4>(w, centre, fwd, bwd, lut, cf, cb, accum, wsum, maxw, pair, conf, s);
    default: return dispatch_rows<C, P, 16>(w, centre, fwd, bwd, lut, cf, cb, accum, wsum, maxw, pair, conf, s);
    }
}

template <int C>
hipError_t dispatch_p(const NlmWindow &w, const float *centre, const float *fwd, const float *bwd, const float *lut,
                      const float *cf, const float *cb, float *accum, float *wsum, float *maxw, bool pair, bool conf,
                      int variant, hipStream_t s)
{
    switch (w.P) {
    case 0: return dispatch<C, 0>(w, centre, fwd, bwd, lut, cf, cb, accum, wsum, maxw, pair, conf, variant, s);
    case 1: return dispatch<C, 1>(w, centre, fwd, bwd, lut, cf, cb, accum, wsum, maxw, pair, conf, variant, s);
    case 2: return dispatch<C, 2>(w, centre, fwd, bwd, lut, cf, cb, accum, wsum, maxw, pair, conf, variant, s);
    case 3: return dispatch<C, 3>(w, centre, fwd, bwd, lut, cf, cb, accum, wsum, maxw, pair, conf, variant, s);
    case 4: return dispatch<C, 4>(w, centre, fwd, bwd, lut, cf, cb, accum, wsum, maxw, pair, conf, variant, s);
    case 5: return dispatch<C, 5>(w, centre, fwd, bwd, lut, cf, cb, accum, wsum, maxw, pair, conf, variant, s);
    case 6: return dispatch<C, 6>(w, centre, fwd, bwd, lut, cf, cb, accum, wsum, maxw, pair, conf, variant, s);
    case 7: return dispatch<C, 7>(w, centre, fwd, bwd, lut, cf, cb, accum, wsum, maxw, pair, conf, variant, s);
    case 8: return dispatch<C, 8>(w, centre, fwd, bwd, lut, cf, cb, accum, wsum, maxw, pair, conf, variant, s);
    default: return hipErrorInvalidValue;
    }
}

template <int C, int P, int ROWS, int SPLIT>
hipError_t dispatch_window_split(const NlmWindow &w, const NlmFrames &f, const float *centre, const float *lut,
                                 float self_weight, float *out, hipStream_t s)
{
    constexpr int WV = waves_for(C, true, SPLIT);
    const size_t lds = tile_bytes(C, P, w.S, true, ROWS, SPLIT);
    const dim3 grid(div_up(w.W, 64 - 2 * P), div_up(w.H, WV * ROWS));
    const bool conf = f.pairs > 0 && f.conf_f[0] != nullptr;
    if (conf)
        k_nlm_window<C, P, WV, ROWS, true, SPLIT><<<grid, WV * SPLIT * 64, lds, s>>>(w, f, centre, lut, self_weight,
                                                                                    out);
    else
        k_nlm_window<C, P, WV, ROWS, false, SPLIT><<<grid, WV * SPLIT * 64, lds, s>>>(w, f, centre, lut, self_weight,
                                                                                     out);
    return hipGetLastError();
}

template <int C, int P, int ROWS>
hipError_t dispatch_window(const NlmWindow &w, const NlmFrames &f, const float *centre, const float *lut,
                           float self_weight, float *out, hipStream_t s)
{
    switch (offset_split(w.W, w.H, C)) {
    case 4: return dispatch_window_split<C, P, ROWS, 4>(w, f, centre, lut, self_weight, out, s);
    case 2: return dispatch_window_split<C, P, ROWS, 2>(w, f, centre, lut, self_weight, out, s);
    default: return dispatch_window_split<C, P, ROWS, 1>(w, f, centre, lut, self_weight, out, s);
    }
}

template <int C>
hipError_t dispatch_window_p(const NlmWindow &w, const NlmFrames &f, const float *centre, const float *lut,
                             float self_weight, float *out, int variant, hipStream_t s)
{
#define AVD_WINDOW_P(PP)                                                                                      \
    case PP:                                                                                                  \
        return variant == 1 ? dispatch_window<C, PP, 4>(w, f, centre, lut, self_weight, out, s)              \
                            : dispatch_window<C, PP, 8>(w, f, centre, lut, self_weight, out, s);
    switch (w.P) {
        AVD_WINDOW_P(0)
        AVD_WINDOW_P(1)
        AVD_WINDOW_P(2)
        AVD_WINDOW_P(3)
        AVD_WINDOW_P(4)
        AVD_WINDOW_P(5)
        AVD_WINDOW_P(6)
        AVD_WINDOW_P(7)
        AVD_WINDOW_P(8)
    default: return hipErrorInvalidValue;
    }
#undef AVD_WINDOW_P
}

}  // namespace

bool nlm_supported(int C, int P, int S)
{
    return (C == 1 || C == 2) && P >= 0 && P <= kMaxP && S >= 0 && S <= kMaxS &&
           tile_bytes(C, P, S, true, 16, 2) <= (size_t)device_lds_limit() &&
           tile_bytes(C, P, S, false, 16, 2) <= (size_t)device_lds_limit() &&
           tile_bytes(C, P, S, true, 16, 4) <= (size_t)device_lds_limit() &&
           tile_bytes(C, P, S, false, 16, 4) <= (size_t)device_lds_limit();
}
int nlm_variant_rows(int variant) { return variant == 1 ? 8 : (variant == 2 ? 4 : 16); }

hipError_t launch_nlm_single(const NlmWindow &w, const float *frame, const float *lut, float *accum, float *wsum,
                             float *maxw, hipStream_t s, int variant)
{
    if (!nlm_supported(w.C, w.P, w.S) || variant < 0 || variant >= kNlmVariants) return hipErrorInvalidValue;
    if (w.C == 1)
        return dispatch_p<1>(w, frame, frame, frame, lut, nullptr, nullptr, accum, wsum, maxw, false, false, variant, s);
    return dispatch_p<2>(w, frame, frame, frame, lut, nullptr, nullptr, accum, wsum, maxw, false, false, variant, s);
}

hipError_t launch_nlm_pair(const NlmWindow &w, const float *centre, const float *fwd, const float *bwd,
                           const float *conf_f, const float *conf_b, float *accum, float *wsum, float *maxw,
                           hipStream_t s, int variant)
{
    if (!nlm_supported(w.C, w.P, w.S) || variant < 0 || variant >= kNlmVariants) return hipErrorInvalidValue;
    const bool conf = conf_f != nullptr && conf_b != nullptr;
    if (w.C == 1)
        return dispatch_p<1>(w, centre, fwd, bwd, nullptr, conf_f, conf_b, accum, wsum, maxw, true, conf, variant, s);
    return dispatch_p<2>(w, centre, fwd, bwd, nullptr, conf_f, conf_b, accum, wsum, maxw, true, conf, variant, s);
}

hipError_t launch_nlm_window(const NlmWindow &w, const NlmFrames &f, const float *centre, const float *lut,
                             float self_weight, float *out, hipStream_t s, int variant)
{
    if (!nlm_supported(w.C, w.P, w.S) || variant < 0 || variant >= kNlmBatchedVariants || f.pairs < 0 ||
        f.pairs > kNlmMaxPairs || !centre || !lut || !out)
        return hipErrorInvalidValue;
    for (int k = 0; k < f.pairs; ++k)
        if (!f.fwd[k] || !f.bwd[k] || (f.conf_f[0] != nullptr) != (f.conf_f[k] != nullptr && f.conf_b[k] != nullptr))
            return hipErrorInvalidValue;
    if (w.C == 1) return dispatch_window_p<1>(w, f, centre, lut, self_weight, out, variant, s);
    return dispatch_window_p<2>(w, f, centre, lut, self_weight, out, variant, s);
}

hipError_t launch_nlm_finish(const float *frame, const float *accum, const float *wsum, const float *maxw,
                             float self_weight, uint32_t pixels, int C, float *out, hipStream_t s)
{
    unsigned g = div_up(pixels, 256);
    if (g > 4096) g = 4096;
    if (g < 1) g = 1;
    if (C == 1) k_finish<1><<<g, 256, 0, s>>>(frame, accum, wsum, maxw, self_weight, pixels, out);
    else k_finish<2><<<g, 256, 0, s>>>(frame, accum, wsum, maxw, self_weight, pixels, out);
    return hipGetLastError();
}

}  // namespace avd
