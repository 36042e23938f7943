/*
 * avd_oracle.c -- CPU restatement of av-denoise's denoising core.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product links, loads or calls this file: the parity
 * tests, the smoke check and the CPU baseline of the benchmark run the HIP path against it.  Every function cites the reference file:line it restates.
 *
 * Arithmetic contract
 *   - Sample codes, block-match SAD/SSD, member positions, group words and the fixed-point
 *     accumulators are integers and are exact.
 *   - The NL4D filter is written twice over one template: `_f32` repeats the reference's f32
 *     operation order (fmaf in the DCT sums, round-half-away in to_fixed), `_f64` evaluates the
 *     same formulas in double.  The HIP FP32 form is checked bit-exact against `_f32`, and the
 *     `_f64` form is the accuracy yardstick (PSNR).
 *   - The NLM window sums are restated in both precisions; they are continuous, so the HIP path
 *     is checked against them to a stated tolerance.
 *
 * Where the reference scores block-match candidates as f32 sums of normalised samples, this
 * restatement scores the exact integer sums (SSD for NL4D grouping, SAD for motion).  The f32
 * value is the integer sum over max^2 (or max) up to rounding, so both rank candidates alike
 * except at f32 near-ties; the integer form is the contract the HIP kernels meet bit for bit.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#ifdef _OPENMP
#include <omp.h>
#endif

#define PATCH 8
#define MAX_K 8
/* Group lists hold up to 16 members: the stride-3 / 16-member grid extends the reference's
 * MAX_K = 8 (collab/mod.rs:16-20); every size below is pp->k_max (8 or 16). */
#define KMAX 16
#define UNFILLED_D 0xFFFFFFFFFFFFFFFFull

static inline int clampi(int v, int lo, int hi) { return v < lo ? lo : (v > hi ? hi : v); }

/* ------------------------------------------------------------------------------------------ */
/* Plane ingest / egress: engine/kernels.rs:25-63 (ingest_words), :103-158 (egress_words).     */
/* ------------------------------------------------------------------------------------------ */

/* ring[p*C + c] = ((word[p/spw] >> (p%spw)*bits) & mask) / max, for channels planes. */
void avdo_ingest_words(const uint32_t *plane0, const uint32_t *plane1, float *slot, uint32_t pixels,
                       uint32_t channels, uint32_t spw, float maxv)
{
    uint32_t bits = 32u / spw, mask = (bits == 32u) ? 0xFFFFFFFFu : ((1u << bits) - 1u);
    for (uint32_t p = 0; p < pixels; ++p) {
        uint32_t w = p / spw, sh = (p % spw) * bits;
        slot[(size_t)p * channels] = (float)((plane0[w] >> sh) & mask) / maxv;
        if (channels > 1) slot[(size_t)p * channels + 1] = (float)((plane1[w] >> sh) & mask) / maxv;
    }
}

/* code = u32(clamp(v,0,1)*max + 0.5), packed per word, tail lanes zero. */
void avdo_egress_words(const float *frame, uint32_t *plane0, uint32_t *plane1, uint32_t pixels,
                       uint32_t channels, uint32_t spw, float maxv)
{
    uint32_t bits = 32u / spw;
    uint32_t words = (pixels + spw - 1) / spw;
    for (uint32_t w = 0; w < words; ++w) {
        uint32_t p0 = 0, p1 = 0;
        for (uint32_t l = 0; l < spw; ++l) {
            uint32_t p = w * spw + l;
            if (p >= pixels) continue;
            float v0 = frame[(size_t)p * channels];
            v0 = v0 < 0.f ? 0.f : (v0 > 1.f ? 1.f : v0);
            volatile float t0 = v0 * maxv;
            p0 |= ((uint32_t)(t0 + 0.5f)) << (l * bits);
            if (channels > 1) {
                float v1 = frame[(size_t)p * channels + 1];
                v1 = v1 < 0.f ? 0.f : (v1 > 1.f ? 1.f : v1);
                volatile float t1 = v1 * maxv;
                p1 |= ((uint32_t)(t1 + 0.5f)) << (l * bits);
            }
        }
        plane0[w] = p0;
        if (channels > 1) plane1[w] = p1;
    }
}

/* ------------------------------------------------------------------------------------------ */
/* NLMeans window sums: nlmeans/kernels/fused.rs:138-357 (pair), :364-530 (single),            */
/* helpers.rs:36-72 (line_sum_sq, channel_scale, welsch_weight), accumulate.rs:54-101 (finish). */
/* ------------------------------------------------------------------------------------------ */

#define NLM_TEMPLATE(R, SFX, EXPF)                                                                  \
    /* d(t,q): squared difference of centre pixel t and neighbour t+q, both clamped per axis. */   \
    static inline R nlm_pix_dist##SFX(const float *cf, const float *nf, int W, int H, int C, int tx, \
                                      int ty, int qx, int qy, R scale)                              \
    {                                                                                               \
        int cx = clampi(tx, 0, W - 1), cy = clampi(ty, 0, H - 1);                                   \
        int nx = clampi(tx + qx, 0, W - 1), ny = clampi(ty + qy, 0, H - 1);                         \
        const float *a = cf + ((size_t)cy * W + cx) * C, *b = nf + ((size_t)ny * W + nx) * C;       \
        R s = 0;                                                                                    \
        for (int c = 0; c < C; ++c) {                                                               \
            R d = (R)a[c] - (R)b[c];                                                                \
            s += d * d;                                                                             \
        }                                                                                           \
        return s * scale;                                                                           \
    }                                                                                               \
    /* Patch distance at pixel (x,y) for offset q, summed row-major over the patch. */              \
    static R nlm_patch_dist##SFX(const float *cf, const float *nf, int W, int H, int C, int P, int x, \
                                 int y, int qx, int qy, R scale)                                    \
    {                                                                                               \
        R sum = 0;                                                                                  \
        for (int uy = -P; uy <= P; ++uy)                                                            \
            for (int ux = -P; ux <= P; ++ux)                                                        \
                sum += nlm_pix_dist##SFX(cf, nf, W, H, C, x + ux, y + uy, qx, qy, scale);           \
        return sum;                                                                                 \
    }                                                                                               \
    /* Spatial window of frame f against itself, q != 0, offsets from lut (fused.rs:364-530). */    \
    void avdo_nlm_single##SFX(const float *f, int W, int H, int C, int P, int S, float h2,          \
                              const float *lut, R *accum, R *wsum, R *maxw)                         \
    {                                                                                               \
        R scale = (C == 1) ? (R)3.0 : (R)1.5;                                                       \
        int side = 2 * S + 1;                                                                       \
        _Pragma("omp parallel for schedule(dynamic, 4)") for (int y = 0; y < H; ++y)               \
        {                                                                                           \
            for (int x = 0; x < W; ++x) {                                                           \
                R acc[2] = {0, 0}, ws = 0, mw = 0;                                                  \
                for (int qy = -S; qy <= S; ++qy)                                                    \
                    for (int qx = -S; qx <= S; ++qx) {                                              \
                        if (qx == 0 && qy == 0) continue;                                           \
                        R d = nlm_patch_dist##SFX(f, f, W, H, C, P, x, y, qx, qy, scale);           \
                        R off = (R)lut[(qy + S) * side + (qx + S)];                                 \
                        R e = d - off;                                                              \
                        if (e < 0) e = 0;                                                           \
                        R w = EXPF(-e * (R)h2);                                                     \
                        int nx = clampi(x + qx, 0, W - 1), ny = clampi(y + qy, 0, H - 1);           \
                        const float *nb = f + ((size_t)ny * W + nx) * C;                            \
                        for (int c = 0; c < C; ++c) acc[c] += (R)nb[c] * w;                         \
                        ws += w;                                                                    \
                        if (w > mw) mw = w;                                                         \
                    }                                                                               \
                size_t p = (size_t)y * W + x;                                                       \
                for (int c = 0; c < C; ++c) accum[p * C + c] += acc[c];                             \
                wsum[p] += ws;                                                                      \
                if (mw > maxw[p]) maxw[p] = mw;                                                     \
            }                                                                                       \
        }                                                                                           \
    }                                                                                               \
    /* One temporal pair: fwd frame at +q, bwd frame at -q, whole window incl. q=0               */ \
    /* (fused.rs:138-347).  conf_fwd/conf_bwd may be NULL (no confidence).                        */ \
    void avdo_nlm_pair##SFX(const float *cf, const float *ff, const float *bf, int W, int H, int C,  \
                            int P, int S, float h2, float noise_offset, const float *conf_fwd,      \
                            const float *conf_bwd, int step, int blocks_x, int blocks_y, R *accum,  \
                            R *wsum, R *maxw)                                                       \
    {                                                                                               \
        R scale = (C == 1) ? (R)3.0 : (R)1.5;                                                       \
        _Pragma("omp parallel for schedule(dynamic, 4)") for (int y = 0; y < H; ++y)               \
        {                                                                                           \
            for (int x = 0; x < W; ++x) {                                                           \
                R acc[2] = {0, 0}, ws = 0, mw = 0;                                                  \
                int bcol = x / step, brow = y / step;                                               \
                if (bcol > blocks_x - 1) bcol = blocks_x - 1;                                       \
                if (brow > blocks_y - 1) brow = blocks_y - 1;                                       \
                int bidx = brow * blocks_x + bcol;                                                  \
                for (int qy = -S; qy <= S; ++qy)                                                    \
                    for (int qx = -S; qx <= S; ++qx) {                                              \
                        R df = nlm_patch_dist##SFX(cf, ff, W, H, C, P, x, y, qx, qy, scale);        \
                        R db = nlm_patch_dist##SFX(cf, bf, W, H, C, P, x, y, -qx, -qy, scale);      \
                        R ef = df - (R)noise_offset, eb = db - (R)noise_offset;                     \
                        if (ef < 0) ef = 0;                                                         \
                        if (eb < 0) eb = 0;                                                         \
                        R wf = EXPF(-ef * (R)h2), wb = EXPF(-eb * (R)h2);                           \
                        if (conf_fwd) {                                                             \
                            wf *= (R)conf_fwd[bidx];                                                \
                            wb *= (R)conf_bwd[bidx];                                                \
                        }                                                                           \
                        int fx = clampi(x + qx, 0, W - 1), fy = clampi(y + qy, 0, H - 1);           \
                        int bx = clampi(x - qx, 0, W - 1), by = clampi(y - qy, 0, H - 1);           \
                        const float *fp = ff + ((size_t)fy * W + fx) * C;                           \
                        const float *bp = bf + ((size_t)by * W + bx) * C;                           \
                        for (int c = 0; c < C; ++c) acc[c] = acc[c] + (R)fp[c] * wf + (R)bp[c] * wb; \
                        ws += wf + wb;                                                              \
                        R m2 = wf > wb ? wf : wb;                                                   \
                        if (m2 > mw) mw = m2;                                                       \
                    }                                                                               \
                size_t p = (size_t)y * W + x;                                                       \
                for (int c = 0; c < C; ++c) accum[p * C + c] += acc[c];                             \
                wsum[p] += ws;                                                                      \
                if (mw > maxw[p]) maxw[p] = mw;                                                     \
            }                                                                                       \
        }                                                                                           \
    }                                                                                               \
    /* out = (x*sw + accum)/(sw + wsum), sw = wref*maxw; x where the denominator <= 1e-30.       */ \
    void avdo_nlm_finish##SFX(const float *f, const R *accum, const R *wsum, const R *maxw,         \
                              float wref, int pixels, int C, float *out)                            \
    {                                                                                               \
        for (int p = 0; p < pixels; ++p) {                                                          \
            R sw = (R)wref * maxw[p];                                                               \
            R den = sw + wsum[p];                                                                   \
            for (int c = 0; c < C; ++c) {                                                           \
                size_t i = (size_t)p * C + c;                                                       \
                out[i] = (den > (R)1e-30) ? (float)(((R)f[i] * sw + accum[i]) * ((R)1 / den)) : f[i]; \
            }                                                                                       \
        }                                                                                           \
    }

NLM_TEMPLATE(float, _f32, expf)
NLM_TEMPLATE(double, _f64, exp)

/* ------------------------------------------------------------------------------------------ */
/* Motion: motion/kernels/pyramid.rs:11-44 (mc_downscale), block_match.rs:117-211 (coarse),    */
/* :306-369 (fine), :489-535 (write_fine_result, match_confidence), :653-927 (staging, SADs),   */
/* :940-1000 (argmin), compensate.rs:12-44 (mc_warp).                                          */
/* Level 0 holds codes; level 1 holds the sum of each 2x2 group of codes, which is the          */
/* reference's f32 mean times 4*max, so every SAD ranks exactly as the reference's f32 SAD.     */
/* ------------------------------------------------------------------------------------------ */

void avdo_downscale_sum(const int32_t *src, int sw, int sh, int32_t *dst, int dw, int dh)
{
    for (int y = 0; y < dh; ++y)
        for (int x = 0; x < dw; ++x) {
            int sx = 2 * x, sy = 2 * y;
            int sx1 = sx + 1 < sw ? sx + 1 : sx, sy1 = sy + 1 < sh ? sy + 1 : sy;
            dst[(size_t)y * dw + x] = src[(size_t)sy * sw + sx] + src[(size_t)sy * sw + sx1] +
                                      src[(size_t)sy1 * sw + sx] + src[(size_t)sy1 * sw + sx1];
        }
}

/* Best offset of one block over a (2R+1)^2 window centred on (cdx,cdy).  Lowest SAD, then the
 * lowest row-major index, and the window centre wins a tie with that pick.  Returns the SAD. */
static int64_t bm_search(const int32_t *cen, const int32_t *nb, int w, int h, int ox, int oy, int blk,
                         int R, int cdx, int cdy, int *bdx, int *bdy)
{
    int side = 2 * R + 1;
    int64_t best = INT64_MAX, centre_sad = 0;
    int bx = cdx, by = cdy;
    for (int dy = 0; dy < side; ++dy)
        for (int dx = 0; dx < side; ++dx) {
            int64_t sad = 0;
            for (int py = 0; py < blk; ++py)
                for (int px = 0; px < blk; ++px) {
                    int cx = clampi(ox + px, 0, w - 1), cy = clampi(oy + py, 0, h - 1);
                    int nx = clampi(ox + cdx - R + dx + px, 0, w - 1);
                    int ny = clampi(oy + cdy - R + dy + py, 0, h - 1);
                    int64_t d = (int64_t)cen[(size_t)cy * w + cx] - nb[(size_t)ny * w + nx];
                    sad += d < 0 ? -d : d;
                }
            if (dy == R && dx == R) centre_sad = sad;
            if (sad < best) {
                best = sad;
                bx = cdx + dx - R;
                by = cdy + dy - R;
            }
        }
    if (centre_sad <= best) {
        best = centre_sad;
        bx = cdx;
        by = cdy;
    }
    *bdx = bx;
    *bdy = by;
    return best;
}

/* Coarse pass on the level-1 sums: one block per (col,row), blksize 8 at step 4 for the default
 * 16/8 grid, winner scaled by 2 and written to every fine block the coarse block covers. */
void avdo_block_match_coarse(const int32_t *cen1, const int32_t *nb1, int lw, int lh, int blk, int step,
                             int R, int level_scale, int fine_blocks_x, int fine_blocks_y, int fine_step,
                             int32_t *mv)
{
    int cbx = (lw + step - 1) / step, cby = (lh + step - 1) / step;
    if (cbx < 1) cbx = 1;
    if (cby < 1) cby = 1;
    for (int br = 0; br < cby; ++br)
        for (int bc = 0; bc < cbx; ++bc) {
            int dx, dy;
            bm_search(cen1, nb1, lw, lh, bc * step, br * step, blk, R, 0, 0, &dx, &dy);
            int c0 = bc * step * level_scale / fine_step, c1 = (bc + 1) * step * level_scale / fine_step;
            int r0 = br * step * level_scale / fine_step, r1 = (br + 1) * step * level_scale / fine_step;
            if (c0 > fine_blocks_x) c0 = fine_blocks_x;
            if (r0 > fine_blocks_y) r0 = fine_blocks_y;
            c1 = (bc == cbx - 1) ? fine_blocks_x : (c1 > fine_blocks_x ? fine_blocks_x : c1);
            r1 = (br == cby - 1) ? fine_blocks_y : (r1 > fine_blocks_y ? fine_blocks_y : r1);
            for (int fr = r0; fr < r1; ++fr)
                for (int fc = c0; fc < c1; ++fc) {
                    mv[((size_t)fr * fine_blocks_x + fc) * 2] = dx * level_scale;
                    mv[((size_t)fr * fine_blocks_x + fc) * 2 + 1] = dy * level_scale;
                }
        }
}

/* match_confidence (block_match.rs:523-535) on the SAD in normalised units. */
float avdo_match_confidence(float best_sad, float floor_sad, float thsad)
{
    float excess = best_sad - floor_sad;
    if (excess < 0.f) excess = 0.f;
    float t2 = thsad * thsad, e2 = excess * excess;
    float c = (t2 - e2) / (t2 + e2);
    return c < 0.f ? 0.f : c;
}

/* Fine pass on level 0: seeded by mv (use_seed) and refined in place; optional confidence from
 * the SAD over max (normalised units), the derived SAD floor and thsad. */
void avdo_block_match_fine(const int32_t *cen0, const int32_t *nb0, int w, int h, int blk, int step,
                           int R, int blocks_x, int blocks_y, int use_seed, int32_t *mv, float *conf,
                           float maxv, float sad_floor, float thsad)
{
    for (int br = 0; br < blocks_y; ++br)
        for (int bc = 0; bc < blocks_x; ++bc) {
            size_t b = (size_t)br * blocks_x + bc;
            int sdx = use_seed ? mv[b * 2] : 0, sdy = use_seed ? mv[b * 2 + 1] : 0;
            int dx, dy;
            int64_t sad = bm_search(cen0, nb0, w, h, bc * step, br * step, blk, R, sdx, sdy, &dx, &dy);
            mv[b * 2] = dx;
            mv[b * 2 + 1] = dy;
            if (conf) conf[b] = avdo_match_confidence((float)sad / maxv, sad_floor, thsad);
        }
}

/* Luma motion onto the chroma block grid: map_luma_fields (motion/chroma_map.rs:56-98) and
 * mc_map_luma_fields (motion/kernels/chroma_map.rs:57-123).  A chroma block covers the luma area
 * (chroma_blk << shift) from (col * chroma_step) << shift; it reads covers = (area - luma_blk) /
 * luma_step + 1 luma blocks a side from start = origin / luma_step, those past the luma grid
 * dropped (the first always kept), takes the per-component lower median of what is left
 * (chroma_map.rs:22-44) and divides it by 1 << shift, rounding half away from zero (:6-16). */
static int map_scale_down(int v, int shift)
{
    if (shift == 0) return v;
    int m = v < 0 ? -v : v;
    int r = (m + (1 << (shift - 1))) / (1 << shift);
    return v < 0 ? -r : r;
}

static int map_lower_median(const int *v, int n)
{
    /* The (n-1)/2-th smallest: the value with at most that many below it and more at or below. */
    int mid = (n - 1) / 2;
    for (int i = 0; i < n; ++i) {
        int below = 0, equal = 0;
        for (int j = 0; j < n; ++j) {
            below += v[j] < v[i];
            equal += v[j] == v[i];
        }
        if (below <= mid && mid < below + equal) return v[i];
    }
    return v[0];
}

void avdo_map_luma_fields(const int32_t *luma, int lbx, int lby, int lstep, int lblk, int lstride, int32_t *chroma,
                          int cbx, int cby, int cstep, int cblk, int cstride, int shift_x, int shift_y, int fields)
{
    int area_x = cblk << shift_x, area_y = cblk << shift_y;
    int covers_x = (area_x > lblk ? area_x - lblk : 0) / lstep + 1;
    int covers_y = (area_y > lblk ? area_y - lblk : 0) / lstep + 1;
    int *vx = (int *)malloc(sizeof(int) * (size_t)covers_x * covers_y);
    int *vy = (int *)malloc(sizeof(int) * (size_t)covers_x * covers_y);
    for (int f = 0; f < fields; ++f)
        for (int row = 0; row < cby; ++row)
            for (int col = 0; col < cbx; ++col) {
                int sc = ((col * cstep) << shift_x) / lstep, sr = ((row * cstep) << shift_y) / lstep;
                int last_c = sc + covers_x - 1 < lbx - 1 ? sc + covers_x - 1 : lbx - 1;
                int last_r = sr + covers_y - 1 < lby - 1 ? sr + covers_y - 1 : lby - 1;
                int first_c = sc < last_c ? sc : last_c, first_r = sr < last_r ? sr : last_r;
                int n = 0;
                for (int r = first_r; r <= last_r && r < first_r + covers_y; ++r)
                    for (int c = first_c; c <= last_c && c < first_c + covers_x; ++c) {
                        const int32_t *src = luma + (size_t)f * lstride + ((size_t)r * lbx + c) * 2;
                        vx[n] = src[0];
                        vy[n] = src[1];
                        ++n;
                    }
                int32_t *dst = chroma + (size_t)f * cstride + ((size_t)row * cbx + col) * 2;
                dst[0] = map_scale_down(map_lower_median(vx, n), shift_x);
                dst[1] = map_scale_down(map_lower_median(vy, n), shift_y);
            }
    free(vx);
    free(vy);
}

/* mc_warp: dst(x,y) = src(clamp(x+mv.x), clamp(y+mv.y)) with block min(x/step, bx-1). */
void avdo_warp(const float *src, float *dst, const int32_t *mv, int W, int H, int C, int step,
               int blocks_x, int blocks_y)
{
    for (int y = 0; y < H; ++y)
        for (int x = 0; x < W; ++x) {
            int bc = x / step, br = y / step;
            if (bc > blocks_x - 1) bc = blocks_x - 1;
            if (br > blocks_y - 1) br = blocks_y - 1;
            const int32_t *v = mv + ((size_t)br * blocks_x + bc) * 2;
            int sx = clampi(x + v[0], 0, W - 1), sy = clampi(y + v[1], 0, H - 1);
            for (int c = 0; c < C; ++c) dst[((size_t)y * W + x) * C + c] = src[((size_t)sy * W + sx) * C + c];
        }
}

/* ------------------------------------------------------------------------------------------ */
/* NL4D collaborative pass: collab/kernels/fused/mod.rs:170-553 (search_group), :565-850       */
/* (filter_group), centre.rs:37-300, trajectory.rs:18-450, plane_ops.rs:36-77 (shift_insert8),  */
/* transforms.rs:39-187 (dct8, haar, variance ladder), grid.rs:119-220, aggregate.rs:147-211   */
/* (scatter_patch, to_fixed), :252-290 (collab_normalise), geometry.rs:6-29, group.rs:9-66.    */
/* ------------------------------------------------------------------------------------------ */

typedef struct {
    int W, H, C;
    int step;           /* reference-grid stride (STEP = 4 in the reference) */
    int refs_x, refs_y;
    int S;              /* spatial_radius */
    int radius;         /* temporal radius */
    int grid_frames;    /* collab/mod.rs:45-51 */
    int refine;
    int k_max;
    int blk_step, blksize, blocks_x, blocks_y;
    float c_min;
    float lambda[2];    /* per-channel hard-threshold multiplier */
    float sigma[2];     /* per-channel sigma, normalised */
    float weight_scale;
    float accum_scale;
    float kaiser[8];
    float profile[8];   /* dct_noise_profile, all ones at rho = 0 (transforms.rs:196-221) */
    float pool_ratio;   /* > 0: pooled threshold at lambda * pool_ratio (nl4d/options.rs:109-117) */
} avdo_collab_params;

static inline uint32_t pack_pos_t(uint32_t x, uint32_t y, uint32_t t) { return (t << 26) | (y << 13) | x; }

/* Distance between the 8x8 patch at (ax,ay) of frame a and (bx,by) of frame b over C channels.
 * Codes: the exact integer SSD.  Floats: the reference's f32 form (distance.rs:160-209) -- each
 * column sums its squared differences row by row, the 8-lane butterfly of plane_ops.rs:17-23 adds
 * the columns, and the channel scale multiplies the total. */
static inline uint64_t dist_u16(const uint16_t *a, const uint16_t *b, int W, int C, int ax, int ay, int bx, int by)
{
    uint64_t s = 0;
    for (int r = 0; r < PATCH; ++r) {
        const uint16_t *pa = a + ((size_t)(ay + r) * W + ax) * C;
        const uint16_t *pb = b + ((size_t)(by + r) * W + bx) * C;
        for (int i = 0; i < PATCH * C; ++i) {
            int64_t d = (int64_t)pa[i] - (int64_t)pb[i];
            s += (uint64_t)(d * d);
        }
    }
    return s;
}

static inline float dist_f32(const float *a, const float *b, int W, int C, int ax, int ay, int bx, int by)
{
    float part[8];
    for (int q = 0; q < PATCH; ++q) {
        float acc = 0.f;
        for (int r = 0; r < PATCH; ++r)
            for (int c = 0; c < C; ++c) {
                float d = a[((size_t)(ay + r) * W + ax + q) * C + c] - b[((size_t)(by + r) * W + bx + q) * C + c];
                acc = fmaf(d, d, acc);
            }
        part[q] = acc;
    }
    float sum = ((part[0] + part[1]) + (part[2] + part[3])) + ((part[4] + part[5]) + (part[6] + part[7]));
    return sum * (C == 1 ? 3.0f : 1.5f);
}

/* covering_lo (trajectory.rs:18-22). */
static inline int covering_lo(int p, int blksize, int step)
{
    int over = (p + PATCH > blksize ? p + PATCH : blksize) - blksize;
    return (over + step - 1) / step;
}

/* search_group (fused/mod.rs:407-553): the group of one reference as 8 pack_pos_t words and the
 * group word.  T is the ring sample type, D the distance type, SELF_D pins the self-match below
 * every distance (centre.rs:274-290) and UNF is the unfilled distance (plane_ops.rs:79). */
#define SEARCH_TEMPLATE(T, D, SFX, DIST, SELF_D, UNF)                                                  \
    static void collab_search_one##SFX(const T *ring, size_t slot_len, const avdo_collab_params *pp,      \
                                       int centre_slot, const uint32_t *nslots, const int32_t *mv,          \
                                       const float *conf, int ix, int iy, uint32_t *out9)                  \
    {                                                                                                    \
        const int W = pp->W, H = pp->H, C = pp->C, S = pp->S;                                            \
        const int maxx = W - PATCH, maxy = H - PATCH;                                                    \
        const int rx = (ix * pp->step < maxx) ? ix * pp->step : maxx;                                    \
        const int ry = (iy * pp->step < maxy) ? iy * pp->step : maxy;                                    \
        const T *cf = ring + (size_t)centre_slot * slot_len;                                             \
        /* centre_search, clipped walk, shift_insert8: a tie never displaces an earlier visit. */        \
        const int K = pp->k_max;                                                                         \
        D bd[KMAX];                                                                                      \
        uint32_t bp[KMAX];                                                                               \
        for (int i = 0; i < K; ++i) {                                                                    \
            bd[i] = UNF;                                                                                 \
            bp[i] = 0;                                                                                   \
        }                                                                                                \
        int left = clampi(rx - S, 0, maxx), right = clampi(rx + S, 0, maxx);                             \
        int top = clampi(ry - S, 0, maxy), bottom = clampi(ry + S, 0, maxy);                             \
        for (int y = top; y <= bottom; ++y)                                                              \
            for (int x = left; x <= right; ++x) {                                                        \
                D d = (x == rx && y == ry) ? SELF_D : DIST(cf, cf, W, C, rx, ry, x, y);                  \
                if (!(d < bd[K - 1])) continue;                                                          \
                int k = K - 1;                                                                           \
                while (k > 0 && d < bd[k - 1]) {                                                         \
                    bd[k] = bd[k - 1];                                                                   \
                    bp[k] = bp[k - 1];                                                                   \
                    --k;                                                                                 \
                }                                                                                        \
                bd[k] = d;                                                                               \
                bp[k] = pack_pos_t((uint32_t)x, (uint32_t)y, 0);                                         \
            }                                                                                            \
        uint32_t n_live = (uint32_t)(right - left + 1) * (uint32_t)(bottom - top + 1);                   \
        uint32_t k_use = 1;                                                                              \
        while (k_use * 2 <= n_live && k_use * 2 <= (uint32_t)pp->k_max) k_use *= 2;                      \
        uint32_t members[KMAX];                                                                          \
        for (int m = 0; m < K; ++m) members[m] = bp[m];                                                  \
        int use_grid = 0;                                                                                \
        /* The volume grid of the reference's 8 members (grid_frames x 8 / grid_frames), and of 16  */  \
        /* in the extension (grid_frames x 16 / grid_frames volumes).                                */  \
        if (pp->grid_frames > 1 && (K == MAX_K || K == KMAX)) {                                          \
            const int gf = pp->grid_frames, volumes = K / gf, tail = gf - 1;                             \
            const int covers = (pp->blksize + pp->blk_step - 1) / pp->blk_step;                          \
            const int nnb = 2 * pp->radius;                                                              \
            const size_t nblk = (size_t)pp->blocks_x * pp->blocks_y;                                     \
            uint32_t gpos[KMAX];                                                                         \
            D gd[KMAX];                                                                                  \
            for (int v = 0; v < volumes; ++v) {                                                          \
                int first = v * gf;                                                                      \
                uint32_t anchor = members[v];                                                            \
                int ax = (int)(anchor & 0x1FFF), ay = (int)((anchor >> 13) & 0x1FFF);                    \
                gpos[first] = anchor;                                                                    \
                gd[first] = 0;                                                                           \
                for (int j = 0; j < tail; ++j) {                                                         \
                    gpos[first + 1 + j] = 0;                                                             \
                    gd[first + 1 + j] = UNF;                                                             \
                }                                                                                        \
                int held_n = first + 1;                                                                  \
                int bx_hi = ax / pp->blk_step, by_hi = ay / pp->blk_step;                                \
                if (bx_hi > pp->blocks_x - 1) bx_hi = pp->blocks_x - 1;                                  \
                if (by_hi > pp->blocks_y - 1) by_hi = pp->blocks_y - 1;                                  \
                int bx_lo = covering_lo(ax, pp->blksize, pp->blk_step);                                  \
                int by_lo = covering_lo(ay, pp->blksize, pp->blk_step);                                  \
                if (bx_lo > bx_hi) bx_lo = bx_hi;                                                        \
                if (by_lo > by_hi) by_lo = by_hi;                                                        \
                for (int t = 0; t < nnb; ++t) {                                                          \
                    const T *nf = ring + (size_t)nslots[t] * slot_len;                                   \
                    uint32_t packed_t = (uint32_t)t + 1;                                                 \
                    D fd = UNF;                                                                          \
                    uint32_t fpos = 0;                                                                   \
                    int sl[16], sr[16], st[16], sb[16];                                                  \
                    for (int s2 = 0; s2 < covers * covers; ++s2) {                                       \
                        sl[s2] = 1;                                                                      \
                        sr[s2] = 0;                                                                      \
                        st[s2] = 1;                                                                      \
                        sb[s2] = 0;                                                                      \
                    }                                                                                    \
                    for (int iyc = 0; iyc < covers; ++iyc)                                               \
                        for (int ixc = 0; ixc < covers; ++ixc) {                                         \
                            int wbx = bx_lo + ixc, wby = by_lo + iyc;                                    \
                            if (wbx > bx_hi || wby > by_hi) continue;                                    \
                            int block = wby * pp->blocks_x + wbx;                                        \
                            float bconf = conf[(size_t)t * nblk + block];                                \
                            if (!(bconf >= pp->c_min)) continue;                                         \
                            const int32_t *v2 = mv + ((size_t)t * nblk + block) * 2;                     \
                            int px = ax + v2[0], py = ay + v2[1];                                        \
                            int tl = clampi(px - pp->refine, 0, maxx), tr = clampi(px + pp->refine, 0, maxx); \
                            int tt = clampi(py - pp->refine, 0, maxy), tb = clampi(py + pp->refine, 0, maxy); \
                            for (int cy = tt; cy <= tb; ++cy)                                            \
                                for (int cx = tl; cx <= tr; ++cx) {                                      \
                                    int skipped = 0;                                                     \
                                    for (int s2 = 0; s2 < covers * covers; ++s2)                         \
                                        if (cx >= sl[s2] && cx <= sr[s2] && cy >= st[s2] && cy <= sb[s2]) \
                                            skipped = 1;                                                 \
                                    uint32_t packed = pack_pos_t((uint32_t)cx, (uint32_t)cy, packed_t);  \
                                    for (int k = 0; k < held_n; ++k)                                     \
                                        if (gpos[k] == packed) skipped = 1;                              \
                                    if (skipped) continue;                                               \
                                    D d = DIST(cf, nf, W, C, ax, ay, cx, cy);                            \
                                    if (d < fd) {                                                        \
                                        fd = d;                                                          \
                                        fpos = packed;                                                   \
                                    }                                                                    \
                                }                                                                        \
                            int ri = iyc * covers + ixc;                                                 \
                            sl[ri] = tl;                                                                 \
                            sr[ri] = tr;                                                                 \
                            st[ri] = tt;                                                                 \
                            sb[ri] = tb;                                                                 \
                        }                                                                                \
                    /* The frame's match joins the volume's ascending list, earlier frames win. */       \
                    D cd = fd;                                                                           \
                    uint32_t cp = fpos;                                                                  \
                    for (int j = 0; j < tail; ++j) {                                                     \
                        int si = first + 1 + j;                                                          \
                        if (cd < gd[si]) {                                                               \
                            D hd = gd[si];                                                               \
                            uint32_t hp = gpos[si];                                                      \
                            gd[si] = cd;                                                                 \
                            gpos[si] = cp;                                                               \
                            cd = hd;                                                                     \
                            cp = hp;                                                                     \
                        }                                                                                \
                    }                                                                                    \
                }                                                                                        \
            }                                                                                            \
            use_grid = (k_use == (uint32_t)K);                                                           \
            for (int v = 0; v < volumes; ++v) use_grid = use_grid && (gd[v * gf + tail] < UNF);          \
            if (use_grid) {                                                                              \
                for (int m = 0; m < K; ++m) members[m] = gpos[m];                                        \
                k_use = (uint32_t)K;                                                                     \
            }                                                                                            \
        }                                                                                                \
        for (int m = 0; m < K; ++m) out9[m] = members[m];                                                \
        out9[K] = k_use | (use_grid ? 0x100u : 0u);                                                      \
    }                                                                                                    \
    void avdo_collab_search##SFX(const T *ring, const avdo_collab_params *pp, int centre_slot,            \
                                 const uint32_t *nslots, const int32_t *mv, const float *conf,            \
                                 uint32_t *members)                                                      \
    {                                                                                                    \
        size_t slot_len = (size_t)pp->W * pp->H * pp->C;                                                 \
        _Pragma("omp parallel for schedule(dynamic, 2)") for (int iy = 0; iy < pp->refs_y; ++iy)         \
        {                                                                                                \
            for (int ix = 0; ix < pp->refs_x; ++ix)                                                      \
                collab_search_one##SFX(ring, slot_len, pp, centre_slot, nslots, mv, conf, ix, iy,          \
                                       members + ((size_t)iy * pp->refs_x + ix) * (pp->k_max + 1));      \
        }                                                                                                \
    }                                                                                                    \
    /* The same for the `n` references listed in `refs` (row-major ids), into members[n][K + 1]. */     \
    void avdo_collab_search_refs##SFX(const T *ring, const avdo_collab_params *pp, int centre_slot,       \
                                      const uint32_t *nslots, const int32_t *mv, const float *conf,       \
                                      const uint32_t *refs, int n, uint32_t *members)                     \
    {                                                                                                    \
        size_t slot_len = (size_t)pp->W * pp->H * pp->C;                                                 \
        _Pragma("omp parallel for schedule(dynamic, 8)") for (int i = 0; i < n; ++i)                      \
        {                                                                                                \
            const int ix = (int)(refs[i] % (uint32_t)pp->refs_x), iy = (int)(refs[i] / (uint32_t)pp->refs_x); \
            collab_search_one##SFX(ring, slot_len, pp, centre_slot, nslots, mv, conf, ix, iy,              \
                                   members + (size_t)i * (pp->k_max + 1));                               \
        }                                                                                                \
    }

/* Codes: the self-match is distance 0 and every other candidate its SSD plus one. */
static inline uint64_t dist_u16_shift(const uint16_t *a, const uint16_t *b, int W, int C, int ax, int ay, int bx,
                                      int by)
{
    return dist_u16(a, b, W, C, ax, ay, bx, by) + 1;
}

SEARCH_TEMPLATE(uint16_t, uint64_t, _u16, dist_u16_shift, 0ull, UNFILLED_D)
SEARCH_TEMPLATE(float, float, _f32, dist_f32, -1.0e38f, 3.0e38f)

/* dct8_basis (transforms.rs:39-43) evaluated in double and rounded to f32 as the kernel taps. */
static double dct8_basis(int row, int col)
{
    double rs = row == 0 ? 1.0 / sqrt(8.0) : 0.5;
    return rs * cos(M_PI * (2.0 * col + 1.0) * row / 16.0);
}

#define COLLAB_TEMPLATE(R, SFX, FMA, SQRT, ROUND)                                                   \
    static void dct8_fwd##SFX(R *l, const R *te, const R *to)                                       \
    {                                                                                               \
        R s[4], d[4];                                                                               \
        for (int k = 0; k < 4; ++k) {                                                               \
            s[k] = l[k] + l[7 - k];                                                                 \
            d[k] = l[k] - l[7 - k];                                                                 \
        }                                                                                           \
        for (int m = 0; m < 4; ++m) {                                                               \
            R e = 0, o = 0;                                                                         \
            for (int k = 0; k < 4; ++k) {                                                           \
                e = FMA(s[k], te[m * 4 + k], e);                                                    \
                o = FMA(d[k], to[m * 4 + k], o);                                                    \
            }                                                                                       \
            l[2 * m] = e;                                                                           \
            l[2 * m + 1] = o;                                                                       \
        }                                                                                           \
    }                                                                                               \
    static void dct8_inv##SFX(R *l, const R *te, const R *to)                                       \
    {                                                                                               \
        R sn[8];                                                                                    \
        for (int j = 0; j < 8; ++j) sn[j] = l[j];                                                   \
        for (int k = 0; k < 4; ++k) {                                                               \
            R e = 0, o = 0;                                                                         \
            for (int m = 0; m < 4; ++m) {                                                           \
                e = FMA(sn[2 * m], te[m * 4 + k], e);                                               \
                o = FMA(sn[2 * m + 1], to[m * 4 + k], o);                                           \
            }                                                                                       \
            l[k] = e + o;                                                                           \
            l[7 - k] = e - o;                                                                       \
        }                                                                                           \
    }                                                                                               \
    /* In-place Haar butterfly over members a (low) and b (high): (a+b)/sqrt2, (a-b)/sqrt2.      */ \
    static inline void haar_pair##SFX(R *stk, int a, int b, R f)                                    \
    {                                                                                               \
        for (int i = 0; i < 64; ++i) {                                                              \
            R x = stk[a * 64 + i], y = stk[b * 64 + i];                                             \
            stk[a * 64 + i] = (x + y) * f;                                                          \
            stk[b * 64 + i] = (x - y) * f;                                                          \
        }                                                                                           \
    }                                                                                               \
    static inline void ihaar_pair##SFX(R *stk, int a, int b, R f)                                   \
    {                                                                                               \
        for (int i = 0; i < 64; ++i) {                                                              \
            R lo = stk[a * 64 + i], hi = stk[b * 64 + i];                                           \
            stk[a * 64 + i] = (lo + hi) * f;                                                        \
            stk[b * 64 + i] = (lo - hi) * f;                                                        \
        }                                                                                           \
    }                                                                                               \
    /* The group's forward Haar, written as in-place butterflies: the reference's reordering     */ \
    /* (approximations first) only permutes member slots, and with one variance per member slot  */ \
    /* (every member carries sigma^2) the threshold does not depend on the slot.  Slot 0 holds    */ \
    /* the group approximation in both layouts.                                                  */ \
    static void group_fwd##SFX(R *stk, int k_use, int grid, int gf, R f)                            \
    {                                                                                               \
        if (grid) {                                                                                 \
            int vols = k_use / gf;                                                                  \
            for (int v = 0; v < vols; ++v)                                                          \
                for (int len = gf, st = 1; len >= 2; len /= 2, st *= 2)                             \
                    for (int p = 0; p < gf; p += 2 * st) haar_pair##SFX(stk, v * gf + p, v * gf + p + st, f); \
            for (int fr = 0; fr < gf; ++fr)                                                         \
                for (int len = vols, st = 1; len >= 2; len /= 2, st *= 2)                           \
                    for (int p = 0; p < vols; p += 2 * st) haar_pair##SFX(stk, fr + p * gf, fr + (p + st) * gf, f); \
        } else {                                                                                    \
            for (int len = k_use, st = 1; len >= 2; len /= 2, st *= 2)                              \
                for (int p = 0; p < k_use; p += 2 * st) haar_pair##SFX(stk, p, p + st, f);          \
        }                                                                                           \
    }                                                                                               \
    static void group_inv##SFX(R *stk, int k_use, int grid, int gf, R f)                            \
    {                                                                                               \
        if (grid) {                                                                                 \
            int vols = k_use / gf;                                                                  \
            int top = 1;                                                                            \
            while (top * 2 < vols) top *= 2;                                                        \
            for (int fr = 0; fr < gf; ++fr)                                                         \
                for (int st = top; st >= 1 && vols >= 2; st /= 2)                                   \
                    for (int p = 0; p < vols; p += 2 * st) ihaar_pair##SFX(stk, fr + p * gf, fr + (p + st) * gf, f); \
            int topf = 1;                                                                           \
            while (topf * 2 < gf) topf *= 2;                                                        \
            for (int v = 0; v < vols; ++v)                                                          \
                for (int st = topf; st >= 1 && gf >= 2; st /= 2)                                    \
                    for (int p = 0; p < gf; p += 2 * st) ihaar_pair##SFX(stk, v * gf + p, v * gf + p + st, f); \
        } else {                                                                                    \
            int top = 1;                                                                            \
            while (top * 2 < k_use) top *= 2;                                                       \
            for (int st = top; st >= 1 && k_use >= 2; st /= 2)                                      \
                for (int p = 0; p < k_use; p += 2 * st) ihaar_pair##SFX(stk, p, p + st, f);         \
        }                                                                                           \
    }                                                                                               \
    /* to_fixed / to_fixed_weight (aggregate.rs:147-159). */                                        \
    static inline int32_t to_fixed##SFX(R v, R scale)                                               \
    {                                                                                               \
        R c = v < (R)-5 ? (R)-5 : (v > (R)5 ? (R)5 : v);                                            \
        return (int32_t)ROUND(c * scale);                                                           \
    }                                                                                               \
    static inline int32_t to_fixed_weight##SFX(R w, R scale)                                        \
    {                                                                                               \
        R c = w < (R)0 ? (R)0 : (w > (R)1 ? (R)1 : w);                                              \
        return (int32_t)ROUND(c * scale * (R)5);                                                    \
    }                                                                                               \
    /* pooled_threshold (collab/kernels/fused/pooled.rs:24-147) over one group in the layout      */ \
    /* stk[j * 64 + v * 8 + u]: lane v walks each member row j left to right; a coefficient's      */ \
    /* energy is c^2 / max(var_j * prof[v], 1e-20) / max(prof[u], 1e-20), the spatial DC counts as  */ \
    /* zero while it stands in as a neighbour, the pooled sum (self + down + up, then left, then   */ \
    /* right) meets threshold^2 times the in-range neighbour count, the spatial DC takes its own   */ \
    /* test at dc_lambda and the group DC is always kept.  retained[v] gains, per member in order, */ \
    /* the member's kept profile sum times its lane variance.                                     */ \
    static void pooled##SFX(R *stk, const R *var, const float *prof, int k_use, R threshold, R dc_lambda,   \
                            R *retained)                                                            \
    {                                                                                               \
        const R bar = threshold * threshold;                                                        \
        R orig[KMAX * 64]; /* every lane reads the coefficients before any is dropped */            \
        const int rows = k_use > MAX_K ? k_use : MAX_K;                                             \
        for (int i = 0; i < rows * 64; ++i) orig[i] = stk[i];                                       \
        for (int v = 0; v < 8; ++v) {                                                               \
            const int dc_lane = v == 0, has_up = v > 0, has_down = v < 7;                           \
            const R vertical = (R)1 + (has_up ? (R)1 : (R)0) + (has_down ? (R)1 : (R)0);           \
            const R bars[4] = {bar * (vertical + (v == 1 ? (R)0 : (R)1)), bar * (vertical + (dc_lane ? (R)1 : (R)2)), \
                               bar * (vertical + (R)2), bar * (vertical + (R)1)};                  \
            R acc = 0;                                                                              \
            for (int j = 0; j < rows; ++j) {                                                        \
                R e[3][8]; /* lanes v-1, v, v+1 */                                                  \
                for (int d = -1; d <= 1; ++d) {                                                     \
                    const int vv = v + d;                                                           \
                    for (int u = 0; u < 8; ++u) {                                                   \
                        R en = 0;                                                                   \
                        if (vv >= 0 && vv < 8) {                                                    \
                            R lv = var[j] * (R)prof[vv];                                            \
                            R li = (R)1 / (lv > (R)1e-20f ? lv : (R)1e-20f);                        \
                            R pi = (R)1 / ((R)prof[u] > (R)1e-20f ? (R)prof[u] : (R)1e-20f);        \
                            R cf = orig[j * 64 + vv * 8 + u];                                       \
                            en = cf * cf * li * pi;                                                 \
                            if (vv == 0 && u == 0) en = 0;                                          \
                        }                                                                           \
                        e[d + 1][u] = en;                                                           \
                    }                                                                               \
                }                                                                                   \
                const R lane_variance = var[j] * (R)prof[v];                                        \
                R kept_profile = 0;                                                                 \
                for (int u = 0; u < 8; ++u) {                                                       \
                    R up = has_up ? e[0][u] : (R)0, down = has_down ? e[2][u] : (R)0;              \
                    R sum = e[1][u] + down + up;                                                    \
                    if (u > 0) sum += e[1][u - 1];                                                  \
                    if (u < 7) sum += e[1][u + 1];                                                  \
                    const R cb = u == 0 ? bars[0] : (u == 1 ? bars[1] : (u == 7 ? bars[3] : bars[2])); \
                    int keep = sum >= cb;                                                           \
                    R *cf = &stk[j * 64 + v * 8 + u];                                               \
                    if (u == 0) {                                                                   \
                        R dcv = lane_variance * (R)prof[0];                                         \
                        int own = (*cf < 0 ? -*cf : *cf) >= dc_lambda * SQRT(dcv);                  \
                        if (dc_lane) keep = own;                                                    \
                        if (j == 0) keep = keep || dc_lane;                                         \
                    }                                                                               \
                    kept_profile += keep ? (R)prof[u] : (R)0;                                       \
                    if (j < k_use && !keep) *cf = 0;                                                \
                }                                                                                   \
                if (j < k_use) acc += kept_profile * lane_variance;                                 \
            }                                                                                       \
            retained[v] = acc;                                                                      \
        }                                                                                           \
    }                                                                                               \
    void avdo_pooled_threshold##SFX(R *stk, const R *var, const float *prof, int k_use, R threshold,  \
                                    R dc_lambda, R *retained)                                       \
    {                                                                                               \
        pooled##SFX(stk, var, prof, k_use, threshold, dc_lambda, retained);                         \
    }                                                                                               \
    /* filter_group for every reference: DCT, group Haar, hard threshold, weight, inverses and    */ \
    /* the fixed-point scatter into each member's own frame region.                               */ \
    void avdo_collab_filter##SFX(const float *values, const avdo_collab_params *pp, int centre_slot,   \
                                 const uint32_t *nslots, const uint32_t *members, int32_t *accum,    \
                                 int32_t *wsum, float *group_weight)                                \
    {                                                                                               \
        const int W = pp->W, H = pp->H, C = pp->C;                                                  \
        const size_t px = (size_t)W * H;                                                            \
        R te[16], to[16];                                                                           \
        for (int m = 0; m < 4; ++m)                                                                 \
            for (int k = 0; k < 4; ++k) {                                                           \
                te[m * 4 + k] = (R)(float)dct8_basis(2 * m, k);                                     \
                to[m * 4 + k] = (R)(float)dct8_basis(2 * m + 1, k);                                 \
            }                                                                                       \
        const R f = (R)0.70710678118654752440f;                                                     \
        const R scale = (R)pp->accum_scale;                                                         \
        _Pragma("omp parallel for schedule(dynamic, 2)") for (int iy = 0; iy < pp->refs_y; ++iy)     \
        {                                                                                           \
            R stk[KMAX * 64];                                                                       \
            for (int ix = 0; ix < pp->refs_x; ++ix) {                                               \
                size_t ref = (size_t)iy * pp->refs_x + ix;                                          \
                const int K = pp->k_max;                                                            \
                const uint32_t *mem = members + ref * (size_t)(K + 1);                              \
                uint32_t gw = mem[K];                                                               \
                int k_use = (int)(gw & 0xFF), grid = (gw & 0x100) != 0;                             \
                int slot_of[KMAX];                                                                  \
                for (int m = 0; m < K; ++m) {                                                       \
                    uint32_t t = mem[m] >> 26;                                                      \
                    slot_of[m] = t > 0 ? (int)nslots[t - 1] : centre_slot;                          \
                }                                                                                   \
                R scaled_weight = 0;                                                                \
                for (int c = 0; c < C; ++c) {                                                       \
                    R sig2 = (R)pp->sigma[c] * (R)pp->sigma[c];                                     \
                    for (int m = 0; m < k_use; ++m) {                                               \
                        int mx = (int)(mem[m] & 0x1FFF), my = (int)((mem[m] >> 13) & 0x1FFF);       \
                        const float *src = values + (size_t)slot_of[m] * px * C;                    \
                        for (int r = 0; r < 8; ++r)                                                 \
                            for (int q = 0; q < 8; ++q)                                             \
                                stk[m * 64 + r * 8 + q] = (R)src[((size_t)(my + r) * W + mx + q) * C + c]; \
                        R line[8];                                                                  \
                        for (int q = 0; q < 8; ++q) { /* vertical pass, column q */                 \
                            for (int r = 0; r < 8; ++r) line[r] = stk[m * 64 + r * 8 + q];          \
                            dct8_fwd##SFX(line, te, to);                                            \
                            for (int r = 0; r < 8; ++r) stk[m * 64 + r * 8 + q] = line[r];          \
                        }                                                                           \
                        for (int r = 0; r < 8; ++r) { /* horizontal pass, vertical frequency r */   \
                            dct8_fwd##SFX(stk + m * 64 + r * 8, te, to);                            \
                        }                                                                           \
                    }                                                                               \
                    group_fwd##SFX(stk, k_use, grid, pp->grid_frames, f);                           \
                    R part[8];                                                                      \
                    if (pp->pool_ratio > 0.0f) {                                                    \
                        R var[KMAX];                                                                \
                        for (int j = 0; j < KMAX; ++j) var[j] = sig2;                               \
                        const R lam_c = (R)pp->lambda[c];                                           \
                        pooled##SFX(stk, var, pp->profile, k_use, lam_c * (R)pp->pool_ratio, lam_c, part); \
                    } else {                                                                        \
                    R cv[64], thr[64];                                                              \
                    for (int v = 0; v < 8; ++v)                                                     \
                        for (int u = 0; u < 8; ++u) {                                               \
                            cv[v * 8 + u] = sig2 * ((R)pp->profile[u] * (R)pp->profile[v]);         \
                            thr[v * 8 + u] = (R)pp->lambda[c] * SQRT(cv[v * 8 + u]);                \
                        }                                                                           \
                    int kept[64];                                                                   \
                    for (int i = 0; i < 64; ++i) kept[i] = 0;                                       \
                    for (int m = 0; m < k_use; ++m)                                                 \
                        for (int v = 0; v < 8; ++v)                                                 \
                            for (int u = 0; u < 8; ++u) {                                           \
                                R *cf = &stk[m * 64 + v * 8 + u];                                   \
                                int keep = (*cf < 0 ? -*cf : *cf) >= thr[v * 8 + u];               \
                                if (m == 0 && v == 0 && u == 0) keep = 1;                           \
                                if (keep) kept[v * 8 + u]++;                                        \
                                else *cf = 0;                                                       \
                            }                                                                       \
                    /* retained_v: lane v (vertical frequency) adds each kept coefficient's        */ \
                    /* variance, horizontal frequency outermost, then the 8-lane butterfly         */ \
                    /* (fused/mod.rs:747-781, plane_ops.rs:17-23).                                 */ \
                    for (int v = 0; v < 8; ++v) {                                                   \
                        R acc = 0;                                                                  \
                        for (int u = 0; u < 8; ++u)                                                 \
                            for (int n = 0; n < kept[v * 8 + u]; ++n) acc += cv[v * 8 + u];         \
                        part[v] = acc;                                                              \
                    }                                                                               \
                    }                                                                               \
                    R sum = ((part[0] + part[1]) + (part[2] + part[3])) + ((part[4] + part[5]) + (part[6] + part[7])); \
                    R weight = (isnan(sum) || isinf(sum)) ? (R)0 : (R)1 / (sum > (R)1e-12f ? sum : (R)1e-12f); \
                    if (c == 0) {                                                                   \
                        group_weight[ref] = (float)weight;                                          \
                        scaled_weight = weight * (R)pp->weight_scale;                               \
                    }                                                                               \
                    group_inv##SFX(stk, k_use, grid, pp->grid_frames, f);                           \
                    for (int m = 0; m < k_use; ++m) {                                               \
                        for (int r = 0; r < 8; ++r) dct8_inv##SFX(stk + m * 64 + r * 8, te, to);    \
                        R line[8];                                                                  \
                        for (int q = 0; q < 8; ++q) {                                               \
                            for (int r = 0; r < 8; ++r) line[r] = stk[m * 64 + r * 8 + q];          \
                            dct8_inv##SFX(line, te, to);                                            \
                            for (int r = 0; r < 8; ++r) stk[m * 64 + r * 8 + q] = line[r];          \
                        }                                                                           \
                        int mx = (int)(mem[m] & 0x1FFF), my = (int)((mem[m] >> 13) & 0x1FFF);       \
                        size_t base = (size_t)slot_of[m] * px;                                      \
                        for (int r = 0; r < 8; ++r)                                                 \
                            for (int q = 0; q < 8; ++q) {                                           \
                                R win = (R)pp->kaiser[r] * (R)pp->kaiser[q];                        \
                                R w = scaled_weight * win;                                          \
                                int32_t a = to_fixed##SFX(stk[m * 64 + r * 8 + q] * w, scale);      \
                                size_t pix = base + (size_t)(my + r) * W + mx + q;                  \
                                _Pragma("omp atomic") accum[pix * C + c] += a;                      \
                                if (c == 0) {                                                       \
                                    int32_t b = to_fixed_weight##SFX(w, scale);                     \
                                    _Pragma("omp atomic") wsum[pix] += b;                           \
                                }                                                                   \
                            }                                                                       \
                    }                                                                               \
                }                                                                                   \
            }                                                                                       \
        }                                                                                           \
    }

static inline float fma_f32(float a, float b, float c) { return fmaf(a, b, c); }
static inline double fma_f64(double a, double b, double c) { return a * b + c; }

COLLAB_TEMPLATE(float, _f32, fma_f32, sqrtf, roundf)
COLLAB_TEMPLATE(double, _f64, fma_f64, sqrt, round)

/* collab_normalise (aggregate.rs:252-290): out = accum*5/wsum, or accum when wsum is 0; the
 * region is zeroed afterwards. */
void avdo_collab_normalise(int32_t *accum, int32_t *wsum, float *out, size_t pixels, int C)
{
    for (size_t p = 0; p < pixels; ++p) {
        int32_t w = wsum[p];
        for (int c = 0; c < C; ++c) {
            float a = (float)accum[p * C + c];
            out[p * C + c] = (w != 0) ? (a * 5.0f) / (float)w : a;
            accum[p * C + c] = 0;
        }
        wsum[p] = 0;
    }
}

int avdo_max_threads(void)
{
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

void avdo_set_threads(int n)
{
#ifdef _OPENMP
    if (n > 0) omp_set_num_threads(n);
#else
    (void)n;
#endif
}

/* Exported single-step helpers, so the oracle's own tests can pin each piece against the
 * reference's known answers (aggregate.rs:305-330, transforms.rs tests). */
int32_t avdo_to_fixed_f32(float v, float scale) { return to_fixed_f32(v, scale); }
int32_t avdo_to_fixed_weight_f32(float w, float scale) { return to_fixed_weight_f32(w, scale); }

static void dct_taps_f32(float *te, float *to)
{
    for (int m = 0; m < 4; ++m)
        for (int k = 0; k < 4; ++k) {
            te[m * 4 + k] = (float)dct8_basis(2 * m, k);
            to[m * 4 + k] = (float)dct8_basis(2 * m + 1, k);
        }
}

void avdo_dct8_fwd_f32(float *line)
{
    float te[16], to[16];
    dct_taps_f32(te, to);
    dct8_fwd_f32(line, te, to);
}

void avdo_dct8_inv_f32(float *line)
{
    float te[16], to[16];
    dct_taps_f32(te, to);
    dct8_inv_f32(line, te, to);
}

void avdo_group_fwd_f32(float *stk, int k_use, int grid, int gf) { group_fwd_f32(stk, k_use, grid, gf, 0.70710678118654752440f); }
void avdo_group_inv_f32(float *stk, int k_use, int grid, int gf) { group_inv_f32(stk, k_use, grid, gf, 0.70710678118654752440f); }

/* ------------------------------------------------------------------------------------------ */
/* Spatial noise measurement (the Immerkaer estimate), noise/kernels/stats.rs:24-140.         */
/* ------------------------------------------------------------------------------------------ */

/* nlm_noise_partial (stats.rs:31-99): per 32x8 tile, the sum over the tile's threads in thread
 * order (row * 32 + column) of |4 c - 2 (l + r + u + d) + (ul + ur + dl + dr)| per channel;
 * border pixels, pixels past the frame and lanes past the channel count add zero. */
void avdo_noise_partial(const float *ring, uint32_t W, uint32_t H, uint32_t C, uint32_t frame, float *partials)
{
    const uint32_t bx = 32, by = 8;
    const uint32_t cx = (W + bx - 1) / bx, cy = (H + by - 1) / by;
    const float *f = ring + (size_t)frame * W * H * C;
    for (uint32_t ty = 0; ty < cy; ++ty)
        for (uint32_t tx = 0; tx < cx; ++tx) {
            float sum[2] = {0.0f, 0.0f};
            for (uint32_t t = 0; t < bx * by; ++t) {
                const uint32_t x = tx * bx + t % bx, y = ty * by + t / bx;
                const int interior = x >= 1 && x + 1 < W && y >= 1 && y + 1 < H;
                for (uint32_t lane = 0; lane < 2; ++lane) {
                    float v = 0.0f;
                    if (interior && lane < C) {
#define AT(xx, yy) f[((size_t)(yy) * W + (xx)) * C + lane]
                        const float centre = AT(x, y), left = AT(x - 1, y), right = AT(x + 1, y);
                        const float up = AT(x, y - 1), down = AT(x, y + 1);
                        const float ul = AT(x - 1, y - 1), ur = AT(x + 1, y - 1), dl = AT(x - 1, y + 1),
                                    dr = AT(x + 1, y + 1);
#undef AT
                        const float a = centre * 4.0f;
                        const float b = (((left + right) + up) + down) * 2.0f;
                        const float d = ((ul + ur) + dl) + dr;
                        v = fabsf((a - b) + d);
                    }
                    sum[lane] += v;
                }
            }
            partials[((size_t)ty * cx + tx) * 2] = sum[0];
            partials[((size_t)ty * cx + tx) * 2 + 1] = sum[1];
        }
}

/* nlm_noise_reduce (stats.rs:101-140): thread t (of 256) adds partials t, t + 256, ... of each
 * lane in turn; each lane's total then adds the 256 thread sums in thread order. */
void avdo_noise_reduce(const float *partials, uint32_t count, float *results, uint32_t slot)
{
    float s0[256], s1[256];
    for (uint32_t t = 0; t < 256; ++t) {
        float a = 0.0f, b = 0.0f;
        for (uint32_t i = t; i < count; i += 256) {
            a += partials[(size_t)i * 2];
            b += partials[(size_t)i * 2 + 1];
        }
        s0[t] = a;
        s1[t] = b;
    }
    float r0 = 0.0f, r1 = 0.0f;
    for (uint32_t t = 0; t < 256; ++t) {
        r0 += s0[t];
        r1 += s1[t];
    }
    results[(size_t)slot * 2] = r0;
    results[(size_t)slot * 2 + 1] = r1;
}

/* ------------------------------------------------------------------------------------------ */
/* Temporal block statistics, nlm_temporal_noise_stats (noise/kernels/stats.rs:142-660) with  */
/* the record layout of noise/temporal.rs:18-131.                                            */
/* ------------------------------------------------------------------------------------------ */

/* One 16x16 block's pixels are numbered t = local_y * 16 + local_x.  Every f32 sum below runs in the
 * reference's order: the lane totals add t = 0..255 in turn, the halving trees add slot s + stride
 * into slot s for stride = n/2 .. 1, the cell means add their pixels row by row and divide. */
static void tree_sum(float *v, uint32_t n)
{
    for (uint32_t stride = n / 2; stride >= 1; stride /= 2)
        for (uint32_t i = 0; i < stride; ++i) v[i] += v[i + stride];
}

/* The 16 means of the 4x4 cells of a 16x16 tile (write_cell_means, stats.rs:606-624). */
static void cell_means16(const float *tile, float *cells)
{
    for (uint32_t c = 0; c < 16; ++c) {
        const uint32_t cx = c % 4, cy = c / 4;
        float cell = 0.0f;
        for (uint32_t dy = 0; dy < 4; ++dy)
            for (uint32_t dx = 0; dx < 4; ++dx) cell += tile[(4 * cy + dy) * 16 + 4 * cx + dx];
        cells[c] = cell / 16.0f;
    }
}

/* The mean squared step between adjacent cells that lie fully inside the block, or 0 with no
 * such pair (cell_texture, stats.rs:626-656). */
static float cell_texture16(const float *cells, uint32_t bw, uint32_t bh)
{
    const uint32_t cols = bw / 4, rows = bh / 4;
    float energy = 0.0f;
    uint32_t pairs = 0;
    for (uint32_t cy = 0; cy < rows; ++cy)
        for (uint32_t cx = 0; cx < cols; ++cx) {
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

uint32_t avdo_temporal_record_len(uint32_t C) { return 2 * C + 3 + 4 * 6 + C - 1; }

/* Every block's record of frame slot_new against slot_prev of an f32 ring [slot][H][W][C].
 * mv (may be NULL: uncompensated) holds whole-pixel vectors (x, y) of an mv_step grid of
 * mv_bx x mv_by blocks from mv[mv_offset]; a block reads the one covering its origin. */
void avdo_temporal_stats(const float *ring, uint32_t W, uint32_t H, uint32_t C, uint32_t slot_new,
                         uint32_t slot_prev, const int32_t *mv, uint32_t mv_offset, uint32_t mv_step,
                         uint32_t mv_bx, uint32_t mv_by, int luma_fields, float *stats)
{
    const uint32_t B = 16, T = 256, bx_n = (W + B - 1) / B, by_n = (H + B - 1) / B;
    const uint32_t rec = avdo_temporal_record_len(C);
    const float *fn = ring + (size_t)slot_new * W * H * C, *fp = ring + (size_t)slot_prev * W * H * C;
#pragma omp parallel for schedule(dynamic, 4)
    for (uint32_t b = 0; b < bx_n * by_n; ++b) {
        const uint32_t ox = (b % bx_n) * B, oy = (b / bx_n) * B;
        const uint32_t bw = W - ox < B ? W - ox : B, bh = H - oy < B ? H - oy : B;
        float *out = stats + (size_t)b * rec;
        int sx = 0, sy = 0, matched = 1;
        float moved = 0.0f;
        if (mv) {
            /* The gain check (stats.rs:257-341): zero-vector and vector residual variances of
             * channel 0 by halving trees over the 256 pixels. */
            uint32_t col = ox / mv_step, row = oy / mv_step;
            if (col > mv_bx - 1) col = mv_bx - 1;
            if (row > mv_by - 1) row = mv_by - 1;
            const int32_t vx = mv[mv_offset + (row * mv_bx + col) * 2], vy = mv[mv_offset + (row * mv_bx + col) * 2 + 1];
            float r[4][256];
            for (uint32_t t = 0; t < T; ++t) {
                const uint32_t x = ox + t % B, y = oy + t / B;
                float zr = 0.0f, vr = 0.0f;
                if (x < W && y < H) {
                    const float nl = fn[((size_t)y * W + x) * C];
                    zr = nl - fp[((size_t)y * W + x) * C];
                    const int64_t mx = (int64_t)x + vx, my = (int64_t)y + vy;
                    if (mx >= 0 && mx < (int64_t)W && my >= 0 && my < (int64_t)H)
                        vr = nl - fp[((size_t)my * W + (size_t)mx) * C];
                }
                r[0][t] = zr;
                r[1][t] = zr * zr;
                r[2][t] = vr;
                r[3][t] = vr * vr;
            }
            for (int k = 0; k < 4; ++k) tree_sum(r[k], T);
            const int64_t left = (int64_t)ox + vx, top = (int64_t)oy + vy, right = left + bw, bottom = top + bh;
            int64_t span_w = (right < (int64_t)W ? right : (int64_t)W) - (left > 0 ? left : 0);
            int64_t span_h = (bottom < (int64_t)H ? bottom : (int64_t)H) - (top > 0 ? top : 0);
            if (span_w < 0) span_w = 0;
            if (span_h < 0) span_h = 0;
            const float matched_px = (float)(span_w * span_h), block_px = (float)(bw * bh);
            const float zmean = r[0][0] / block_px;
            const float zvar = r[1][0] / block_px - zmean * zmean;
            const float safe = matched_px > 1.0f ? matched_px : 1.0f;
            const float vmean = r[2][0] / safe;
            const float vvar = r[3][0] / safe - vmean * vmean;
            const int has_vector = vx != 0 || vy != 0;
            const int follows = has_vector && (matched_px == 0.0f || vvar <= 0.5f * zvar);
            if (follows) {
                sx = vx;
                sy = vy;
                matched = left >= 0 && right <= (int64_t)W && top >= 0 && bottom <= (int64_t)H;
                if (matched) moved = 1.0f;
            }
        }
        /* Per pixel: the residual of each channel, its square, the lag-1 product of channel 0 with
         * the next pixel of the row, the new frame's values and the luma fields. */
        float res[2][256], nv[2][256], lane[5][256], mean_tile[256];
        float qsum[256], qmin[256], qmax[256], qd[256], qd2[256];
        for (uint32_t t = 0; t < T; ++t) {
            const uint32_t lx = t % B, ly = t / B, x = ox + lx, y = oy + ly;
            const int valid = x < W && y < H && matched;
            float nl_raw = 0.0f, mean_raw = 0.0f, mn = 1.0e30f, mx = -1.0e30f;
            for (uint32_t c = 0; c < 2; ++c) res[c][t] = nv[c][t] = 0.0f;
            if (valid) {
                const size_t pn = ((size_t)y * W + x) * C;
                const size_t pp = ((size_t)(y + sy) * W + (size_t)(x + sx)) * C;
                for (uint32_t c = 0; c < C; ++c) {
                    res[c][t] = fn[pn + c] - fp[pp + c];
                    nv[c][t] = fn[pn + c];
                }
                nl_raw = fn[pn];
                mean_raw = 0.5f * (fn[pn] + fp[pp]);
                mn = mx = fn[pn];
            }
            for (uint32_t c = 0; c < C; ++c) {
                lane[c][t] = res[c][t];
                lane[C + c][t] = res[c][t] * res[c][t];
            }
            const uint32_t slot = ((ly / 8) * 2 + lx / 8) * 64 + (ly % 8) * 8 + lx % 8;
            qsum[slot] = nl_raw;
            qmin[slot] = mn;
            qmax[slot] = mx;
            qd[slot] = res[0][t];
            qd2[slot] = res[0][t] * res[0][t];
            mean_tile[t] = mean_raw;
        }
        for (uint32_t t = 0; t < T; ++t) {
            const uint32_t lx = t % B, y = oy + t / B, x = ox + lx;
            const int valid = x < W && y < H && matched;
            lane[2 * C][t] = valid && lx + 1 < bw ? res[0][t] * res[0][t + 1] : 0.0f;
        }
        for (uint32_t l = 0; l < 2 * C + 1; ++l) {
            float total = 0.0f;
            for (uint32_t t = 0; t < T; ++t) total += lane[l][t];
            out[l] = total;
        }
        float cells[16];
        for (uint32_t c = 0; c < C; ++c) {
            cell_means16(nv[c], cells);
            out[c == 0 ? 2 * C + 1 : 2 * C + 3 + 24 + c - 1] = cell_texture16(cells, bw, bh);
        }
        out[2 * C + 2] = moved;
        float *quarters = out + 2 * C + 3;
        if (!luma_fields) {
            for (int k = 0; k < 24; ++k) quarters[k] = 0.0f;
            continue;
        }
        for (uint32_t q = 0; q < 4; ++q) {
            tree_sum(qsum + 64 * q, 64);
            tree_sum(qd + 64 * q, 64);
            tree_sum(qd2 + 64 * q, 64);
            for (uint32_t stride = 32; stride >= 1; stride /= 2)
                for (uint32_t i = 0; i < stride; ++i) {
                    qmin[64 * q + i] = fminf(qmin[64 * q + i], qmin[64 * q + i + stride]);
                    qmax[64 * q + i] = fmaxf(qmax[64 * q + i], qmax[64 * q + i + stride]);
                }
        }
        /* Flatness (stats.rs:520-585): an 8x8 grid of 2x2 means of the temporal mean, each cell's
         * squared steps to its right and lower neighbour inside its quarter, 16 scores a quarter
         * summed by a halving tree, over the quarter's 24 pairs. */
        float grid[64], pair[64];
        for (uint32_t c = 0; c < 64; ++c) {
            const uint32_t cy = c / 8, cx = c % 8;
            float cell = 0.0f;
            for (uint32_t dy = 0; dy < 2; ++dy)
                for (uint32_t dx = 0; dx < 2; ++dx) cell += mean_tile[(2 * cy + dy) * 16 + 2 * cx + dx];
            grid[c] = cell / 4.0f;
        }
        for (uint32_t c = 0; c < 64; ++c) {
            const uint32_t cy = c / 8, cx = c % 8;
            float e = 0.0f;
            if (cx % 4 != 3) e += (grid[c] - grid[c + 1]) * (grid[c] - grid[c + 1]);
            if (cy % 4 != 3) e += (grid[c] - grid[c + 8]) * (grid[c] - grid[c + 8]);
            pair[((cy / 4) * 2 + cx / 4) * 16 + (cy % 4) * 4 + cx % 4] = e;
        }
        for (uint32_t q = 0; q < 4; ++q) tree_sum(pair + 16 * q, 16);
        for (uint32_t q = 0; q < 4; ++q) {
            float *f = quarters + 6 * q;
            if (!matched) {
                for (int k = 0; k < 6; ++k) f[k] = 0.0f;
                continue;
            }
            const int full = bw >= (q % 2 + 1) * 8 && bh >= (q / 2 + 1) * 8;
            f[0] = qd[64 * q];
            f[1] = qd2[64 * q];
            f[2] = qsum[64 * q];
            f[3] = full ? pair[16 * q] / 24.0f : 3.0e38f;
            f[4] = qmin[64 * q];
            f[5] = qmax[64 * q];
        }
    }
}

/* ------------------------------------------------------------------------------------------ */
/* Film-grain measurement, grain_measure (nl4d/kernels/grain.rs:243-513, constants             */
/* nl4d/grain/consts.rs:1-45).                                                                */
/* ------------------------------------------------------------------------------------------ */

#define GR_CELL 8
#define GR_TILE_W 20
#define GR_TILE_H 11
#define GR_HALO 6
#define GR_LAGS 46
#define GR_PARTIAL 48
#define GR_BUCKETS 64
#define GR_BINS 16
#define GR_HIST (GR_BINS * GR_BUCKETS)

/* Rust's saturating float-to-u32 cast. */
static uint32_t gr_u32(float v)
{
    if (!(v >= 0.0f)) return 0;
    if (v >= 4294967296.0f) return 0xFFFFFFFFu;
    return (uint32_t)v;
}

static uint32_t gr_shift(uint32_t v, int32_t s, uint32_t limit)
{
    int64_t m = (int64_t)v + s;
    if (m < 0) m = 0;
    if (m > (int64_t)limit - 1) m = (int64_t)limit - 1;
    return (uint32_t)m;
}

static uint32_t gr_block(uint32_t cx, uint32_t cy, uint32_t bx, uint32_t by, uint32_t step)
{
    uint32_t x = cx * GR_CELL / step, y = cy * GR_CELL / step;
    if (x > bx - 1) x = bx - 1;
    if (y > by - 1) y = by - 1;
    return y * bx + x;
}

static uint32_t gr_bucket(float s, const float *edges)
{
    uint32_t b = 0;
    for (uint32_t i = 1; i < GR_BUCKETS; ++i)
        if (s >= edges[i]) b = i;
    return b;
}

static uint32_t gr_luma_bin(float mean)
{
    const uint32_t b = gr_u32(mean * (float)GR_BINS);
    return b < GR_BINS - 1 ? b : GR_BINS - 1;
}

/* The 64-slot halving tree of grain.rs:362-385 (slot s + stride into slot s, stride 32 .. 1). */
static float gr_tree_sum(float *v)
{
    for (uint32_t stride = 32; stride >= 1; stride /= 2)
        for (uint32_t i = 0; i < stride; ++i) v[i] += v[i + stride];
    return v[0];
}

/* Every 8x8 cell of channel `ch` of ring slots t / next ([slot][H][W][C] f32) against the denoised
 * frame `out_t` ([H][W][C] f32): source-grain and one-frame histograms added into hist[0..2048)
 * (i32 counts) and the cell's 48-float partial (46 lag sums, pixel count, strength group). */
void avdo_grain_measure(const float *ring, const float *out_t, uint32_t W, uint32_t H, uint32_t C, uint32_t ch,
                        uint32_t slot_t, uint32_t slot_next, const int32_t *saved_mv, const float *saved_conf,
                        uint32_t bx, uint32_t by, uint32_t step, uint32_t entry, int has_source,
                        const float *luma_means, int has_luma_means, uint32_t means_bx, int luma_band,
                        const float *edges, int32_t *hist, float *partials)
{
    const uint32_t cells_x = W / GR_CELL, cells_y = H / GR_CELL, blocks = bx * by;
    const float flat_range = 5.0f / 255.0f, luma_low = 20.0f / 255.0f, luma_high = 235.0f / 255.0f;
    const float clip_low = 4.0f / 255.0f, clip_high = 251.0f / 255.0f, std_min = 0.05f / 255.0f;
    const float r2 = 0.70710677f;
#define GR_AT(xx, yy, slot) ring[(((size_t)(slot) * H + (yy)) * W + (xx)) * C + ch]
    for (uint32_t cy = 0; cy < cells_y; ++cy)
        for (uint32_t cx = 0; cx < cells_x; ++cx) {
            const uint32_t ox = cx * GR_CELL, oy = cy * GR_CELL;
            const uint32_t blk = gr_block(cx, cy, bx, by, step), base = entry * blocks * 2;
            const int32_t dx = saved_mv[base + blk * 2], dy = saved_mv[base + blk * 2 + 1];
            float tile[GR_TILE_W * GR_TILE_H], fv[64], s0[64], s1[64], s2[64], s3[64], lo0[64], lo1[64], hi0[64],
                hi1[64];
            const float origin = GR_AT(ox, oy, slot_t);
            for (uint32_t t = 0; t < 64; ++t) {
                const uint32_t lx = t % 8, ly = t / 8, x = ox + lx, y = oy + ly;
                const float cur = GR_AT(x, y, slot_t);
                const float nxt = GR_AT(gr_shift(x, dx, W), gr_shift(y, dy, H), slot_next);
                const float g = (nxt - cur) * r2;
                const float clean = out_t[((size_t)y * W + x) * C + ch];
                tile[ly * GR_TILE_W + lx + GR_HALO] = g;
                fv[t] = cur - origin;
                s0[t] = g;
                s1[t] = g * g;
                s2[t] = clean;
                lo0[t] = hi0[t] = clean;
                lo1[t] = fminf(cur, nxt);
                hi1[t] = fmaxf(cur, nxt);
            }
            for (uint32_t t = 0; t < 64; ++t) {
                const uint32_t lx = t % 8, ly = t / 8;
                float hp = 0.0f;
                if (lx + 1 < 8) {
                    const float r = fv[t + 1] - fv[t];
                    hp += r * r;
                }
                if (ly + 1 < 8) {
                    const float d = fv[t + 8] - fv[t];
                    hp += d * d;
                }
                s3[t] = hp;
            }
            const float sum = gr_tree_sum(s0), sum_sq = gr_tree_sum(s1), clean_sum = gr_tree_sum(s2),
                        hp_sum = gr_tree_sum(s3);
            for (uint32_t stride = 32; stride >= 1; stride /= 2)
                for (uint32_t i = 0; i < stride; ++i) {
                    lo0[i] = fminf(lo0[i], lo0[i + stride]);
                    lo1[i] = fminf(lo1[i], lo1[i + stride]);
                    hi0[i] = fmaxf(hi0[i], hi0[i + stride]);
                    hi1[i] = fmaxf(hi1[i], hi1[i + stride]);
                }
            const float count = 64.0f;
            float var = (sum_sq - sum * sum / count) / (count - 1.0f);
            const float grain_std = sqrtf(var > 0.0f ? var : 0.0f);
            const float mean = clean_sum / count;
            const float frame_std = sqrtf(hp_sum / 224.0f);
            const float range = hi0[0] - lo0[0];
            const int interior = cx >= 1 && cx + 2 <= cells_x && cy + 2 <= cells_y;
            const float conf = saved_conf[entry * blocks + blk];
            const int band = !luma_band || (mean > luma_low && mean < luma_high);
            const int usable = has_source && range < flat_range && band && lo1[0] >= clip_low && hi1[0] <= clip_high;
            const int passes = usable && interior && conf >= 0.8f && grain_std > std_min;
            float group = 0.0f;
            if (passes) {
                const uint32_t b = gr_bucket(grain_std, edges);
                hist[gr_luma_bin(mean) * GR_BUCKETS + b] += 1;
                group = (float)(b / 4);
            }
            if (usable && (means_bx == 0 || has_luma_means)) {
                const uint32_t b = gr_bucket(frame_std, edges);
                uint32_t bin = gr_luma_bin(mean);
                if (means_bx > 0) {
                    const uint32_t mb = (cy * GR_CELL / 16) * means_bx + cx * GR_CELL / 16;
                    const uint32_t v = gr_u32(luma_means[mb * 3] / 16.0f);
                    bin = v < GR_BINS - 1 ? v : GR_BINS - 1;
                }
                hist[GR_HIST + bin * GR_BUCKETS + b] += 1;
            }
            float *part = partials + ((size_t)cy * cells_x + cx) * GR_PARTIAL;
            if (passes) {
                for (uint32_t ty = 0; ty < GR_TILE_H; ++ty)
                    for (uint32_t tx = 0; tx < GR_TILE_W; ++tx) {
                        if (ty < GR_CELL && tx >= GR_HALO && tx < GR_HALO + GR_CELL) continue;
                        const uint32_t x = gr_shift(0, (int32_t)ox + (int32_t)tx - GR_HALO, W);
                        const uint32_t y = gr_shift(oy + ty, 0, H);
                        const uint32_t b2 = gr_block(x / GR_CELL, y / GR_CELL, bx, by, step);
                        const int32_t ex = saved_mv[base + b2 * 2], ey = saved_mv[base + b2 * 2 + 1];
                        const float nxt = GR_AT(gr_shift(x, ex, W), gr_shift(y, ey, H), slot_next);
                        tile[ty * GR_TILE_W + tx] = (nxt - GR_AT(x, y, slot_t)) * r2;
                    }
            }
            const float gmean = sum / 64.0f;
            for (uint32_t l = 0; l < GR_LAGS; ++l) {
                const uint32_t row = l > GR_HALO ? 1 + (l - GR_HALO - 1) / 13 : 0;
                const uint32_t col = l > GR_HALO ? (l - GR_HALO - 1) % 13 : l + GR_HALO;
                float acc = 0.0f;
                if (passes)
                    for (uint32_t py = 0; py < 8; ++py)
                        for (uint32_t px = 0; px < 8; ++px) {
                            const float c = tile[py * GR_TILE_W + px + GR_HALO] - gmean;
                            const float n = tile[(py + row) * GR_TILE_W + px + col] - gmean;
                            acc += c * n;
                        }
                part[l] = acc;
            }
            part[GR_LAGS] = passes ? 64.0f : 0.0f;
            part[GR_LAGS + 1] = group;
        }
#undef GR_AT
}

/* grain_reduce_partials (nl4d/kernels/grain.rs:515-560): for each of the 47 record lanes, thread t
 * of 128 adds lane `lane` of cells t, t + 128, ... in turn into its strength group's column (the
 * group clamped to 15), each group's 128 columns then fold by a halving tree, and group g's total
 * is added into chunk[record_offset + 47 g + lane]. */
void avdo_grain_reduce(const float *partials, uint32_t cells, float *chunk, uint32_t record_offset)
{
    for (uint32_t lane = 0; lane < 47; ++lane) {
        float col[16][128];
        memset(col, 0, sizeof col);
        for (uint32_t t = 0; t < 128; ++t)
            for (uint32_t c = t; c < cells; c += 128) {
                const float gv = partials[(size_t)c * 48 + 47];
                uint32_t g = gr_u32(gv);
                if (g > 15) g = 15;
                col[g][t] += partials[(size_t)c * 48 + lane];
            }
        for (uint32_t g = 0; g < 16; ++g) {
            for (uint32_t stride = 64; stride >= 1; stride /= 2)
                for (uint32_t i = 0; i < stride; ++i) col[g][i] += col[g][i + stride];
            chunk[record_offset + g * 47 + lane] += col[g][0];
        }
    }
}

/* ------------------------------------------------------------------------------------------ */
/* The AR(3) grain fit in f64: undo_mean_removal, yule_walker, solve and quantise_ar           */
/* (nl4d/grain/fit.rs:79-300).                                                                */
/* ------------------------------------------------------------------------------------------ */

/* measured_lane (fit.rs:79-95): the record lane of a lag folded onto the measured half plane. */
static int gf_lane(int dy, int dx)
{
    if (dy < 0 || (dy == 0 && dx < 0)) {
        dy = -dy;
        dx = -dx;
    }
    if (dy > 3 || dx > 6 || dx < -6) return -1;
    return dy == 0 ? dx : 7 + (dy - 1) * 13 + dx + 6;
}

/* cell_mean_term (fit.rs:97-118). */
static double gf_cell_term(const double *rho, int dy, int dx)
{
    double total = 0.0;
    for (int oy = -7; oy <= 7; ++oy)
        for (int ox = -7; ox <= 7; ++ox) {
            const int lane = gf_lane(dy + oy, dx + ox);
            if (lane < 0) continue;
            const double pairs = (double)((8 - abs(oy)) * (8 - abs(ox)));
            total += pairs / 4096.0 * rho[lane];
        }
    return total;
}

static const int GF_LAG_Y[46] = {0, 0, 0, 0, 0, 0, 0, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 2, 2, 2, 2, 2, 2, 2,
                                 2, 2, 2, 2, 2, 2, 3, 3, 3, 3, 3, 3, 3, 3, 3, 3, 3, 3, 3};
static int gf_lag_x(int l) { return l < 7 ? l : (l - 7) % 13 - 6; }

/* quantise_ar (fit.rs:272-305): the finest shift 9..7 at which every weight rounds (half away from
 * zero) into a signed byte, else 6, and the weights rounded at it and clamped. */
void avdo_grain_quantise(const double *sol, int32_t *q, uint32_t *shift)
{
    uint32_t s = 6;
    for (uint32_t cand = 9; cand >= 7; --cand) {
        int fits = 1;
        for (int i = 0; i < 24; ++i) {
            const double v = round(sol[i] * (double)(1u << cand));
            if (v < -128.0 || v > 127.0) fits = 0;
        }
        if (fits) {
            s = cand;
            break;
        }
    }
    for (int i = 0; i < 24; ++i) {
        double v = round(sol[i] * (double)(1u << s));
        if (v < -128.0) v = -128.0;
        if (v > 127.0) v = 127.0;
        q[i] = (int32_t)v;
    }
    *shift = s;
}

/* yule_walker and solve (fit.rs:160-216): AV1's 24 lag-3 weights from a 47-value record by
 * Gaussian elimination with partial pivoting (the last of equal maxima, as Iterator::max_by keeps);
 * 0 for an empty record, a pivot below 1e-18 or a non-finite solution. */
int avdo_grain_yule_walker(const double *cor, double *sol)
{
    int oy[24], ox[24], k = 0;
    for (int dy = -3; dy <= 0; ++dy)
        for (int dx = -3; dx <= 3; ++dx)
            if (dy < 0 || dx < 0) {
                oy[k] = dy;
                ox[k] = dx;
                ++k;
            }
    const double pixels = cor[46];
    if (!(pixels > 0.0)) return 0;
    double m[24][25];
    for (int i = 0; i < 24; ++i) {
        for (int j = 0; j < 24; ++j) m[i][j] = cor[gf_lane(oy[i] - oy[j], ox[i] - ox[j])] / pixels;
        m[i][24] = cor[gf_lane(oy[i], ox[i])] / pixels;
    }
    for (int col = 0; col < 24; ++col) {
        int pivot = col;
        for (int i = col + 1; i < 24; ++i)
            if (fabs(m[i][col]) >= fabs(m[pivot][col])) pivot = i;
        if (fabs(m[pivot][col]) < 1e-18) return 0;
        for (int j = 0; j < 25; ++j) {
            const double t = m[col][j];
            m[col][j] = m[pivot][j];
            m[pivot][j] = t;
        }
        for (int i = col + 1; i < 24; ++i) {
            const double factor = m[i][col] / m[col][col];
            for (int j = col; j < 25; ++j) m[i][j] -= factor * m[col][j];
        }
    }
    for (int i = 23; i >= 0; --i) {
        double v = m[i][24];
        for (int j = i + 1; j < 24; ++j) v -= m[i][j] * sol[j];
        sol[i] = v / m[i][i];
    }
    for (int i = 0; i < 24; ++i)
        if (!isfinite(sol[i])) return 0;
    return 1;
}

/* Per record (47 f64: 46 lag sums, pixel count): the corrected record, the 24 solved weights in
 * AR_OFFSETS order, their AV1 integers and shift, and whether the system solved. */
void avdo_grain_fit(const double *records, uint32_t n, double *corrected, double *weights, int32_t *quantised,
                    uint32_t *shift, int32_t *solved)
{
    for (uint32_t r = 0; r < n; ++r) {
        const double *rec = records + (size_t)r * 47;
        double *cor = corrected + (size_t)r * 47, *w = weights + (size_t)r * 24;
        int32_t *q = quantised + (size_t)r * 24;
        for (int l = 0; l < 47; ++l) cor[l] = rec[l];
        const double zero_lag = rec[0];
        if (zero_lag > 0.0) {
            double measured[46], rho[46], next[46];
            for (int l = 0; l < 46; ++l) measured[l] = rho[l] = rec[l] / zero_lag;
            rho[0] = 1.0;
            for (int round = 0; round < 20; ++round) {
                const double zt = gf_cell_term(rho, 0, 0);
                for (int l = 0; l < 46; ++l) next[l] = measured[l] * (1.0 - zt) + gf_cell_term(rho, GF_LAG_Y[l], gf_lag_x(l));
                for (int l = 0; l < 46; ++l) rho[l] = next[l];
            }
            for (int l = 0; l < 46; ++l) cor[l] = rho[l] * zero_lag;
        }
        for (int i = 0; i < 24; ++i) {
            w[i] = 0.0;
            q[i] = 0;
        }
        shift[r] = 0;
        solved[r] = 0;
        double sol[24];
        if (!avdo_grain_yule_walker(cor, sol)) continue;
        for (int i = 0; i < 24; ++i) w[i] = sol[i];
        avdo_grain_quantise(sol, q, shift + r);
        solved[r] = 1;
    }
}

/* ------------------------------------------------------------------------------------------ */
/* Grain persistence and the vector save (nl4d/kernels/grain_persistence.rs:54-279,           */
/* nl4d/kernels/grain.rs:50-75).                                                              */
/* ------------------------------------------------------------------------------------------ */

/* grain_save_vectors: one neighbour's vectors and confidences into saved entry `entry`. */
void avdo_grain_save_vectors(const int32_t *mv_field, const float *confidence, int32_t *saved_mv, float *saved_conf,
                             uint32_t mv_offset, uint32_t conf_offset, uint32_t entry, uint32_t blocks)
{
    for (uint32_t b = 0; b < blocks; ++b) {
        const uint32_t s = entry * blocks + b;
        saved_mv[s * 2] = mv_field[mv_offset + b * 2];
        saved_mv[s * 2 + 1] = mv_field[mv_offset + b * 2 + 1];
        saved_conf[s] = confidence[conf_offset + b];
    }
}

/* grain_project (grain_persistence.rs:54-153): per 8x8 cell the 8 pattern projections of channel
 * `ch` of ring slot `frame_slot` (pixel y * 8 + x added where the pattern bit is set, taken away
 * where clear; `patterns` two u32 words per pattern, low first) and the flat flag of the denoised
 * cell, into history entry (history_slot * cells + cell) * 9; 64-slot halving trees. */
void avdo_grain_project(const float *ring, const float *out, const uint32_t *patterns, float *history, uint32_t W,
                        uint32_t H, uint32_t C, uint32_t ch, uint32_t frame_slot, uint32_t history_slot, int luma_band)
{
    const uint32_t cells_x = W / 8, cells = cells_x * (H / 8);
    for (uint32_t cell = 0; cell < cells; ++cell) {
        const uint32_t ox = cell % cells_x * 8, oy = cell / cells_x * 8;
        float s[9][64], lo0[64], lo1[64], hi0[64], hi1[64];
        for (uint32_t t = 0; t < 64; ++t) {
            const uint32_t x = ox + t % 8, y = oy + t / 8;
            const float sample = ring[(((size_t)frame_slot * H + y) * W + x) * C + ch];
            const float clean = out[((size_t)y * W + x) * C + ch];
            for (uint32_t p = 0; p < 8; ++p) {
                const uint32_t bit = (patterns[p * 2 + t / 32] >> (t % 32)) & 1u;
                s[p][t] = ((float)bit * 2.0f - 1.0f) * sample;
            }
            s[8][t] = clean;
            lo0[t] = hi0[t] = clean;
            lo1[t] = hi1[t] = sample;
        }
        for (uint32_t k = 0; k < 9; ++k) gr_tree_sum(s[k]);
        for (uint32_t stride = 32; stride >= 1; stride /= 2)
            for (uint32_t i = 0; i < stride; ++i) {
                lo0[i] = fminf(lo0[i], lo0[i + stride]);
                lo1[i] = fminf(lo1[i], lo1[i + stride]);
                hi0[i] = fmaxf(hi0[i], hi0[i + stride]);
                hi1[i] = fmaxf(hi1[i], hi1[i + stride]);
            }
        float *e = history + ((size_t)history_slot * cells + cell) * 9;
        for (uint32_t p = 0; p < 8; ++p) e[p] = s[p][0];
        const float mean = s[8][0] / 64.0f, range = hi0[0] - lo0[0];
        const int band = !luma_band || (mean > 20.0f / 255.0f && mean < 235.0f / 255.0f);
        const int flat = range < 5.0f / 255.0f && band && lo1[0] >= 4.0f / 255.0f && hi1[0] <= 251.0f / 255.0f;
        e[8] = flat ? 1.0f : 0.0f;
    }
}

/* grain_persist (grain_persistence.rs:155-230): per block of 1024 cells and lag n + 1, the count,
 * auto sum and cross sum of the cells flagged flat in both slots; thread t of 256 walks cells
 * t, t + 256, t + 512, t + 768 of the block, then a 256-slot halving tree per lane. */
void avdo_grain_persist(const float *history, float *partials, uint32_t now_slot, uint32_t cells, uint32_t lags)
{
    const uint32_t blocks = (cells + 1023) / 1024;
    const float auto_scale = 1.0f / 1024.0f, cross_scale = 1.0f / 512.0f;
    for (uint32_t lag = 0; lag < lags; ++lag)
        for (uint32_t b = 0; b < blocks; ++b) {
            const uint32_t earlier = (now_slot + 9 - 1 - lag) % 9;
            float sc[3][256];
            for (uint32_t t = 0; t < 256; ++t) {
                float count = 0.0f, a = 0.0f, x = 0.0f;
                for (uint32_t step = 0; step < 4; ++step) {
                    const uint32_t cell = b * 1024 + t + step * 256;
                    if (cell >= cells) continue;
                    const float *now = history + ((size_t)now_slot * cells + cell) * 9;
                    const float *ear = history + ((size_t)earlier * cells + cell) * 9;
                    if (now[8] * ear[8] > 0.5f) {
                        float sq = 0.0f, pr = 0.0f;
                        for (uint32_t p = 0; p < 8; ++p) {
                            sq += now[p] * now[p] + ear[p] * ear[p];
                            pr += now[p] * ear[p];
                        }
                        count += 1.0f;
                        a += sq * auto_scale;
                        x += pr * cross_scale;
                    }
                }
                sc[0][t] = count;
                sc[1][t] = a;
                sc[2][t] = x;
            }
            for (uint32_t stride = 128; stride >= 1; stride /= 2)
                for (uint32_t i = 0; i < stride; ++i)
                    for (uint32_t l = 0; l < 3; ++l) sc[l][i] += sc[l][i + stride];
            for (uint32_t l = 0; l < 3; ++l) partials[((size_t)lag * blocks + b) * 3 + l] = sc[l][0];
        }
}

/* grain_persist_reduce (grain_persistence.rs:232-279): per lag, thread t of 128 adds blocks t,
 * t + 128, ... in turn, a 128-slot tree per lane, added into record[record_offset + 3 lag + lane]. */
void avdo_grain_persist_reduce(const float *partials, float *record, uint32_t record_offset, uint32_t blocks,
                               uint32_t lags)
{
    for (uint32_t lag = 0; lag < lags; ++lag) {
        float sc[3][128];
        for (uint32_t t = 0; t < 128; ++t)
            for (uint32_t l = 0; l < 3; ++l) {
                float acc = 0.0f;
                for (uint32_t b = t; b < blocks; b += 128) acc += partials[((size_t)lag * blocks + b) * 3 + l];
                sc[l][t] = acc;
            }
        for (uint32_t stride = 64; stride >= 1; stride /= 2)
            for (uint32_t i = 0; i < stride; ++i)
                for (uint32_t l = 0; l < 3; ++l) sc[l][i] += sc[l][i + stride];
        for (uint32_t l = 0; l < 3; ++l) record[record_offset + lag * 3 + l] += sc[l][0];
    }
}
