// NL4D grouping, collaborative filter and aggregation for gfx950.
//
// collab_search follows search_group (collab/kernels/fused/mod.rs:407-553): the centre search
// (centre.rs:37-300) scores the clipped (2S+1)^2 rectangle around each reference and keeps the
// best eight by distance, a tie never displacing an earlier row-major visit (plane_ops.rs:36-77),
// the self-match pinned to slot 0; the trajectory search (trajectory.rs:142-450) follows each
// volume anchor into every neighbour frame through the refine rectangles of its covering motion
// blocks, skipping blocks under c_min, positions inside a rectangle already searched for that
// frame and positions the group already holds, keeps the best per frame and the best grid_frames-1
// frames per volume.  The group word is the member count with 0x100 set for a volume grid.
//
// Distances: with u16 codes the exact integer SSD (packed 16-bit products, v_dot2_u32_u16),
// with f32 values the reference's F32Search form (per-column sums of squared differences, the
// 8-lane butterfly over columns, times the channel scale).
//
// collab_filter follows filter_group (fused/mod.rs:565-850): per channel a vertical then a
// horizontal 8-point DCT (transforms.rs:45-111, fused multiply-adds with f32 taps), the group
// Haar (stack of k_use members or time-by-volume grid, grid.rs:119-220), hard threshold at
// lambda*sqrt(sigma^2 * prof[u] * prof[v]) with the group DC always kept, the inverse-variance
// group weight from channel 0, the inverses, and the fixed-point scatter of every member into its
// own frame's region with the Kaiser taper (aggregate.rs:147-211).
//
// MI355X layout: one wavefront per reference.  In the search the 64 lanes split the candidate
// positions and a wave-wide (distance, visit-order) minimum picks the members; in the filter lane
// (m, q) = (lane / 8, lane % 8) owns column q of member m, so the 8x8 transposes stay inside 8-lane
// groups (through LDS) and the Haar butterflies between members are lane-xor shuffles.  The group
// Haar is written as in-place butterflies: the reference's reordering of approximations only
// permutes member slots, every member carries the same variance, and slot 0 holds the group
// approximation in both layouts, so every kept coefficient and every output value is identical.
#include "common.h"
#include <cstring>
#include <mutex>
#include <type_traits>

namespace avd {

namespace {

constexpr int kWaves = 4;
constexpr int kMaxS = 16;
constexpr uint64_t kNoKey = ~0ull;

typedef short short2_t __attribute__((ext_vector_type(2)));
typedef int int4_t __attribute__((ext_vector_type(4)));

__device__ __forceinline__ int clampi(int v, int lo, int hi) { return v < lo ? lo : (v > hi ? hi : v); }

__device__ __forceinline__ uint32_t pack_pos_t(uint32_t x, uint32_t y, uint32_t t) { return (t << 26) | (y << 13) | x; }

__device__ __forceinline__ uint64_t shfl_u64(uint64_t v, int src)
{
    const uint32_t lo = (uint32_t)__shfl((int)(uint32_t)v, src);
    const uint32_t hi = (uint32_t)__shfl((int)(uint32_t)(v >> 32), src);
    return ((uint64_t)hi << 32) | lo;
}

__device__ __forceinline__ uint64_t wave_min_u64(uint64_t v)
{
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) {
        const uint64_t o = shfl_u64(v, (int)(threadIdx.x & 63) ^ off);
        v = o < v ? o : v;
    }
    return v;
}

// One DPP step of a 64-bit minimum: every lane takes the smaller of its own key and the key of
// the lane CTRL selects within its row of 16 (quad permutes and row rotations read no other row).
template <int CTRL>
__device__ __forceinline__ uint64_t dpp_min_step(uint64_t v)
{
    const uint32_t lo = (uint32_t)__builtin_amdgcn_update_dpp(0, (int)(uint32_t)v, CTRL, 0xF, 0xF, false);
    const uint32_t hi = (uint32_t)__builtin_amdgcn_update_dpp(0, (int)(uint32_t)(v >> 32), CTRL, 0xF, 0xF, false);
    const uint64_t o = ((uint64_t)hi << 32) | lo;
    return o < v ? o : v;
}

__device__ __forceinline__ uint64_t readlane_u64(uint64_t v, int lane)
{
    const uint32_t lo = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)v, lane);
    const uint32_t hi = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)(v >> 32), lane);
    return ((uint64_t)hi << 32) | lo;
}

// Minimum of every row of 16 lanes, left in all of the row's lanes: quad swaps (xor 1, xor 2),
// then row rotations by 4 and 8.  Needs every lane active.
__device__ __forceinline__ uint64_t row16_min_u64(uint64_t v)
{
    v = dpp_min_step<0xB1>(v);   // quad_perm [1, 0, 3, 2]
    v = dpp_min_step<0x4E>(v);   // quad_perm [2, 3, 0, 1]
    v = dpp_min_step<0x124>(v);  // row_ror 4
    v = dpp_min_step<0x128>(v);  // row_ror 8
    return v;
}

// Orders a non-negative f32 distance as an unsigned integer; 0 stays for the pinned self-match.
__device__ __forceinline__ uint32_t float_key(float d) { return __float_as_uint(d) + 1u; }

// ---------------------------------------------------------------------------------------------
// Distances.  `ring` is the slot's base; (ax, ay) the reference/anchor patch and (bx, by) the
// candidate, both in frame `a` / `b`.
// ---------------------------------------------------------------------------------------------

template <int C>
__device__ __forceinline__ uint32_t ssd_codes(const uint16_t *__restrict__ a, const uint16_t *__restrict__ b, int W,
                                              int ax, int ay, int bx, int by)
{
    int acc = 0;
#pragma unroll
    for (int r = 0; r < 8; ++r) {
        const uint16_t *pa = a + ((size_t)(ay + r) * W + ax) * C;
        const uint16_t *pb = b + ((size_t)(by + r) * W + bx) * C;
#pragma unroll
        for (int i = 0; i < 8 * C; i += 2) {
            short2_t va = {(short)pa[i], (short)pa[i + 1]};
            short2_t vb = {(short)pb[i], (short)pb[i + 1]};
            short2_t d = va - vb;
            acc = __builtin_amdgcn_sdot2(d, d, acc, false);
        }
    }
    return (uint32_t)acc;
}

template <int C>
__device__ __forceinline__ float dist_values(const float *__restrict__ a, const float *__restrict__ b, int W, int ax,
                                             int ay, int bx, int by)
{
    float part[8];
#pragma unroll
    for (int q = 0; q < 8; ++q) {
        float acc = 0.0f;
#pragma unroll
        for (int r = 0; r < 8; ++r)
#pragma unroll
            for (int c = 0; c < C; ++c) {
                const float d = a[((size_t)(ay + r) * W + ax + q) * C + c] - b[((size_t)(by + r) * W + bx + q) * C + c];
                acc = __builtin_fmaf(d, d, acc);
            }
        part[q] = acc;
    }
    const float sum = ((part[0] + part[1]) + (part[2] + part[3])) + ((part[4] + part[5]) + (part[6] + part[7]));
    return sum * (C == 1 ? 3.0f : 1.5f);
}

// Distance key of a candidate: codes give the SSD plus one (the self-match is 0); values give the
// f32 distance's bits plus one.  Both order exactly as the reference's distances.
template <int C, bool CODES>
__device__ __forceinline__ uint32_t dist_key(const uint16_t *ca, const uint16_t *cb, const float *fa, const float *fb,
                                             int W, int ax, int ay, int bx, int by)
{
    if (CODES) return ssd_codes<C>(ca, cb, W, ax, ay, bx, by) + 1u;
    return float_key(dist_values<C>(fa, fb, W, ax, ay, bx, by));
}

// ---------------------------------------------------------------------------------------------
// collab_search
// ---------------------------------------------------------------------------------------------

// Inserts `k` into the ascending N-entry list `a`, dropping the largest (keys are unique).
template <int N>
__device__ __forceinline__ void insert_n(uint64_t *a, uint64_t k)
{
#pragma unroll
    for (int i = N - 1; i > 0; --i) a[i] = (k < a[i - 1]) ? a[i - 1] : ((k < a[i]) ? k : a[i]);
    a[0] = (k < a[0]) ? k : a[0];
}

// Makes this wave's LDS writes visible to its other lanes before they read.
__device__ __forceinline__ void wave_sync()
{
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    __builtin_amdgcn_wave_barrier();
}

// KM: the group's member count (8, the reference's MAX_K, or 16); a reference writes KM member
// words and then the group word.
template <int C, bool CODES, int KM>
__global__ __launch_bounds__(kWaves * 64) void k_collab_search(avd_collab_pass p, const uint16_t *__restrict__ codes,
                                                               const float *__restrict__ values,
                                                               const uint32_t *__restrict__ nslots,
                                                               const int32_t *__restrict__ mv,
                                                               const float *__restrict__ conf,
                                                               uint32_t *__restrict__ members,
                                                               unsigned long long *__restrict__ scored_count)
{
    __shared__ int s_rect[kWaves][16][5];  // left, right, top, bottom, positions (0 when skipped)
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int W = p.width, H = p.height;
    const int refs_x = (W - 8 + p.step - 1) / p.step + 1, refs_y = (H - 8 + p.step - 1) / p.step + 1;
    const int refs = refs_x * refs_y;
    int ref = blockIdx.x * kWaves + wave;
    const bool live = ref < refs;
    if (!live) ref = refs - 1;
    const int iy = ref / refs_x, ix = ref - iy * refs_x;
    const int maxx = W - 8, maxy = H - 8;
    const int rx = min(ix * (int)p.step, maxx), ry = min(iy * (int)p.step, maxy);
    const size_t slot_len = (size_t)W * H * C;
    const uint16_t *cc = CODES ? codes + (size_t)p.centre_slot * slot_len : nullptr;
    const float *cv = CODES ? nullptr : values + (size_t)p.centre_slot * slot_len;
    const int S = p.spatial_radius;

    // ---- centre search: each lane keeps the best KM (distance, visit order) keys of its share of
    // the clipped rectangle; KM rounds of a wave-wide minimum merge them in order.
    const int left = clampi(rx - S, 0, maxx), right = clampi(rx + S, 0, maxx);
    const int top = clampi(ry - S, 0, maxy), bottom = clampi(ry + S, 0, maxy);
    const int rw = right - left + 1, rh = bottom - top + 1;
    const int n_live = rw * rh;
    uint64_t best8[KM];
#pragma unroll
    for (int i = 0; i < KM; ++i) best8[i] = kNoKey;
    uint32_t scored = 0;  // candidate patches this lane scores, for the stage's work count
    for (int idx = lane; idx < n_live; idx += 64) {
        const int y = top + idx / rw, x = left + idx % rw;
        const bool self = x == rx && y == ry;
        scored += self ? 0u : 1u;
        const uint32_t d = self ? 0u : dist_key<C, CODES>(cc, cc, cv, cv, W, rx, ry, x, y);
        const uint64_t key = ((uint64_t)d << 16) | (uint64_t)idx;
        if (key < best8[KM - 1]) insert_n<KM>(best8, key);
    }
    uint32_t mem[KM];
#pragma unroll
    for (int m = 0; m < KM; ++m) {
        const uint64_t best = wave_min_u64(best8[0]);
        if (best == kNoKey) {
            mem[m] = 0;  // an unfilled slot keeps position word 0
        } else {
            const int idx = (int)(best & 0xFFFF);
            mem[m] = pack_pos_t((uint32_t)(left + idx % rw), (uint32_t)(top + idx / rw), 0);
            if (best8[0] == best) {
#pragma unroll
                for (int i = 0; i < KM - 1; ++i) best8[i] = best8[i + 1];
                best8[KM - 1] = kNoKey;
            }
        }
    }
    uint32_t k_use = 1;
    while (k_use * 2 <= (uint32_t)n_live && k_use * 2 <= p.group_size) k_use *= 2;

    // ---- trajectory search into the neighbour frames.
    const int radius = p.temporal_radius;
    const int gf = radius == 0 ? 1 : (radius == 1 ? 2 : 4);
    bool use_grid = false;
    if (gf > 1) {
        const int volumes = KM / gf, tail = gf - 1;
        const int bstep = p.blk_step, bsz = p.blksize, bx_n = p.blocks_x, by_n = p.blocks_y;
        const int covers = (bsz + bstep - 1) / bstep;
        const int nrect = covers * covers;
        const size_t nblk = (size_t)bx_n * by_n;
        const int refine = p.refine;
        uint32_t gpos[KM];
        uint64_t gd[KM];
#pragma unroll
        for (int i = 0; i < KM; ++i) {
            gpos[i] = 0;
            gd[i] = kNoKey;
        }
        for (int v = 0; v < volumes; ++v) {
            const int first = v * gf;
            uint32_t anchor = 0;
#pragma unroll
            for (int i = 0; i < KM; ++i)
                if (i == v) anchor = mem[i];
            const int ax = (int)(anchor & 0x1FFF), ay = (int)((anchor >> 13) & 0x1FFF);
#pragma unroll
            for (int i = 0; i < KM; ++i) {
                if (i == first) {
                    gpos[i] = anchor;
                    gd[i] = 0;
                } else if (i > first && i < first + gf) {
                    gpos[i] = 0;
                    gd[i] = kNoKey;
                }
            }
            const int held_n = first + 1;
            const int bx_hi = min(ax / bstep, bx_n - 1), by_hi = min(ay / bstep, by_n - 1);
            const int ovx = max(ax + 8, bsz) - bsz, ovy = max(ay + 8, bsz) - bsz;
            const int bx_lo = min((ovx + bstep - 1) / bstep, bx_hi), by_lo = min((ovy + bstep - 1) / bstep, by_hi);
            for (int t = 0; t < 2 * radius; ++t) {
                const uint16_t *nc = CODES ? codes + (size_t)nslots[t] * slot_len : nullptr;
                const float *nv = CODES ? nullptr : values + (size_t)nslots[t] * slot_len;
                // Lane b describes rectangle b of this frame, in the reference's (iy, ix) walk order.
                if (lane < nrect) {
                    const int wbx = bx_lo + lane % covers, wby = by_lo + lane / covers;
                    int l = 1, r = 0, tp = 1, bt = 0;
                    if (wbx <= bx_hi && wby <= by_hi) {
                        const int block = wby * bx_n + wbx;
                        if (conf[(size_t)t * nblk + block] >= p.c_min) {
                            const int32_t *v2 = mv + ((size_t)t * nblk + block) * 2;
                            const int px = ax + v2[0], py = ay + v2[1];
                            l = clampi(px - refine, 0, maxx);
                            r = clampi(px + refine, 0, maxx);
                            tp = clampi(py - refine, 0, maxy);
                            bt = clampi(py + refine, 0, maxy);
                        }
                    }
                    s_rect[wave][lane][0] = l;
                    s_rect[wave][lane][1] = r;
                    s_rect[wave][lane][2] = tp;
                    s_rect[wave][lane][3] = bt;
                    s_rect[wave][lane][4] = (r - l + 1) * (bt - tp + 1);
                }
                wave_sync();
                // Lanes split each rectangle's positions; the walk index orders the first visit.
                uint64_t best = kNoKey;
                int walk = 0;
                for (int b = 0; b < nrect; ++b) {
                    const int n = s_rect[wave][b][4];
                    if (n <= 0) continue;
                    const int l = s_rect[wave][b][0], tp = s_rect[wave][b][2];
                    const int w = s_rect[wave][b][1] - l + 1;
                    for (int i = lane; i < n; i += 64) {
                        const int cx = l + i % w, cy = tp + i / w;
                        bool skipped = false;
                        for (int s2 = 0; s2 < b; ++s2)
                            if (s_rect[wave][s2][4] > 0 && cx >= s_rect[wave][s2][0] && cx <= s_rect[wave][s2][1] &&
                                cy >= s_rect[wave][s2][2] && cy <= s_rect[wave][s2][3])
                                skipped = true;
                        const uint32_t packed = pack_pos_t((uint32_t)cx, (uint32_t)cy, (uint32_t)t + 1);
#pragma unroll
                        for (int k = 0; k < KM; ++k)
                            if (k < held_n && gpos[k] == packed) skipped = true;
                        if (skipped) continue;
                        ++scored;
                        const uint32_t d = dist_key<C, CODES>(cc, nc, cv, nv, W, ax, ay, cx, cy);
                        const uint64_t key = ((uint64_t)d << 32) | (uint32_t)(walk + i);
                        best = key < best ? key : best;
                    }
                    walk += n;
                }
                best = wave_min_u64(best);
                uint64_t cd = kNoKey;
                uint32_t cp = 0;
                if (best != kNoKey) {
                    int i = (int)(best & 0xFFFFFFFFu);
                    for (int b = 0; b < nrect; ++b) {
                        const int n = s_rect[wave][b][4];
                        if (n <= 0) continue;
                        if (i < n) {
                            const int l = s_rect[wave][b][0], tp = s_rect[wave][b][2];
                            const int w = s_rect[wave][b][1] - l + 1;
                            cp = pack_pos_t((uint32_t)(l + i % w), (uint32_t)(tp + i / w), (uint32_t)t + 1);
                            break;
                        }
                        i -= n;
                    }
                    cd = best >> 32;
                }
                wave_sync();
                // The frame's match joins the volume's ascending list; earlier frames win ties.
#pragma unroll
                for (int i = 0; i < KM; ++i) {
                    if (i > first && i < first + gf && cd < gd[i]) {
                        const uint64_t hd = gd[i];
                        const uint32_t hp = gpos[i];
                        gd[i] = cd;
                        gpos[i] = cp;
                        cd = hd;
                        cp = hp;
                    }
                }
            }
        }
        use_grid = (k_use == KM);
#pragma unroll
        for (int i = 0; i < KM; ++i)
            if (i % gf == tail) use_grid = use_grid && (gd[i] != kNoKey);
        if (use_grid) {
#pragma unroll
            for (int m = 0; m < KM; ++m) mem[m] = gpos[m];
            k_use = KM;
        }
    }
    if (scored_count) {
        uint32_t total = live ? scored : 0u;
#pragma unroll
        for (int off = 32; off >= 1; off >>= 1) total += (uint32_t)__shfl_xor((int)total, off);
        if (lane == 0) atomicAdd(scored_count, (unsigned long long)total * C);
    }
    if (live && lane < KM + 1) {
        uint32_t word = k_use | (use_grid ? 0x100u : 0u);
#pragma unroll
        for (int m = 0; m < KM; ++m)
            if (lane == m) word = mem[m];
        members[(size_t)ref * (KM + 1) + lane] = word;
    }
}

// ---------------------------------------------------------------------------------------------
// collab_search on u16 codes: exact integer SSD with column sliding.
//
// Every lane holds the reference (anchor) patch in registers.  A lane owns one candidate column x
// and a run of k consecutive candidate rows y0..y0+k-1: it streams the k+7 frame rows of that
// column once, adds each row's packed cross products with the eight patch rows into the eight
// open candidates (v_dot2_u32_u16) and its energy into a window sum, and hands every finished
// candidate's SSD = |P|^2 + |B|^2 - 2<P, B> to `emit`.  Each frame row is loaded once per column
// instead of once per candidate.
// ---------------------------------------------------------------------------------------------

__device__ __forceinline__ short2_t as_s2(uint32_t v) { return __builtin_bit_cast(short2_t, v); }

// Eight pixels starting at absolute code index `a` of the ring, as 4*C packed 16-bit pairs.
// One-channel rows may start on an odd code: two aligned loads and a byte funnel shift.
template <int C>
__device__ __forceinline__ void load_row(const uint16_t *__restrict__ ring, size_t a, uint32_t (&out)[4 * C])
{
    if (C == 1) {
        const uint32_t *p = reinterpret_cast<const uint32_t *>(ring + (a & ~(size_t)1));
        const uint32_t d0 = p[0], d1 = p[1], d2 = p[2], d3 = p[3];
        const uint32_t e = ring[a + 7];
        const uint32_t sh = (uint32_t)(a & 1) * 2u;
        out[0] = __builtin_amdgcn_alignbyte(d1, d0, sh);
        out[1] = __builtin_amdgcn_alignbyte(d2, d1, sh);
        out[2] = __builtin_amdgcn_alignbyte(d3, d2, sh);
        out[3] = __builtin_amdgcn_alignbyte(e, d3, sh);
    } else {
        const uint32_t *p = reinterpret_cast<const uint32_t *>(ring + a);
#pragma unroll
        for (int i = 0; i < 4 * C; ++i) out[i] = p[i];
    }
}

typedef unsigned short ushort2_t __attribute__((ext_vector_type(2)));

__device__ __forceinline__ ushort2_t as_u2(uint32_t v) { return __builtin_bit_cast(ushort2_t, v); }

// Sum over the row of products of packed code pairs (v_dot2_u32_u16).
template <int C>
__device__ __forceinline__ uint32_t row_dot(const uint32_t (&a)[4 * C], const uint32_t (&b)[4 * C], uint32_t acc)
{
#pragma unroll
    for (int i = 0; i < 4 * C; ++i) acc = __builtin_amdgcn_udot2(as_u2(a[i]), as_u2(b[i]), acc, false);
    return acc;
}

// The 8x8 patch at (x, y) of the frame starting at code `frame`; every lane holds a copy (in
// vector registers: scalar registers cannot hold it beside the kernel's other uniform state).
template <int C>
__device__ __forceinline__ void load_patch(const uint16_t *__restrict__ ring, size_t frame, int W, int x, int y,
                                           uint32_t (&P)[8][4 * C])
{
#pragma unroll
    for (int r = 0; r < 8; ++r) load_row<C>(ring, frame + ((size_t)(y + r) * W + x) * C, P[r]);
}

// Candidates (x, y0 + j) for j < K against patch P, whose energy sum(P^2) is A2; emit(j, ssd)
// for j < kk (kk <= K).  SSD = A2 + B2 - 2 * <P, B>, exact in u32 for codes of at most 12 bits
// (2 * 64 * C * 4095^2 < 2^32): the cross products accumulate per candidate with
// v_dot2_u32_u16, and B2 is the window sum of per-row energies.  Candidate j uses frame row
// y0 + t with patch row t - j, and its sums sit in slot j % 8; unrolled, so every (row, patch
// row) pair is a compile-time slot.  Rows past the frame (a lane with kk < K) are clamped and
// their sums dropped.
template <int C, int K, typename Emit>
__device__ __forceinline__ void slide_fixed(const uint16_t *__restrict__ ring, size_t frame, int W, int H, int x,
                                            int y0, int kk, const uint32_t (&P)[8][4 * C], uint32_t A2, Emit &&emit)
{
    uint32_t acc[8], energy[8];
#pragma unroll
    for (int i = 0; i < 8; ++i) acc[i] = 0;
    uint32_t window = 0;
    const size_t col = frame + (size_t)x * C;
#pragma unroll
    for (int t = 0; t < K + 7; ++t) {
        uint32_t row[4 * C];
        load_row<C>(ring, col + (size_t)min(y0 + t, H - 1) * W * C, row);
        // Pair index outermost: consecutive dot ops feed different sums (a dependent dot op
        // waits a cycle).
        uint32_t e = 0;
#pragma unroll
        for (int i = 0; i < 4 * C; ++i) {
            e = __builtin_amdgcn_udot2(as_u2(row[i]), as_u2(row[i]), e, false);
#pragma unroll
            for (int r = 0; r < 8; ++r)
                if (t - r >= 0 && t - r < K)
                    acc[(t - r) & 7] = __builtin_amdgcn_udot2(as_u2(P[r][i]), as_u2(row[i]), acc[(t - r) & 7], false);
        }
        energy[t & 7] = e;
        window += e;
        if (t >= 7) {
            if (t - 7 < kk) emit(t - 7, A2 + window - 2u * acc[(t - 7) & 7]);
            acc[(t - 7) & 7] = 0;
            window -= energy[(t - 7) & 7];
        }
    }
}

// slide_fixed for the trajectory search, over the two adjacent columns x0 (even) and x0 + 1 at
// once: candidates (x0 + c, y0 + j), j < kk <= K, of a frame with H >= K + 7, against patch P.
// One set of row loads serves both columns: a row of the pair spans the codes of
// nine pixels, from the first pixel of column x0, which is dword-aligned (one-channel frames of
// even width, every ring slot and row then starting on an even code).  The K + 7 rows are read
// from y0s = min(y0, H - K - 7), so none passes the frame (a rectangle clipped at the bottom slides
// from higher up and skips its first y0 - y0s candidates).  Rows are read as raw buffer loads from
// the ring (`ring_bytes` long): a scalar resource and one 32-bit byte offset per row; emit(c, j, ok,
// ssd) runs for every row, `ok` false for a j outside [0, kk).  The ninth
// pixel is read alone (one code, or one dword of two): past the last column of a row it is the next
// row's first, and past the ring's end the load returns 0; column x0 + 1 reads it only when it lies
// on the frame.
template <int C, int K, typename Emit>
__device__ __forceinline__ void slide_pair(const uint16_t *__restrict__ ring, uint32_t ring_bytes, size_t frame, int W,
                                           int H, int x0, int y0, int kk, const uint32_t (&P)[8][4 * C], uint32_t A2,
                                           Emit &&emit)
{
    const int y0s = min(y0, H - K - 7);
    const int delta = y0 - y0s;
    const size_t a0 = frame + ((size_t)y0s * W + x0) * C;  // first code of row 0, even
    const __amdgpu_buffer_rsrc_t rsrc =
        __builtin_amdgcn_make_buffer_rsrc(const_cast<uint16_t *>(ring), (short)0, (int)ring_bytes, 0x00020000);
    uint32_t acc0[8], acc1[8], en0[8], en1[8];
#pragma unroll
    for (int i = 0; i < 8; ++i) acc0[i] = acc1[i] = 0;
    uint32_t win0 = 0, win1 = 0;
    const uint32_t base = (uint32_t)(a0 * 2), pitch = (uint32_t)(2 * W * C);
#pragma unroll
    for (int t = 0; t < K + 7; ++t) {
        const uint32_t off = base + (uint32_t)t * pitch;
        uint32_t ra[4 * C], rb[4 * C];
        const int4_t v0 = __builtin_amdgcn_raw_buffer_load_b128(rsrc, (int)off, 0, 0);
        if (C == 1) {
            const uint32_t d4 = (uint32_t)__builtin_amdgcn_raw_buffer_load_b16(rsrc, (int)off + 16, 0, 0);
#pragma unroll
            for (int i = 0; i < 4; ++i) ra[i] = (uint32_t)v0[i];
            rb[0] = __builtin_amdgcn_alignbyte((uint32_t)v0[1], (uint32_t)v0[0], 2);
            rb[1] = __builtin_amdgcn_alignbyte((uint32_t)v0[2], (uint32_t)v0[1], 2);
            rb[2] = __builtin_amdgcn_alignbyte((uint32_t)v0[3], (uint32_t)v0[2], 2);
            rb[3] = __builtin_amdgcn_alignbyte(d4, (uint32_t)v0[3], 2);
        } else {
            const int4_t v1 = __builtin_amdgcn_raw_buffer_load_b128(rsrc, (int)off + 16, 0, 0);
            const uint32_t d8 = (uint32_t)__builtin_amdgcn_raw_buffer_load_b32(rsrc, (int)off + 32, 0, 0);
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                ra[i] = (uint32_t)v0[i];
                ra[4 + i] = (uint32_t)v1[i];
            }
#pragma unroll
            for (int i = 0; i < 7; ++i) rb[i] = ra[i + 1];
            rb[7] = d8;
        }
        uint32_t e0 = 0, e1 = 0;
#pragma unroll
        for (int i = 0; i < 4 * C; ++i) {
            e0 = __builtin_amdgcn_udot2(as_u2(ra[i]), as_u2(ra[i]), e0, false);
            e1 = __builtin_amdgcn_udot2(as_u2(rb[i]), as_u2(rb[i]), e1, false);
#pragma unroll
            for (int r = 0; r < 8; ++r)
                if (t - r >= 0 && t - r < K) {
                    acc0[(t - r) & 7] = __builtin_amdgcn_udot2(as_u2(P[r][i]), as_u2(ra[i]), acc0[(t - r) & 7], false);
                    acc1[(t - r) & 7] = __builtin_amdgcn_udot2(as_u2(P[r][i]), as_u2(rb[i]), acc1[(t - r) & 7], false);
                }
        }
        en0[t & 7] = e0;
        en1[t & 7] = e1;
        win0 += e0;
        win1 += e1;
        if (t >= 7) {
            const int j = t - 7 - delta;
            const bool ok = (unsigned)j < (unsigned)kk;  // no branch: the emit selects
            emit(0, j, ok, A2 + win0 - 2u * acc0[(t - 7) & 7]);
            emit(1, j, ok, A2 + win1 - 2u * acc1[(t - 7) & 7]);
            acc0[(t - 7) & 7] = acc1[(t - 7) & 7] = 0;
            win0 -= en0[(t - 7) & 7];
            win1 -= en1[(t - 7) & 7];
        }
    }
}

// The same for a run length k known only at run time (k up to 33 with spatial_radius 16), in
// blocks of eight rows so that every sum slot stays a compile-time register.  After the first block
// every row adds its cross products into all eight slots: a slot of a candidate past k collects
// sums nobody reads (candidate c and c + 8 never overlap), which spares a condition per product.
template <int C, typename Emit>
__device__ __forceinline__ void slide_blocks(const uint16_t *__restrict__ ring, size_t frame, int W, int H, int x,
                                             int y0, int k, int kk, const uint32_t (&P)[8][4 * C], uint32_t A2,
                                             Emit &&emit)
{
    uint32_t acc[8], energy[8];
#pragma unroll
    for (int i = 0; i < 8; ++i) acc[i] = 0;
    uint32_t window = 0;
    const size_t col = frame + (size_t)x * C;
    const int rows = k + 7;
#pragma unroll
    for (int t = 0; t < 8; ++t) {  // rows < 8 start candidates 0..t only
        uint32_t row[4 * C];
        load_row<C>(ring, col + (size_t)min(y0 + t, H - 1) * W * C, row);
        uint32_t e = 0;
#pragma unroll
        for (int i = 0; i < 4 * C; ++i) {
            e = __builtin_amdgcn_udot2(as_u2(row[i]), as_u2(row[i]), e, false);
#pragma unroll
            for (int r = 0; r <= t; ++r)
                acc[(t - r) & 7] = __builtin_amdgcn_udot2(as_u2(P[r][i]), as_u2(row[i]), acc[(t - r) & 7], false);
        }
        energy[t] = e;
        window += e;
        if (t == 7) {
            if (kk > 0) emit(0, A2 + window - 2u * acc[0]);
            acc[0] = 0;
            window -= energy[0];
        }
    }
    for (int t0 = 8; t0 < rows; t0 += 8) {
#pragma unroll
        for (int u = 0; u < 8; ++u) {
            const int t = t0 + u;  // t % 8 == u; candidate t - 7 sits in slot (u + 1) % 8
            uint32_t row[4 * C];
            load_row<C>(ring, col + (size_t)min(y0 + t, H - 1) * W * C, row);
            uint32_t e = 0;
#pragma unroll
            for (int i = 0; i < 4 * C; ++i) {
                e = __builtin_amdgcn_udot2(as_u2(row[i]), as_u2(row[i]), e, false);
#pragma unroll
                for (int r = 0; r < 8; ++r)
                    acc[(u - r) & 7] = __builtin_amdgcn_udot2(as_u2(P[r][i]), as_u2(row[i]), acc[(u - r) & 7], false);
            }
            energy[u] = e;
            window += e;
            if (t - 7 < kk) emit(t - 7, A2 + window - 2u * acc[(u + 1) & 7]);
            acc[(u + 1) & 7] = 0;
            window -= energy[(u + 1) & 7];
        }
    }
}

// Dispatches a wave-uniform run length to its unrolled form.
template <int C, typename Emit>
__device__ __forceinline__ void slide(const uint16_t *__restrict__ ring, size_t frame, int W, int H, int x, int y0,
                                      int k, int kk, const uint32_t (&P)[8][4 * C], uint32_t A2, Emit &&emit)
{
    switch (k) {
#define AVD_SLIDE_CASE(n) \
    case n: slide_fixed<C, n>(ring, frame, W, H, x, y0, kk, P, A2, emit); break;
        AVD_SLIDE_CASE(1) AVD_SLIDE_CASE(2) AVD_SLIDE_CASE(3) AVD_SLIDE_CASE(4) AVD_SLIDE_CASE(5) AVD_SLIDE_CASE(6)
        AVD_SLIDE_CASE(7) AVD_SLIDE_CASE(8) AVD_SLIDE_CASE(9) AVD_SLIDE_CASE(10) AVD_SLIDE_CASE(11)
        AVD_SLIDE_CASE(12) AVD_SLIDE_CASE(13) AVD_SLIDE_CASE(14) AVD_SLIDE_CASE(15) AVD_SLIDE_CASE(16)
        AVD_SLIDE_CASE(17)
#undef AVD_SLIDE_CASE
    default: slide_blocks<C>(ring, frame, W, H, x, y0, k, kk, P, A2, emit); break;
    }
}

template <int C>
__device__ __forceinline__ uint32_t patch_energy(const uint32_t (&P)[8][4 * C])
{
    uint32_t e[8];
#pragma unroll
    for (int r = 0; r < 8; ++r) e[r] = row_dot<C>(P[r], P[r], 0u);
    return ((e[0] + e[1]) + (e[2] + e[3])) + ((e[4] + e[5]) + (e[6] + e[7]));
}

__device__ __forceinline__ uint32_t med3u(uint32_t a, uint32_t b, uint32_t c)
{
    uint32_t r;
    asm("v_med3_u32 %0, %1, %2, %3" : "=v"(r) : "v"(a), "v"(b), "v"(c));
    return r;
}

// Inserts `k` into the ascending list `a` of N distinct keys, dropping the largest: slot i
// becomes clamp(k, a[i-1], a[i]), one v_med3_u32 per slot.
template <int N>
__device__ __forceinline__ void insert_n_u32(uint32_t *a, uint32_t k)
{
#pragma unroll
    for (int i = N - 1; i > 0; --i) a[i] = med3u(a[i - 1], k, a[i]);
    a[0] = min(a[0], k);
}

__device__ __forceinline__ int wave_inclusive_scan(int x)
{
    const int lane = (int)(threadIdx.x & 63);
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        const int o = __shfl_up(x, off);
        if (lane >= off) x += o;
    }
    return x;
}

// blockIdx -> work block, so that each XCD (blocks are dealt round-robin over eight) walks a
// contiguous run of references and keeps their shared search windows in its own L2.
__device__ __forceinline__ int xcd_block()
{
    const int n = (int)gridDim.x, b = (int)blockIdx.x;
    const int xcd = b & 7, i = b >> 3, q = n >> 3, rem = n & 7;
    return xcd * q + min(xcd, rem) + i;
}


// Adds this workgroup's scored candidates (times the channels) to one of kScoreSlots counters, so
// the atomics of many workgroups do not queue on one address.  Every thread of the block calls it.
__device__ __forceinline__ void add_scored(unsigned long long *scored_count, uint32_t scored, bool live, int C)
{
    __shared__ uint32_t s_sum[kWaves];
    if (!scored_count) return;
    uint32_t total = live ? scored : 0u;
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) total += (uint32_t)__shfl_xor((int)total, off);
    if ((threadIdx.x & 63) == 0) s_sum[threadIdx.x >> 6] = total;
    __syncthreads();
    if (threadIdx.x == 0) {
        unsigned long long sum = 0;
#pragma unroll
        for (int w = 0; w < kWaves; ++w) sum += s_sum[w];
        atomicAdd(scored_count + (size_t)(blockIdx.x % kScoreSlots) * kScoreStride, sum * (unsigned long long)C);
    }
}

// Centre search (centre.rs:37-300) of every reference: the best KM of the clipped rectangle by
// (distance, row-major visit), the self-match pinned first.  Writes the KM positions and the
// stack word k_use; the trajectory kernel may then replace them with a volume grid.
// NARROW: codes of at most 10 bits and spatial_radius <= 15, so a lane's keys are
// (distance << 5 | candidate) in 32 bits; otherwise (distance << 32 | visit index) in 64.
template <int C, bool NARROW, int KM>
__device__ __forceinline__ uint32_t centre_search_wave(const avd_collab_pass &p, const uint16_t *__restrict__ ring,
                                                       uint32_t *__restrict__ members, int ref_raw)
{
    const int lane = threadIdx.x & 63;
    const int W = p.width, H = p.height, maxx = W - 8, maxy = H - 8;
    const int refs_x = (W - 8 + p.step - 1) / p.step + 1, refs_y = (H - 8 + p.step - 1) / p.step + 1;
    const int refs = refs_x * refs_y;
    const bool live = ref_raw < refs;
    const int ref = min(ref_raw, refs - 1);
    const int iy = ref / refs_x, ix = ref - iy * refs_x;
    const int rx = min(ix * (int)p.step, W - 8), ry = min(iy * (int)p.step, H - 8);
    const size_t cframe = (size_t)p.centre_slot * ((size_t)W * H * C);
    const int S = p.spatial_radius;

    uint32_t P[8][4 * C];
    load_patch<C>(ring, cframe, W, rx, ry, P);
    const uint32_t A2 = patch_energy<C>(P);

    const int left = clampi(rx - S, 0, maxx), right = clampi(rx + S, 0, maxx);
    const int top = clampi(ry - S, 0, maxy), bottom = clampi(ry + S, 0, maxy);
    const int rw = right - left + 1, rh = bottom - top + 1;
    const int n_live = rw * rh;
    const int per_col = max(1, 64 / rw);
    const int k = (rh + per_col - 1) / per_col;
    const int tasks = rw * ((rh + k - 1) / k);  // <= 64 for rw <= 33
    const int col = lane % rw, y0 = top + (lane / rw) * k;
    uint32_t L[KM];
    uint64_t LW[KM];
#pragma unroll
    for (int i = 0; i < KM; ++i) {
        L[i] = ~0u;
        LW[i] = kNoKey;
    }
    if (lane < tasks) {
        const int x = left + col;
        const int kk = min(k, bottom - y0 + 1);
        slide<C>(ring, cframe, W, H, x, y0, k, kk, P, A2, [&](int j, uint32_t ssd) {
            const bool self = x == rx && y0 + j == ry;
            const uint32_t d = self ? 0u : ssd + 1u;
            if (NARROW) insert_n_u32<KM>(L, (d << 5) | (uint32_t)j);
            else insert_n<KM>(LW, ((uint64_t)d << 32) | (uint32_t)((y0 + j - top) * rw + col));
        });
    }
    uint32_t word = 0;  // lane m < KM keeps member m, lane KM the group word
#pragma unroll
    for (int m = 0; m < KM; ++m) {
        uint64_t head;
        if (NARROW)
            head = L[0] == ~0u ? kNoKey
                               : (((uint64_t)(L[0] >> 5) << 32) | (uint32_t)((y0 + (int)(L[0] & 31) - top) * rw + col));
        else
            head = LW[0];
        // The wave minimum: rows of 16 by DPP, then the four rows read out.
        const uint64_t rows = row16_min_u64(head);
        const uint64_t a = readlane_u64(rows, 0), b = readlane_u64(rows, 16);
        const uint64_t c = readlane_u64(rows, 32), d = readlane_u64(rows, 48);
        const uint64_t ab = a < b ? a : b, cd = c < d ? c : d;
        const uint64_t best = ab < cd ? ab : cd;
        if (best != kNoKey) {  // an unfilled slot keeps position word 0
            const int idx = (int)(best & 0xFFFFFFFFu);
            if (lane == m) word = pack_pos_t((uint32_t)(left + idx % rw), (uint32_t)(top + idx / rw), 0);
            if (head == best) {
#pragma unroll
                for (int i = 0; i < KM - 1; ++i) {
                    L[i] = L[i + 1];
                    LW[i] = LW[i + 1];
                }
                L[KM - 1] = ~0u;
                LW[KM - 1] = kNoKey;
            }
        }
    }
    uint32_t k_use = 1;
    while (k_use * 2 <= (uint32_t)n_live && k_use * 2 <= p.group_size) k_use *= 2;
    if (lane == KM) word = k_use;
    if (live && lane < KM + 1) members[(size_t)ref * (KM + 1) + lane] = word;
    return live && lane == 0 ? (uint32_t)(n_live - 1) : 0u;  // the self-match is not scored
}

template <int C, bool NARROW, int KM>
__global__ __launch_bounds__(kWaves * 64) void k_centre_codes(avd_collab_pass p, const uint16_t *__restrict__ ring,
                                                              uint32_t *__restrict__ members,
                                                              unsigned long long *__restrict__ scored_count)
{
    const int wave = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
    const uint32_t scored = centre_search_wave<C, NARROW, KM>(p, ring, members, xcd_block() * kWaves + wave);
    add_scored(scored_count, scored, true, C);
}

// ---------------------------------------------------------------------------------------------
// Centre search on the matrix cores.
//
// The cross term <P, B> of every (reference, candidate) pair is a matrix product: the references'
// patches are rows of A, the candidates' patches columns of B, patch elements the shared K.  With
// codes below 1024 it is exact on v_mfma_f32_16x16x32_f16: B holds 1024 + code (the f16 bit
// pattern 0x6400 | code, prepared once per pass), A holds the reference codes split into their
// high and low five bits, so every product and partial sum is an integer below 2^24 and the fp32
// accumulation is exact in any order; <P, B> = 32 * C_hi + C_lo - 1024 * sum(P).
//
// A wave takes a block of 2 x 4 references (2 grid columns, 4 grid rows).  A's 16 rows are the
// (reference, half) pairs; B's columns are 16 consecutive candidate columns of one candidate row
// cy, two tiles covering every column any of the 8 references searches, and the wave walks cy down
// the union of their row ranges, so the 16 lanes of a k-block read consecutive codes of one frame
// row.  Lane (n, g) receives rows 4g..4g+3 of each product, both halves of references 2g and 2g+1
// (grid row g) for candidate column n of the tile, forms their keys and keeps a sorted top KM per
// reference (KM = 8 or 16 group members); the 16 lanes of a row then merge each reference's lists
// by DPP.
//
// Keys are (min(d, kKeyCap) << 10 | row-major index) in 32 bits.  A reference whose last chosen
// key reaches the cap (a distance that large could tie-break differently) is listed and searched
// again with the exact 64-bit keys of the vector search (k_centre_redo).
// ---------------------------------------------------------------------------------------------

typedef _Float16 half8_t __attribute__((ext_vector_type(8)));
typedef float float4_t __attribute__((ext_vector_type(4)));

constexpr uint32_t kKeyCap = (1u << 22) - 1u;
constexpr int kPrepRows = 32;

// The centre frame as f16 patterns of 1024 + code (`plane`, and for one channel `plane_odd`, the
// same shifted by one code so that an 8-code row starting on an odd code is a 4-byte-aligned read)
// and the energy sum(code^2) of the 8x8 patch at every position x <= W - 8, y <= H - 8.  Thread x
// of the grid walks kPrepRows rows, keeping the last 8 rows' energies.
template <int C>
__global__ __launch_bounds__(256) void k_centre_prep(const uint16_t *__restrict__ frame, int W, int H,
                                                     uint16_t *__restrict__ plane, uint16_t *__restrict__ plane_odd,
                                                     uint32_t *__restrict__ energy)
{
    const int x = blockIdx.x * 256 + threadIdx.x;
    const int y0 = blockIdx.y * kPrepRows;
    if (x >= W) return;
    const size_t n_codes = (size_t)W * H * C;
    const int y_end = min(y0 + kPrepRows, H);
    for (int y = y0; y < y_end; ++y) {
#pragma unroll
        for (int c = 0; c < C; ++c) {
            const size_t i = ((size_t)y * W + x) * C + c;
            plane[i] = (uint16_t)(0x6400u | frame[i]);
            if (C == 1) plane_odd[i] = (uint16_t)(0x6400u | (i + 1 < n_codes ? frame[i + 1] : 0u));
        }
    }
    if (x > W - 8) return;
    uint32_t rows[8];
    uint32_t window = 0;
    const int last = min(y0 + kPrepRows, H - 7) + 7;  // energy rows y0 .. min(y0 + kPrepRows, H - 7) - 1
    for (int y = y0; y < last; ++y) {
        uint32_t e = 0;
#pragma unroll
        for (int u = 0; u < 8 * C; ++u) {
            const uint32_t v = frame[((size_t)y * W + x) * C + u];
            e += v * v;
        }
        const int slot = (y - y0) & 7;
        if (y - y0 >= 8) window -= rows[slot];
        rows[slot] = e;
        window += e;
        if (y - y0 >= 7) energy[(size_t)(y - 7) * W + x] = window;
    }
}

// Row minimum of 32-bit keys over the 16 lanes of each DPP row (the value lands in every lane).
__device__ __forceinline__ uint32_t row16_min_u32(uint32_t v)
{
    v = min(v, (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0xB1, 0xF, 0xF, false));   // quad_perm [1,0,3,2]
    v = min(v, (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x4E, 0xF, 0xF, false));   // quad_perm [2,3,0,1]
    v = min(v, (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x124, 0xF, 0xF, false));  // row_ror 4
    v = min(v, (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x128, 0xF, 0xF, false));  // row_ror 8
    return v;
}

// Eight consecutive codes of the f16 plane starting at code `a`, as 4-byte-aligned reads: for one
// channel an odd start reads the shifted copy, `odd_at` elements past the plane.
template <int C>
__device__ __forceinline__ half8_t load_b8(const uint16_t *__restrict__ plane, uint32_t odd_at, uint32_t a)
{
    const uint32_t i = (C == 1 && (a & 1u)) ? odd_at + a - 1u : a;
    const uint32_t *q = reinterpret_cast<const uint32_t *>(plane) + (i >> 1);
    uint32_t w[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) w[k] = q[k];
    return __builtin_bit_cast(half8_t, w);
}

template <int C, int KM>
__global__ __launch_bounds__(kWaves * 64) void k_centre_mfma(avd_collab_pass p, const uint16_t *__restrict__ ring,
                                                             const uint16_t *__restrict__ plane, uint32_t odd_at,
                                                             const uint32_t *__restrict__ energy,
                                                             uint32_t *__restrict__ members,
                                                             uint32_t *__restrict__ redo_list,
                                                             unsigned long long *__restrict__ scored_count)
{
    constexpr int KS = 2 * C;  // K steps of 32 patch elements
    const int lane = threadIdx.x & 63, wave = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
    const int n = lane & 15, g = lane >> 4;
    const int W = p.width, H = p.height, maxx = W - 8, maxy = H - 8;
    const int step = p.step, S = p.spatial_radius;
    const int refs_x = (W - 8 + step - 1) / step + 1, refs_y = (H - 8 + step - 1) / step + 1;
    const int blocks_x = (refs_x + 1) >> 1, blocks_y = (refs_y + 3) >> 2;
    const int blk_raw = xcd_block() * kWaves + wave;
    const bool blk_live = blk_raw < blocks_x * blocks_y;
    const int blk = min(blk_raw, blocks_x * blocks_y - 1);
    const int by = blk / blocks_x, bx = blk - by * blocks_x;
    const size_t cframe = (size_t)p.centre_slot * ((size_t)W * H * C);

    // Reference j = 2q + i of the block sits in grid column 2bx + i and grid row 4by + q (both
    // clamped to the grid; a clamped one is searched but not written).
    auto ref_col = [&](int i) { return min(2 * bx + i, refs_x - 1); };
    auto ref_row = [&](int q) { return min(4 * by + q, refs_y - 1); };
    auto pos_x = [&](int ix) { return min(ix * step, W - 8); };
    auto pos_y = [&](int iy) { return min(iy * step, H - 8); };

    // ---- A: lane (n, g) holds row n (reference n / 2; its high five bits when n is even) at the
    // k-block g of every K step: patch row 4s + g (one channel) or half g & 1 of row 2s + g / 2.
    half8_t A[KS];
    {
        const int j = n >> 1;
        const int ax = pos_x(ref_col(j & 1)), ay = pos_y(ref_row(j >> 1));
        const bool hi = (n & 1) == 0;
#pragma unroll
        for (int s = 0; s < KS; ++s) {
            uint32_t w[4];
            if (C == 1) {
                load_row<1>(ring, cframe + (size_t)(ay + 4 * s + g) * W + ax, w);
            } else {
                const uint32_t *q = reinterpret_cast<const uint32_t *>(
                    ring + cframe + ((size_t)(ay + 2 * s + (g >> 1)) * W + ax) * 2 + 8 * (g & 1));
#pragma unroll
                for (int i = 0; i < 4; ++i) w[i] = q[i];
            }
#pragma unroll
            for (int e = 0; e < 8; ++e) {
                const uint32_t code = (w[e >> 1] >> (16 * (e & 1))) & 0xFFFFu;
                A[s][e] = (_Float16)(float)(hi ? (code >> 5) : (code & 31u));
            }
        }
    }

    // ---- the two references whose keys this lane forms: j = 2g + r, grid row g, column r.
    int rx_[2], ry_[2], left_[2], top_[2], bottom_[2], rw_[2], nlive_[2];
    uint32_t kc_[2];
    bool live_[2];
    int ref_[2];
#pragma unroll
    for (int r = 0; r < 2; ++r) {
        const int ix = ref_col(r), iy = ref_row(g);
        live_[r] = blk_live && 2 * bx + r < refs_x && 4 * by + g < refs_y;
        ref_[r] = iy * refs_x + ix;
        rx_[r] = pos_x(ix);
        ry_[r] = pos_y(iy);
        left_[r] = clampi(rx_[r] - S, 0, maxx);
        const int right = clampi(rx_[r] + S, 0, maxx);
        top_[r] = clampi(ry_[r] - S, 0, maxy);
        bottom_[r] = clampi(ry_[r] + S, 0, maxy);
        rw_[r] = right - left_[r] + 1;
        nlive_[r] = rw_[r] * (bottom_[r] - top_[r] + 1);
        // kc = |P|^2 + 2 * 1024 * sum(P): SSD = kc + |B|^2 - 2 * (32 * C_hi + C_lo).
        uint32_t e2 = 0, sum = 0;
#pragma unroll
        for (int v = 0; v < 8; ++v) {
            uint32_t w[4 * C];
            load_row<C>(ring, cframe + ((size_t)(ry_[r] + v) * W + rx_[r]) * C, w);
            e2 = row_dot<C>(w, w, e2);
#pragma unroll
            for (int t = 0; t < 4 * C; ++t) sum = __builtin_amdgcn_udot2(as_u2(w[t]), as_u2(0x00010001u), sum, false);
        }
        kc_[r] = e2 + 2048u * sum;
    }
    // Per tile t and reference r: candidate column cx = cx_base + 16 t + n, whether it is searched,
    // its offset from the rectangle's left edge and whether it is the reference's own column.  A
    // column past the frame reads the last one (its keys are dropped).
    const int cx_base = clampi(pos_x(ref_col(0)) - S, 0, maxx);
    bool colok[2][2], selfcol[2][2];
    int colidx[2][2];
    uint32_t ecol[2], bcol[KS][2];
#pragma unroll
    for (int t = 0; t < 2; ++t) {
        const int cx = cx_base + 16 * t + n;
#pragma unroll
        for (int r = 0; r < 2; ++r) {
            colok[t][r] = cx >= left_[r] && cx < left_[r] + rw_[r];
            selfcol[t][r] = cx == rx_[r];
            colidx[t][r] = cx - left_[r];
        }
        const int lx = min(cx, maxx);
        ecol[t] = (uint32_t)lx;
#pragma unroll
        for (int s = 0; s < KS; ++s)
            bcol[s][t] = C == 1 ? (uint32_t)((4 * s + g) * W + lx) : (uint32_t)(((2 * s + (g >> 1)) * W + lx) * 2 + 8 * (g & 1));
    }
    const int cy_lo = clampi(pos_y(ref_row(0)) - S, 0, maxy), cy_hi = clampi(pos_y(ref_row(3)) + S, 0, maxy);

    // Each lane keeps its best KL = 8 keys per reference (and counts the keys it was offered): with
    // sixteen members a reference's best sixteen rarely crowd one lane's two columns (about 2 % of
    // references put more than eight there on textured 4K content), and a lane that gave up keys
    // and runs out of kept ones during the merge sends its reference to the exact search.
    constexpr int KL = 8;
    uint32_t L[2][KL], offered[2] = {0u, 0u};
#pragma unroll
    for (int r = 0; r < 2; ++r)
#pragma unroll
        for (int i = 0; i < KL; ++i) L[r][i] = ~0u;
    // The next row's operands load while this row's products and keys are formed (two operand
    // sets in turn, so no copy waits on a load in flight).  cy <= maxy, so every patch row exists.
    auto load_row_ops = [&](int cy, half8_t (&B)[KS][2], uint32_t (&eb)[2]) {
        const uint32_t base = (uint32_t)(cy * W * C);
#pragma unroll
        for (int s = 0; s < KS; ++s)
#pragma unroll
            for (int t = 0; t < 2; ++t) B[s][t] = load_b8<C>(plane, odd_at, base + bcol[s][t]);
#pragma unroll
        for (int t = 0; t < 2; ++t) eb[t] = energy[(uint32_t)(cy * W) + ecol[t]];
    };
    auto score_row = [&](int cy, const half8_t (&B)[KS][2], const uint32_t (&eb)[2]) {
        float4_t acc[2] = {{0.f, 0.f, 0.f, 0.f}, {0.f, 0.f, 0.f, 0.f}};
#pragma unroll
        for (int s = 0; s < KS; ++s)
#pragma unroll
            for (int t = 0; t < 2; ++t) acc[t] = __builtin_amdgcn_mfma_f32_16x16x32_f16(A[s], B[s][t], acc[t], 0, 0, 0);
#pragma unroll
        for (int r = 0; r < 2; ++r) {
            const bool rowok = cy >= top_[r] && cy <= bottom_[r];
            const int rowbase = (cy - top_[r]) * rw_[r];
            const bool selfrow = cy == ry_[r];
#pragma unroll
            for (int t = 0; t < 2; ++t) {
                const bool ok = rowok && colok[t][r];
                const uint32_t cross = ((uint32_t)acc[t][2 * r] << 5) + (uint32_t)acc[t][2 * r + 1];
                const uint32_t ssd = kc_[r] + eb[t] - 2u * cross;
                const uint32_t d = (selfrow && selfcol[t][r]) ? 0u : ssd + 1u;
                const uint32_t key = (min(d, kKeyCap) << 10) | (uint32_t)(rowbase + colidx[t][r]);
                insert_n_u32<KL>(L[r], key | (ok ? 0u : ~0u));
                offered[r] += ok ? 1u : 0u;
            }
        }
    };
    half8_t Bx[KS][2], By[KS][2];
    uint32_t ex[2], ey[2];
    load_row_ops(cy_lo, Bx, ex);
    for (int cy = cy_lo; cy <= cy_hi; cy += 2) {
        load_row_ops(min(cy + 1, cy_hi), By, ey);
        score_row(cy, Bx, ex);
        if (cy + 1 > cy_hi) break;
        load_row_ops(min(cy + 2, cy_hi), Bx, ex);
        score_row(cy + 1, By, ey);
    }

    // ---- merge each reference's 16 lists (one DPP row) into its best KM.
    uint32_t word[2] = {0u, 0u};
    bool redo[2];
#pragma unroll
    for (int r = 0; r < 2; ++r) {
        const uint32_t inv = (65536u + (uint32_t)rw_[r] - 1u) / (uint32_t)rw_[r];  // idx / rw exact below 1024
        uint32_t last = ~0u;
        bool short_list = false;  // this lane ran out of kept keys while a later pick could need one
#pragma unroll
        for (int m = 0; m < KM; ++m) {
            const uint32_t head = L[r][0];
            const uint32_t best = row16_min_u32(head);
            if (best != ~0u) {
                last = best;
                const uint32_t idx = best & 1023u;
                const uint32_t qy = (idx * inv) >> 16, qx = idx - qy * (uint32_t)rw_[r];
                if (n == m) word[r] = pack_pos_t((uint32_t)left_[r] + qx, (uint32_t)top_[r] + qy, 0);
                if (head == best) {
#pragma unroll
                    for (int i = 0; i < KL - 1; ++i) L[r][i] = L[r][i + 1];
                    L[r][KL - 1] = ~0u;
                    if (m < KM - 1 && L[r][0] == ~0u && offered[r] > (uint32_t)KL) short_list = true;
                }
            }
        }
        // Keys sort by the capped distance; once a chosen one is capped, their order may not be
        // the distances' order.  A lane of the reference's row that ran short leaves the picks
        // unproven.
        const uint32_t row_short = (uint32_t)(__ballot(short_list) >> (lane & 48)) & 0xFFFFu;
        redo[r] = (last != ~0u && (last >> 10) >= kKeyCap) || row_short != 0u;
        uint32_t k_use = 1;
        while (k_use * 2 <= (uint32_t)nlive_[r] && k_use * 2 <= p.group_size) k_use *= 2;
        // Lane n < KM writes member n; the group word follows them (lane KM for eight members,
        // lane 0 again for sixteen, which fill the row).
        if (live_[r] && !redo[r]) {
            if (n < KM) members[(size_t)ref_[r] * (KM + 1) + n] = word[r];
            if (n == (KM < 16 ? KM : 0)) members[(size_t)ref_[r] * (KM + 1) + KM] = k_use;
        }
    }

    // ---- references whose keys reached the cap go to the exact vector search (k_centre_redo).
#pragma unroll
    for (int r = 0; r < 2; ++r)
        if (n == 0 && redo[r] && live_[r]) redo_list[1 + atomicAdd(redo_list, 1u)] = (uint32_t)ref_[r];
    uint32_t scored = 0;
    if (n == 0) {
#pragma unroll
        for (int r = 0; r < 2; ++r)
            if (live_[r]) scored += (uint32_t)(nlive_[r] - 1);  // the self-match is not scored
    }
    add_scored(scored_count, scored, true, C);
}

// The exact vector centre search (64-bit keys) of the references in `redo_list` (count, then
// references); every wave walks the list with the grid's stride.
template <int C, int KM>
__global__ __launch_bounds__(kWaves * 64) void k_centre_redo(avd_collab_pass p, const uint16_t *__restrict__ ring,
                                                             const uint32_t *__restrict__ redo_list,
                                                             uint32_t *__restrict__ members)
{
    const int wave = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
    const uint32_t count = __builtin_amdgcn_readfirstlane(redo_list[0]);
    for (uint32_t i = blockIdx.x * kWaves + wave; i < count; i += gridDim.x * kWaves)
        centre_search_wave<C, false, KM>(p, ring, members, (int)__builtin_amdgcn_readfirstlane(redo_list[1 + i]));
}

// Trajectory search (trajectory.rs:142-450) and the grid decision (fused/mod.rs:407-553) of every
// reference whose centre search is in `members`: each volume anchor is followed into every
// neighbour frame through the refine rectangles of its covering motion blocks, skipping blocks
// under c_min, positions inside a rectangle already searched for that frame and positions the
// group already holds; the best per frame joins the volume.  When every volume fills and the
// stack holds the group size, the members become the volume grid.  Needs temporal_radius >= 1.
//
// K = 2 * refine + 1.  A wave serves NR references (lanes r * 64 / NR onwards keep reference r's
// lists).  The rectangles of every volume are built up front (entry (v, t, b), b the covering
// block in row-major order, padded to a power-of-two square); the non-empty ones of all volumes
// and all NR references are compacted into one list and lane task k * KP + jp slides column pair
// jp of rectangle k (two adjacent columns from one set of row loads), so the references and
// volumes share the lanes their own few rectangles would leave idle.  Each (reference, volume,
// frame) best (distance, walk index) is an LDS atomic min; the volumes then resolve in order, lane
// t of a reference's lanes decoding frame t (a best an earlier volume already holds is searched
// again without the held positions), and the frames join the volume through the reference's
// ordered insertion.
// LDS of one trajectory workgroup with NR references per wave: the rectangle tables (dynamic) and
// the lists, frame slots, per-frame bests, anchor patches, energies and held picks, and the scored
// counters' per-wave sums (static).
constexpr size_t kLdsBudget = 64 * 1024;

// Volumes of a group at most (two frames per volume) for KM group members; an earlier volume
// holds at most one pick per neighbour frame, so a frame holds at most VOLS - 1 picks.
// The static LDS of the trajectory kernel for NR references per wave in groups of KM with volume
// grids of GF frames (KM / GF volumes).
__host__ __device__ constexpr size_t trajectory_static_lds(int C, int NR, int KM, int GF)
{
    return (size_t)kWaves * (16 * 4 + NR * (KM / GF) * 8 * 4 * C * 4 + NR * (KM / GF) * 4 +
                             NR * 16 * (KM / GF - 1) * 4 + NR * 16 + NR * KM * 4 + 4);
}

// x / d for x < 2^16 and a divisor fixed for the kernel: one high multiply by m = ceil(2^32 / d)
// (exact over that range for every d < 2^16; d = 1 has no 32-bit m and passes x through).
struct FastDiv {
    uint32_t m;
    __device__ explicit FastDiv(uint32_t d) : m(d > 1 ? (uint32_t)((0x100000000ull + d - 1) / d) : 0u) {}
    __device__ __forceinline__ int operator()(int x) const
    {
        return m ? (int)__umulhi((uint32_t)x, m) : x;
    }
};

__device__ __forceinline__ uint32_t lane_mask_lt(uint64_t bal)
{
    return __builtin_amdgcn_mbcnt_hi((uint32_t)(bal >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)bal, 0u));
}

// Entries of one reference's rectangle table per volume (neighbour frames times covering blocks),
// and over all its volumes.
__host__ __device__ __forceinline__ int trajectory_volume_entries(const avd_collab_pass &p)
{
    const int covers = ((int)p.blksize + (int)p.blk_step - 1) / (int)p.blk_step;
    const int cshift = covers <= 1 ? 0 : (covers <= 2 ? 1 : 2);
    return (2 * (int)p.temporal_radius) << (2 * cshift);
}

__host__ __device__ __forceinline__ int trajectory_entries(const avd_collab_pass &p)
{
    const int gf = p.temporal_radius == 1 ? 2 : 4;
    return ((int)p.group_size / gf) * trajectory_volume_entries(p);
}

// The dynamic LDS of NR references per wave: each wave's best keys ([NR][volumes][frames] u64),
// rectangle tables ([NR * entries][2] u32), list of non-empty entries ([NR * entries] u16) and
// list of entries searched again ([NR * volume entries] u16).
__host__ __device__ __forceinline__ size_t trajectory_dynamic_lds(const avd_collab_pass &p, int NR)
{
    const int gf = p.temporal_radius == 1 ? 2 : 4;
    const size_t bests = (size_t)kWaves * NR * ((int)p.group_size / gf) * 2 * p.temporal_radius * sizeof(uint64_t);
    const size_t rects = (size_t)kWaves * NR * trajectory_entries(p) * 2 * sizeof(uint32_t);
    const size_t lists =
        (size_t)kWaves * NR * (trajectory_entries(p) + trajectory_volume_entries(p)) * sizeof(uint16_t);
    return (bests + rects + lists + 15) & ~(size_t)15;
}

// Waves per SIMD the trajectory kernel is built for: four where its registers fit 128 (one channel
// up to seven rows of candidates, two channels up to five), three otherwise.
constexpr int trajectory_waves(int C, int K, bool SHORT) { return !SHORT && K <= (C == 1 ? 7 : 5) ? 4 : 3; }

template <int C, int K, int NR, bool SHORT, int KM, int GF>
__global__ __launch_bounds__(kWaves * 64) __attribute__((amdgpu_waves_per_eu(trajectory_waves(C, K, SHORT))))
void k_trajectory_codes(avd_collab_pass p, const uint16_t *__restrict__ ring,
                                                                  const uint32_t *__restrict__ nslots,
                                                                  const int32_t *__restrict__ mv,
                                                                  const float *__restrict__ conf,
                                                                  uint32_t *__restrict__ members,
                                                                  unsigned long long *__restrict__ scored_count)
{
    constexpr int LPR = 64 / NR;  // lanes per reference
    constexpr int VOLS = KM / GF;
    static_assert(LPR >= KM, "a reference's lanes write its member words");
    // Dynamic: [kWaves][NR][volumes][frames] the best key of each (reference, volume, frame); then
    // [kWaves][NR * entries][2] rectangles, (left | right << 16), (top | bottom << 16); then
    // [kWaves][NR * entries] the non-empty entries, e | r << 10; then [kWaves][NR * volume entries]
    // the entries searched again.
    extern __shared__ unsigned long long s_dyn[];
    __shared__ uint32_t s_slot[kWaves][16];
    __shared__ uint32_t s_patch[kWaves][NR][VOLS][8][4 * C];  // every volume's anchor patch
    __shared__ uint32_t s_energy[kWaves][NR][VOLS];
    __shared__ uint32_t s_held[kWaves][NR][16][VOLS - 1];  // earlier volumes' picks, by neighbour frame
    __shared__ uint8_t s_hcnt[kWaves][NR][16];  // held picks per frame (at most VOLS - 1)
    __shared__ uint32_t s_final[kWaves][NR][KM];  // the grid's members, volume by volume
    constexpr int refine = (K - 1) / 2;
    const int lane = threadIdx.x & 63, wave = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
    const int hr = lane / LPR, hl = lane % LPR, hbase = hr * LPR;
    const int W = p.width, H = p.height, maxx = W - 8, maxy = H - 8;
    const int refs_x = (W - 8 + p.step - 1) / p.step + 1, refs_y = (H - 8 + p.step - 1) / p.step + 1;
    const int refs = refs_x * refs_y;
    const int ref0 = (xcd_block() * kWaves + wave) * NR;  // the wave's first reference
    const int my_ref = min(ref0 + hr, refs - 1);
    const bool live = ref0 + hr < refs;
    const size_t slot_len = (size_t)W * H * C;
    const size_t cframe = (size_t)p.centre_slot * slot_len;
    const uint32_t ring_bytes = (uint32_t)(slot_len * p.slots * 2);
    const uint32_t mine = members[(size_t)my_ref * (KM + 1) + min(hl, KM - 1)];
    const uint32_t k_use = members[(size_t)my_ref * (KM + 1) + KM];
    uint32_t anchors[VOLS];
#pragma unroll
    for (int v = 0; v < VOLS; ++v) anchors[v] = (uint32_t)__shfl((int)mine, hbase + v);

    const int radius = p.temporal_radius;
    constexpr int gf = GF;
    constexpr int volumes = VOLS;
    const int bstep = p.blk_step, bsz = p.blksize, bx_n = p.blocks_x, by_n = p.blocks_y;
    const int covers = (bsz + bstep - 1) / bstep;
    const int cshift = covers <= 1 ? 0 : (covers <= 2 ? 1 : 2);
    const int nshift = 2 * cshift, cmask = (1 << cshift) - 1;
    const int frames_n = 2 * radius;
    const int per_vol = frames_n << nshift;
    const int ents = volumes * per_vol;
    const size_t nblk = (size_t)bx_n * by_n;
    const FastDiv div_vol((uint32_t)per_vol), div_ents((uint32_t)ents), div_bstep((uint32_t)bstep);
    const int fb_n = volumes * frames_n;  // bests per reference
    unsigned long long *const s_fbest = s_dyn + (size_t)wave * NR * fb_n;
    uint32_t *const rect_base = reinterpret_cast<uint32_t *>(s_dyn + (size_t)kWaves * NR * fb_n);
    uint32_t(*const rect_tab)[2] = reinterpret_cast<uint32_t(*)[2]>(rect_base) + (size_t)wave * NR * ents;
    uint16_t *const list_base = reinterpret_cast<uint16_t *>(rect_base + (size_t)kWaves * NR * ents * 2);
    uint16_t *const s_list = list_base + (size_t)wave * NR * ents;
    uint16_t *const s_redo = list_base + (size_t)kWaves * NR * ents + (size_t)wave * NR * per_vol;
    if (lane < frames_n) s_slot[wave][lane] = nslots[lane];
    // Lane (v, r) stages row r of volume v's anchor patch: its loads go out now, its LDS stores
    // after the rectangle tables' loads have gone out too.
    constexpr int PROWS = (VOLS * 8 + LPR - 1) / LPR;
    uint32_t prow[PROWS][4 * C];
#pragma unroll
    for (int u = 0; u < PROWS; ++u) {
        const int i = min(hl + u * LPR, VOLS * 8 - 1), v = i >> 3, r = i & 7;
        uint32_t anchor = anchors[0];
#pragma unroll
        for (int j = 1; j < VOLS; ++j)
            if (v == j) anchor = anchors[j];
        load_row<C>(ring, cframe + ((size_t)(((anchor >> 13) & 0x1FFF) + r) * W + (anchor & 0x1FFF)) * C, prow[u]);
    }

    // ---- every volume's rectangles, per reference.  A lane takes six entries at a time and issues
    // the loads of all six fields (block index 0 for an entry without a block) before it reads any,
    // so the wave waits for memory once per six entries.  Six is a lane's share at sixteen members
    // and temporal radius 3 (2 r entries for each of its 16 lanes): one round there, no idle slots
    // (eight measured 1.7 % slower for the search).
    constexpr int EPL = 6;
    for (int base = hl; base < ents; base += EPL * LPR) {
        size_t blk[EPL];
        bool has[EPL];
        int px[EPL], py[EPL];
#pragma unroll
        for (int u = 0; u < EPL; ++u) {
            const int idx = base + u * LPR;
            const int v = div_vol(idx), e = idx - v * per_vol;
            const int t = e >> nshift, b = e & ((1 << nshift) - 1);
            uint32_t anchor = anchors[0];
#pragma unroll
            for (int i = 1; i < VOLS; ++i)
                if (v == i) anchor = anchors[i];
            const int ax = (int)(anchor & 0x1FFF), ay = (int)((anchor >> 13) & 0x1FFF);
            const int bx_hi = min(div_bstep(ax), bx_n - 1), by_hi = min(div_bstep(ay), by_n - 1);
            const int ovx = max(ax + 8, bsz) - bsz, ovy = max(ay + 8, bsz) - bsz;
            const int bx_lo = min(div_bstep(ovx + bstep - 1), bx_hi), by_lo = min(div_bstep(ovy + bstep - 1), by_hi);
            const int wbx = bx_lo + (b & cmask), wby = by_lo + (b >> cshift);
            has[u] = idx < ents && wbx <= bx_hi && wby <= by_hi;
            blk[u] = has[u] ? (size_t)t * nblk + (size_t)wby * bx_n + wbx : 0;
            px[u] = ax;
            py[u] = ay;
        }
        float c[EPL];
        int32_t dx[EPL], dy[EPL];
#pragma unroll
        for (int u = 0; u < EPL; ++u) {
            c[u] = conf[blk[u]];
            dx[u] = mv[blk[u] * 2];
            dy[u] = mv[blk[u] * 2 + 1];
        }
#pragma unroll
        for (int u = 0; u < EPL; ++u) {
            const int idx = base + u * LPR;
            uint32_t xr = 1u, yr = 1u;  // left 1 > right 0: empty
            if (has[u] && c[u] >= p.c_min) {
                const int qx = px[u] + dx[u], qy = py[u] + dy[u];
                xr = (uint32_t)clampi(qx - refine, 0, maxx) | ((uint32_t)clampi(qx + refine, 0, maxx) << 16);
                yr = (uint32_t)clampi(qy - refine, 0, maxy) | ((uint32_t)clampi(qy + refine, 0, maxy) << 16);
            }
            if (idx < ents) {
                rect_tab[hr * ents + idx][0] = xr;
                rect_tab[hr * ents + idx][1] = yr;
            }
        }
    }
#pragma unroll
    for (int u = 0; u < PROWS; ++u) {
        const int i = hl + u * LPR;
        if (i < VOLS * 8) {
#pragma unroll
            for (int j = 0; j < 4 * C; ++j) s_patch[wave][hr][i >> 3][i & 7][j] = prow[u][j];
        }
    }
    if (hl < frames_n) s_hcnt[wave][hr][hl] = 0;
    if (hl < KM) {  // each volume's anchor first, its picks to come
        uint32_t anchor = 0;
#pragma unroll
        for (int i = 0; i < VOLS; ++i)
            if (hl == i * GF) anchor = anchors[i];
        s_final[wave][hr][hl] = anchor;
    }
    wave_sync();
    if (hl < volumes) {
        uint32_t e = 0;
#pragma unroll
        for (int r = 0; r < 8; ++r)
#pragma unroll
            for (int i = 0; i < 4 * C; ++i)
                e = __builtin_amdgcn_udot2(as_u2(s_patch[wave][hr][hl][r][i]), as_u2(s_patch[wave][hr][hl][r][i]), e,
                                           false);
        s_energy[wave][hr][hl] = e;
    }

    // Every volume's frames are searched in one task loop first, with no position excluded for
    // the group: a volume's list then takes its frames' bests in volume order, and a best that is
    // already held (an earlier volume picked it in that frame) is searched again, that reference
    // and frame alone, with the held positions skipped.  The minimum over the positions not held
    // is the unconstrained minimum whenever that is not held, so the picks are the ones of a search
    // that skips held positions from the start.  A volume's list holds its gf - 1 frame picks and
    // their distances (SSD + 1 fits 32 bits; ~0u marks an empty slot); s_final collects the grid's
    // members.  All of this is the same on every lane of one reference.
    uint32_t scored = 0;
    for (int i = lane; i < NR * fb_n; i += 64) s_fbest[i] = kNoKey;
    // Compact the rectangles that hold a position no earlier rectangle of their (volume, frame)
    // covers (an empty or fully covered one adds nothing but its share of the walk index), in
    // (reference, volume, entry) order.
    int n_valid = 0;
    for (int base = 0; base < NR * ents; base += 64) {
        const int idx = base + lane;
        const int r_of = min(div_ents(idx), NR - 1), e = idx - r_of * ents;
        const uint32_t(*rects)[2] = rect_tab + r_of * ents;
        bool valid = idx < NR * ents && (rects[e][0] & 0xFFFFu) <= (rects[e][0] >> 16);
        if (valid) {
            const int l = (int)(rects[e][0] & 0xFFFFu), r = (int)(rects[e][0] >> 16);
            const int tp = (int)(rects[e][1] & 0xFFFFu), bt = (int)(rects[e][1] >> 16);
            const int f0 = e & ~((1 << nshift) - 1);
            for (int e2 = f0; e2 < e; ++e2) {
                const int l2 = (int)(rects[e2][0] & 0xFFFFu), r2 = (int)(rects[e2][0] >> 16);
                const int tp2 = (int)(rects[e2][1] & 0xFFFFu), bt2 = (int)(rects[e2][1] >> 16);
                if (l2 <= l && r <= r2 && tp2 <= tp && bt <= bt2) valid = false;
            }
        }
        const uint64_t bal = __ballot(valid);
        if (valid) s_list[n_valid + lane_mask_lt(bal)] = (uint16_t)(e | (r_of << 10));
        n_valid += __popcll(bal);
    }
    wave_sync();

    // Column pair `jp` of list entry `ent`'s rectangle, columns x0 = (left & ~1) + 2 jp and x0 + 1
    // (either may lie outside the rectangle): their best (distance, walk index) key joins the
    // (reference, volume, frame) best.  Positions inside an earlier rectangle of the frame are
    // skipped (`walk` counts the positions of those rectangles), and with `held` so are the
    // positions the group holds in that frame.  Returns the positions scored.
    auto run_task = [&](uint32_t ent, int jp, bool held) -> uint32_t {
        const int tr = (int)(ent >> 10), ei = (int)(ent & 0x3FFu);
        const int v = div_vol(ei), e = ei - v * per_vol;
        const uint32_t(*rects)[2] = rect_tab + tr * ents + v * per_vol;
        const int t = e >> nshift, b = e & ((1 << nshift) - 1);
        const int l = (int)(rects[e][0] & 0xFFFFu), r = (int)(rects[e][0] >> 16);
        const int tp = (int)(rects[e][1] & 0xFFFFu), bt = (int)(rects[e][1] >> 16);
        const int x0 = (l & ~1) + 2 * jp;
        const bool on0 = x0 >= l && x0 <= r, on1 = x0 + 1 <= r;
        if (!on0 && !on1) return 0u;  // a rectangle clipped narrower than K
        const int w = r - l + 1, h = bt - tp + 1;
        uint32_t skip0 = on0 ? 0u : ~0u, skip1 = on1 ? 0u : ~0u;
        int walk = 0;
        for (int b2 = 0; b2 < b; ++b2) {
            const uint32_t x2 = rects[(t << nshift) + b2][0], y2 = rects[(t << nshift) + b2][1];
            const int l2 = (int)(x2 & 0xFFFFu), r2 = (int)(x2 >> 16);
            if (l2 > r2) continue;
            const int tp2 = (int)(y2 & 0xFFFFu), bt2 = (int)(y2 >> 16);
            walk += (r2 - l2 + 1) * (bt2 - tp2 + 1);
            const int ylo = max(tp2, tp), yhi = min(bt2, bt);
            if (ylo <= yhi) {
                const uint32_t bits = ((2u << (yhi - ylo)) - 1u) << (ylo - tp);
                if (x0 >= l2 && x0 <= r2) skip0 |= bits;
                if (x0 + 1 >= l2 && x0 + 1 <= r2) skip1 |= bits;
            }
        }
        if (held) {
            const int n_held = (int)s_hcnt[wave][tr][t];
            for (int i = 0; i < n_held; ++i) {
                const uint32_t g = s_held[wave][tr][t][i];
                const int gx = (int)(g & 0x1FFF), gy = (int)((g >> 13) & 0x1FFF);
                if (gy >= tp && gy <= bt) {
                    if (gx == x0) skip0 |= 1u << (gy - tp);
                    if (gx == x0 + 1) skip1 |= 1u << (gy - tp);
                }
            }
        }
        const uint32_t rows = (1u << h) - 1u;
        const uint32_t live0 = rows & ~skip0, live1 = rows & ~skip1;
        if (!(live0 | live1)) return 0u;  // every position of these columns was searched or is held
        uint32_t P[8][4 * C];
#pragma unroll
        for (int rr = 0; rr < 8; ++rr)
#pragma unroll
            for (int i = 0; i < 4 * C; ++i) P[rr][i] = s_patch[wave][tr][v][rr][i];
        uint64_t best = kNoKey;
        uint32_t best_d = ~0u, best_i = 0;
        const int c0 = x0 - l;  // column x0's index in the rectangle (-1 left of it)
        // Every lane runs K rows of candidates; a rectangle clipped shorter drops the rest.
        // Selects only, so the slide stays one basic block and its row loads can all be issued ahead.
        // The paired slide emits in ascending walk index (row by row, column x0 before x0 + 1), so
        // a strictly smaller distance alone decides there; the clamping slide runs its columns one
        // after the other and compares whole keys.
        auto keep = [&](int c, int j, bool ok, uint32_t ssd) {
            const uint32_t sk = c ? skip1 : skip0;
            const bool take = ok & !((sk >> (j & 31)) & 1u);
            const uint32_t idx = (uint32_t)(walk + j * w + c0 + c);
            if constexpr (SHORT) {
                const uint64_t key = ((uint64_t)(ssd + 1u) << 32) | idx;
                const uint64_t k2 = take ? key : kNoKey;
                best = k2 < best ? k2 : best;
            } else {
                const bool better = take & (ssd + 1u < best_d);
                best_d = better ? ssd + 1u : best_d;
                best_i = better ? idx : best_i;
            }
        };
        const size_t frame = (size_t)s_slot[wave][t] * slot_len;
        if (SHORT) {  // rows past the frame are clamped; one column at a time
            if (on0)
                slide_fixed<C, K>(ring, frame, W, H, x0, tp, h, P, s_energy[wave][tr][v],
                                  [&](int j, uint32_t ssd) { keep(0, j, true, ssd); });
            if (on1)
                slide_fixed<C, K>(ring, frame, W, H, x0 + 1, tp, h, P, s_energy[wave][tr][v],
                                  [&](int j, uint32_t ssd) { keep(1, j, true, ssd); });
        } else {
            slide_pair<C, K>(ring, ring_bytes, frame, W, H, x0, tp, h, P, s_energy[wave][tr][v], keep);
            if (best_d != ~0u) best = ((uint64_t)best_d << 32) | best_i;
        }
        if (best != kNoKey) atomicMin(&s_fbest[(tr * volumes + v) * frames_n + t], (unsigned long long)best);
        return ref0 + tr < refs ? (uint32_t)(__popc(live0) + __popc(live1)) : 0u;
    };
    constexpr int KP = (K + 2) / 2;  // column pairs per rectangle
    // One task loop serves the first search (the whole list, nothing held) and each search again
    // (a volume's clashing entries, held positions skipped), so the slide is inlined once.  After a
    // search again every frame of the volume decodes its best again (only the clashing ones
    // changed; keeping the others' decoded bests in LDS across the loop cost the two-channel kernel
    // a workgroup per CU).
    const uint16_t *tlist = s_list;
    int tn = n_valid, v = 0;
    bool held = false;
    bool filled = true;
    for (;;) {
        for (int task = lane; task < tn * KP; task += 64) {
            const int k = task / KP, jp = task - k * KP;
            const uint32_t n = run_task(tlist[k], jp, held);
            if (!held) scored += n;
        }
        wave_sync();
        for (; v < volumes; ++v) {
            const uint32_t(*rects)[2] = rect_tab + hr * ents + v * per_vol;
            // Lane t of each reference's lanes decodes that reference's frame t.
            auto decode = [&](uint32_t &cd, uint32_t &cp) {
                cd = ~0u;
                cp = 0;
                const uint64_t best = s_fbest[(hr * volumes + v) * frames_n + hl];
                if (best != kNoKey) {
                    int i = (int)(best & 0xFFFFFFFFu);
                    for (int b = 0; b < (1 << nshift); ++b) {
                        const uint32_t x2 = rects[(hl << nshift) + b][0], y2 = rects[(hl << nshift) + b][1];
                        const int l2 = (int)(x2 & 0xFFFFu), r2 = (int)(x2 >> 16);
                        if (l2 > r2) continue;
                        const int tp2 = (int)(y2 & 0xFFFFu), bt2 = (int)(y2 >> 16);
                        const int w2 = r2 - l2 + 1, n2 = w2 * (bt2 - tp2 + 1);
                        if (i < n2) {
                            // i / w2 as (i + 0.5) times the reciprocal: exact for i < 2^12, w2 <= 32
                            // (the half keeps the product clear of an integer by more than the
                            // reciprocal's error).
                            const int qy = (int)(((float)i + 0.5f) * __builtin_amdgcn_rcpf((float)w2));
                            cp = pack_pos_t((uint32_t)(l2 + i - qy * w2), (uint32_t)(tp2 + qy), (uint32_t)hl + 1);
                            break;
                        }
                        i -= n2;
                    }
                    cd = (uint32_t)(best >> 32);
                }
            };
            uint32_t cd = ~0u, cp = 0;
            if (held) {
                held = false;  // back from searching this volume again
                if (hl < frames_n) decode(cd, cp);
            } else {
                bool clash = false;
                if (hl < frames_n) {
                    decode(cd, cp);
                    // Held positions of this frame: a clash when one is the best; and each one
                    // inside the frame's rectangles was counted above but is not scored.
                    const int n_held = (int)s_hcnt[wave][hr][hl];
                    for (int i = 0; i < n_held; ++i) {
                        const uint32_t g = s_held[wave][hr][hl][i];
                        clash = clash || g == cp;
                        const int gx = (int)(g & 0x1FFF), gy = (int)((g >> 13) & 0x1FFF);
                        bool inside = false;
                        for (int b = 0; b < (1 << nshift); ++b) {
                            const uint32_t x2 = rects[(hl << nshift) + b][0], y2 = rects[(hl << nshift) + b][1];
                            inside = inside || ((int)(x2 & 0xFFFFu) <= gx && gx <= (int)(x2 >> 16) &&
                                                (int)(y2 & 0xFFFFu) <= gy && gy <= (int)(y2 >> 16));
                        }
                        if (inside && ref0 + hr < refs) scored -= 1u;
                    }
                }
                const uint64_t clashes = __ballot(clash);
                if (clashes) {
                    if (clash) s_fbest[(hr * volumes + v) * frames_n + hl] = kNoKey;
                    // This volume's entries of the clashing (reference, frame) pairs.
                    int n_redo = 0;
                    for (int base = 0; base < n_valid; base += 64) {
                        const int i = base + lane;
                        bool take = false;
                        uint16_t ent = 0;
                        if (i < n_valid) {
                            ent = s_list[i];
                            const int tr = (int)(ent >> 10), ei = (int)(ent & 0x3FFu), vv = div_vol(ei);
                            const int t = (ei - vv * per_vol) >> nshift;
                            take = vv == v && ((clashes >> (tr * LPR + t)) & 1u);
                        }
                        const uint64_t bal = __ballot(take);
                        if (take) s_redo[n_redo + lane_mask_lt(bal)] = ent;
                        n_redo += __popcll(bal);
                    }
                    wave_sync();
                    tlist = s_redo;
                    tn = n_redo;
                    held = true;
                    break;  // search again, then come back to this volume
                }
            }
            // Frames join the volume's ascending list in frame order; a later frame never displaces
            // an equal one, and a displaced entry moves on past equal ones (the reference's ordered
            // insertion, which is not a stable sort when distances tie).
            uint32_t vd[3] = {~0u, ~0u, ~0u}, vp[3] = {0u, 0u, 0u};
            for (int u = 0; u < frames_n; ++u) {
                uint32_t d = (uint32_t)__shfl((int)cd, hbase + u);
                uint32_t q = (uint32_t)__shfl((int)cp, hbase + u);
#pragma unroll
                for (int j = 0; j < 3; ++j) {
                    if (j < gf - 1 && d < vd[j]) {
                        const uint32_t hd = vd[j], hp = vp[j];
                        vd[j] = d;
                        vp[j] = q;
                        d = hd;
                        q = hp;
                    }
                }
            }
            filled = filled && (gf == 2 ? vd[0] : vd[2]) != ~0u;
            // Lanes 1 .. gf - 1 of each reference place members first + 1 .. (the anchor is in
            // place already).
            const int first = v * gf;
            if (hl >= 1 && hl < gf) s_final[wave][hr][first + hl] = hl == 1 ? vp[0] : (hl == 2 ? vp[1] : vp[2]);
            // This volume's picks are held from the next volume on, each in its own frame's list
            // (the picks of one volume sit in distinct frames, so lane j appends pick j alone).
            if (hl < gf - 1) {
                const uint32_t pick = hl == 0 ? vp[0] : (hl == 1 ? vp[1] : vp[2]);
                if (pick != 0u) {
                    const int tf = (int)(pick >> 26) - 1;
                    const uint32_t at = s_hcnt[wave][hr][tf];
                    s_held[wave][hr][tf][at] = pick;
                    s_hcnt[wave][hr][tf] = (uint8_t)(at + 1);
                }
            }
            wave_sync();
        }
        if (v == volumes) break;
    }
    const bool use_grid = k_use == KM && filled;
    if (use_grid && live) {
        // Lane hl < KM writes member hl; the group word follows (lane KM, or lane 0 again when the
        // reference has only KM lanes).
        if (hl < KM) members[(size_t)my_ref * (KM + 1) + hl] = s_final[wave][hr][hl];
        if (hl == (LPR > KM ? KM : 0)) members[(size_t)my_ref * (KM + 1) + KM] = (uint32_t)KM | 0x100u;
    }
    add_scored(scored_count, scored, true, C);
}

// ---------------------------------------------------------------------------------------------
// collab_filter
// ---------------------------------------------------------------------------------------------

struct DctTaps {
    float e[16], o[16];
};

__constant__ DctTaps c_taps;

__device__ __forceinline__ void dct8_fwd(float *l)
{
    float s[4], d[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        s[k] = l[k] + l[7 - k];
        d[k] = l[k] - l[7 - k];
    }
#pragma unroll
    for (int m = 0; m < 4; ++m) {
        float e = 0.0f, o = 0.0f;
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            e = __builtin_fmaf(s[k], c_taps.e[m * 4 + k], e);
            o = __builtin_fmaf(d[k], c_taps.o[m * 4 + k], o);
        }
        l[2 * m] = e;
        l[2 * m + 1] = o;
    }
}

__device__ __forceinline__ void dct8_inv(float *l)
{
    float sn[8];
#pragma unroll
    for (int j = 0; j < 8; ++j) sn[j] = l[j];
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        float e = 0.0f, o = 0.0f;
#pragma unroll
        for (int m = 0; m < 4; ++m) {
            e = __builtin_fmaf(sn[2 * m], c_taps.e[m * 4 + k], e);
            o = __builtin_fmaf(sn[2 * m + 1], c_taps.o[m * 4 + k], o);
        }
        l[k] = e + o;
        l[7 - k] = e - o;
    }
}

// One Haar butterfly level between member lanes XOR (8, 16 or 32) apart.  `active` lanes take
// part; a low lane keeps (low + high) * f, a high lane (low - high) * f, operands in the
// reference's order.
template <int XOR>
__device__ __forceinline__ void haar_level(float (&v)[8], bool active, bool low)
{
    const float f = 0.70710678118654752440f;
#pragma unroll
    for (int u = 0; u < 8; ++u) {
        const float other = __shfl_xor(v[u], XOR);
        if (active) v[u] = low ? (v[u] + other) * f : (other - v[u]) * f;
    }
}

// The group Haar over the members (grid.rs:119-220, stack_haar_*, fused/mod.rs:854-892).  Level i
// pairs members 2^i apart in every layout: lanes 8 << i apart for i < 3, and for sixteen members
// (held as two halves per lane, members m and m + 8) level 3 pairs the two halves inside the lane.
// A stack of k_use members runs log2(k_use) levels; a grid of gf frames per volume runs the time
// levels (pairs inside a volume, only approximations continue) and then the volume levels (pairs
// of volumes, every time coefficient), log2(members) in all.
struct HaarPlan {
    int levels;
    bool act[4], low[4];
};

__device__ __forceinline__ HaarPlan haar_plan(int m, int k_use, bool grid, int gf)
{
    HaarPlan h;
    if (grid) {  // k_use members: gf frames per volume, k_use / gf volumes
        const int f = m % gf, vol = m / gf;
        const int tl = gf == 4 ? 2 : 1;  // time levels
        h.levels = k_use >= 16 ? 4 : 3;
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            if (i < tl) {
                h.act[i] = (f & ((1 << i) - 1)) == 0;
                h.low[i] = ((f >> i) & 1) == 0;
            } else {
                const int j = i - tl;  // volume level j pairs volumes 2^j apart
                h.act[i] = (vol & ((1 << j) - 1)) == 0;
                h.low[i] = ((vol >> j) & 1) == 0;
            }
        }
    } else {
        h.levels = k_use >= 16 ? 4 : (k_use >= 8 ? 3 : (k_use >= 4 ? 2 : (k_use >= 2 ? 1 : 0)));
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            h.act[i] = (m & ((1 << i) - 1)) == 0 && m < k_use;
            h.low[i] = ((m >> i) & 1) == 0;
        }
    }
    return h;
}

// Level 3 of sixteen members: member m (lo) and m + 8 (hi) in one lane.
__device__ __forceinline__ void haar_level_halves(float (&lo)[8], float (&hi)[8], bool active)
{
    const float f = 0.70710678118654752440f;
    if (!active) return;
#pragma unroll
    for (int u = 0; u < 8; ++u) {
        const float a = lo[u], b = hi[u];
        lo[u] = (a + b) * f;
        hi[u] = (a - b) * f;
    }
}

template <int HM>
__device__ __forceinline__ void group_forward(float (&v)[HM][8], const HaarPlan (&h)[HM])
{
#pragma unroll
    for (int s = 0; s < HM; ++s) {
        if (h[s].levels > 0) haar_level<8>(v[s], h[s].act[0], h[s].low[0]);
        if (h[s].levels > 1) haar_level<16>(v[s], h[s].act[1], h[s].low[1]);
        if (h[s].levels > 2) haar_level<32>(v[s], h[s].act[2], h[s].low[2]);
    }
    if (HM == 2 && h[0].levels > 3) haar_level_halves(v[0], v[HM - 1], h[0].act[3]);
}

template <int HM>
__device__ __forceinline__ void group_inverse(float (&v)[HM][8], const HaarPlan (&h)[HM])
{
    if (HM == 2 && h[0].levels > 3) haar_level_halves(v[0], v[HM - 1], h[0].act[3]);
#pragma unroll
    for (int s = 0; s < HM; ++s) {
        if (h[s].levels > 2) haar_level<32>(v[s], h[s].act[2], h[s].low[2]);
        if (h[s].levels > 1) haar_level<16>(v[s], h[s].act[1], h[s].low[1]);
        if (h[s].levels > 0) haar_level<8>(v[s], h[s].act[0], h[s].low[0]);
    }
}

__device__ __forceinline__ int32_t to_fixed(float v, float scale)
{
    const float c = fminf(fmaxf(v, -5.0f), 5.0f);
    return (int32_t)roundf(c * scale);
}

__device__ __forceinline__ int32_t to_fixed_weight(float w, float scale)
{
    const float c = fminf(fmaxf(w, 0.0f), 1.0f);
    return (int32_t)roundf(c * scale * 5.0f);
}

// One pixel's (weight sum, first-channel sum) pair as one 64-bit addend: the weight in the high
// word, the signed sum sign-extended below it.  Both sums stay within i32 (the accumulation scale
// is chosen for that), so the total's low word is the sum and its high word the weight minus the
// borrow of a negative sum.
__device__ __forceinline__ unsigned long long pack_aggregate(int32_t accum, int32_t weight)
{
    return ((unsigned long long)(uint32_t)weight << 32) + (unsigned long long)(long long)accum;
}

__device__ __forceinline__ void unpack_aggregate(unsigned long long v, int32_t &accum, int32_t &weight)
{
    accum = (int32_t)(uint32_t)v;
    weight = (int32_t)(uint32_t)(v >> 32) + (accum < 0 ? 1 : 0);
}

// STAGED: instead of atomics into the accumulators, each member's fixed-point values go to
// `stage` ([ref][member][c][row][column] i32: a member's patch is 256 contiguous bytes per channel)
// and the member enters the list of every 32x32 tile its patch overlaps; k_tile_gather then sums
// them per tile.  KM group members: lane (m, q) holds column q of members m + 8 h, h < KM / 8.
template <int C, bool STAGED, int KM>
__global__ __launch_bounds__(kWaves * 64) void k_collab_filter(avd_collab_pass p, const float *__restrict__ values,
                                                               const uint32_t *__restrict__ nslots,
                                                               const uint32_t *__restrict__ members,
                                                               unsigned long long *__restrict__ aggregate,
                                                               int32_t *__restrict__ accum_second,
                                                               float *__restrict__ group_weight, TileBins tiles)
{
    constexpr int HM = KM / 8;
    constexpr int KSHIFT = KM == 16 ? 4 : 3;
    __shared__ float s_t[kWaves][8][8][9];  // [wave][member][row][col], padded
    // Per channel and coefficient (v, u): the variance sigma^2 prof[u] prof[v] and the hard
    // threshold lambda sqrt(variance), the same f32 expressions for every reference, once per
    // workgroup.
    __shared__ float s_cvs[C][64], s_thr[C][64];
    if (threadIdx.x < 64 * C) {
        const int c = threadIdx.x >> 6, v = (threadIdx.x >> 3) & 7, u = threadIdx.x & 7;
        const float cvs = (p.sigma[c] * p.sigma[c]) * (p.profile[u] * p.profile[v]);
        s_cvs[c][v * 8 + u] = cvs;
        s_thr[c][v * 8 + u] = p.lambda[c] * sqrtf(cvs);
    }
    __syncthreads();
    const int lane = threadIdx.x & 63, wave = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
    const int m = lane >> 3, q = lane & 7;
    const int W = p.width, H = p.height;
    const int refs_x = (W - 8 + p.step - 1) / p.step + 1, refs_y = (H - 8 + p.step - 1) / p.step + 1;
    const int refs = refs_x * refs_y;
    int ref = xcd_block() * kWaves + wave;
    const bool live = ref < refs;
    if (!live) ref = refs - 1;
    const uint32_t *mw = members + (size_t)ref * (KM + 1);
    const uint32_t gword = mw[KM];
    const int k_use = (int)(gword & 0xFF);
    const bool grid = (gword & 0x100u) != 0;
    const int gf = p.temporal_radius == 0 ? 1 : (p.temporal_radius == 1 ? 2 : 4);
    const size_t px = (size_t)W * H;
    HaarPlan plan[HM];
    int slot[HM], mx[HM], my[HM];
    bool in_group[HM];
#pragma unroll
    for (int h = 0; h < HM; ++h) {
        const int mm = m + 8 * h;
        plan[h] = haar_plan(mm, k_use, grid, gf);
        const uint32_t pos = mw[mm];
        const uint32_t t = pos >> 26;
        slot[h] = t > 0 ? (int)nslots[t - 1] : (int)p.centre_slot;
        mx[h] = (int)(pos & 0x1FFF);
        my[h] = (int)((pos >> 13) & 0x1FFF);
        in_group[h] = mm < k_use;
    }
    const float scale = p.accum_scale;
    const int v = q;  // after the first transpose lane (m, q) holds vertical frequency v = q
    float scaled_weight = 0.0f;

#pragma unroll
    for (int c = 0; c < C; ++c) {
        float row[HM][8];  // lane (m, v): coefficients (v, u) of members m + 8 h, u = 0..7
#pragma unroll
        for (int h = 0; h < HM; ++h) {
            const float *src = values + (size_t)slot[h] * px * C;
            float col[8];
#pragma unroll
            for (int r = 0; r < 8; ++r) col[r] = src[((size_t)(my[h] + r) * W + mx[h] + q) * C + c];
            dct8_fwd(col);  // vertical: col[v]
#pragma unroll
            for (int vv = 0; vv < 8; ++vv) s_t[wave][m][vv][q] = col[vv];
            wave_sync();
#pragma unroll
            for (int u = 0; u < 8; ++u) row[h][u] = s_t[wave][m][q][u];
            wave_sync();
            dct8_fwd(row[h]);
        }
        group_forward<HM>(row, plan);

        const float sig2 = p.sigma[c] * p.sigma[c];
        const float lam = p.lambda[c];
        const float pv = p.profile[v];
        float part = 0.0f;  // retained variance of vertical frequency v (every member lane agrees)
        if (p.pool_ratio > 0.0f) {
            // pooled_threshold (fused/pooled.rs:24-147): a coefficient's energy is
            // c^2 / max(sigma^2 prof[v], 1e-20) / max(prof[u], 1e-20); the spatial DC counts as zero
            // while it stands in as a neighbour; the pooled sum (self + below + above, then left,
            // then right) must reach threshold^2 times the in-range neighbour count; the spatial DC
            // takes its own test at lambda and the group DC is always kept.
            const float lv = sig2 * pv;
            const float li = 1.0f / fmaxf(lv, 1e-20f);
            const bool has_up = v > 0, has_down = v < 7;
            const float vertical = 1.0f + (has_up ? 1.0f : 0.0f) + (has_down ? 1.0f : 0.0f);
            const float threshold = lam * p.pool_ratio;
            const float bar = threshold * threshold;
            const float bar_first = bar * (vertical + (v == 1 ? 0.0f : 1.0f));
            const float bar_second = bar * (vertical + (v == 0 ? 1.0f : 2.0f));
            const float bar_middle = bar * (vertical + 2.0f), bar_last = bar * (vertical + 1.0f);
            float kept_profile[HM];
#pragma unroll
            for (int h = 0; h < HM; ++h) {
                float e[8];
#pragma unroll
                for (int u = 0; u < 8; ++u)
                    e[u] = row[h][u] * row[h][u] * li * (1.0f / fmaxf(p.profile[u], 1e-20f));
                if (v == 0) e[0] = 0.0f;
                kept_profile[h] = 0.0f;
#pragma unroll
                for (int u = 0; u < 8; ++u) {
                    const float above = __shfl_up(e[u], 1), below = __shfl_down(e[u], 1);
                    float sum = e[u] + (has_down ? below : 0.0f) + (has_up ? above : 0.0f);
                    if (u > 0) sum += e[u - 1];
                    if (u < 7) sum += e[u + 1];
                    const float bar_u = u == 0 ? bar_first : (u == 1 ? bar_second : (u == 7 ? bar_last : bar_middle));
                    bool keep = sum >= bar_u;
                    if (u == 0) {
                        const bool own = fabsf(row[h][0]) >= lam * sqrtf(lv * p.profile[0]);
                        if (v == 0) keep = own;
                        if (m == 0 && h == 0) keep = keep || v == 0;
                    }
                    kept_profile[h] += keep ? p.profile[u] : 0.0f;
                    if (in_group[h] && !keep) row[h][u] = 0.0f;
                }
            }
            // The reference's lane v adds each member's kept profile times its variance, in member
            // order; lane (j % 8, v) holds member j's in half j / 8.
#pragma unroll
            for (int j = 0; j < KM; ++j) {
                const float kp = __shfl(kept_profile[j / 8], ((j & 7) << 3) | v);
                if (j < k_use) part += kp * lv;
            }
        } else {
            // Kept coefficients per u, one nibble each, per half (at most eight members a half).
            uint32_t counts[HM];
            float cvs[8];
#pragma unroll
            for (int u = 0; u < 8; ++u) cvs[u] = s_cvs[c][v * 8 + u];
#pragma unroll
            for (int h = 0; h < HM; ++h) {
                counts[h] = 0;
#pragma unroll
                for (int u = 0; u < 8; ++u) {
                    bool keep = fabsf(row[h][u]) >= s_thr[c][v * 8 + u];
                    if (u == 0 && m == 0 && h == 0 && v == 0) keep = true;
                    keep = keep && in_group[h];
                    if (keep) counts[h] += 1u << (4 * u);
                    else if (in_group[h]) row[h][u] = 0.0f;
                }
                counts[h] += (uint32_t)__shfl_xor((int)counts[h], 8);
                counts[h] += (uint32_t)__shfl_xor((int)counts[h], 16);
                counts[h] += (uint32_t)__shfl_xor((int)counts[h], 32);
            }
            // Retained variance: each lane (m, v) adds, for every u, the variance of (u, v) once per
            // member that kept it; the eight v lanes then sum by butterfly.
#pragma unroll
            for (int u = 0; u < 8; ++u) {
                int n = 0;
#pragma unroll
                for (int h = 0; h < HM; ++h) n += (counts[h] >> (4 * u)) & 15;
                for (int i = 0; i < n; ++i) part += cvs[u];
            }
        }
        float sum = part;
        sum += __shfl_xor(sum, 1);
        sum += __shfl_xor(sum, 2);
        sum += __shfl_xor(sum, 4);
        const float weight = (isnan(sum) || isinf(sum)) ? 0.0f : 1.0f / fmaxf(sum, 1e-12f);
        if (c == 0) {
            if (live && lane == 0) group_weight[ref] = weight;
            scaled_weight = weight * p.weight_scale;
        }

        group_inverse<HM>(row, plan);
#pragma unroll
        for (int h = 0; h < HM; ++h) {
            const int mm = m + 8 * h;
            dct8_inv(row[h]);  // lane (m, v): values at columns k for vertical frequency v
#pragma unroll
            for (int k = 0; k < 8; ++k) s_t[wave][m][v][k] = row[h][k];
            wave_sync();
            float col[8];
#pragma unroll
            for (int vv = 0; vv < 8; ++vv) col[vv] = s_t[wave][m][vv][q];
            wave_sync();
            dct8_inv(col);  // rows of column q

            if (STAGED) {
                if (live) {
                    int32_t *st = tiles.stage + (((size_t)ref * KM + mm) * C + c) * 64 + q;
                    const float kq = p.kaiser[q];
#pragma unroll
                    for (int r = 0; r < 8; ++r) {
                        const float w = scaled_weight * (p.kaiser[r] * kq);
                        st[r * 8] = in_group[h] ? to_fixed(col[r] * w, scale) : 0;
                    }
                }
            } else if (live && in_group[h]) {
                const size_t base = (size_t)slot[h] * px;
                const float kq = p.kaiser[q];
#pragma unroll
                for (int r = 0; r < 8; ++r) {
                    const float w = scaled_weight * (p.kaiser[r] * kq);
                    const size_t pix = base + (size_t)(my[h] + r) * W + mx[h] + q;
                    const int32_t a = to_fixed(col[r] * w, scale);
                    if (c == 0) atomicAdd(&aggregate[pix], pack_aggregate(a, to_fixed_weight(w, scale)));
                    else atomicAdd(&accum_second[pix], a);
                }
            }
        }
        if (STAGED && c == 0) {
            // Enter each member in the list of every tile its patch overlaps.  Lane (m, q < 4)
            // holds claim q of members m + 8 h.  The workgroup counts its claims per tile in an LDS
            // hash table (rank = the bucket's count before the claim), one atomic per distinct tile
            // reserves the entries, and a full list sends the (member, tile) pair to the overflow
            // list.
            constexpr int kBuckets = kWaves * 64 * HM;
            __shared__ uint32_t s_key[kBuckets], s_cnt[kBuckets], s_base[kBuckets];
#pragma unroll
            for (int h = 0; h < HM; ++h) {
                s_key[threadIdx.x + h * kWaves * 64] = ~0u;
                s_cnt[threadIdx.x + h * kWaves * 64] = 0;
            }
            __syncthreads();
            uint32_t tile[HM], bucket[HM], rank[HM];
#pragma unroll
            for (int h = 0; h < HM; ++h) {
                const int tx0 = mx[h] >> kTileShift, ty0 = my[h] >> kTileShift;
                const int nx = ((mx[h] + 7) >> kTileShift) - tx0 + 1, ny = ((my[h] + 7) >> kTileShift) - ty0 + 1;
                tile[h] = ~0u, bucket[h] = 0, rank[h] = 0;
                if (live && in_group[h] && q < nx * ny) {
                    tile[h] = (uint32_t)slot[h] * tiles.per_slot + (ty0 + q / nx) * tiles.tiles_x + tx0 + q % nx;
                    uint32_t b = (tile[h] * 2654435761u) >> (KM == 16 ? 23 : 24);
                    for (;;) {  // at most half as many keys as buckets: a free or equal bucket exists
                        const uint32_t old = atomicCAS(&s_key[b], ~0u, tile[h]);
                        if (old == ~0u || old == tile[h]) break;
                        b = (b + 1) & (kBuckets - 1);
                    }
                    bucket[h] = b;
                    rank[h] = atomicAdd(&s_cnt[b], 1u);
                }
            }
            __syncthreads();
#pragma unroll
            for (int h = 0; h < HM; ++h) {
                const int i = threadIdx.x + h * kWaves * 64;
                if (s_key[i] != ~0u) s_base[i] = atomicAdd(&tiles.count[s_key[i]], s_cnt[i]);
            }
            __syncthreads();
#pragma unroll
            for (int h = 0; h < HM; ++h) {
                if (tile[h] != ~0u) {
                    const uint32_t k = s_base[bucket[h]] + rank[h];
                    const uint32_t id = ((uint32_t)ref << KSHIFT) | (uint32_t)(m + 8 * h);
                    if (k < tiles.cap) {
                        tiles.list[(size_t)tile[h] * tiles.cap + k] = id;
                    } else {
                        const uint32_t o = atomicAdd(tiles.overflow_count, 1u);
                        tiles.overflow[2 * (size_t)o] = id;
                        tiles.overflow[2 * (size_t)o + 1] = tile[h];
                    }
                }
            }
        }
    }
}

// Adds member `id`'s staged patch (sums, and weights recomputed from the group weight in the
// filter's own arithmetic) for the pixels of row r that fall in [x0, x0 + T) x [y0, y0 + T).
// MIXED: the stage holds the filtered samples as f16; the window and the group weight are applied
// here, as the FP32 form's filter applies them.
// Member ids are (reference << log2 KM) | member, KM the group size.
template <int C, bool MIXED, int KM, typename Add>
__device__ __forceinline__ void member_row(const avd_collab_pass &p, const uint32_t *__restrict__ members,
                                           const float *__restrict__ group_weight, const TileBins &tiles,
                                           uint32_t id, int r, int x0, int y0, int T, Add &&add)
{
    constexpr int KSHIFT = KM == 16 ? 4 : 3;
    const uint32_t ref = id >> KSHIFT, m = id & (KM - 1);
    const uint32_t pos = members[(size_t)ref * (KM + 1) + m];
    const int y = (int)((pos >> 13) & 0x1FFF) + r - y0;
    if (y < 0 || y >= T) return;
    const int xb = (int)(pos & 0x1FFF) - x0;
    const float sw = group_weight[ref] * p.weight_scale;
    const float kr = p.kaiser[r];
    const int32_t *st = tiles.stage + (size_t)id * C * 64 + r * 8;
    const _Float16 *st16 = reinterpret_cast<const _Float16 *>(tiles.stage) + (size_t)id * C * 64 + r * 8;
#pragma unroll
    for (int q = 0; q < 8; ++q) {
        const int x = xb + q;
        if (x < 0 || x >= T) continue;
        const float w = sw * (kr * p.kaiser[q]);
        if (MIXED) {  // the mixed gather's arithmetic
            const float ws = sw * (kr * p.kaiser[q] * p.accum_scale);
            const float lim = 5.0f * p.accum_scale;
            const int32_t a = (int32_t)roundf(fminf(fmaxf((float)st16[q] * ws, -lim), lim));
            const int32_t a2 = C == 2 ? (int32_t)roundf(fminf(fmaxf((float)st16[64 + q] * ws, -lim), lim)) : 0;
            add(y * T + x, pack_aggregate(a, (int32_t)roundf(fminf(ws, p.accum_scale) * 5.0f)), a2);
        } else {
            add(y * T + x, pack_aggregate(st[q], to_fixed_weight(w, p.accum_scale)), C == 2 ? st[64 + q] : 0);
        }
    }
}

// ---------------------------------------------------------------------------------------------
// The mixed-precision filter form: every transform of a group on the matrix cores.
//
// A wave filters four references (32 members).  The FP32 form's 3-D transform (the 2-D DCT of each
// patch, then the group Haar across members) and its inverse become five steps, four of them
// products on v_mfma_f32_32x32x16_f16 (f16 operands, f32 accumulation):
//   1. X1[member][coef]  = P[member][pixel] T[pixel][coef]           2-D DCT, T = (D (x) D)^T
//   2. X2[member][coef]  = H[member][member'] X1[member'][coef]      group Haar: block diagonal,
//                                                                     one 8x8 block per reference
//   3. the hard threshold of X2 and the group weight (the retained variance)
//   4. X3[coef][member]  = X2^T[coef][member'] Hi^T[member'][member]  inverse Haar
//   5. X4[pixel][member] = Ti[pixel][coef] X3[coef][member]           inverse DCT, Ti = T^T
// An accumulator tile keeps its rows in registers (register 4 j + i holds row 8 j + 4 (lane / 32)
// + i) and its columns on lanes, while each operand takes its rows (or columns) on lanes and its K
// values in registers.  So a tile is, as it lies, the operand of the next product contracting over
// its rows: K slot i' of lane half h in K step s is row 16 s + 8 (i' / 4) + 4 h + i' % 4, and the
// tables (T, Ti, H, Hi) are laid out in that order.  No lane exchanges the transforms.  The Haar
// blocks are the linear maps of each reference's Haar plan (the FP32 form's levels, applied to
// unit vectors), split into exact signs and an f32 scale (below).  Samples of up to 11-bit codes
// are exact in f16; the DCT tables and the inverse DCT's input round to f16 (2^-11 relative) and
// the Haar's inputs enter as hi + lo pairs, which keeps PSNR within a thousandth of a dB of the
// FP32 form; the threshold and the retained variance follow the FP32 form's rules, summed in
// another order.  Hard threshold only.  The filtered samples are staged as f16 (half the FP32 form's bytes); the
// gather applies the window and the group weight in fixed point.
// ---------------------------------------------------------------------------------------------

typedef float float16_t __attribute__((ext_vector_type(16)));
typedef float float4u_t __attribute__((ext_vector_type(4))) __attribute__((aligned(4)));
typedef _Float16 half4_t __attribute__((ext_vector_type(4)));

// DCT tables in operand order, uploaded once per device.  Fragment 2 s + mt (s < 4, mt < 2), the
// B operand of step 1: lane (n, h) holds T[pixel 16 s + 8 h + i'][coef 32 mt + n] for i' < 8.
// Fragment 8 + 2 s + pt, the A operand of step 5: lane (n, h) holds Ti[pixel 32 pt + n][coef 16 s +
// 8 (i' / 4) + 4 h + i' % 4].
__device__ uint4 g_mixed_frags[16][64];

__device__ __forceinline__ half8_t frag_from(const float16_t &acc, int base)
{
    half8_t f;
#pragma unroll
    for (int i = 0; i < 8; ++i) f[i] = (_Float16)acc[base + i];
    return f;
}

__device__ __forceinline__ half8_t frag_from(const float (&acc)[16], int base)
{
    half8_t f;
#pragma unroll
    for (int i = 0; i < 8; ++i) f[i] = (_Float16)acc[base + i];
    return f;
}

// f32 values as two f16 fragments, hi = f16(v) and lo = f16(v - hi): hi + lo holds v to about
// 2^-22 relative, so a product with an exact f16 operand (the Haar signs, +-2^-L) keeps f32-like
// precision in two matrix-core steps.
template <typename V>
__device__ __forceinline__ void split_frag(const V &acc, int base, half8_t &hi, half8_t &lo)
{
#pragma unroll
    for (int i = 0; i < 8; ++i) {
        const _Float16 h = (_Float16)acc[base + i];
        hi[i] = h;
        lo[i] = (_Float16)(acc[base + i] - (float)h);
    }
}

// One Haar butterfly level, scaled by f, applied to four unit-vector columns held across the
// member lanes (a reference's members are 8 consecutive lanes; partners are XOR apart).
template <int XOR, int N>
__device__ __forceinline__ void haar_columns(float (&v)[N], bool active, bool low, float f)
{
#pragma unroll
    for (int k = 0; k < N; ++k) {
        const float other = __shfl_xor(v[k], XOR);
        if (active) v[k] = low ? (v[k] + other) * f : (other - v[k]) * f;
    }
}

// The filter of the wave's 32 members (four references of eight, or two of sixteen: KM), from
// their samples converted to f16 (`pa[c][s]`: rows 2 s + h of member n, channel c), writing the
// group weights and the staged samples.
template <int C, int KM>
__device__ __forceinline__ void mixed_group(const avd_collab_pass &p, const uint4 (*s_frag)[64],
                                            const float (*s_cvs)[64], const float (*s_thr)[64],
                                            const half8_t (&pa)[C][4], float inv_s, int first, int refs,
                                            const int (&k_use)[32 / KM], uint32_t gword, int lane,
                                            float *__restrict__ group_weight, const TileBins &tiles)
{
    constexpr int NREF = 32 / KM, KSHIFT = KM == 16 ? 4 : 3;
    const int n = lane & 31, h = lane >> 5, m = n & (KM - 1), jn = n >> KSHIFT;
    const int ref = min(first + jn, refs - 1);
    const bool live = first + jn < refs;
    const int k_use_n = (int)(gword & 0xFF);
    const bool grid = (gword & 0x100u) != 0;
    const int gf = p.temporal_radius == 0 ? 1 : (p.temporal_radius == 1 ? 2 : 4);
    const bool in_group = m < k_use_n;

    // The Haar blocks.  A member that passes L active levels of its reference's plan leaves the
    // forward transform as f^L (f = 1/sqrt 2) times a sum of its group's coefficients with signs
    // +-1, so H = D S with S of +-1 and 0 (exact in f16) and D = diag(f^L); the inverse is
    // H^T = S^T D.  Step 2 multiplies by S only, step 3 compares |S X1| f^L with the threshold, and
    // step 4 multiplies by S^T D^2, whose entries +-2^-L are exact too.  K slot i of lane half h in
    // K step s is wave member 16 s + 8 (i / 4) + 4 h + i % 4: for eight members that is member
    // 4 h + i % 4 of reference 2 s + i / 4, for sixteen member 8 (i / 4) + 4 h + i % 4 of reference
    // s.  Lane (n, h) holds S[m][col] and (S^T D^2)[m][col] of its reference for those columns (the
    // butterflies run on unit vectors with factors 1 and 1/2), in the step-2 A and step-4 B
    // operands.
    constexpr int NC = KM / 2;  // columns of its reference's blocks a lane holds
    half8_t hfrag[2], gfrag[2];
    // f^L of wave member 8 j + 4 h + i (the rows this lane holds after step 2), from the lane of
    // that member.
    float row_scale[4][4];
    {
        const HaarPlan plan = haar_plan(m, k_use_n, grid, gf);
        float hs[NC], hg[NC];
#pragma unroll
        for (int k = 0; k < NC; ++k) hs[k] = hg[k] = m == 8 * (k >> 2) + 4 * h + (k & 3) ? 1.0f : 0.0f;
        haar_columns<1>(hs, plan.levels > 0 && plan.act[0], plan.low[0], 1.0f);
        haar_columns<2>(hs, plan.levels > 1 && plan.act[1], plan.low[1], 1.0f);
        haar_columns<4>(hs, plan.levels > 2 && plan.act[2], plan.low[2], 1.0f);
        if (KM == 16) haar_columns<8>(hs, plan.levels > 3 && plan.act[3], plan.low[3], 1.0f);
        if (KM == 16) haar_columns<8>(hg, plan.levels > 3 && plan.act[3], plan.low[3], 0.5f);
        haar_columns<4>(hg, plan.levels > 2 && plan.act[2], plan.low[2], 0.5f);
        haar_columns<2>(hg, plan.levels > 1 && plan.act[1], plan.low[1], 0.5f);
        haar_columns<1>(hg, plan.levels > 0 && plan.act[0], plan.low[0], 0.5f);
#pragma unroll
        for (int s = 0; s < 2; ++s)
#pragma unroll
            for (int i = 0; i < 8; ++i) {
                const bool mine = KM == 16 ? s == jn : (s == (jn >> 1) && (i >> 2) == (jn & 1));
                const int k = KM == 16 ? i : (i & 3);
                hfrag[s][i] = (_Float16)(mine ? hs[k] : 0.0f);
                gfrag[s][i] = (_Float16)(mine ? hg[k] : 0.0f);
            }
        const int L = (plan.levels > 0 && plan.act[0]) + (plan.levels > 1 && plan.act[1]) +
                      (plan.levels > 2 && plan.act[2]) + (plan.levels > 3 && plan.act[3]);
        const float own = ((L & 1) ? 0.70710678118654752440f : 1.0f) * __builtin_ldexpf(1.0f, -(L >> 1));
#pragma unroll
        for (int j = 0; j < 4; ++j)
#pragma unroll
            for (int i = 0; i < 4; ++i) row_scale[j][i] = __shfl(own, 8 * j + 4 * h + i) * inv_s;
    }

#pragma unroll
    for (int c = 0; c < C; ++c) {
        // One channel at a time: overlapping the two would hold both channels' tiles at once.
        if (c > 0) __builtin_amdgcn_sched_barrier(0);
        // ---- 1. 2-D DCT: X1 tile mt holds coefficients 32 mt + n of members 8 j + 4 h + i.
        float16_t x1[2], x2[2];
#pragma unroll
        for (int mt = 0; mt < 2; ++mt) {
            x1[mt] = float16_t{};
#pragma unroll
            for (int s = 0; s < 4; ++s)
                x1[mt] = __builtin_amdgcn_mfma_f32_32x32x16_f16(
                    pa[c][s], __builtin_bit_cast(half8_t, s_frag[2 * s + mt][lane]), x1[mt], 0, 0, 0);
        }
        // ---- 2. group Haar (S), on X1 split into hi + lo: the members' large DC terms would
        // otherwise leave their f16 rounding in the small differences the Haar forms.
#pragma unroll
        for (int mt = 0; mt < 2; ++mt) {
            x2[mt] = float16_t{};
#pragma unroll
            for (int s = 0; s < 2; ++s) {
                half8_t hi, lo;
                split_frag(x1[mt], 8 * s, hi, lo);
                x2[mt] = __builtin_amdgcn_mfma_f32_32x32x16_f16(hfrag[s], hi, x2[mt], 0, 0, 0);
                x2[mt] = __builtin_amdgcn_mfma_f32_32x32x16_f16(hfrag[s], lo, x2[mt], 0, 0, 0);
            }
        }
        // ---- 3. hard threshold (the group DC always kept) and the retained variance of each
        // reference, from the first channel.  Groups hold 1, 2, 4, 8 or 16 members and the Haar
        // pairs members inside the group, so the members past k_use (zeroed here like any dropped
        // coefficient) never reach the group's; they only leave the variance sum.
        float y[2][16];
        float kept[4][4];  // [reference j][register i]: retained variance over both tiles
#pragma unroll
        for (int mt = 0; mt < 2; ++mt) {
            const float thr = s_thr[c][32 * mt + n], cvs = s_cvs[c][32 * mt + n];
            const float thr_dc = mt == 0 && lane == 0 ? -INFINITY : thr;  // coefficient 0 of member 0
#pragma unroll
            for (int j = 0; j < 4; ++j)
#pragma unroll
                for (int i = 0; i < 4; ++i) {
                    const float v = x2[mt][4 * j + i];
                    // Row 8 j (i = 0, lane half 0) is member 0 of a reference for every j with eight
                    // members, for even j with sixteen.
                    const bool first_member = i == 0 && (KM == 8 || (j & 1) == 0);
                    const bool keep = fabsf(v) * row_scale[j][i] >= (first_member ? thr_dc : thr);
                    y[mt][4 * j + i] = keep ? v : 0.0f;
                    const float add = keep ? cvs : 0.0f;
                    kept[j][i] = mt == 0 ? add : kept[j][i] + add;
                }
        }
        if (c == 0) {
            // Reference R holds register blocks j = R (KM / 8) + jj, rows 8 jj + 4 h + i of its group.
#pragma unroll
            for (int R = 0; R < NREF; ++R) {
                float part = 0.0f;
#pragma unroll
                for (int jj = 0; jj < KM / 8; ++jj)
#pragma unroll
                    for (int i = 0; i < 4; ++i)
                        part += 8 * jj + 4 * h + i < k_use[R] ? kept[R * (KM / 8) + jj][i] : 0.0f;
#pragma unroll
                for (int o = 1; o < 64; o <<= 1) part += __shfl_xor(part, o);
                const float weight = (isnan(part) || isinf(part)) ? 0.0f : 1.0f / fmaxf(part, 1e-12f);
                if (lane == 0 && first + R < refs) group_weight[first + R] = weight;
            }
        }
        // ---- 4. inverse Haar (S^T D^2, on hi + lo): X3 tile mt holds coefficients 32 mt + 8 j + 4 h + i of member n.
        float16_t x3[2];
#pragma unroll
        for (int mt = 0; mt < 2; ++mt) {
            x3[mt] = float16_t{};
#pragma unroll
            for (int s = 0; s < 2; ++s) {
                half8_t hi, lo;
                split_frag(y[mt], 8 * s, hi, lo);
                x3[mt] = __builtin_amdgcn_mfma_f32_32x32x16_f16(hi, gfrag[s], x3[mt], 0, 0, 0);
                x3[mt] = __builtin_amdgcn_mfma_f32_32x32x16_f16(lo, gfrag[s], x3[mt], 0, 0, 0);
            }
        }
        // ---- 5. inverse DCT: X4 tile pt holds pixel (4 pt + j, 4 h + i) of member n.
        float16_t x4[2];
#pragma unroll
        for (int pt = 0; pt < 2; ++pt) {
            x4[pt] = float16_t{};
#pragma unroll
            for (int s = 0; s < 4; ++s)
                x4[pt] = __builtin_amdgcn_mfma_f32_32x32x16_f16(
                    __builtin_bit_cast(half8_t, s_frag[8 + 2 * s + pt][lane]), frag_from(x3[s >> 1], 8 * (s & 1)), x4[pt],
                    0, 0, 0);
        }
        if (live && in_group) {
            _Float16 *st = reinterpret_cast<_Float16 *>(tiles.stage) + (((size_t)ref * KM + m) * C + c) * 64 + 4 * h;
#pragma unroll
            for (int pt = 0; pt < 2; ++pt)
#pragma unroll
                for (int j = 0; j < 4; ++j) {
                    half4_t o;
#pragma unroll
                    for (int i = 0; i < 4; ++i) o[i] = (_Float16)(x4[pt][4 * j + i] * inv_s);
                    *reinterpret_cast<half4_t *>(st + (4 * pt + j) * 8) = o;
                }
        }
    }
}

// One group of 32 members per wave (four references of eight, two of sixteen).  The workgroup
// first enters each member in the list of
// every tile its patch overlaps (lane (n, h) holds claims 2 h, 2 h + 1 of member n): its claims are
// counted per tile in an LDS hash table (rank = the bucket's count before the claim) and one
// atomic per distinct tile reserves the entries.  Those atomics, the sample loads and the table
// copy are then in flight together, and the reserved positions are only read after the filter,
// when the list entries are written (a full list sends the (member, tile) pair to the overflow
// list).
// CODES: the members' samples come from the ring's u16 code mirror (half the bytes of the values),
// for codes of at most 11 bits, which enter the transforms exactly either way.
template <int C, int KM, bool CODES>
__global__ __launch_bounds__(kWaves * 64) void k_collab_filter_mixed(avd_collab_pass p, const float *__restrict__ values,
                                                                     const uint16_t *__restrict__ codes,
                                                                     const uint32_t *__restrict__ nslots,
                                                                     const uint32_t *__restrict__ members,
                                                                     float *__restrict__ group_weight, TileBins tiles)
{
    constexpr int kBuckets = kWaves * 256;
    constexpr int NREF = 32 / KM, KSHIFT = KM == 16 ? 4 : 3;  // references per wave
    __shared__ uint4 s_frag[16][64];
    __shared__ float s_cvs[C][64], s_thr[C][64];
    __shared__ uint32_t s_key[kBuckets], s_cnt[kBuckets];
    const int lane = threadIdx.x & 63, wave = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
    const int n = lane & 31, h = lane >> 5, m = n & (KM - 1), jn = n >> KSHIFT;
    const int W = p.width, H = p.height;
    const int refs_x = (W - 8 + p.step - 1) / p.step + 1, refs_y = (H - 8 + p.step - 1) / p.step + 1;
    const int refs = refs_x * refs_y;
    const int first = (xcd_block() * kWaves + wave) * NREF;  // the wave's first reference
    for (int i = threadIdx.x; i < kBuckets; i += kWaves * 64) {
        s_key[i] = ~0u;
        s_cnt[i] = 0;
    }
    int k_use[NREF];
#pragma unroll
    for (int j = 0; j < NREF; ++j) k_use[j] = (int)(members[(size_t)min(first + j, refs - 1) * (KM + 1) + KM] & 0xFF);
    const uint32_t ref = (uint32_t)min(first + jn, refs - 1);
    const bool live = first + jn < refs;
    const uint32_t *mw = members + (size_t)ref * (KM + 1);
    const uint32_t gword = mw[KM], pos = mw[m];
    const bool in_group = m < (int)(gword & 0xFF);
    // The ring slot of trajectory index t (0: the centre frame) from the 2 r neighbour slots.
    const uint32_t t = pos >> 26;
    uint32_t slot = p.centre_slot;
    for (int k = 0; k < 2 * (int)p.temporal_radius; ++k) slot = t == (uint32_t)(k + 1) ? nslots[k] : slot;
    const int mx = (int)(pos & 0x1FFF), my = (int)((pos >> 13) & 0x1FFF);
    __syncthreads();

    // ---- claims: hash, then one reserving atomic per distinct tile.
    const int tx0 = mx >> kTileShift, ty0 = my >> kTileShift;
    const int nx = ((mx + 7) >> kTileShift) - tx0 + 1, ny = ((my + 7) >> kTileShift) - ty0 + 1;
    uint32_t tile[2] = {~0u, ~0u}, bucket[2] = {0, 0}, rank[2] = {0, 0};
#pragma unroll
    for (int k = 0; k < 2; ++k) {
        const int q = 2 * h + k;
        if (live && in_group && q < nx * ny) {
            tile[k] = slot * tiles.per_slot + (ty0 + q / nx) * tiles.tiles_x + tx0 + q % nx;
            uint32_t bk = (tile[k] * 2654435761u) >> 22;
            for (;;) {  // at most 512 keys in 1024 buckets: a free or equal bucket exists
                const uint32_t old = atomicCAS(&s_key[bk], ~0u, tile[k]);
                if (old == ~0u || old == tile[k]) break;
                bk = (bk + 1) & (kBuckets - 1);
            }
            bucket[k] = bk;
            rank[k] = atomicAdd(&s_cnt[bk], 1u);
        }
    }
    __syncthreads();
    uint32_t base[kBuckets / (kWaves * 64)];  // reserved positions of this thread's buckets
#pragma unroll
    for (int k = 0; k < kBuckets / (kWaves * 64); ++k) {
        const int i = (int)threadIdx.x + k * kWaves * 64;
        base[k] = s_key[i] != ~0u ? atomicAdd(&tiles.count[s_key[i]], s_cnt[i]) : 0u;
    }

    // ---- samples: rows 2 s + h of the member's patch, 8 C values (or codes) each; the tables.
    const size_t at = (size_t)slot * W * H * C + ((size_t)(my + h) * W + mx) * C;
    float4u_t raw[4][2 * C];
    uint32_t cw[4][4 * C];
    if constexpr (CODES) {
#pragma unroll
        for (int s = 0; s < 4; ++s) load_row<C>(codes, at + (size_t)2 * s * W * C, cw[s]);
    } else {
        const float *src = values + at;
#pragma unroll
        for (int s = 0; s < 4; ++s)
#pragma unroll
            for (int k = 0; k < 2 * C; ++k)
                raw[s][k] = *reinterpret_cast<const float4u_t *>(src + (size_t)2 * s * W * C + 4 * k);
    }
    for (int i = threadIdx.x; i < 16 * 64; i += kWaves * 64) s_frag[i >> 6][i & 63] = g_mixed_frags[i >> 6][i & 63];
    if (threadIdx.x < 64 * C) {
        const int c = threadIdx.x >> 6, v = (threadIdx.x >> 3) & 7, u = threadIdx.x & 7;
        const float cvs = (p.sigma[c] * p.sigma[c]) * (p.profile[u] * p.profile[v]);
        s_cvs[c][v * 8 + u] = cvs;
        s_thr[c][v * 8 + u] = p.lambda[c] * sqrtf(cvs);
    }
    __syncthreads();

    // Samples of up to 11-bit codes enter exactly, as code 2^-bits = v s (s = (2^bits - 1) / 2^bits);
    // the threshold and the output take 1 / s back in f32.  Deeper samples round to f16.
    const bool exact = p.code_bits >= 1 && p.code_bits <= 11;
    const float maxv = exact ? (float)((1u << p.code_bits) - 1u) : 1.0f;
    const float pow2 = exact ? __builtin_ldexpf(1.0f, -(int)p.code_bits) : 1.0f;
    const float inv_s = exact ? (float)(1u << p.code_bits) / maxv : 1.0f;
    half8_t pa[C][4];
#pragma unroll
    for (int c = 0; c < C; ++c)
#pragma unroll
        for (int s = 0; s < 4; ++s)
#pragma unroll
            for (int i = 0; i < 8; ++i) {
                if constexpr (CODES) {
                    const int j = i * C + c;
                    const uint16_t code = (uint16_t)(cw[s][j >> 1] >> (16 * (j & 1)));
                    pa[c][s][i] = (_Float16)code * (_Float16)pow2;
                } else {
                    const float v = raw[s][(i * C + c) >> 2][(i * C + c) & 3];
                    pa[c][s][i] = (_Float16)(exact ? rintf(v * maxv) * pow2 : v);
                }
            }
    mixed_group<C, KM>(p, s_frag, s_cvs, s_thr, pa, inv_s, first, refs, k_use, gword, lane, group_weight, tiles);

    // ---- list entries, at the reserved positions (s_cnt now holds the bases).
#pragma unroll
    for (int k = 0; k < kBuckets / (kWaves * 64); ++k) s_cnt[threadIdx.x + k * kWaves * 64] = base[k];
    __syncthreads();
#pragma unroll
    for (int k = 0; k < 2; ++k) {
        if (tile[k] == ~0u) continue;
        const uint32_t e = s_cnt[bucket[k]] + rank[k];
        const uint32_t id = (ref << KSHIFT) | (uint32_t)m;
        if (e < tiles.cap) {
            tiles.list[(size_t)tile[k] * tiles.cap + e] = id;
        } else {
            const uint32_t o = atomicAdd(tiles.overflow_count, 1u);
            tiles.overflow[2 * (size_t)o] = id;
            tiles.overflow[2 * (size_t)o + 1] = tile[k];
        }
    }
}

// One workgroup per (slot, 32x32 tile): the listed members' positions and group weights are
// fetched 256 at a time into LDS, then each wave walks members with lane (r, q) on pixel (r, q)
// of the patch, adding its staged sum and its weight (recomputed from the group weight in the
// filter's own arithmetic) to the tile in LDS.  The tile then adds to the accumulators once, as
// the only writer of those pixels this pass (overflowing pairs are added by k_tile_overflow).
// MIXED: the stage holds f16 samples, weighted and converted to fixed point here.
template <int C, bool MIXED, int KM>
__global__ __launch_bounds__(256) void k_tile_gather(avd_collab_pass p, const uint32_t *__restrict__ members,
                                                     const float *__restrict__ group_weight, TileBins tiles,
                                                     unsigned long long *__restrict__ aggregate,
                                                     int32_t *__restrict__ accum_second)
{
    constexpr int T = 1 << kTileShift;
    // Rows padded to 40 cells: the eight rows of a patch then fall on distinct banks (two passes
    // per 64-bit add, the minimum, instead of eight).
    constexpr int kRow = T + 8;
    // Past the tile's cells, one trash cell per lane: a member pixel outside the tile adds there, so
    // the walk has no branch (a branch let the compiler sink a member's staged load into it, where
    // the lane then waited on it alone).  The trash is never read.
    constexpr int kTrash = T * kRow;
    __shared__ unsigned long long s_agg[T * kRow + 64];
    __shared__ int32_t s_second[C == 2 ? T * kRow + 64 : 1];
    __shared__ uint32_t s_id[256], s_pos[256];
    __shared__ float s_sw[256];
    const uint32_t tile = blockIdx.x;
    const uint32_t slot = tile / tiles.per_slot, in_slot = tile - slot * tiles.per_slot;
    const int ty = (int)(in_slot / tiles.tiles_x), tx = (int)(in_slot - ty * tiles.tiles_x);
    const int x0 = tx * T, y0 = ty * T;
    for (int i = threadIdx.x; i < T * kRow; i += 256) {
        s_agg[i] = 0;
        if (C == 2) s_second[i] = 0;
    }
    const uint32_t *list = tiles.list + (size_t)tile * tiles.cap;
    // The first entries are read with the count (a list always holds cap >= 256 entries, so the
    // read stays in bounds; entries past the count are ignored).
    const uint32_t id0 = list[threadIdx.x];
    const uint32_t n = min(tiles.count[tile], tiles.cap);
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int r = lane >> 3, q = lane & 7;
    const float kr_kq = p.kaiser[r] * p.kaiser[q];
    const float scale = p.accum_scale;
    const float kk_scale = kr_kq * scale;  // MIXED: the window and the fixed-point scale in one factor
    for (uint32_t base = 0; base < n; base += 256) {
        __syncthreads();
        const uint32_t k = base + threadIdx.x;
        if (k < n) {
            const uint32_t id = base == 0 ? id0 : list[k];
            s_id[threadIdx.x] = id;
            constexpr int KSHIFT = KM == 16 ? 4 : 3;
            s_pos[threadIdx.x] = members[(size_t)(id >> KSHIFT) * (KM + 1) + (id & (KM - 1))];
            s_sw[threadIdx.x] = group_weight[id >> KSHIFT] * p.weight_scale;
        }
        __syncthreads();
        const uint32_t batch = min(256u, n - base);
        // kPer members per wave per step: every staged load is issued before the first add (the
        // loads are unconditional, a lane past the batch reading the first member's row again, so
        // none waits behind a branch).
        constexpr int kPer = 8;
        for (uint32_t e0 = wave; e0 < batch; e0 += 4 * kPer) {
            int32_t a[kPer], a2[kPer], wq[kPer];
            int cell[kPer];
            float sw[kPer];
            using Staged = typename std::conditional<MIXED, _Float16, int32_t>::type;
            Staged v0[kPer], v1[kPer];
#pragma unroll
            for (int u = 0; u < kPer; ++u) {
                const uint32_t e = e0 + 4 * u;
                const uint32_t ec = e < batch ? e : e0;
                const uint32_t id = s_id[ec], pos = s_pos[ec];
                sw[u] = s_sw[ec];
                const int lx = (int)(pos & 0x1FFF) + q - x0, ly = (int)((pos >> 13) & 0x1FFF) + r - y0;
                const bool in = e < batch && (unsigned)lx < (unsigned)T && (unsigned)ly < (unsigned)T;
                cell[u] = in ? ly * kRow + lx : kTrash + lane;
                const Staged *st = reinterpret_cast<const Staged *>(tiles.stage) + (size_t)id * C * 64 + lane;
                v0[u] = st[0];
                if (C == 2) v1[u] = st[64];
            }
#pragma unroll
            for (int u = 0; u < kPer; ++u) {
                if (MIXED) {
                    // The FP32 form's weight and clamp, to_fixed(v w) = round(clamp(v w scale,
                    // +-5 scale)) and the weight term round(min(w scale, scale) 5), w = sw kaiser[r]
                    // kaiser[q], rounded half away from zero by adding +-0.5 and truncating (which
                    // differs from f32::round only for values within an ulp below a half, one unit in
                    // ~2^20; this form is not bit-exact anyway).
                    const float ws = sw[u] * kk_scale;
                    const float lim = 5.0f * scale;
                    const float y = __builtin_amdgcn_fmed3f((float)v0[u] * ws, -lim, lim);
                    a[u] = (int32_t)(y + __builtin_copysignf(0.5f, y));
                    if (C == 2) {
                        const float y2 = __builtin_amdgcn_fmed3f((float)v1[u] * ws, -lim, lim);
                        a2[u] = (int32_t)(y2 + __builtin_copysignf(0.5f, y2));
                    }
                    wq[u] = (int32_t)(fminf(ws, scale) * 5.0f + 0.5f);
                } else {
                    a[u] = (int32_t)v0[u];
                    if (C == 2) a2[u] = (int32_t)v1[u];
                    wq[u] = to_fixed_weight(sw[u] * kr_kq, scale);
                }
            }
#pragma unroll
            for (int u = 0; u < kPer; ++u) {
                atomicAdd(&s_agg[cell[u]], pack_aggregate(a[u], wq[u]));
                if (C == 2) atomicAdd(&s_second[cell[u]], a2[u]);
            }
        }
    }
    __syncthreads();
    const size_t px = (size_t)p.width * p.height;
    for (int i = threadIdx.x; i < T * T; i += 256) {
        const int lx = i & (T - 1), ly = i >> kTileShift;
        const int x = x0 + lx, y = y0 + ly;
        if (x >= (int)p.width || y >= (int)p.height) continue;
        // The only writer of these pixels this pass: the adds need no result, so they are atomics
        // (the L2 does the read-modify-write and the workgroup does not wait for it).
        const size_t pix = (size_t)slot * px + (size_t)y * p.width + x;
        const unsigned long long v = s_agg[ly * kRow + lx];
        if (v) atomicAdd(&aggregate[pix], v);
        if (C == 2 && s_second[ly * kRow + lx]) atomicAdd(&accum_second[pix], s_second[ly * kRow + lx]);
    }
}

// The (member, tile) pairs that did not fit a tile's list: thread (pair, row) adds its row's pixels
// inside the pair's tile with global atomics.
template <int C, bool MIXED, int KM>
__global__ __launch_bounds__(256) void k_tile_overflow(avd_collab_pass p, const uint32_t *__restrict__ members,
                                                       const float *__restrict__ group_weight, TileBins tiles,
                                                       unsigned long long *__restrict__ aggregate,
                                                       int32_t *__restrict__ accum_second)
{
    constexpr int T = 1 << kTileShift;
    const uint32_t pairs = *tiles.overflow_count;
    const size_t px = (size_t)p.width * p.height;
    for (uint32_t k = blockIdx.x * 256 + threadIdx.x; k < pairs * 8; k += gridDim.x * 256) {
        const uint32_t pair = k >> 3;
        const int r = (int)(k & 7);
        const uint32_t id = tiles.overflow[2 * (size_t)pair], tile = tiles.overflow[2 * (size_t)pair + 1];
        const uint32_t slot = tile / tiles.per_slot, in_slot = tile - slot * tiles.per_slot;
        const int ty = (int)(in_slot / tiles.tiles_x), tx = (int)(in_slot - ty * tiles.tiles_x);
        const int x0 = tx * T, y0 = ty * T;
        member_row<C, MIXED, KM>(p, members, group_weight, tiles, id, r, x0, y0, T,
                      [&](int i, unsigned long long v, int32_t v2) {
                          const size_t pix = (size_t)slot * px + (size_t)(y0 + (i >> kTileShift)) * p.width +
                                             x0 + (i & (T - 1));
                          atomicAdd(&aggregate[pix], v);
                          if (C == 2) atomicAdd(&accum_second[pix], v2);
                      });
    }
}

// collab_normalise (aggregate.rs:252-290): out = accum * 5 / wsum per channel (accum when the
// weight is zero), then the region is zeroed for its next use.
template <int C>
__global__ __launch_bounds__(256) void k_normalise(unsigned long long *__restrict__ aggregate,
                                                   int32_t *__restrict__ accum_second, float *__restrict__ out,
                                                   uint32_t pixels)
{
    for (uint32_t i = blockIdx.x * 256 + threadIdx.x; i < pixels; i += gridDim.x * 256) {
        int32_t a, w;
        unpack_aggregate(aggregate[i], a, w);
        out[(size_t)i * C] = (w != 0) ? ((float)a * 5.0f) / (float)w : (float)a;
        aggregate[i] = 0;
        if (C == 2) {
            const float a2 = (float)accum_second[i];
            out[(size_t)i * C + 1] = (w != 0) ? (a2 * 5.0f) / (float)w : a2;
            accum_second[i] = 0;
        }
    }
}

double dct8_basis(int row, int col)
{
    const double rs = row == 0 ? 1.0 / sqrt(8.0) : 0.5;
    return rs * cos(M_PI * (2.0 * col + 1.0) * row / 16.0);
}

// The mixed form's DCT tables in operand order (see g_mixed_frags): products of two DCT basis
// values, rounded once to f16.
void mixed_frags(uint4 (&out)[16][64])
{
    for (int lane = 0; lane < 64; ++lane) {
        const int h = lane >> 5, n = lane & 31;
        for (int f = 0; f < 16; ++f) {
            _Float16 v[8];
            for (int i = 0; i < 8; ++i) {
                int coef, pixel;
                if (f < 8) {
                    coef = 32 * (f & 1) + n;
                    pixel = 16 * (f >> 1) + 8 * h + i;
                } else {
                    pixel = 32 * ((f - 8) & 1) + n;
                    coef = 16 * ((f - 8) >> 1) + 8 * (i >> 2) + 4 * h + (i & 3);
                }
                v[i] = (_Float16)(dct8_basis(coef >> 3, pixel >> 3) * dct8_basis(coef & 7, pixel & 7));
            }
            std::memcpy(&out[f][lane], v, sizeof(v));
        }
    }
}

hipError_t upload_taps(hipStream_t s)
{
    // Once per device (the constants live in each device's copy of the code object), under a lock:
    // engines on several host threads may reach here together.
    constexpr int kMaxDevices = 64;
    static std::mutex lock;
    static bool done[kMaxDevices] = {};
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return e;
    if (dev < 0 || dev >= kMaxDevices) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> guard(lock);
    if (done[dev]) return hipSuccess;
    DctTaps t;
    for (int m = 0; m < 4; ++m)
        for (int k = 0; k < 4; ++k) {
            t.e[m * 4 + k] = (float)dct8_basis(2 * m, k);
            t.o[m * 4 + k] = (float)dct8_basis(2 * m + 1, k);
        }
    static uint4 frags[16][64];
    mixed_frags(frags);
    e = hipMemcpyToSymbolAsync(HIP_SYMBOL(c_taps), &t, sizeof(t), 0, hipMemcpyHostToDevice, s);
    if (e == hipSuccess)
        e = hipMemcpyToSymbolAsync(HIP_SYMBOL(g_mixed_frags), frags, sizeof(frags), 0, hipMemcpyHostToDevice, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (e == hipSuccess) done[dev] = true;
    return e;
}

inline int ref_count(const avd_collab_pass &p)
{
    const int rx = ((int)p.width - 8 + (int)p.step - 1) / (int)p.step + 1;
    const int ry = ((int)p.height - 8 + (int)p.step - 1) / (int)p.step + 1;
    return rx * ry;
}

bool pass_ok(const avd_collab_pass &p)
{
    return p.width >= 8 && p.height >= 8 && p.width < 8192 && p.height < 8192 && (p.channels == 1 || p.channels == 2) &&
           p.step >= 1 && p.step <= 8 && p.spatial_radius >= 1 && p.spatial_radius <= kMaxS &&
           (p.group_size == 8 || p.group_size == 16) &&
           p.temporal_radius <= 8 && p.refine >= 1 && p.refine <= 4 && p.blk_step >= 1 && p.blksize >= 1 &&
           (p.blksize + p.blk_step - 1) / p.blk_step <= 4;
}

}  // namespace

size_t centre_scratch_bytes(uint32_t W, uint32_t H, uint32_t C)
{
    const size_t px = (size_t)W * H;
    const size_t plane = (px * C * 2 + 255) & ~(size_t)255, odd = C == 1 ? (px * 2 + 255) & ~(size_t)255 : 0;
    // the redo list holds a count and at most one entry per reference (refs < px)
    return plane + odd + px * 4 + (px + 1) * 4;
}

CentreScratch centre_scratch_at(void *base, uint32_t W, uint32_t H, uint32_t C)
{
    const size_t px = (size_t)W * H;
    const size_t plane = (px * C * 2 + 255) & ~(size_t)255, odd = C == 1 ? (px * 2 + 255) & ~(size_t)255 : 0;
    char *b = static_cast<char *>(base);
    return CentreScratch{reinterpret_cast<uint16_t *>(b), C == 1 ? reinterpret_cast<uint16_t *>(b + plane) : nullptr,
                         reinterpret_cast<uint32_t *>(b + plane + odd),
                         reinterpret_cast<uint32_t *>(b + plane + odd + px * 4)};
}

bool centre_on_matrix_cores(const avd_collab_pass &p)
{
    return p.code_bits >= 1 && p.code_bits <= 10 && p.step + 2 * p.spatial_radius <= 31;
}

namespace {

// Workgroups of the exact centre search over the redo list (a grid-stride loop over the list's
// count; a few percent of a 4K frame's references when a lane's short list runs out).
constexpr unsigned kRedoBlocks = 1024;

template <int KM>
hipError_t launch_centre_km(const avd_collab_pass &p, const uint16_t *codes, uint32_t *members, hipStream_t s,
                            const CentreScratch *centre)
{
    const dim3 blk(kWaves * 64);
    if (centre && centre_on_matrix_cores(p)) {
        if (!centre->plane || !centre->energy || !centre->redo || (p.channels == 1 && !centre->plane_odd))
            return hipErrorInvalidValue;
        const int W = (int)p.width, H = (int)p.height, st = (int)p.step;
        const uint16_t *frame = codes + (size_t)p.centre_slot * W * H * p.channels;
        const dim3 pg(div_up(W, 256), div_up(H, kPrepRows));
        const int refs_x = (W - 8 + st - 1) / st + 1, refs_y = (H - 8 + st - 1) / st + 1;
        const unsigned gm = div_up(((refs_x + 1) / 2) * ((refs_y + 3) / 4), kWaves);
        const hipError_t e = launch_fill_u32(centre->redo, 0, 1, s);
        if (e != hipSuccess) return e;
        if (p.channels == 1) {
            k_centre_prep<1><<<pg, 256, 0, s>>>(frame, W, H, centre->plane, centre->plane_odd, centre->energy);
            k_centre_mfma<1, KM><<<gm, blk, 0, s>>>(p, codes, centre->plane,
                                                    (uint32_t)(centre->plane_odd - centre->plane), centre->energy,
                                                    members, centre->redo, nullptr);
            k_centre_redo<1, KM><<<kRedoBlocks, blk, 0, s>>>(p, codes, centre->redo, members);
        } else {
            k_centre_prep<2><<<pg, 256, 0, s>>>(frame, W, H, centre->plane, nullptr, centre->energy);
            k_centre_mfma<2, KM><<<gm, blk, 0, s>>>(p, codes, centre->plane, 0u, centre->energy, members,
                                                    centre->redo, nullptr);
            k_centre_redo<2, KM><<<kRedoBlocks, blk, 0, s>>>(p, codes, centre->redo, members);
        }
        return hipGetLastError();
    }
    // 32-bit keys hold (SSD + 1) << 5: SSD < 64 * C * 1023^2 + 1 < 2^27 for codes of at most 10
    // bits, and a lane's run of candidates stays below 32 while rw <= 32.
    const bool narrow = p.code_bits >= 1 && p.code_bits <= 10 && p.spatial_radius <= 15;
    const unsigned g = div_up(ref_count(p), kWaves);
    if (p.channels == 1) {
        if (narrow) k_centre_codes<1, true, KM><<<g, blk, 0, s>>>(p, codes, members, nullptr);
        else k_centre_codes<1, false, KM><<<g, blk, 0, s>>>(p, codes, members, nullptr);
    } else {
        if (narrow) k_centre_codes<2, true, KM><<<g, blk, 0, s>>>(p, codes, members, nullptr);
        else k_centre_codes<2, false, KM><<<g, blk, 0, s>>>(p, codes, members, nullptr);
    }
    return hipGetLastError();
}

// One trajectory launch: NR references per wave (4, or fewer where the rectangle tables outgrow
// the LDS budget: 2 for eight members, 2 then 1 for sixteen), SHORT for frames shorter than K + 7
// rows or rings past a buffer load's offset.
template <int C, int K, int KM, int GF>
hipError_t launch_trajectory_ck(const avd_collab_pass &p, const uint16_t *codes, const uint32_t *nslots,
                                const int32_t *mv, const float *conf, uint32_t *members, hipStream_t s,
                                unsigned long long *scored_count, int nr, bool short_frame, size_t lds)
{
    const dim3 blk(kWaves * 64);
    const unsigned g = div_up(ref_count(p), kWaves * nr);
#define AVD_TRAJ_LAUNCH(NRn, SHORTn) \
    k_trajectory_codes<C, K, NRn, SHORTn, KM, GF><<<g, blk, lds, s>>>(p, codes, nslots, mv, conf, members, scored_count)
    if constexpr (KM == 8) {
        if (short_frame) AVD_TRAJ_LAUNCH(2, true);
        else if (nr == 4) AVD_TRAJ_LAUNCH(4, false);
        else AVD_TRAJ_LAUNCH(2, false);
    } else {
        if (short_frame) AVD_TRAJ_LAUNCH(1, true);
        else if (nr == 4) AVD_TRAJ_LAUNCH(4, false);
        else if (nr == 2) AVD_TRAJ_LAUNCH(2, false);
        else AVD_TRAJ_LAUNCH(1, false);
    }
#undef AVD_TRAJ_LAUNCH
    return hipGetLastError();
}

template <int KM, int GF>
hipError_t launch_trajectory_gf(const avd_collab_pass &p, const uint16_t *codes, const uint32_t *nslots,
                                const int32_t *mv, const float *conf, uint32_t *members, hipStream_t s,
                                unsigned long long *scored_count)
{
    // References per wave: the most whose LDS (static arrays and rectangle tables) fits the budget.
    const int C = (int)p.channels;
    auto lds_of = [&](int nr) {
        return (C == 1 ? trajectory_static_lds(1, nr, KM, GF) : trajectory_static_lds(2, nr, KM, GF)) +
               trajectory_dynamic_lds(p, nr);
    };
    const int nr_lo = KM == 8 ? 2 : 1;
    int nr = lds_of(4) <= kLdsBudget ? 4 : (KM == 16 && lds_of(2) <= kLdsBudget ? 2 : nr_lo);
    // Frames shorter than K + 7 rows, and rings past 2^31 bytes (beyond a buffer load's offset),
    // take the clamping slide (at two references per wave for eight members, one for sixteen).
    const bool short_frame = (int)p.height < 2 * (int)p.refine + 8 ||
                             (uint64_t)p.width * p.height * p.channels * p.slots * 2 >= (1ull << 31) ||
                             (C == 1 && (p.width & 1u));
    if (short_frame) nr = nr_lo;
    if (lds_of(nr) > 160 * 1024) return hipErrorInvalidValue;  // past the LDS of a workgroup
    const size_t lds = trajectory_dynamic_lds(p, nr);
#define AVD_TRAJ(Cn, Kn) \
    launch_trajectory_ck<Cn, Kn, KM, GF>(p, codes, nslots, mv, conf, members, s, scored_count, nr, short_frame, lds)
    switch (p.refine * 2 + 1 + (p.channels == 2 ? 100 : 0)) {
    case 3: return AVD_TRAJ(1, 3);
    case 5: return AVD_TRAJ(1, 5);
    case 7: return AVD_TRAJ(1, 7);
    case 9: return AVD_TRAJ(1, 9);
    case 103: return AVD_TRAJ(2, 3);
    case 105: return AVD_TRAJ(2, 5);
    case 107: return AVD_TRAJ(2, 7);
    default: return AVD_TRAJ(2, 9);
    }
#undef AVD_TRAJ
}

// Volume grids of 2 frames at temporal radius 1 and 4 above it (collab/mod.rs:45-51).
template <int KM>
hipError_t launch_trajectory_km(const avd_collab_pass &p, const uint16_t *codes, const uint32_t *nslots,
                                const int32_t *mv, const float *conf, uint32_t *members, hipStream_t s,
                                unsigned long long *scored_count)
{
    return p.temporal_radius == 1 ? launch_trajectory_gf<KM, 2>(p, codes, nslots, mv, conf, members, s, scored_count)
                                  : launch_trajectory_gf<KM, 4>(p, codes, nslots, mv, conf, members, s, scored_count);
}

}  // namespace

hipError_t launch_collab_centre(const avd_collab_pass &p, const uint16_t *codes, uint32_t *members, hipStream_t s,
                                const CentreScratch *centre)
{
    if (!pass_ok(p) || !codes) return hipErrorInvalidValue;
    return p.group_size == 16 ? launch_centre_km<16>(p, codes, members, s, centre)
                              : launch_centre_km<8>(p, codes, members, s, centre);
}

hipError_t launch_collab_trajectory(const avd_collab_pass &p, const uint16_t *codes, const uint32_t *nslots,
                                    const int32_t *mv, const float *conf, uint32_t *members, hipStream_t s,
                                    unsigned long long *scored_count)
{
    if (!pass_ok(p) || !codes) return hipErrorInvalidValue;
    if (p.temporal_radius < 1) return hipSuccess;
    return p.group_size == 16 ? launch_trajectory_km<16>(p, codes, nslots, mv, conf, members, s, scored_count)
                              : launch_trajectory_km<8>(p, codes, nslots, mv, conf, members, s, scored_count);
}

hipError_t launch_collab_search(const avd_collab_pass &p, const uint16_t *codes, const float *values,
                                const uint32_t *nslots, const int32_t *mv, const float *conf, uint32_t *members,
                                hipStream_t s, unsigned long long *scored_count, const CentreScratch *centre)
{
    if (!pass_ok(p)) return hipErrorInvalidValue;
    if (codes) {
        const hipError_t e = launch_collab_centre(p, codes, members, s, centre);
        return e != hipSuccess ? e : launch_collab_trajectory(p, codes, nslots, mv, conf, members, s, scored_count);
    }
    const unsigned g = div_up(ref_count(p), kWaves);
    const dim3 blk(kWaves * 64);
#define AVD_F32_SEARCH(Cn, KMn) \
    k_collab_search<Cn, false, KMn><<<g, blk, 0, s>>>(p, codes, values, nslots, mv, conf, members, scored_count)
    if (p.group_size == 16) {
        if (p.channels == 1) AVD_F32_SEARCH(1, 16);
        else AVD_F32_SEARCH(2, 16);
    } else {
        if (p.channels == 1) AVD_F32_SEARCH(1, 8);
        else AVD_F32_SEARCH(2, 8);
    }
#undef AVD_F32_SEARCH
    return hipGetLastError();
}

hipError_t launch_collab_filter(const avd_collab_pass &p, const float *values, const uint32_t *nslots,
                                const uint32_t *members, unsigned long long *aggregate, int32_t *accum_second,
                                float *group_weight, hipStream_t s)
{
    if (!pass_ok(p) || (p.channels == 2 && !accum_second)) return hipErrorInvalidValue;
    hipError_t e = upload_taps(s);
    if (e != hipSuccess) return e;
    const unsigned g = div_up(ref_count(p), kWaves);
    const TileBins none{};
    const dim3 blk(kWaves * 64);
#define AVD_DIRECT(Cn, KMn)                                                                                 \
    k_collab_filter<Cn, false, KMn><<<g, blk, 0, s>>>(p, values, nslots, members, aggregate, accum_second, \
                                                      group_weight, none)
    if (p.group_size == 16) {
        if (p.channels == 1) AVD_DIRECT(1, 16);
        else AVD_DIRECT(2, 16);
    } else {
        if (p.channels == 1) AVD_DIRECT(1, 8);
        else AVD_DIRECT(2, 8);
    }
#undef AVD_DIRECT
    return hipGetLastError();
}

namespace {

template <int C, int KM>
void binned_launches(const avd_collab_pass &p, const float *values, const uint16_t *codes, const uint32_t *nslots,
                     const uint32_t *members, const TileBins &tiles, unsigned long long *aggregate,
                     int32_t *accum_second, float *group_weight, hipStream_t s, bool mixed, uint32_t refs,
                     uint32_t n_tiles)
{
    const dim3 blk(kWaves * 64);
    if (mixed) {
        const unsigned gm = div_up(refs, kWaves * (32 / KM));  // 32 members per wave
        if (codes && p.code_bits >= 1 && p.code_bits <= 11)
            k_collab_filter_mixed<C, KM, true><<<gm, blk, 0, s>>>(p, values, codes, nslots, members, group_weight,
                                                                  tiles);
        else
            k_collab_filter_mixed<C, KM, false><<<gm, blk, 0, s>>>(p, values, nullptr, nslots, members,
                                                                   group_weight, tiles);
        k_tile_gather<C, true, KM><<<n_tiles, 256, 0, s>>>(p, members, group_weight, tiles, aggregate, accum_second);
        k_tile_overflow<C, true, KM><<<64, 256, 0, s>>>(p, members, group_weight, tiles, aggregate, accum_second);
    } else {
        const unsigned g = div_up(refs, kWaves);
        k_collab_filter<C, true, KM><<<g, blk, 0, s>>>(p, values, nslots, members, aggregate, accum_second,
                                                       group_weight, tiles);
        k_tile_gather<C, false, KM><<<n_tiles, 256, 0, s>>>(p, members, group_weight, tiles, aggregate, accum_second);
        k_tile_overflow<C, false, KM><<<64, 256, 0, s>>>(p, members, group_weight, tiles, aggregate, accum_second);
    }
}

}  // namespace

hipError_t launch_collab_filter_binned(const avd_collab_pass &p, const float *values, const uint32_t *nslots,
                                       const uint32_t *members, const TileBins &tiles,
                                       unsigned long long *aggregate, int32_t *accum_second, float *group_weight,
                                       hipStream_t s, bool mixed, const uint16_t *codes)
{
    if (mixed && p.pool_ratio > 0.0f) return hipErrorInvalidValue;  // the mixed form thresholds hard
    if (!pass_ok(p) || (p.channels == 2 && !accum_second) || !tiles.stage || !tiles.count || !tiles.list ||
        !tiles.overflow || !tiles.overflow_count || tiles.cap < kTileCapMin ||
        tiles.tiles_x != (p.width + (1u << kTileShift) - 1) >> kTileShift ||
        tiles.per_slot != tiles.tiles_x * ((p.height + (1u << kTileShift) - 1) >> kTileShift))
        return hipErrorInvalidValue;
    hipError_t e = upload_taps(s);
    if (e != hipSuccess) return e;
    const uint32_t n_tiles = tiles.per_slot * p.slots;
    e = launch_fill_u32(tiles.count, 0, n_tiles, s);
    if (e == hipSuccess) e = launch_fill_u32(tiles.overflow_count, 0, 1, s);
    if (e != hipSuccess) return e;
    const uint32_t refs = (uint32_t)ref_count(p);
    if (p.group_size == 16) {
        if (p.channels == 1) binned_launches<1, 16>(p, values, codes, nslots, members, tiles, aggregate, accum_second,
                                                    group_weight, s, mixed, refs, n_tiles);
        else binned_launches<2, 16>(p, values, codes, nslots, members, tiles, aggregate, accum_second, group_weight, s,
                                    mixed, refs, n_tiles);
    } else {
        if (p.channels == 1) binned_launches<1, 8>(p, values, codes, nslots, members, tiles, aggregate, accum_second,
                                                   group_weight, s, mixed, refs, n_tiles);
        else binned_launches<2, 8>(p, values, codes, nslots, members, tiles, aggregate, accum_second, group_weight, s,
                                   mixed, refs, n_tiles);
    }
    return hipGetLastError();
}

hipError_t launch_collab_normalise(unsigned long long *aggregate, int32_t *accum_second, float *out, uint32_t pixels,
                                   int C, hipStream_t s)
{
    if (C == 2 && !accum_second) return hipErrorInvalidValue;
    unsigned g = div_up(pixels, 256);
    g = g < 1 ? 1 : (g > 4096 ? 4096 : g);
    if (C == 1) k_normalise<1><<<g, 256, 0, s>>>(aggregate, accum_second, out, pixels);
    else k_normalise<2><<<g, 256, 0, s>>>(aggregate, accum_second, out, pixels);
    return hipGetLastError();
}

}  // namespace avd
