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

namespace avd {

namespace {

constexpr int kMaxR = 8;
constexpr int kMaxBlk = 32;
constexpr int kThreads = 128;

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
    __device__ __forceinline__ int32_t at(int x, int y) const
    {
        x = clampi(x, w - 1);
        y = clampi(y, h - 1);
        return codes ? (int32_t)codes[((size_t)y * w + x) * C] : plane[(size_t)y * w + x];
    }
};

// One window search for this workgroup's block: stages the block and the window in LDS, scores
// every offset, and returns (thread 0 only) the winner relative to (cdx, cdy) and its SAD.
__device__ void search_window(const Level &cen, const Level &nb, int ox, int oy, int blk, int R, int cdx, int cdy,
                              int32_t *s_blk, int32_t *s_win, int32_t *s_sad, int *bdx, int *bdy, int32_t *bsad)
{
    const int side = 2 * R + 1;
    const int ww = blk + 2 * R;
    for (int i = threadIdx.x; i < blk * blk; i += kThreads) {
        const int py = i / blk, px = i - py * blk;
        s_blk[i] = cen.at(ox + px, oy + py);
    }
    const int wx0 = ox + cdx - R, wy0 = oy + cdy - R;
    for (int i = threadIdx.x; i < ww * ww; i += kThreads) {
        const int wy = i / ww, wx = i - wy * ww;
        s_win[i] = nb.at(wx0 + wx, wy0 + wy);
    }
    __syncthreads();
    for (int c = threadIdx.x; c < side * side; c += kThreads) {
        const int dy = c / side, dx = c - dy * side;
        int32_t sad = 0;
        for (int py = 0; py < blk; ++py) {
            const int32_t *a = s_blk + py * blk;
            const int32_t *b = s_win + (py + dy) * ww + dx;
            for (int px = 0; px < blk; ++px) sad += abs(a[px] - b[px]);
        }
        s_sad[c] = sad;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        int32_t best = s_sad[0];
        int bi = 0;
        for (int c = 1; c < side * side; ++c)
            if (s_sad[c] < best) {
                best = s_sad[c];
                bi = c;
            }
        const int32_t centre_sad = s_sad[R * side + R];
        if (centre_sad <= best) {
            *bdx = cdx;
            *bdy = cdy;
            *bsad = centre_sad;
        } else {
            *bdx = cdx + (bi % side) - R;
            *bdy = cdy + (bi / side) - R;
            *bsad = best;
        }
    }
}

__global__ __launch_bounds__(kThreads) void k_bm_coarse(const int32_t *__restrict__ pyr1, int lw, int lh,
                                                        int blk, int step, int R, int centre_slot,
                                                        const uint32_t *__restrict__ nslots,
                                                        int32_t *__restrict__ mv, int fine_bx, int fine_by,
                                                        int fine_step)
{
    __shared__ int32_t s_blk[kMaxBlk * kMaxBlk];
    __shared__ int32_t s_win[(kMaxBlk + 2 * kMaxR) * (kMaxBlk + 2 * kMaxR)];
    __shared__ int32_t s_sad[(2 * kMaxR + 1) * (2 * kMaxR + 1)];
    const int bc = blockIdx.x, br = blockIdx.y, n = blockIdx.z;
    const size_t lp = (size_t)lw * lh;
    const Level cen{nullptr, pyr1 + (size_t)centre_slot * lp, lw, lh, 1};
    const Level nb{nullptr, pyr1 + (size_t)nslots[n] * lp, lw, lh, 1};
    int dx = 0, dy = 0;
    int32_t sad = 0;
    search_window(cen, nb, bc * step, br * step, blk, R, 0, 0, s_blk, s_win, s_sad, &dx, &dy, &sad);
    if (threadIdx.x != 0) return;
    int c0 = min(bc * step * 2 / fine_step, fine_bx), r0 = min(br * step * 2 / fine_step, fine_by);
    int c1 = (bc == (int)gridDim.x - 1) ? fine_bx : min((bc + 1) * step * 2 / fine_step, fine_bx);
    int r1 = (br == (int)gridDim.y - 1) ? fine_by : min((br + 1) * step * 2 / fine_step, fine_by);
    int32_t *m = mv + (size_t)n * fine_bx * fine_by * 2;
    for (int fr = r0; fr < r1; ++fr)
        for (int fc = c0; fc < c1; ++fc) {
            m[((size_t)fr * fine_bx + fc) * 2] = dx * 2;
            m[((size_t)fr * fine_bx + fc) * 2 + 1] = dy * 2;
        }
}

__global__ __launch_bounds__(kThreads) void k_bm_fine(const uint16_t *__restrict__ codes, int W, int H, int C,
                                                      int blk, int step, int R, int use_seed, int centre_slot,
                                                      const uint32_t *__restrict__ nslots, int32_t *__restrict__ mv,
                                                      float *__restrict__ conf, float max_value, float sad_floor,
                                                      float thsad)
{
    __shared__ int32_t s_blk[kMaxBlk * kMaxBlk];
    __shared__ int32_t s_win[(kMaxBlk + 2 * kMaxR) * (kMaxBlk + 2 * kMaxR)];
    __shared__ int32_t s_sad[(2 * kMaxR + 1) * (2 * kMaxR + 1)];
    const int bc = blockIdx.x, br = blockIdx.y, n = blockIdx.z;
    const int bx = gridDim.x, by = gridDim.y;
    const size_t sp = (size_t)W * H * C;
    const Level cen{codes + (size_t)centre_slot * sp, nullptr, W, H, C};
    const Level nb{codes + (size_t)nslots[n] * sp, nullptr, W, H, C};
    const size_t b = (size_t)n * bx * by + (size_t)br * bx + bc;
    const int sdx = use_seed ? mv[b * 2] : 0, sdy = use_seed ? mv[b * 2 + 1] : 0;
    int dx = 0, dy = 0;
    int32_t sad = 0;
    search_window(cen, nb, bc * step, br * step, blk, R, sdx, sdy, s_blk, s_win, s_sad, &dx, &dy, &sad);
    if (threadIdx.x != 0) return;
    mv[b * 2] = dx;
    mv[b * 2 + 1] = dy;
    if (conf) {
        float excess = (float)sad / max_value - sad_floor;
        if (excess < 0.0f) excess = 0.0f;
        const float t2 = thsad * thsad, e2 = excess * excess;
        const float c = (t2 - e2) / (t2 + e2);
        conf[b] = c < 0.0f ? 0.0f : c;
    }
}

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
    if (g.radius > kMaxR || g.blksize > kMaxBlk || g.blksize < 2 || g.levels < 1 || g.levels > 2)
        return hipErrorInvalidValue;
    int use_seed = 0;
    if (g.levels == 2) {
        const int lw = max(g.W / 2, 1), lh = max(g.H / 2, 1);
        const int cblk = max(g.blksize / 2, 2), cstep = max(g.step / 2, 1);
        const dim3 grid(max((int)div_up(lw, cstep), 1), max((int)div_up(lh, cstep), 1), neighbours);
        k_bm_coarse<<<grid, kThreads, 0, s>>>(pyr1, lw, lh, cblk, cstep, g.radius, centre_slot, nslots_dev, mv,
                                              g.blocks_x, g.blocks_y, g.step);
        use_seed = 1;
        hipError_t e = hipGetLastError();
        if (e != hipSuccess) return e;
    }
    const dim3 grid(g.blocks_x, g.blocks_y, neighbours);
    k_bm_fine<<<grid, kThreads, 0, s>>>(codes, g.W, g.H, g.C, g.blksize, g.step, g.radius, use_seed, centre_slot,
                                        nslots_dev, mv, conf, max_value, sad_floor, thsad);
    return hipGetLastError();
}

hipError_t launch_warp(const float *src, float *dst, const int32_t *mv, int W, int H, int C, int step, int blocks_x,
                       int blocks_y, hipStream_t s)
{
    k_warp<<<grid_for((uint64_t)W * H), 256, 0, s>>>(src, dst, mv, W, H, C, step, blocks_x, blocks_y);
    return hipGetLastError();
}

}  // namespace avd
