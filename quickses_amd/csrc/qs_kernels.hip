/* HIP kernels for the SES mesh path on CDNA4 (gfx950, wave64).
 *
 * Floating-point expressions that decide labels, cube indices or vertex positions are written
 * in the evaluation order of the reference kernel they follow and the library is compiled with
 * -ffp-contract=off, so each operation rounds once (IEEE single precision, correctly rounded
 * division and sqrt).  Integer outputs (cell ids, atom order, labels, cube indices, vertex and
 * triangle ids) are therefore bit-identical to a host evaluation of the same formulas.
 */
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "qs_common.h"
#include "qs_kernels.h"
#include <algorithm>
#include <cmath>
#include <type_traits>

#include "qs_mc_table.h"

/* field-pass-2 tiles: x, y, z extents (P2R_TX in x for the fixed-radius kernel, whose halo costs
   (16 + 2R) / 16 in x) */
#define P2_TX 8
#define P2_TY 8
#define P2_TZ 32
#define P2R_TX 16 /* x extent of the fixed-radius kernel's tiles (halo cost (16 + 2R) / 16 in x) */

__constant__ uint64_t c_tri[256] = QS_TRI_PACKED_DATA;

/* cube corners and edges (see qs_mc_table.h): edge e runs from corner offset kEdgeLow[e] along
   axis kEdgeAxis[e]; packed into bit masks so that per-lane lookups are shifts, not loads */
constexpr int8_t kEdgeLow[12][3] = {{0, 0, 0}, {1, 0, 0}, {0, 1, 0}, {0, 0, 0}, {0, 0, 1}, {1, 0, 1},
                                    {0, 1, 1}, {0, 0, 1}, {0, 0, 0}, {1, 0, 0}, {1, 1, 0}, {0, 1, 0}};
constexpr int8_t kEdgeAxis[12] = {0, 1, 0, 1, 0, 1, 0, 1, 2, 2, 2, 2};
constexpr uint32_t edge_low_bits(int c)
{
    uint32_t m = 0;
    for (int e = 0; e < 12; e++) m |= (uint32_t)(kEdgeLow[e][c] != 0) << e;
    return m;
}
constexpr uint32_t edge_axis_bits()
{
    uint32_t m = 0;
    for (int e = 0; e < 12; e++) m |= (uint32_t)kEdgeAxis[e] << (2 * e);
    return m;
}
constexpr uint32_t kEdgeLX = edge_low_bits(0), kEdgeLY = edge_low_bits(1), kEdgeLZ = edge_low_bits(2);
constexpr uint32_t kEdgeAX = edge_axis_bits();
static_assert(kEdgeLX == 0x622u && kEdgeLY == 0xC44u && kEdgeLZ == 0xF0u && kEdgeAX == 0xAA4444u, "edge tables");

/* ------------------------------------------------------------------------------------------ */
/* small helpers                                                                                */

__device__ __forceinline__ int iclamp(int f, int a, int b) { return max(a, min(f, b)); }

/* gridToSpace Kernels.cu:124-127: origin + (float)i * dx, two roundings */
__device__ __forceinline__ float grid_pos(float o, int i, float dx)
{
    float t = (float)i * dx;
    return o + t;
}

/* spaceToGrid Kernels.cu:119-123: truncation of (p - o) / dx */
__device__ __forceinline__ int cell_of(float p, float o, float dx) { return (int)((p - o) / dx); }

/* fast_distance Kernels.cu:149-155, left to right, IEEE sqrt */
/* the squared sum under dist3's root, with dist3's roundings */
__device__ __forceinline__ float dist3_sq(float ax, float ay, float az, float bx, float by, float bz)
{
    float ex = ax - bx, ey = ay - by, ez = az - bz;
    float x = ex * ex;
    float y = ey * ey;
    float z = ez * ez;
    float s = x + y;
    return s + z;
}

/* sqrtf, correctly rounded.  For a normal x or +inf the hardware root is within one ulp of the
   correctly rounded one, and Tuckerman's rounding test on the residuals of its two neighbours
   (exact FMAs) selects it; zero, denormal and NaN inputs take the general routine. */
__device__ __forceinline__ float sqrt_rn(float x)
{
    if (!(x >= 1.17549435e-38f)) return sqrtf(x);
    const float s = __builtin_amdgcn_sqrtf(x);
    const float sm = __uint_as_float(__float_as_uint(s) - 1u), sp = __uint_as_float(__float_as_uint(s) + 1u);
    float r = s;
    if (__builtin_fmaf(-sm, s, x) <= 0.0f) r = sm;
    if (__builtin_fmaf(-sp, s, x) > 0.0f) r = sp;
    return r;
}

__device__ __forceinline__ float dist3(float ax, float ay, float az, float bx, float by, float bz)
{
    float ex = ax - bx, ey = ay - by, ez = az - bz;
    float x = ex * ex;
    float y = ey * ey;
    float z = ez * ez;
    float s = x + y;
    s = s + z;
    return sqrtf(s);
}

__device__ __forceinline__ int tri_entry(uint64_t packed, int k) { return (int)((packed >> (4 * k)) & 0xF); }

/* number of triangles of a packed case: nibble k is 0xF exactly for k >= 3 * ntri (k < 16) */
__device__ __forceinline__ int tri_count_packed(uint64_t p)
{
    const uint64_t f = p & (p >> 1) & (p >> 2) & (p >> 3) & 0x1111111111111111ull; /* bit 4k: nibble k is 0xF */
    return (int)(__builtin_ctzll(f | (1ull << 63)) >> 2) / 3;
}

__device__ __forceinline__ int tri_count(int ci) { return tri_count_packed(c_tri[ci]); }

/* Segment lookups.  A prefix array pre (n + 1 non-decreasing entries, pre[n] = total) comes with
   a coarse index: coarse[k] = the segment holding item k << shift.  Segment of item v: start at
   coarse[v >> shift] and step forward past the segments that end at or before v (segments are
   long next to 2^shift items, so this is usually no step at all). */
template <int SHIFT, typename T>
__device__ __forceinline__ int seg_coarse(const T *__restrict__ pre, const uint32_t *__restrict__ coarse, T v)
{
    int j = (int)coarse[v >> SHIFT];
    while (pre[j + 1] <= v) j++;
    return j;
}

/* Workgroups are dealt to the eight XCDs round-robin (block b runs on XCD b % 8), and each XCD
   has its own L2.  This renumbers the blocks of an nb-block launch in groups of 8*C: inside a group
   each XCD runs C consecutive logical blocks, so neighbouring tiles share an L2 (their halo reads
   hit there and their partial output lines merge there before going to HBM), while the groups
   still spread every region of the grid over all eight XCDs.  A bijection on [0, nb): the last,
   incomplete group keeps its order.  Grid tiles use C = 8; per-vertex gathers, whose neighbours
   lie about one grid plane (a few thousand vertices) away, use C = 64. */
template <uint32_t C>
__device__ __forceinline__ uint32_t xcd_group(uint32_t b, uint32_t nb)
{
    constexpr uint32_t G = 8 * C;
    if (b >= nb / G * G) return b;
    const uint32_t w = b % G;
    return b - w + (w & 7u) * C + (w >> 3);
}
constexpr uint32_t XCD_CHUNK = 8;
__device__ __forceinline__ uint32_t xcd_block(uint32_t b, uint32_t nb) { return xcd_group<XCD_CHUNK>(b, nb); }
/* the same renumbering with a group size chosen at launch (c >= 1) */
__device__ __forceinline__ uint32_t xcd_group_rt(uint32_t b, uint32_t nb, uint32_t c)
{
    const uint32_t G = 8 * c;
    if (b >= nb / G * G) return b;
    const uint32_t w = b % G;
    return b - w + (w & 7u) * c + (w >> 3);
}

/* molecule of block b of a batched launch, and the block's index inside the molecule */
__device__ __forceinline__ int batch_mol(const QsBatch &B, uint32_t b, uint32_t &local)
{
    const int j = seg_coarse<QS_COARSE_BLOCKS>(B.bpre, B.coarse, b);
    local = b - B.bpre[j];
    return B.mid ? B.mid[j] : j;
}

/* index of the segment of a sorted prefix array (n + 1 entries) that holds item v */
__device__ __forceinline__ int seg_of(const uint32_t *__restrict__ pre, int n, uint32_t v)
{
    int lo = 0, hi = n - 1;
    while (lo < hi) {
        int mid = (lo + hi + 1) >> 1;
        if (pre[mid] <= v) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

/* segment of v for a wavefront whose items usually share one: the first active lane's segment
   is found with scalar loads, the others fall back to their own lookup */
__device__ __forceinline__ int seg_of_wave(const uint32_t *__restrict__ pre, const uint32_t *__restrict__ coarse, uint32_t v)
{
    const uint32_t v0 = __builtin_amdgcn_readfirstlane(v);
    const int m0 = __builtin_amdgcn_readfirstlane(seg_coarse<QS_COARSE_ITEMS>(pre, coarse, v0));
    if (v >= pre[m0] && v < pre[m0 + 1]) return m0;
    return seg_coarse<QS_COARSE_ITEMS>(pre, coarse, v);
}

/* exclusive scan over the block; returns this thread's prefix, *total the block sum */
template <int NT>
__device__ __forceinline__ uint32_t block_excl_scan(uint32_t v, uint32_t *total, uint32_t *lds)
{
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    uint32_t x = v;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        uint32_t y = __shfl_up(x, d, 64);
        if (lane >= d) x += y;
    }
    if (lane == 63) lds[wid] = x;
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t s = 0;
        for (int w = 0; w < NT / 64; w++) {
            uint32_t t = lds[w];
            lds[w] = s;
            s += t;
        }
        lds[NT / 64] = s;
    }
    __syncthreads();
    uint32_t res = x - v + lds[wid];
    *total = lds[NT / 64];
    __syncthreads();
    return res;
}

template <int NT>
__device__ __forceinline__ uint32_t block_sum(uint32_t v, uint32_t *lds)
{
    uint32_t t;
    block_excl_scan<NT>(v, &t, lds);
    return t;
}

/* sum of sums[0 .. b) for b <= 64, in every lane of the wave (a scan whose tile count is small
   takes its tiles' offsets this way instead of from a launch that scans the tile sums) */
__device__ __forceinline__ uint32_t wave_sum_below(const uint32_t *__restrict__ sums, uint32_t b)
{
    const uint32_t lane = threadIdx.x & 63;
    uint32_t x = (lane < b) ? sums[lane] : 0u;
#pragma unroll
    for (int d = 32; d >= 1; d >>= 1) x += (uint32_t)__shfl_xor((int)x, d, 64);
    return x;
}

/* ------------------------------------------------------------------------------------------ */
/* device-wide exclusive scan of uint32 (reduce tiles, scan tile sums, scan tiles)              */

#define SCAN_NT 256
#define SCAN_IPT 8
#define SCAN_TILE (SCAN_NT * SCAN_IPT)

__global__ void __launch_bounds__(SCAN_NT) k_tile_reduce(const uint32_t *__restrict__ in, uint64_t n,
                                                         uint32_t *__restrict__ sums)
{
    __shared__ uint32_t lds[SCAN_NT / 64 + 1];
    uint64_t base = (uint64_t)blockIdx.x * SCAN_TILE;
    uint32_t s = 0;
#pragma unroll
    for (int k = 0; k < SCAN_IPT; k++) {
        uint64_t i = base + (uint64_t)k * SCAN_NT + threadIdx.x;
        if (i < n) s += in[i];
    }
    uint32_t t = block_sum<SCAN_NT>(s, lds);
    if (threadIdx.x == 0) sums[blockIdx.x] = t;
}

/* out[i] = offsets[tile] + exclusive prefix inside the tile; in may alias out.  With self_sums
   (at most QS_SCAN_SELF_TILES tiles) the tile's offset is the sum of the tile sums before it. */
__global__ void __launch_bounds__(SCAN_NT) k_tile_scan(const uint32_t *in, uint32_t *out, uint64_t n,
                                                       const uint32_t *__restrict__ offsets,
                                                       uint32_t *__restrict__ total, const uint32_t *__restrict__ self_sums)
{
    __shared__ uint32_t lds[SCAN_NT / 64 + 1];
    uint64_t base = (uint64_t)blockIdx.x * SCAN_TILE + (uint64_t)threadIdx.x * SCAN_IPT;
    uint32_t v[SCAN_IPT];
    uint32_t s = 0;
#pragma unroll
    for (int k = 0; k < SCAN_IPT; k++) {
        uint64_t i = base + k;
        v[k] = (i < n) ? in[i] : 0u;
        s += v[k];
    }
    uint32_t tot;
    uint32_t pre = block_excl_scan<SCAN_NT>(s, &tot, lds);
    uint32_t off = offsets ? offsets[blockIdx.x] : (self_sums ? wave_sum_below(self_sums, blockIdx.x) : 0u);
    uint32_t run = off + pre;
#pragma unroll
    for (int k = 0; k < SCAN_IPT; k++) {
        uint64_t i = base + k;
        if (i < n) out[i] = run;
        run += v[k];
    }
    if (total && blockIdx.x == gridDim.x - 1 && threadIdx.x == 0) *total = off + tot;
}

/* Shell ranks: spre[i] = shell voxels (set bits of shellm) in words [0, i).  Field pass 2 stores
   the shell voxels' values only, densely in that order: voxel bit b of word i at
   spre[i] + popcount(shellm[i] below b). */
__global__ void __launch_bounds__(SCAN_NT) k_shell_reduce(const uint64_t *__restrict__ shellm, uint64_t n,
                                                          uint32_t *__restrict__ sums)
{
    __shared__ uint32_t lds[SCAN_NT / 64 + 1];
    const uint64_t base = (uint64_t)blockIdx.x * SCAN_TILE;
    uint32_t s = 0;
#pragma unroll
    for (int k = 0; k < SCAN_IPT; k++) {
        const uint64_t i = base + (uint64_t)k * SCAN_NT + threadIdx.x;
        if (i < n) s += (uint32_t)__popcll(shellm[i]);
    }
    const uint32_t t = block_sum<SCAN_NT>(s, lds);
    if (threadIdx.x == 0) sums[blockIdx.x] = t;
}

/* (self: offsets holds the raw tile sums of at most QS_SCAN_SELF_TILES tiles, each tile adds up
   the ones before it, and the last tile writes totals = {shell voxels, 0}) */
__global__ void __launch_bounds__(SCAN_NT) k_shell_scan(const uint64_t *__restrict__ shellm, uint64_t n,
                                                        const uint32_t *__restrict__ offsets, uint32_t *__restrict__ spre,
                                                        int self, uint32_t *__restrict__ totals)
{
    /* wave w of the block takes words [w * 512, (w + 1) * 512) of the tile, 64 consecutive words
       per load, and scans them in order: lane prefixes per group of 64, carried across groups */
    __shared__ uint32_t wsum[SCAN_NT / 64];
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    const uint64_t base = (uint64_t)blockIdx.x * SCAN_TILE + (uint64_t)wid * (64 * SCAN_IPT) + lane;
    uint32_t v[SCAN_IPT];
#pragma unroll
    for (int k = 0; k < SCAN_IPT; k++) {
        const uint64_t i = base + 64 * k;
        v[k] = (i < n) ? (uint32_t)__popcll(shellm[i]) : 0u;
    }
    uint32_t run = 0; /* words of the wave's earlier groups */
#pragma unroll
    for (int k = 0; k < SCAN_IPT; k++) {
        uint32_t x = v[k];
#pragma unroll
        for (int d = 1; d < 64; d <<= 1) {
            const uint32_t y = __shfl_up(x, d, 64);
            if (lane >= d) x += y;
        }
        const uint32_t tot = __shfl(x, 63, 64);
        v[k] = run + x - v[k]; /* exclusive prefix inside the wave's range */
        run += tot;
    }
    if (lane == 0) wsum[wid] = run;
    __syncthreads();
    uint32_t off = self ? wave_sum_below(offsets, blockIdx.x) : offsets[blockIdx.x];
    if (self && blockIdx.x == gridDim.x - 1 && threadIdx.x == 0) {
        uint32_t t = off;
        for (int w = 0; w < SCAN_NT / 64; w++) t += wsum[w];
        totals[0] = t;
        totals[1] = 0u;
    }
    for (int w = 0; w < wid; w++) off += wsum[w];
#pragma unroll
    for (int k = 0; k < SCAN_IPT; k++) {
        const uint64_t i = base + 64 * k;
        if (i < n) spre[i] = off + v[k];
    }
}

/* Marching-cubes word scans in one pass: per row word its vertex count, triangle count and
   "active" flag (either nonzero) are scanned together.  Plane 0 = triangles, 1 = vertices,
   2 = active words; tile sums are stored plane-major (3 x ntiles). */
__global__ void __launch_bounds__(SCAN_NT) k_mc_reduce3(const uint32_t *__restrict__ cnt, uint64_t n, uint32_t *__restrict__ sums,
                                                        uint64_t ntiles)
{
    __shared__ uint32_t lds[SCAN_NT / 64 + 1];
    const uint64_t base = (uint64_t)blockIdx.x * SCAN_TILE;
    uint32_t sv = 0, st = 0, sa = 0;
#pragma unroll
    for (int k = 0; k < SCAN_IPT; k++) {
        const uint64_t i = base + (uint64_t)k * SCAN_NT + threadIdx.x;
        if (i < n) {
            const uint32_t c = cnt[i], v = c & 0xFFFFu, t = c >> 16;
            sv += v;
            st += t;
            sa += (v | t) ? 1u : 0u;
        }
    }
    const uint32_t tt = block_sum<SCAN_NT>(st, lds);
    const uint32_t tv = block_sum<SCAN_NT>(sv, lds);
    const uint32_t ta = block_sum<SCAN_NT>(sa, lds);
    if (threadIdx.x == 0) {
        sums[blockIdx.x] = tt;
        sums[ntiles + blockIdx.x] = tv;
        sums[2 * ntiles + blockIdx.x] = ta;
    }
}

/* block p: exclusive scan of plane p of the tile sums in place, total to totals[p]; the slot
   after the plane totals (the degenerate-triangle count, accumulated later) starts at zero */
__global__ void __launch_bounds__(SCAN_NT) k_scan_planes(uint32_t *__restrict__ sums, uint64_t ntiles,
                                                         uint32_t *__restrict__ totals)
{
    __shared__ uint32_t lds[SCAN_NT / 64 + 1];
    uint32_t *plane = sums + (uint64_t)blockIdx.x * ntiles;
    uint32_t carry = 0;
    for (uint64_t b0 = 0; b0 < ntiles; b0 += SCAN_TILE) {
        const uint64_t base = b0 + (uint64_t)threadIdx.x * SCAN_IPT;
        uint32_t v[SCAN_IPT], s = 0;
#pragma unroll
        for (int k = 0; k < SCAN_IPT; k++) {
            v[k] = (base + k < ntiles) ? plane[base + k] : 0u;
            s += v[k];
        }
        uint32_t tot;
        uint32_t run = carry + block_excl_scan<SCAN_NT>(s, &tot, lds);
#pragma unroll
        for (int k = 0; k < SCAN_IPT; k++) {
            if (base + k < ntiles) plane[base + k] = run;
            run += v[k];
        }
        carry += tot;
    }
    if (threadIdx.x == 0) totals[blockIdx.x] = carry;
    if (blockIdx.x == 0 && threadIdx.x == 0) totals[gridDim.x] = 0u;
}

/* An active word with items [first, first + count) is where every 256-item block whose first
   item lies in that range starts: it writes those blocks' start (its index in the active list),
   below the start array's capacity cap. */
__device__ __forceinline__ void mark_block_starts(uint32_t *__restrict__ start, uint32_t cap, uint32_t first, uint32_t count,
                                                  uint32_t r)
{
    for (uint32_t b = (first + 255u) >> 8; b < cap && (b << 8) < first + count; b++) start[b] = r;
}

/* per word: first vertex (vbw) and first triangle (tbw); active words listed in order; with
   bs.vstart, the first active word of every 256-vertex and 256-triangle block, and the triangle
   blocks' counters cleared */
__global__ void __launch_bounds__(SCAN_NT) k_mc_scan3(const uint32_t *__restrict__ cnt, uint64_t n, const uint32_t *__restrict__ sums, uint64_t ntiles,
                                                      uint32_t *__restrict__ vbw, uint32_t *__restrict__ tbw,
                                                      uint32_t *__restrict__ list, uint32_t *__restrict__ avb,
                                                      uint32_t *__restrict__ atb, uint32_t list_base, int self,
                                                      uint32_t *__restrict__ totals, QsBlockStarts bs)
{
    if (bs.zero)
        for (uint32_t i = blockIdx.x * SCAN_NT + threadIdx.x; i < bs.nzero; i += gridDim.x * SCAN_NT) bs.zero[i] = 0u;
    /* (self: sums holds the raw tile sums of at most QS_SCAN_SELF_TILES tiles, each tile adds up
       the ones before it, and the last tile writes totals = {triangles, vertices, active, 0}) */
    /* wave w of the block takes words [w * 512, (w + 1) * 512) of the tile, 64 consecutive words
       per load (as k_shell_scan); the three scans run per group of 64 in order, carried across
       groups, and each group's active words are listed in order by ballot */
    __shared__ uint32_t wsum[3][SCAN_NT / 64];
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    const uint64_t base = (uint64_t)blockIdx.x * SCAN_TILE + (uint64_t)wid * (64 * SCAN_IPT) + lane;
    const uint64_t below = (1ull << lane) - 1ull;
    uint32_t v[SCAN_IPT], t[SCAN_IPT], pv[SCAN_IPT], pt[SCAN_IPT], pa[SCAN_IPT];
    uint64_t am[SCAN_IPT];
#pragma unroll
    for (int k = 0; k < SCAN_IPT; k++) {
        const uint64_t i = base + 64 * k;
        const uint32_t c = (i < n) ? cnt[i] : 0u;
        v[k] = c & 0xFFFFu;
        t[k] = c >> 16;
    }
    uint32_t rv = 0, rt = 0, ra = 0; /* the wave's earlier groups */
#pragma unroll
    for (int k = 0; k < SCAN_IPT; k++) {
        uint32_t xv = v[k], xt = t[k];
#pragma unroll
        for (int d = 1; d < 64; d <<= 1) {
            const uint32_t yv = __shfl_up(xv, d, 64), yt = __shfl_up(xt, d, 64);
            if (lane >= d) {
                xv += yv;
                xt += yt;
            }
        }
        am[k] = __ballot((v[k] | t[k]) != 0u);
        pv[k] = rv + xv - v[k];
        pt[k] = rt + xt - t[k];
        pa[k] = ra + (uint32_t)__popcll(am[k] & below);
        rv += __shfl(xv, 63, 64);
        rt += __shfl(xt, 63, 64);
        ra += (uint32_t)__popcll(am[k]);
    }
    if (lane == 0) {
        wsum[0][wid] = rv;
        wsum[1][wid] = rt;
        wsum[2][wid] = ra;
    }
    __syncthreads();
    uint32_t ov, ot, oa;
    if (self) {
        ot = wave_sum_below(sums, blockIdx.x);
        ov = wave_sum_below(sums + ntiles, blockIdx.x);
        oa = wave_sum_below(sums + 2 * ntiles, blockIdx.x);
        if (blockIdx.x == gridDim.x - 1 && threadIdx.x == 0) {
            uint32_t tv = ov, tt = ot, ta = oa;
            for (int w = 0; w < SCAN_NT / 64; w++) {
                tv += wsum[0][w];
                tt += wsum[1][w];
                ta += wsum[2][w];
            }
            totals[0] = tt;
            totals[1] = tv;
            totals[2] = ta;
            totals[3] = 0u;
        }
    } else {
        ov = sums[ntiles + blockIdx.x];
        ot = sums[blockIdx.x];
        oa = sums[2 * ntiles + blockIdx.x];
    }
    for (int w = 0; w < wid; w++) {
        ov += wsum[0][w];
        ot += wsum[1][w];
        oa += wsum[2][w];
    }
#pragma unroll
    for (int k = 0; k < SCAN_IPT; k++) {
        const uint64_t i = base + 64 * k;
        if (i < n) {
            vbw[i] = ov + pv[k];
            tbw[i] = ot + pt[k];
            if ((am[k] >> lane) & 1ull) {
                const uint32_t r = oa + pa[k];
                avb[r] = ov + pv[k];
                atb[r] = ot + pt[k];
                list[r] = list_base + (uint32_t)i;
                if (bs.vstart) {
                    mark_block_starts(bs.vstart, bs.nv, ov + pv[k], v[k], r);
                    mark_block_starts(bs.tstart, bs.nt, ot + pt[k], t[k], r);
                }
            }
        }
    }
}

/* The same scans in one block of SCAN1_NT threads for at most QS_SCAN1_MAX entries: one launch
   instead of three for a small grid.  The block runs over chunks of SCAN1_NT * SCAN_IPT entries in
   order, each wave 64 consecutive entries per load as above, carrying the running totals. */
#define SCAN1_NT 1024
#define SCAN1_W (SCAN1_NT / 64)

/* this wave's offset (waves below it) and the chunk total, from the per-wave sums */
__device__ __forceinline__ void scan1_offsets(const uint32_t *wsum, int wid, uint32_t &below, uint32_t &all)
{
    below = 0u;
    all = 0u;
#pragma unroll
    for (int w = 0; w < SCAN1_W; w++) {
        const uint32_t s = wsum[w];
        below += (w < wid) ? s : 0u;
        all += s;
    }
}

/* exclusive prefix of the wave's SCAN_IPT groups of 64 in place; returns the wave's total */
__device__ __forceinline__ uint32_t wave_groups_scan(uint32_t (&v)[SCAN_IPT], int lane)
{
    uint32_t run = 0;
#pragma unroll
    for (int k = 0; k < SCAN_IPT; k++) {
        uint32_t x = v[k];
#pragma unroll
        for (int d = 1; d < 64; d <<= 1) {
            const uint32_t y = __shfl_up(x, d, 64);
            if (lane >= d) x += y;
        }
        const uint32_t tot = __shfl(x, 63, 64);
        v[k] = run + x - v[k];
        run += tot;
    }
    return run;
}

/* generic exclusive scan (in may equal out); total to *total if given */
__global__ void __launch_bounds__(SCAN1_NT) k_scan1(const uint32_t *in, uint32_t *out, uint64_t n, uint32_t *__restrict__ total)
{
    __shared__ uint32_t wsum[SCAN1_W];
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    uint32_t carry = 0;
    for (uint64_t c0 = 0; c0 < n; c0 += SCAN1_NT * SCAN_IPT) {
        const uint64_t base = c0 + (uint64_t)wid * (64 * SCAN_IPT) + lane;
        uint32_t v[SCAN_IPT];
#pragma unroll
        for (int k = 0; k < SCAN_IPT; k++) v[k] = (base + 64 * k < n) ? in[base + 64 * k] : 0u;
        const uint32_t run = wave_groups_scan(v, lane);
        if (lane == 0) wsum[wid] = run;
        __syncthreads();
        uint32_t below, all;
        scan1_offsets(wsum, wid, below, all);
#pragma unroll
        for (int k = 0; k < SCAN_IPT; k++)
            if (base + 64 * k < n) out[base + 64 * k] = carry + below + v[k];
        carry += all;
        __syncthreads(); /* wsum is written again for the next chunk */
    }
    if (total && threadIdx.x == 0) *total = carry;
}

/* shell ranks (as k_shell_reduce, k_scan_planes, k_shell_scan): totals = {shell voxels, 0} */
__global__ void __launch_bounds__(SCAN1_NT) k_shell_scan1(const uint64_t *__restrict__ shellm, uint64_t n,
                                                          uint32_t *__restrict__ totals, uint32_t *__restrict__ spre)
{
    __shared__ uint32_t wsum[SCAN1_W];
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    uint32_t carry = 0;
    for (uint64_t c0 = 0; c0 < n; c0 += SCAN1_NT * SCAN_IPT) {
        const uint64_t base = c0 + (uint64_t)wid * (64 * SCAN_IPT) + lane;
        uint32_t v[SCAN_IPT];
#pragma unroll
        for (int k = 0; k < SCAN_IPT; k++) v[k] = (base + 64 * k < n) ? (uint32_t)__popcll(shellm[base + 64 * k]) : 0u;
        const uint32_t run = wave_groups_scan(v, lane);
        if (lane == 0) wsum[wid] = run;
        __syncthreads();
        uint32_t below, all;
        scan1_offsets(wsum, wid, below, all);
#pragma unroll
        for (int k = 0; k < SCAN_IPT; k++)
            if (base + 64 * k < n) spre[base + 64 * k] = carry + below + v[k];
        carry += all;
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        totals[0] = carry;
        totals[1] = 0u;
    }
}

/* marching-cubes word scans (as k_mc_reduce3, k_scan_planes, k_mc_scan3): totals = {triangles,
   vertices, active words, 0} */
__global__ void __launch_bounds__(SCAN1_NT) k_mc_scan1(const uint32_t *__restrict__ cnt, uint64_t n, uint32_t *__restrict__ vbw, uint32_t *__restrict__ tbw,
                                                       uint32_t *__restrict__ list, uint32_t *__restrict__ avb,
                                                       uint32_t *__restrict__ atb, uint32_t list_base,
                                                       uint32_t *__restrict__ totals, QsBlockStarts bs)
{
    __shared__ uint32_t wsum[3][SCAN1_W];
    if (bs.zero)
        for (uint32_t i = threadIdx.x; i < bs.nzero; i += SCAN1_NT) bs.zero[i] = 0u;
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    const uint64_t lbelow = (1ull << lane) - 1ull;
    uint32_t cv = 0, ct = 0, ca = 0;
    for (uint64_t c0 = 0; c0 < n; c0 += SCAN1_NT * SCAN_IPT) {
        const uint64_t base = c0 + (uint64_t)wid * (64 * SCAN_IPT) + lane;
        uint32_t v[SCAN_IPT], t[SCAN_IPT], pa[SCAN_IPT], raw[SCAN_IPT];
        uint64_t am[SCAN_IPT];
        uint32_t ra = 0;
#pragma unroll
        for (int k = 0; k < SCAN_IPT; k++) {
            const uint64_t i = base + 64 * k;
            const uint32_t c = (i < n) ? cnt[i] : 0u;
            v[k] = c & 0xFFFFu;
            t[k] = c >> 16;
            raw[k] = c;
        }
#pragma unroll
        for (int k = 0; k < SCAN_IPT; k++) {
            am[k] = __ballot((v[k] | t[k]) != 0u);
            pa[k] = ra + (uint32_t)__popcll(am[k] & lbelow);
            ra += (uint32_t)__popcll(am[k]);
        }
        const uint32_t rv = wave_groups_scan(v, lane), rt = wave_groups_scan(t, lane);
        if (lane == 0) {
            wsum[0][wid] = rv;
            wsum[1][wid] = rt;
            wsum[2][wid] = ra;
        }
        __syncthreads();
        uint32_t bv, av, bt, at, ba, aa;
        scan1_offsets(wsum[0], wid, bv, av);
        scan1_offsets(wsum[1], wid, bt, at);
        scan1_offsets(wsum[2], wid, ba, aa);
#pragma unroll
        for (int k = 0; k < SCAN_IPT; k++) {
            const uint64_t i = base + 64 * k;
            if (i < n) {
                const uint32_t fv = cv + bv + v[k], ft = ct + bt + t[k];
                vbw[i] = fv;
                tbw[i] = ft;
                if ((am[k] >> lane) & 1ull) {
                    const uint32_t r = ca + ba + pa[k];
                    avb[r] = fv;
                    atb[r] = ft;
                    list[r] = list_base + (uint32_t)i;
                    if (bs.vstart) {
                        mark_block_starts(bs.vstart, bs.nv, fv, raw[k] & 0xFFFFu, r);
                        mark_block_starts(bs.tstart, bs.nt, ft, raw[k] >> 16, r);
                    }
                }
            }
        }
        cv += av;
        ct += at;
        ca += aa;
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        totals[0] = ct;
        totals[1] = cv;
        totals[2] = ca;
        totals[3] = 0u;
    }
}

/* ------------------------------------------------------------------------------------------ */
/* bounding box (getMinMax CudaSurf.cu:149-174): partial boxes per molecule, then one wave    */

#define BB_NT 256
#define BB_NV 8 /* min x,y,z, max x,y,z, max radius, count of non-finite inputs */
#define BB_PMAX 256 /* partial boxes per molecule at most */

__device__ __forceinline__ void bb_combine(float *v, const float *w)
{
    for (int c = 0; c < 3; c++) v[c] = fminf(v[c], w[c]);
    for (int c = 3; c < 7; c++) v[c] = fmaxf(v[c], w[c]);
    v[7] += w[7];
}

/* block (p, m) of P partial boxes per molecule: atoms off[m] + p*BB_NT + k*P*BB_NT + t of molecule m
   -> part[m][p] (off == nullptr: one molecule, atoms [o0, o1)) */
__global__ void __launch_bounds__(BB_NT) k_bbox_part(const float *__restrict__ pos, const float *__restrict__ rad,
                                                     const uint32_t *__restrict__ off, uint32_t o0, uint32_t o1,
                                                     float *__restrict__ part)
{
    const uint32_t P = gridDim.x;
    __shared__ float s[BB_NV][BB_NT];
    const int m = blockIdx.y, p = blockIdx.x;
    float v[BB_NV] = {100000.0f, 100000.0f, 100000.0f, -100000.0f, -100000.0f, -100000.0f, 0.0f, 0.0f};
    const uint32_t beg = off ? off[m] : o0, end = off ? off[m + 1] : o1;
    for (uint32_t i = beg + p * BB_NT + threadIdx.x; i < end; i += P * BB_NT) {
        for (int c = 0; c < 3; c++) {
            float q = pos[3 * (size_t)i + c];
            v[c] = fminf(v[c], q);
            v[3 + c] = fmaxf(v[3 + c], q);
            if (!isfinite(q)) v[7] += 1.0f;
        }
        float ra = rad[i];
        v[6] = fmaxf(v[6], ra);
        if (!isfinite(ra)) v[7] += 1.0f;
    }
    for (int c = 0; c < BB_NV; c++) s[c][threadIdx.x] = v[c];
    __syncthreads();
    for (int st = BB_NT / 2; st > 0; st >>= 1) {
        if ((int)threadIdx.x < st) {
            float w[BB_NV];
            for (int c = 0; c < BB_NV; c++) {
                v[c] = s[c][threadIdx.x];
                w[c] = s[c][threadIdx.x + st];
            }
            bb_combine(v, w);
            for (int c = 0; c < BB_NV; c++) s[c][threadIdx.x] = v[c];
        }
        __syncthreads();
    }
    if (threadIdx.x < BB_NV) part[((size_t)m * P + p) * BB_NV + threadIdx.x] = s[threadIdx.x][0];
}

/* one wavefront per molecule: combine its P partial boxes, lane k taking partials k, k + 64, ...,
   then across the lanes (minima, maxima and a sum of small integer counts: the same box in any
   order) */
__global__ void __launch_bounds__(64) k_bbox_final(const float *__restrict__ part, int P, float *__restrict__ out)
{
    const int m = blockIdx.x, lane = threadIdx.x;
    float v[BB_NV] = {INFINITY, INFINITY, INFINITY, -INFINITY, -INFINITY, -INFINITY, -INFINITY, 0.0f};
    for (int p = lane; p < P; p += 64) {
        float w[BB_NV];
#pragma unroll
        for (int c = 0; c < BB_NV; c++) w[c] = part[((size_t)m * P + p) * BB_NV + c];
        bb_combine(v, w);
    }
#pragma unroll
    for (int d = 32; d >= 1; d >>= 1) {
        float w[BB_NV];
#pragma unroll
        for (int c = 0; c < BB_NV; c++) w[c] = __shfl_xor(v[c], d, 64);
        bb_combine(v, w);
    }
    if (lane < BB_NV) {
        float r = v[0];
#pragma unroll
        for (int c = 1; c < BB_NV; c++) r = (lane == c) ? v[c] : r;
        out[BB_NV * m + lane] = r;
    }
}

/* ------------------------------------------------------------------------------------------ */
/* atom hashing and cell ordering                                                              */

/* hashAtoms Kernels.cu:185-216: cell id x-major (global over the batch); count atoms per cell.
   key[atom] = (cell, the atom's rank among its cell's atoms in the order the counts were taken):
   the scatter places it with that rank, and the cell's atoms are put in index order afterwards. */
__global__ void __launch_bounds__(256) k_hash_count(QsBatch B, const float *__restrict__ pos, int2 *__restrict__ key,
                                                    uint32_t *__restrict__ count)
{
    uint32_t lb;
    const QsMol &M = B.mols[batch_mol(B, blockIdx.x, lb)];
    const uint32_t i = lb * 256 + threadIdx.x;
    if (i >= M.n_atoms) return;
    const QsGrid G = M.G;
    const size_t ga = (size_t)M.atom_off + i;
    /* in range for every atom by construction of the grid; clamped only as a memory guard */
    int cx = iclamp(cell_of(pos[3 * ga + 0], G.ox, G.dxn), 0, G.gn - 1);
    int cy = iclamp(cell_of(pos[3 * ga + 1], G.oy, G.dxn), 0, G.gn - 1);
    int cz = iclamp(cell_of(pos[3 * ga + 2], G.oz, G.dxn), 0, G.gn - 1);
    int h = (int)M.cell_off + ((G.gn * G.gn * cx) + (G.gn * cy) + cz);
    key[ga] = make_int2(h, (int)atomicAdd(&count[h], 1u));
}

__global__ void __launch_bounds__(256) k_scatter(const int2 *__restrict__ key, int n, const uint32_t *__restrict__ start,
                                                 int *__restrict__ perm)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int2 hr = key[i];
    perm[start[hr.x] + (uint32_t)hr.y] = i;
}

/* Per cell: order its atoms by input index, so the whole order is (cell, index) -- the order a
   stable sort by cell id gives (thrust::sort CudaSurf.cu:610-612) -- and write the cell range
   (sortCell Kernels.cu:306-341). */
__global__ void __launch_bounds__(256) k_cell_finalize(const uint32_t *__restrict__ start, const uint32_t *__restrict__ count,
                                                       int ncell, int *__restrict__ perm, int2 *__restrict__ cse)
{
    int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= ncell) return;
    uint32_t s = start[c], k = count[c];
    if (k == 0) {
        cse[c] = make_int2(QS_EMPTY, QS_EMPTY);
        return;
    }
    for (uint32_t a = s + 1; a < s + k; a++) {
        int v = perm[a];
        uint32_t b = a;
        while (b > s && perm[b - 1] > v) {
            perm[b] = perm[b - 1];
            b--;
        }
        perm[b] = v;
    }
    cse[c] = make_int2((int)s, (int)(s + k));
}

/* ------------------------------------------------------------------------------------------ */
/* field pass 1: probeIntersection + computeInOrOut Kernels.cu:344-475                         */
/*
 * Labels are kept as two bit masks per 64-voxel z-word of every (x, y) row: "outside" (no
 * admissible atom within r + probe) and "shell" (within r + probe of one, within r - dx of none);
 * a voxel in neither is inside.  Bits for z >= g are 0; the last layer of each axis is outside
 * (never written by the reference, CudaSurf.cu:664).
 *
 * k_pass1_rows: one wavefront per row word of up to four adjacent rows.  The lanes take the
 * candidate atoms of the rows (the atoms of the neighbour cells any of their voxels would visit),
 * and each lane computes, for its atom and each row, the z-interval of voxels it reaches.  Along a row the single-precision distance
 * sqrt((x^2 + y^2) + z^2) is non-increasing then non-decreasing in z (every rounding is
 * monotone), so {z : d(z) < T} is an interval; it is located from the closed-form estimate and
 * then pinned by evaluating the reference's float predicate at its ends.  A voxel counts an atom
 * only if the atom's cell is one of the 27 (clamped) cells it would visit, as in the reference.
 * The wave ORs the lanes' interval masks.  Cost per row word: a few hundred instructions,
 * instead of ~44 candidate atoms per voxel.
 */

/* OR of v over the 64 lanes (all active), returned wave-uniform: row shifts inside each 16-lane
   row, then row broadcasts, leave the total in lane 63 */
__device__ __forceinline__ uint32_t wave_or_u32(uint32_t v)
{
    v |= (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x111, 0xf, 0xf, true); /* row_shr:1 */
    v |= (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x112, 0xf, 0xf, true); /* row_shr:2 */
    v |= (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x114, 0xf, 0xf, true); /* row_shr:4 */
    v |= (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x118, 0xf, 0xf, true); /* row_shr:8 */
    v |= (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x142, 0xa, 0xf, false); /* row_bcast:15 */
    v |= (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x143, 0xc, 0xf, false); /* row_bcast:31 */
    return (uint32_t)__builtin_amdgcn_readlane((int)v, 63);
}

__device__ __forceinline__ uint64_t wave_or_u64(uint64_t v)
{
    return ((uint64_t)wave_or_u32((uint32_t)(v >> 32)) << 32) | wave_or_u32((uint32_t)v);
}

/* bits [a, b] (inclusive, 0 <= a, b <= 63) */
__device__ __forceinline__ uint64_t bit_range(int a, int b)
{
    if (a > b) return 0ull;
    uint64_t hi = (b >= 63) ? ~0ull : ((2ull << b) - 1ull);
    uint64_t lo = (1ull << a) - 1ull;
    return hi & ~lo;
}

/* bits of the word starting at z0 for z < L */
__device__ __forceinline__ uint64_t bits_below(int L, int z0)
{
    int n = L - z0;
    return n <= 0 ? 0ull : (n >= 64 ? ~0ull : ((1ull << n) - 1ull));
}

/* smallest float s with sqrtf(s) >= T (T > 0), so that for v >= 0: sqrtf(v) < T <=> v < s.
   Every step uses the correctly rounded sqrtf, so the comparison is the reference's. */
__device__ __forceinline__ float sqrt_threshold(float T)
{
    if (!(T > 0.0f)) return -1.0f;
    float s = T * T;
    for (int i = 0; i < 8 && sqrtf(s) < T; i++) s = __uint_as_float(__float_as_uint(s) + 1u);
    for (int i = 0; i < 8 && s > 0.0f; i++) {
        float q = __uint_as_float(__float_as_uint(s) - 1u);
        if (sqrtf(q) >= T) s = q;
        else break;
    }
    return s;
}

/* one atom seen from one row: fl(S + fl(e(z)^2)) with S = fl(x^2 + y^2), e(z) = az - pz(z) */
struct RowAtom {
    float az, S, dx, oz;
    __device__ __forceinline__ float Zf(int z) const
    {
        float pz = grid_pos(oz, z, dx);
        float e = az - pz;
        return e * e;
    }
    __device__ __forceinline__ bool pred(int z, float ss) const
    {
        float v = S + Zf(z);
        return v < ss;
    }
};

/* exact [L, R] within [zlo, zhi] where the predicate holds; zb = its argmin in [zlo, zhi];
   L > R if none.  zc is the atom's z in voxel units; zc -+ h only start the exact walks, so the
   hardware root (within an ulp) serves for h. */
__device__ __forceinline__ void row_interval(const RowAtom &A, float ss, int zlo, int zhi, int zb, float zc, float inv_dx,
                                             int &L, int &R)
{
    L = 1;
    R = 0;
    if (zlo > zhi || !A.pred(zb, ss)) return;
    float h2 = ss - A.S;
    float h = (h2 > 0.0f) ? __builtin_amdgcn_sqrtf(h2) * inv_dx : 0.0f;
    int l = iclamp((int)ceilf(zc - h), zlo, zb);
    if (A.pred(l, ss)) {
        while (l > zlo && A.pred(l - 1, ss)) l--;
    } else {
        while (l < zb && !A.pred(l, ss)) l++;
    }
    int r = iclamp((int)floorf(zc + h), zb, zhi);
    if (A.pred(r, ss)) {
        while (r < zhi && A.pred(r + 1, ss)) r++;
    } else {
        while (r > zb && !A.pred(r, ss)) r--;
    }
    L = l;
    R = r;
}

/* per atom in cell order, once: the atom (perm gives its input index) as float4 x, y, z, r, the
   exact float thresholds of its near (r + probe) and inside (r - dx) tests, and its neighbour-cell
   y and z (packed) */
__global__ void __launch_bounds__(256) k_gather_aux(QsBatch B, const int *__restrict__ perm, const float *__restrict__ pos,
                                                    const float *__restrict__ rad, float4 *__restrict__ sx,
                                                    float4 *__restrict__ aux)
{
    uint32_t lb;
    const QsMol &M = B.mols[batch_mol(B, blockIdx.x, lb)];
    const uint32_t i = lb * 256 + threadIdx.x;
    if (i >= M.n_atoms) return;
    const QsGrid G = M.G;
    const size_t k = (size_t)M.atom_off + i;
    const int id = perm[k];
    const float4 a = make_float4(pos[3 * (size_t)id], pos[3 * (size_t)id + 1], pos[3 * (size_t)id + 2], rad[id]);
    sx[k] = a;
    const int acy = iclamp(cell_of(a.y, G.oy, G.dxn), 0, G.gn - 1);
    const int acz = iclamp(cell_of(a.z, G.oz, G.dxn), 0, G.gn - 1);
    aux[k] = make_float4(sqrt_threshold(QS_PROBE + a.w), sqrt_threshold(a.w - G.dx), __int_as_float((acy << 16) | acz), 0.0f);
}

/* rows per wavefront: R consecutive y rows (R * dx <= neighbour-cell size, so they span at most
   two neighbour cells in y and the candidate set is at most 3 x 4 cell columns) */
#define P1_RMAX 4
#define P1_QCAP 128

__device__ __forceinline__ void wave_lds_sync()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

__global__ void __launch_bounds__(256) k_pass1_rows(QsBatch B, const uint32_t *__restrict__ cstart_all,
                                                    const float4 *__restrict__ sx, const float4 *__restrict__ aux,
                                                    uint64_t *__restrict__ outm_all, uint64_t *__restrict__ shellm_all)
{
    __shared__ int queue_all[4][P1_QCAP];
    /* wave-uniform values go through readfirstlane so that they live in scalar registers */
    const int lane = threadIdx.x & 63, wid = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    int *queue = queue_all[wid];
    uint32_t lb;
    const QsMol &M = B.mols[batch_mol(B, blockIdx.x, lb)];
    const QsGrid G = M.G;
    const int g = G.g, gn = G.gn, W = (int)M.W, R = M.rows_per_wave;
    const uint32_t *__restrict__ cstart = cstart_all + M.cell_off;
    uint64_t *__restrict__ outm = outm_all + M.word_off;
    uint64_t *__restrict__ shellm = shellm_all + M.word_off;
    /* one wavefront per group of R rows, looping over the rows' W words: the group's setup
       (molecule, rows, x and y cells) is paid once */
    const int nby = (g + 4 * R - 1) / (4 * R);
    const int by = (int)(lb % (uint32_t)nby);
    const int x = M.p1_lo + (int)(lb / (uint32_t)nby);
    const int x1 = M.p1_hi;
    const int y0 = (by * 4 + wid) * R;
    if (y0 >= g || x >= x1) return; /* whole wavefront */
    const int ny = min(R, g - y0);
    if (x >= g - 1) {
        for (int w = 0; w < W; w++) {
            const int nvalid = min(64, g - w * 64);
            const uint64_t valid = (nvalid >= 64) ? ~0ull : ((1ull << nvalid) - 1ull);
            if (lane < ny) {
                size_t widx = ((size_t)x * g + y0 + lane) * W + w;
                outm[widx] = valid;
                shellm[widx] = 0ull;
            }
        }
        return;
    }
    const float dx = G.dx, inv_dx = 1.0f / G.dx;
    const float px = grid_pos(G.ox, x, dx);
    const int cx = cell_of(px, G.ox, G.dxn);
    float py[P1_RMAX];
    int cyl[P1_RMAX], cyh[P1_RMAX];
#pragma unroll
    for (int k = 0; k < P1_RMAX; k++) {
        int yk = min(y0 + min(k, ny - 1), g - 1);
        py[k] = grid_pos(G.oy, yk, dx);
        int cy = cell_of(py[k], G.oy, G.dxn);
        cyl[k] = iclamp(cy - 1, 0, gn - 1);
        cyh[k] = iclamp(cy + 1, 0, gn - 1);
    }
    for (int w = 0; w < W; w++) {
        const int z0 = w * 64;
        const int nvalid = min(64, g - z0);
        const uint64_t valid = (nvalid >= 64) ? ~0ull : ((1ull << nvalid) - 1ull);
        wave_lds_sync(); /* the previous word's queue reads are done */
        /* this lane's voxel z cell; the wave's cells are monotone in the lane */
        const int zl = min(z0 + lane, g - 1);
        const int cz = cell_of(grid_pos(G.oz, zl, dx), G.oz, G.dxn);
        const int czmin = __builtin_amdgcn_readfirstlane(cz), czmax = __builtin_amdgcn_readlane(cz, 63);
        const int D = czmax - czmin + 1; /* distinct z cells + 1 */
        const int zc0 = iclamp(czmin - 1, 0, gn - 1), zc1 = iclamp(czmax + 1, 0, gn - 1);
        const int xa = iclamp(cx - 1, 0, gn - 1), xb = iclamp(cx + 1, 0, gn - 1);
        const int ya = cyl[0], yb = cyh[P1_RMAX - 1];
        /* candidate atoms: up to 3 x 4 contiguous ranges of the cell-ordered array */
        int rs[12], pre[13];
        pre[0] = 0;
    #pragma unroll
        for (int i = 0; i < 3; i++) {
    #pragma unroll
            for (int j = 0; j < 4; j++) {
                int k = 4 * i + j;
                int xc = xa + i, yc = ya + j;
                int len = 0, st = 0;
                if (xc <= xb && yc <= yb) {
                    int h0 = (xc * gn + yc) * gn + zc0;
                    st = (int)cstart[h0];
                    len = (int)cstart[h0 + (zc1 - zc0) + 1] - st;
                }
                rs[k] = st;
                pre[k + 1] = pre[k] + len;
            }
        }
        const int total = pre[12];
        const int zmax = min(z0 + 63, g - 2); /* labelled voxels: z <= g-2 (z = g-1 is outside) */
        if (total == 0) {
            /* no atom can reach these rows: every voxel is outside */
            if (lane < ny) {
                const size_t widx = ((size_t)x * g + y0 + lane) * W + w;
                outm[widx] = valid;
                shellm[widx] = 0ull;
            }
            continue;
        }
        /* tab (lane j) = number of lanes whose z cell is < czmin + j, j = 0..D */
        int tab = (lane >= D) ? 64 : 0;
        if (D < 64) {
            for (int d = 1; d < D; d++) {
                uint64_t m = __ballot(cz < czmin + d);
                if (lane == d) tab = __popcll(m);
            }
        }
        uint64_t m_in[P1_RMAX], m_near[P1_RMAX];
    #pragma unroll
        for (int k = 0; k < P1_RMAX; k++) m_in[k] = m_near[k] = 0ull;

        /* interval work for up to 64 queued atoms, one per lane */
        auto process = [&](int count) {
            const int id = (lane < count) ? queue[lane] : -1;
            float4 a = make_float4(0.f, 0.f, 0.f, 0.f), ax = make_float4(0.f, 0.f, 0.f, 0.f);
            int acz = czmin, acy = 0;
            if (id >= 0) {
                a = sx[id];
                ax = aux[id];
                const int packed = __float_as_int(ax.z);
                acy = packed >> 16;
                acz = packed & 0xFFFF;
            }
            /* voxels whose z cell admits the atom's z cell: lanes [La, Ra] */
            int La, Ra;
            if (D < 64) {
                int j1 = acz - 1 - czmin, j2 = acz + 2 - czmin;
                int t1 = __shfl(tab, iclamp(j1, 0, 63), 64);
                int t2 = __shfl(tab, iclamp(j2, 0, 63), 64);
                La = (j1 <= 0) ? 0 : (j1 >= D ? 64 : t1);
                Ra = ((j2 >= D) || (acz == gn - 1)) ? 63 : ((j2 <= 0) ? -1 : t2 - 1);
            } else {
                const int alo = iclamp(cz - 1, 0, gn - 1), ahi = iclamp(cz + 1, 0, gn - 1);
                int lo_i = 0, hi_i = 64;
                for (int it = 0; it < 7; it++) {
                    int mid = (lo_i + hi_i) >> 1;
                    int v = __shfl(ahi, min(mid, 63), 64);
                    if (lo_i < hi_i) {
                        if (v >= acz) hi_i = mid;
                        else lo_i = mid + 1;
                    }
                }
                La = lo_i;
                lo_i = 0;
                hi_i = 64;
                for (int it = 0; it < 7; it++) {
                    int mid = (lo_i + hi_i) >> 1;
                    int v = __shfl(alo, min(mid, 63), 64);
                    if (lo_i < hi_i) {
                        if (v > acz) hi_i = mid;
                        else lo_i = mid + 1;
                    }
                }
                Ra = lo_i - 1;
            }
            if (id < 0) return;
            const int zlo = z0 + La, zhi = min(z0 + Ra, zmax);
            if (zlo > zhi) return;
            const float sN = ax.x, sD = ax.y;
            const float ex = a.x - px;
            const float X = ex * ex;
            const float zc = (a.z - G.oz) * inv_dx;
            RowAtom A;
            A.az = a.z;
            A.dx = dx;
            A.oz = G.oz;
            const int zb0 = iclamp((int)floorf(zc), zlo, zhi);
    #pragma unroll
            for (int k = 0; k < P1_RMAX; k++) {
                if (k >= ny) break;
                if (acy < cyl[k] || acy > cyh[k]) continue; /* row's y cell does not admit the atom */
                float ey = a.y - py[k];
                float Y = ey * ey;
                A.S = X + Y;
                if (!(A.S < sN)) continue;
                int zb = zb0;
                if (zb + 1 <= zhi && A.Zf(zb + 1) < A.Zf(zb)) zb = zb + 1;
                int L, Rr;
                row_interval(A, sN, zlo, zhi, zb, zc, inv_dx, L, Rr);
                if (L > Rr) continue;
                m_near[k] |= bit_range(L - z0, Rr - z0);
                if (!(A.S < sD)) continue;
                row_interval(A, sD, L, Rr, zb, zc, inv_dx, L, Rr);
                if (L <= Rr) m_in[k] |= bit_range(L - z0, Rr - z0);
            }
        };

        /* candidates that can reach one of the rows (xy distance below r + probe) are queued */
        int qn = 0;
        const float pymin = py[0], pymax = py[ny - 1];
        for (int base = 0; base < total; base += 64) {
            const int kk = base + lane;
            int id = -1;
    #pragma unroll
            for (int q = 0; q < 12; q++)
                if (kk >= pre[q] && kk < pre[q + 1]) id = rs[q] + (kk - pre[q]);
            bool keep = false;
            if (id >= 0) {
                const float4 a = sx[id];
                const float sN = aux[id].x;
                const float ex = a.x - px;
                const float X = ex * ex;
                const float yy = fminf(fmaxf(a.y, pymin), pymax); /* nearest row y */
                const float ey = a.y - yy;
                const float Y = ey * ey;
                keep = (X + Y) < sN * 1.0001f + 1e-6f; /* conservative: exact tests follow */
            }
            const uint64_t km = __ballot(keep);
            if (keep) queue[qn + __popcll(km & ((1ull << lane) - 1ull))] = id;
            qn += __popcll(km);
            wave_lds_sync();
            if (qn >= 64) {
                process(64);
                wave_lds_sync();
                const int rest = qn - 64;
                int moved = (lane < rest) ? queue[64 + lane] : 0;
                wave_lds_sync();
                if (lane < rest) queue[lane] = moved;
                wave_lds_sync();
                qn = rest;
            }
        }
        if (qn > 0) process(qn);

    #pragma unroll
        for (int k = 0; k < P1_RMAX; k++) {
            m_in[k] = wave_or_u64(m_in[k]);
            m_near[k] = wave_or_u64(m_near[k]);
        }
        if (lane < ny) {
            uint64_t mi = 0ull, mn = 0ull;
    #pragma unroll
            for (int k = 0; k < P1_RMAX; k++)
                if (lane == k) {
                    mi = m_in[k];
                    mn = m_near[k];
                }
            const int y = y0 + lane;
            uint64_t lab = (zmax >= z0 && y < g - 1) ? bit_range(0, zmax - z0) : 0ull;
            mi &= lab;
            mn &= lab;
            size_t widx = ((size_t)x * g + y) * W + w;
            outm[widx] = valid & ~(mi | mn);
            shellm[widx] = mn & ~mi;
        }
    }
}

/* OR bits [L, R] (absolute z) into a row's words in LDS */
__device__ __forceinline__ void lds_or_range(uint64_t *row, int L, int R)
{
    for (int w = L >> 6; w <= (R >> 6); w++) {
        const int z0 = w << 6;
        atomicOr((unsigned long long *)&row[w], (unsigned long long)bit_range(max(L, z0) - z0, min(R, z0 + 63) - z0));
    }
}

/* Field pass 1 by scatter into LDS: one block per slab of up to P1S_MAXROWS consecutive y rows of
   one x plane, all z.  Each candidate atom (x cells cx-1..cx+1 times the slab's y cells, every z
   cell) takes the rows whose y cell admits it and, lane by lane, ORs its exact near and inside
   z-intervals (the ones of k_pass1_rows) into the slab's two masks.  The z-intervals are clipped
   to the voxels whose z cell admits the atom, so each voxel sees exactly the atoms of its 27
   neighbour cells, as in probeIntersection (Kernels.cu:394-475); inside wins over near, which
   makes the result independent of the order of the ORs. */
#define P1S_MAXROWS 256   /* rows per slab block at most (row indexes in the queue are 8-bit) */
#define P1S_DEFROWS 64    /* rows per slab block when no tuned count is set */
#define P1S_MAXCOL 256
/* a wave takes its queued atoms as (atom, row) items when K1 * (item rounds of 64) < K2 * (the
   largest row count of one atom), i.e. when walking atoms would leave most lanes idle */
#ifndef P1_FLAT_K1
#define P1_FLAT_K1 3
#define P1_FLAT_K2 2
#endif
__global__ void __launch_bounds__(256) k_pass1_slab(QsBatch B, const uint32_t *__restrict__ cstart_all,
                                                    const float4 *__restrict__ sx, const float4 *__restrict__ aux,
                                                    uint64_t *__restrict__ outm_all, uint64_t *__restrict__ shellm_all,
                                                    uint64_t *__restrict__ negm_all, uint64_t *__restrict__ zerom_all,
                                                    uint8_t *__restrict__ tflag, uint8_t tmark)
{
    extern __shared__ __align__(16) uint64_t p1s_dyn[];
    __shared__ uint32_t colpre[P1S_MAXCOL + 1], colst[P1S_MAXCOL];
    __shared__ int q_id[4][P1_QCAP], q_rows[4][P1_QCAP], q_z[4][P1_QCAP];
    const int lane = threadIdx.x & 63, wid = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    uint32_t lb;
    const QsMol &M = B.mols[batch_mol(B, blockIdx.x, lb)];
    const QsGrid G = M.G;
    const int g = G.g, gn = G.gn, W = (int)M.W, YB = M.slab_rows;
    const uint32_t *__restrict__ cstart = cstart_all + M.cell_off;
    uint64_t *__restrict__ outm = outm_all + M.word_off;
    uint64_t *__restrict__ shellm = shellm_all + M.word_off;
    const int nyb = (g + YB - 1) / YB;
    const int x = M.p1_lo + (int)(lb / (uint32_t)nyb);
    const int y0 = (int)(lb % (uint32_t)nyb) * YB;
    if (x >= M.p1_hi) return; /* whole block */
    const int ny = min(YB, g - y0);
    const size_t wbase = ((size_t)x * g + y0) * W;
    uint64_t *__restrict__ negm = negm_all ? negm_all + M.word_off : nullptr;
    uint64_t *__restrict__ zerom = zerom_all ? zerom_all + M.word_off : nullptr;
    if (x >= g - 1) {
        /* the last plane is not labelled: outside */
        for (int t = threadIdx.x; t < ny * W; t += 256) {
            const int w = t % W;
            outm[wbase + t] = bits_below(g, w * 64);
            shellm[wbase + t] = 0ull;
            if (negm) {
                negm[wbase + t] = 0ull;
                zerom[wbase + t] = 0ull;
            }
        }
        return;
    }
    uint64_t *nearw = p1s_dyn, *insw = p1s_dyn + (size_t)YB * W;
    /* yfirst[c] (zfirst[c]) = first y (z) whose neighbour cell is >= c, g if none; c = 0..gn+1 */
    int *yfirst = (int *)(p1s_dyn + 2 * (size_t)YB * W), *zfirst = yfirst + (gn + 2);
    for (int t = threadIdx.x; t < ny * W; t += 256) nearw[t] = insw[t] = 0ull;
    for (int t = threadIdx.x; t < gn + 2; t += 256) yfirst[t] = zfirst[t] = g;
    __syncthreads();
    const float dx = G.dx, inv_dx = 1.0f / G.dx;
    for (int i = threadIdx.x; i < g; i += 256) {
        const int cy = min(cell_of(grid_pos(G.oy, i, dx), G.oy, G.dxn), gn + 1);
        const int cyp = i ? min(cell_of(grid_pos(G.oy, i - 1, dx), G.oy, G.dxn), gn + 1) : -1;
        for (int c = cyp + 1; c <= cy; c++) yfirst[c] = i;
        const int cz = min(cell_of(grid_pos(G.oz, i, dx), G.oz, G.dxn), gn + 1);
        const int czp = i ? min(cell_of(grid_pos(G.oz, i - 1, dx), G.oz, G.dxn), gn + 1) : -1;
        for (int c = czp + 1; c <= cz; c++) zfirst[c] = i;
    }
    /* labelled rows of the slab: y <= g - 2 */
    const int ylast = min(y0 + ny - 1, g - 2);
    const float px = grid_pos(G.ox, x, dx);
    const int cx = cell_of(px, G.ox, G.dxn);
    const int xa = iclamp(cx - 1, 0, gn - 1), xb = iclamp(cx + 1, 0, gn - 1);
    const int ya = iclamp(cell_of(grid_pos(G.oy, y0, dx), G.oy, G.dxn) - 1, 0, gn - 1);
    const int yb = iclamp(cell_of(grid_pos(G.oy, max(ylast, y0), dx), G.oy, G.dxn) + 1, 0, gn - 1);
    const int nyc = yb - ya + 1, ncol = (xb - xa + 1) * nyc;
    /* candidate columns: every z cell of (xc, yc), contiguous in the cell-ordered atoms */
    uint32_t len = 0;
    if ((int)threadIdx.x < ncol) {
        const int xc = xa + (int)threadIdx.x / nyc, yc = ya + (int)threadIdx.x % nyc;
        const size_t h0 = ((size_t)xc * gn + yc) * gn;
        colst[threadIdx.x] = cstart[h0];
        len = cstart[h0 + gn] - cstart[h0];
    }
    /* exclusive prefix of the column lengths (ncol <= 256) */
    {
        uint32_t v = len;
#pragma unroll
        for (int o = 1; o < 64; o <<= 1) {
            const uint32_t t = __shfl_up(v, o, 64);
            if (lane >= o) v += t;
        }
        __shared__ uint32_t wsum[4];
        if (lane == 63) wsum[wid] = v;
        __syncthreads();
        uint32_t add = 0;
        for (int k = 0; k < wid; k++) add += wsum[k];
        colpre[threadIdx.x + 1] = add + v;
        if (threadIdx.x == 0) colpre[0] = 0;
    }
    __syncthreads();
    const int total = (int)colpre[ncol];
    const int rows_hi = ylast - y0; /* rows 0..rows_hi of the slab are labelled */
    int *qid = q_id[wid], *qrows = q_rows[wid], *qz = q_z[wid];

    /* One (atom, row) pair: the atom's exact near and inside z-intervals in the row, ORed into the
       slab's masks (the ORs are order-independent). */
    auto row_item = [&](int yr, float X, float sN, float sD, float zc, float az, float ay, int zr, int zb) {
        const int zlo = zr & 0xFFFF, zhi = zr >> 16;
        RowAtom A;
        A.az = az;
        A.dx = dx;
        A.oz = G.oz;
        const float ey = ay - grid_pos(G.oy, y0 + yr, dx);
        const float Y = ey * ey;
        A.S = X + Y;
        if (!(A.S < sN)) return;
        int L, Rr;
        row_interval(A, sN, zlo, zhi, zb, zc, inv_dx, L, Rr);
        if (L > Rr) return;
        lds_or_range(nearw + (size_t)yr * W, L, Rr);
        if (!(A.S < sD)) return;
        row_interval(A, sD, L, Rr, zb, zc, inv_dx, L, Rr);
        if (L <= Rr) lds_or_range(insw + (size_t)yr * W, L, Rr);
    };
    /* The wave's queued atoms, lane k < count holding atom k.  Each lane walks its own atom's rows
       r0..r1, unless the row counts are so uneven that most lanes would idle: the atoms' rows are
       then items pre_k .. pre_k + r1 - r0 that the lanes take 64 at a time, each with its atom's
       values from the lane that holds it.  Called by the whole wave. */
    auto process = [&](int count) {
        int rr = 0, nrows = 0, zr = 0, zb = 0;
        float X = 0.0f, sN = 0.0f, sD = 0.0f, zc = 0.0f, az = 0.0f, ay = 0.0f;
        if (lane < count) {
            const int id = qid[lane];
            rr = qrows[lane];
            zr = qz[lane];
            nrows = (rr >> 8) - (rr & 0xFF) + 1;
            const float4 a = sx[id], ax = aux[id];
            sN = ax.x;
            sD = ax.y;
            const float ex = a.x - px;
            X = ex * ex;
            zc = (a.z - G.oz) * inv_dx;
            az = a.z;
            ay = a.y;
            const int zlo = zr & 0xFFFF, zhi = zr >> 16;
            RowAtom A;
            A.az = a.z;
            A.dx = dx;
            A.oz = G.oz;
            zb = iclamp((int)floorf(zc), zlo, zhi);
            if (zb + 1 <= zhi && A.Zf(zb + 1) < A.Zf(zb)) zb = zb + 1;
        }
        int pre = nrows;
#pragma unroll
        for (int d = 1; d < 64; d <<= 1) {
            const int t = __shfl_up(pre, d, 64);
            if (lane >= d) pre += t;
        }
        const int nitems = __shfl(pre, 63, 64);
        pre -= nrows; /* exclusive */
        int maxrows = nrows;
#pragma unroll
        for (int d = 1; d < 64; d <<= 1) maxrows = max(maxrows, __shfl_xor(maxrows, d, 64));
        if (P1_FLAT_K1 * ((nitems + 63) >> 6) >= P1_FLAT_K2 * maxrows) {
            if (lane < count)
                for (int yr = rr & 0xFF; yr <= (rr >> 8); yr++) row_item(yr, X, sN, sD, zc, az, ay, zr, zb);
            return;
        }
        for (int base = 0; base < nitems; base += 64) {
            const int j = base + lane;
            /* the item's atom: the last lane whose first item is at or below j */
            int o = 0;
#pragma unroll
            for (int step = 32; step >= 1; step >>= 1)
                if (__shfl(pre, o + step, 64) <= j) o += step;
            const float oX = __shfl(X, o, 64), osN = __shfl(sN, o, 64), osD = __shfl(sD, o, 64);
            const float ozc = __shfl(zc, o, 64), oaz = __shfl(az, o, 64), oay = __shfl(ay, o, 64);
            const int orr = __shfl(rr, o, 64), ozr = __shfl(zr, o, 64), ozb = __shfl(zb, o, 64), opre = __shfl(pre, o, 64);
            if (j < nitems) row_item((orr & 0xFF) + (j - opre), oX, osN, osD, ozc, oaz, oay, ozr, ozb);
        }
    };

    int qn = 0;
    if (rows_hi >= 0) {
        for (int base = wid * 64; base < total; base += 256) {
            const int k = base + lane;
            bool keep = false;
            int id = 0, rows = 0, zr = 0;
            if (k < total) {
                int lo = 0, hi = ncol - 1;
                while (lo < hi) {
                    const int mid = (lo + hi + 1) >> 1;
                    if ((int)colpre[mid] <= k) lo = mid;
                    else hi = mid - 1;
                }
                id = (int)colst[lo] + (k - (int)colpre[lo]);
                const float4 a = sx[id], ax = aux[id];
                const float ex = a.x - px;
                const float X = ex * ex;
                if (X < ax.x) {
                    const int packed = __float_as_int(ax.z);
                    const int acy = packed >> 16, acz = packed & 0xFFFF;
                    /* rows whose y cell admits the atom: cy >= acy - 1, and cy <= acy + 1 unless acy is the last cell */
                    int ylo = yfirst[max(acy - 1, 0)];
                    int yhi = (acy == gn - 1) ? g - 1 : yfirst[acy + 2] - 1;
                    /* rows the atom can reach in y (a superset with a row of margin: the exact
                       test is per row, so the hardware root serves) */
                    const float h = __builtin_amdgcn_sqrtf(ax.x - X);
                    ylo = max(ylo, (int)floorf((a.y - h - G.oy) * inv_dx) - 1);
                    yhi = min(yhi, (int)ceilf((a.y + h - G.oy) * inv_dx) + 1);
                    const int r0 = max(ylo - y0, 0), r1 = min(yhi - y0, rows_hi);
                    /* voxels whose z cell admits the atom, labelled ones only (z <= g - 2) */
                    const int zlo = zfirst[max(acz - 1, 0)];
                    const int zhi = min((acz == gn - 1) ? g - 1 : zfirst[acz + 2] - 1, g - 2);
                    keep = r0 <= r1 && zlo <= zhi;
                    rows = r0 | (r1 << 8);
                    zr = zlo | (zhi << 16);
                }
            }
            const uint64_t km = __ballot(keep);
            if (keep) {
                const int slot = qn + __popcll(km & ((1ull << lane) - 1ull));
                qid[slot] = id;
                qrows[slot] = rows;
                qz[slot] = zr;
            }
            qn += __popcll(km);
            wave_lds_sync();
            if (qn >= 64) {
                process(64);
                wave_lds_sync();
                const int rest = qn - 64;
                int m0 = 0, m1 = 0, m2 = 0;
                if (lane < rest) {
                    m0 = qid[64 + lane];
                    m1 = qrows[64 + lane];
                    m2 = qz[64 + lane];
                }
                wave_lds_sync();
                if (lane < rest) {
                    qid[lane] = m0;
                    qrows[lane] = m1;
                    qz[lane] = m2;
                }
                wave_lds_sync();
                qn = rest;
            }
        }
        if (qn > 0) process(qn);
    }
    __syncthreads();
    for (int t = threadIdx.x; t < ny * W; t += 256) {
        const int yr = t / W, w = t - yr * W;
        const uint64_t valid = bits_below(g, w * 64);
        uint64_t mi = 0ull, mn = 0ull;
        if (yr <= rows_hi) {
            mi = insw[t];
            mn = nearw[t];
        }
        outm[wbase + t] = valid & ~(mi | mn);
        const uint64_t sh = mn & ~mi;
        shellm[wbase + t] = sh;
        if (negm) {
            /* the sign words of a voxel without a shell voxel in its pass-2 tile: inside (-dx)
               is negative, outside (the probe radius) is not, no field value is zero; field
               pass 2 rewrites the tiles that hold a shell voxel, which it finds marked here */
            negm[wbase + t] = valid & (mi | mn);
            zerom[wbase + t] = 0ull;
            if (sh && tflag && M.tile_off != ~0u && x >= M.p2_lo && x < M.p2_hi) {
                const uint32_t nbz = (uint32_t)(g + P2_TZ - 1) / P2_TZ, nby = (uint32_t)(g + P2_TY - 1) / P2_TY;
                const uint32_t tyx = ((uint32_t)(x - M.p2_lo) / P2R_TX) * nby + (uint32_t)(y0 + yr) / P2_TY;
                if ((uint32_t)sh) tflag[M.tile_off + tyx * nbz + 2 * w] = tmark;
                if (sh >> 32) tflag[M.tile_off + tyx * nbz + 2 * w + 1] = tmark;
            }
        }
    }
}

/* per-voxel form (27 cells per voxel, as the reference); cross-check of k_pass1_rows */
__global__ void __launch_bounds__(256) k_pass1_direct(QsBatch B, const int2 *__restrict__ cse_all,
                                                      const float4 *__restrict__ sx, uint64_t *__restrict__ outm_all,
                                                      uint64_t *__restrict__ shellm_all)
{
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    uint32_t lb;
    const QsMol &M = B.mols[batch_mol(B, blockIdx.x, lb)];
    const QsGrid G = M.G;
    const int g = G.g, W = (int)M.W;
    const int2 *__restrict__ cse = cse_all + M.cell_off;
    uint64_t *__restrict__ outm = outm_all + M.word_off;
    uint64_t *__restrict__ shellm = shellm_all + M.word_off;
    const int nby = (g + 3) / 4;
    const int w = (int)(lb % (uint32_t)W);
    const int y = (int)((lb / (uint32_t)W) % (uint32_t)nby) * 4 + wid;
    const int x = M.p1_lo + (int)(lb / ((uint32_t)W * nby));
    if (y >= g || x >= M.p1_hi) return;
    const int z = w * 64 + lane;
    uint8_t res = QS_OUT;
    if (x < g - 1 && y < g - 1 && z < g - 1) {
        const float dx = G.dx;
        float px = grid_pos(G.ox, x, dx), py = grid_pos(G.oy, y, dx), pz = grid_pos(G.oz, z, dx);
        int cx = cell_of(px, G.ox, G.dxn), cy = cell_of(py, G.oy, G.dxn), cz = cell_of(pz, G.oz, G.dxn);
        const int gn = G.gn;
        bool near = false, inside = false;
        for (int a = -1; a <= 1 && !inside; a++) {
            int ix = iclamp(cx + a, 0, gn - 1);
            for (int b = -1; b <= 1 && !inside; b++) {
                int iy = iclamp(cy + b, 0, gn - 1);
                for (int c = -1; c <= 1 && !inside; c++) {
                    int iz = iclamp(cz + c, 0, gn - 1);
                    int2 se = cse[(gn * gn * ix) + (gn * iy) + iz];
                    if (se.x >= QS_EMPTY) continue;
                    for (int id = se.x; id < se.y; id++) {
                        float4 at = sx[id];
                        float d = dist3(at.x, at.y, at.z, px, py, pz);
                        if (d < at.w - dx) {
                            inside = true;
                            break;
                        }
                        if (d < (QS_PROBE + at.w)) near = true;
                    }
                }
            }
        }
        res = inside ? QS_IN : (near ? QS_SHELL : QS_OUT);
    }
    uint64_t mo = __ballot(z < g && res == QS_OUT);
    uint64_t ms = __ballot(z < g && res == QS_SHELL);
    if (lane == 0) {
        size_t widx = ((size_t)x * g + y) * W + w;
        outm[widx] = mo;
        shellm[widx] = ms;
    }
}

/* ------------------------------------------------------------------------------------------ */
/* field pass 2 (distanceFieldRefine Kernels.cu:478-545)                                        */
/*
 * For a shell voxel v the reference takes the minimum, over "outside" voxels q of the (2r+1)^3
 * window clamped to the grid, of the single-precision distance between their world positions.
 * Every such float distance lies within delta of dx*sqrt(n) for the integer n = |q - v|^2, and
 * norms that differ give distances further apart than 2*delta, so the minimum is attained among
 * the outside voxels whose norm lies within [nmin, nmax(nmin)], nmin being the exact smallest
 * norm.  The tile kernel finds nmin with three separable integer passes in LDS (z by bit scan
 * of the outside mask, then y, then x) and evaluates the reference's float distance only for
 * those candidates.  The result is the reference's value bit for bit.
 */


/* bits [s, s + len) of a row of W mask words (bit i <-> z = s + i), zero outside [0, 64 W) */
__device__ __forceinline__ uint64_t row_bits(const uint64_t *__restrict__ rp, int W, int s, int len)
{
    int w0 = s >> 6; /* floor */
    int sh = s & 63;
    uint64_t a = (w0 >= 0 && w0 < W) ? rp[w0] : 0ull;
    uint64_t v = a >> sh;
    if (sh) {
        uint64_t b = (w0 + 1 >= 0 && w0 + 1 < W) ? rp[w0 + 1] : 0ull;
        v |= b << (64 - sh);
    }
    if (len < 64) v &= (1ull << len) - 1ull;
    return v;
}

size_t qs_pass2_lds_bytes(int r)
{
    size_t LX = P2_TX + 2 * r, LY = P2_TY + 2 * r;
    size_t dt = (3 * r * r < 255) ? 1 : 2;
    return LX * LY * 8 + P2_TX * P2_TY * 4 + LX * LY * P2_TZ * dt + LX * P2_TY * P2_TZ * dt;
}

/* sign (field < 0) and zero (field == 0) bits of the field, one ballot per wave-half: lanes 0-31
   are the 32 voxels (gx, gy, tz0..tz0+31), lanes 32-63 the same z range of row gy + 1 */
__device__ __forceinline__ void emit_sign_bits(uint32_t *__restrict__ negm, uint32_t *__restrict__ zerom, int W, int g,
                                               int gx, int gy, int tz0, bool valid, float res)
{
    const int lane = threadIdx.x & 63;
    uint64_t n = __ballot(valid && res < 0.0f);
    uint64_t zr = __ballot(valid && res == 0.0f);
    if ((lane & 31) == 0 && valid) {
        size_t w32 = ((size_t)gx * g + gy) * (2 * (size_t)W) + (tz0 >> 5);
        negm[w32] = (uint32_t)(lane ? (n >> 32) : n);
        zerom[w32] = (uint32_t)(lane ? (zr >> 32) : zr);
    }
}

template <typename DT>
__global__ void __launch_bounds__(256) k_pass2_tile(QsBatch B, const uint64_t *__restrict__ outm_all,
                                                    const uint64_t *__restrict__ shellm_all, const uint32_t *__restrict__ spre_all,
                                                    float *__restrict__ field,
                                                    uint32_t *__restrict__ negm_all, uint32_t *__restrict__ zerom_all,
                                                    const int *__restrict__ offs, const int *__restrict__ norm_start,
                                                    const int *__restrict__ enum_end)
{
    extern __shared__ __align__(16) uint8_t smem[];
    __shared__ int s_flags;
    __shared__ uint32_t sbase[P2_TX * P2_TY]; /* shell rank of each own row's first tile voxel */
    constexpr int INF = (sizeof(DT) == 1) ? 0xFF : 0xFFFF;
    uint32_t lb;
    const QsMol &M = B.mols[batch_mol(B, blockIdx.x, lb)];
    const QsGrid G = M.G;
    const int r = G.r, g = G.g, W = (int)M.W;
    const uint64_t *__restrict__ outm = outm_all + M.word_off;
    const uint64_t *__restrict__ shellm = shellm_all + M.word_off;
    const uint32_t *__restrict__ spre = spre_all + M.word_off;
    uint32_t *__restrict__ negm = negm_all + 2 * M.word_off;
    uint32_t *__restrict__ zerom = zerom_all + 2 * M.word_off;
    const int x0 = M.p2_lo, x1 = M.p2_hi;
    const uint32_t nbx = (g + P2_TZ - 1) / P2_TZ, nby = (g + P2_TY - 1) / P2_TY;
    const int bidx = (int)(lb % nbx), bidy = (int)((lb / nbx) % nby), bidz = (int)(lb / (nbx * nby));
    const int LX = P2_TX + 2 * r, LY = P2_TY + 2 * r, LZ = P2_TZ + 2 * r;
    uint64_t *rows = (uint64_t *)smem;
    uint32_t *own = (uint32_t *)(smem + (size_t)LX * LY * 8);
    DT *DZ = (DT *)(smem + (size_t)LX * LY * 8 + P2_TX * P2_TY * 4);
    DT *DY = DZ + (size_t)LX * LY * P2_TZ;
    const int tz0 = bidx * P2_TZ, ty0 = bidy * P2_TY, tx0 = x0 + bidz * P2_TX;
    if (threadIdx.x == 0) s_flags = 0;
    __syncthreads();

    /* 1. outside bits of the haloed tile; shell bits of the tile */
    int fl = 0;
    for (int row = threadIdx.x; row < LX * LY; row += 256) {
        int lx = row / LY, ly = row - lx * LY;
        int gx = tx0 - r + lx, gy = ty0 - r + ly;
        uint64_t m = 0ull;
        /* planes outside [p1_lo, p1_hi) are not stored (and are read by no output voxel) */
        if (gx >= M.p1_lo && gx < M.p1_hi && gy >= 0 && gy < g)
            m = row_bits(outm + ((size_t)gx * g + gy) * W, W, tz0 - r, LZ);
        rows[row] = m;
        if (m) fl |= 2;
    }
    for (int t = threadIdx.x; t < P2_TX * P2_TY; t += 256) {
        int tx = t / P2_TY, ty = t - tx * P2_TY;
        int gx = tx0 + tx, gy = ty0 + ty;
        uint32_t m = 0u;
        if (gx < x1 && gx < g && gy < g) m = (uint32_t)row_bits(shellm + ((size_t)gx * g + gy) * W, W, tz0, P2_TZ);
        own[t] = m;
        if (m) {
            fl |= 1;
            const size_t wi = ((size_t)gx * g + gy) * W + (tz0 >> 6);
            sbase[t] = spre[wi] + (uint32_t)__popcll(shellm[wi] & bits_below(tz0 & 63, 0));
        }
    }
    if (fl) atomicOr(&s_flags, fl);
    __syncthreads();
    const int flags = s_flags;

    const int ty = threadIdx.x / P2_TZ, tz = threadIdx.x - ty * P2_TZ;
    const int gy = ty0 + ty, gz = tz0 + tz;
    const bool vyz = gy < g && gz < g;
    const float dx = G.dx;
    if (flags != 3) {
        /* no shell voxel, or no outside voxel within reach: outside -> probe, else -dx */
        for (int tx = 0; tx < P2_TX; tx++) {
            int gx = tx0 + tx;
            if (gx >= x1 || gx >= g) break;
            bool out = (rows[(size_t)(tx + r) * LY + (ty + r)] >> (tz + r)) & 1ull;
            float res = out ? QS_PROBE : -dx;
            const uint32_t om = own[tx * P2_TY + ty];
            if (vyz && ((om >> tz) & 1u)) field[sbase[tx * P2_TY + ty] + (uint32_t)__popc(om & ((1u << tz) - 1u))] = res;
            emit_sign_bits(negm, zerom, W, g, gx, gy, tz0, vyz, res);
        }
        return;
    }

    /* 2. z: squared distance to the nearest outside voxel within |c| <= r along each row */
    for (int w = threadIdx.x; w < LX * LY * P2_TZ; w += 256) {
        int row = w / P2_TZ, z = w - row * P2_TZ;
        uint64_t m = rows[row];
        int best = INF;
        if (m) {
            int c = z + r;
            uint64_t up = m >> c;
            if (up) {
                int d = __builtin_ctzll(up);
                if (d <= r) best = d * d;
            }
            uint64_t dn = m & ((2ull << c) - 1ull);
            if (dn) {
                int d = c - (63 - __builtin_clzll(dn));
                if (d <= r) best = min(best, d * d);
            }
        }
        DZ[w] = (DT)best;
    }
    __syncthreads();

    /* 3. y */
    for (int w = threadIdx.x; w < LX * P2_TY * P2_TZ; w += 256) {
        int lx = w / (P2_TY * P2_TZ);
        int rem = w - lx * (P2_TY * P2_TZ);
        int yy = rem / P2_TZ, z = rem - yy * P2_TZ;
        int best = INF;
        for (int b = -r; b <= r; b++) {
            int v = DZ[((size_t)lx * LY + (yy + r + b)) * P2_TZ + z];
            if (v != INF) best = min(best, v + b * b);
        }
        DY[w] = (DT)best;
    }
    __syncthreads();

    /* 4. x, then the field value of each output voxel */
    const float vy = grid_pos(G.oy, gy, dx), vz = grid_pos(G.oz, gz, dx);
    for (int tx = 0; tx < P2_TX; tx++) {
        int gx = tx0 + tx;
        if (gx >= x1 || gx >= g) break;
        bool out = (rows[(size_t)(tx + r) * LY + (ty + r)] >> (tz + r)) & 1ull;
        bool shell = (own[tx * P2_TY + ty] >> tz) & 1u;
        float res;
        if (!vyz || out) {
            res = QS_PROBE;
        } else if (!shell) {
            res = -dx;
        } else {
            int nmin = INF;
            for (int a = -r; a <= r; a++) {
                int v = DY[((size_t)(tx + r + a) * P2_TY + ty) * P2_TZ + tz];
                if (v != INF) nmin = min(nmin, v + a * a);
            }
            if (nmin == INF) {
                res = -dx;
            } else {
                const float vx = grid_pos(G.ox, gx, dx);
                float min_dist = 100000.0f;
                if (enum_end[nmin] == norm_start[nmin + 1]) {
                    /* only norm nmin can attain the minimum: trace the minimisers back through the
                       separable passes (a with a^2 + DY = nmin, then b with b^2 + DZ = DY, then
                       c = +-sqrt(DZ) where the voxel is outside) */
                    for (int a = -r; a <= r; a++) {
                        int dy = DY[((size_t)(tx + r + a) * P2_TY + ty) * P2_TZ + tz];
                        if (dy == INF || dy + a * a != nmin) continue;
                        for (int b = -r; b <= r; b++) {
                            int dz = DZ[((size_t)(tx + r + a) * LY + (ty + r + b)) * P2_TZ + tz];
                            if (dz == INF || dz + b * b != dy) continue;
                            int c = (int)sqrtf((float)dz);
                            uint64_t m = rows[(size_t)(tx + r + a) * LY + (ty + r + b)];
                            const float qx = grid_pos(G.ox, gx + a, dx), qy = grid_pos(G.oy, gy + b, dx);
                            if ((m >> (tz + r + c)) & 1ull)
                                min_dist = fminf(dist3(qx, qy, grid_pos(G.oz, gz + c, dx), vx, vy, vz), min_dist);
                            if (c && ((m >> (tz + r - c)) & 1ull))
                                min_dist = fminf(dist3(qx, qy, grid_pos(G.oz, gz - c, dx), vx, vy, vz), min_dist);
                        }
                    }
                } else {
                    for (int k = norm_start[nmin]; k < enum_end[nmin]; k++) {
                        int o = offs[k];
                        int a = (int)(int8_t)(o & 0xFF), b = (int)(int8_t)((o >> 8) & 0xFF), c = (int)(int8_t)((o >> 16) & 0xFF);
                        uint64_t m = rows[(size_t)(tx + r + a) * LY + (ty + r + b)];
                        if (!((m >> (tz + r + c)) & 1ull)) continue;
                        float d = dist3(grid_pos(G.ox, gx + a, dx), grid_pos(G.oy, gy + b, dx), grid_pos(G.oz, gz + c, dx),
                                        vx, vy, vz);
                        min_dist = fminf(d, min_dist);
                    }
                }
                res = (min_dist < 999.0f) ? (QS_PROBE - min_dist) : -dx;
            }
        }
        if (vyz && shell)
            field[sbase[tx * P2_TY + ty] + (uint32_t)__popc(own[tx * P2_TY + ty] & ((1u << tz) - 1u))] = res;
        emit_sign_bits(negm, zerom, W, g, gx, gy, tz0, vyz, res);
    }
}

/* squared distance from bit z + R of row m to its nearest set bit within R; z < 32 and 2R < 32,
   so bits z .. z + 2R are the low bits of one 32-bit funnel shift.  Branch free: bits above the
   window are cleared and a sentinel stands at bit 31, so a row with no set bit within R gives a
   distance of at least 31 - R and a square of at least (31 - R)^2 >= 484, above INF = 255 (every
   in-window norm is at most 3R^2 <= 243 for R <= 9); callers clamp it to INF. */
template <int R>
__device__ __forceinline__ uint32_t zdist2(uint64_t m, int z)
{
    static_assert(2 * R < 31 && (31 - R) * (31 - R) > 255, "window wider than a funnel shift");
    const uint32_t w = __builtin_amdgcn_alignbit((uint32_t)(m >> 32), (uint32_t)m, (uint32_t)z); /* bit k = m bit z + k */
    const uint32_t ws = (w & ((1u << (2 * R + 1)) - 1u)) | 0x80000000u;
    const uint32_t up = (uint32_t)__builtin_ctz(ws >> R);               /* nearest set bit at or above the centre */
    const uint32_t dn = (uint32_t)__builtin_clz((ws << (31 - R)) | 1u); /* at or below it (31: none) */
    const uint32_t d = min(up, dn);
    return d * d;
}

/* Field pass 2 for a fixed window half-width R (probeIntersection's second pass,
   Kernels.cu:478-545), one 16 x 8 x 32 tile per block:
   - the tile's own shell and outside bits; a tile without a shell voxel writes its sign words only;
   - the outside rows of the tile and its R halo in LDS;
   - the exact minimal squared norm n_min of every shell voxel by separable minima: z distances
     from the row bits, y minima (with the masks of the attaining y offsets) for the listed
     columns only, x minima two outputs per register;
   - the exact float distance of the minimisers, traced back through the masks (or enumerated
     from a norm-sorted table where another norm could tie), rooted once;
   - the field of shell voxels (the others follow from the masks) and every voxel's sign bits. */

/* entry it of a per-plane bit list: the plane lx with cpre[lx] <= it < cpre[lx + 1] (binary
   search) and the z of the (it - cpre[lx])-th set bit of cmask[lx] */
template <int LX>
__device__ __forceinline__ void list_column(const uint32_t *cpre, const uint32_t *cmask, uint32_t it, int &lx, int &z)
{
    int lo = 0;
#pragma unroll
    for (int step = 32; step >= 1; step >>= 1)
        if (lo + step < LX && cpre[lo + step] <= it) lo += step;
    lx = lo;
    uint32_t w = cmask[lo], k = it - cpre[lo];
    int zz = 0;
#pragma unroll
    for (int h = 16; h >= 1; h >>= 1) {
        const uint32_t c = (uint32_t)__builtin_popcount(w & ((1u << h) - 1u));
        if (k >= c) {
            k -= c;
            w >>= h;
            zz += h;
        }
    }
    z = zz;
}

template <int R>
__global__ void __launch_bounds__(256) k_pass2_r(QsBatch B, const uint64_t *__restrict__ outm_all,
                                                 const uint64_t *__restrict__ shellm_all, const uint32_t *__restrict__ spre_all,
                                                 float *__restrict__ field, uint32_t *__restrict__ negm_all,
                                                 uint32_t *__restrict__ zerom_all, const int *__restrict__ offs,
                                                 const int *__restrict__ norm_start, const int *__restrict__ enum_end,
                                                 QsNormFlags nf, const uint8_t *__restrict__ tflag, uint8_t tmark, uint32_t xchunk)
{
    constexpr int LX = P2R_TX + 2 * R, LY = P2_TY + 2 * R, LZ = P2_TZ + 2 * R;
    constexpr int INF = 255;
    static_assert(LZ <= 64 && 3 * R * R < INF, "window too wide for this kernel");
    __shared__ uint64_t rows[LX * LY];
    /* ownout: the own rows' outside bits, read only by a tile without a shell voxel; after that
       test the array holds the own rows' shell ranks (sbase) instead */
    __shared__ uint32_t own[P2R_TX * P2_TY], ownout[P2R_TX * P2_TY];
    uint32_t *const sbase = ownout;
    __shared__ uint32_t DYw[LX * P2_TY * P2_TZ / 4];
    uint8_t *DY = (uint8_t *)DYw;
    /* y-pass work list: per halo x plane lx, the z columns (bit z) that need y minima, and the
       exclusive prefix of their counts (cpre[LX] = total) */
    __shared__ uint32_t shx[P2R_TX], cmask[LX], cpre[LX + 1];
    /* norms whose minimisers are traced back (bit n), from nf */
    __shared__ uint32_t nfw[8];
    /* y offsets (bit b + R) attaining each DY value */
    using BMask = typename std::conditional<(2 * R + 1 <= 8), uint8_t,
                                            typename std::conditional<(2 * R + 1 <= 16), uint16_t, uint32_t>::type>::type;
    __shared__ BMask BM[LX * P2_TY * P2_TZ];
    /* per wave, the voxels whose minimisers are still to be traced back (tx | row bit << 4 | z << 5),
       and per own (x, y) row of the tile its sign and zero words */
    __shared__ uint16_t tq[4][128];
    __shared__ uint32_t sneg[P2R_TX * P2_TY], szero[P2R_TX * P2_TY];
    uint32_t lb;
    /* each XCD takes xchunk consecutive tiles at a time (whole rows of z tiles: the y and z halos
       of a tile are then mostly read by tiles of the same XCD, from its L2) */
    const uint32_t bt = xcd_group_rt(blockIdx.x, gridDim.x, xchunk);
    /* a tile without a shell voxel keeps the sign words field pass 1 wrote for it (the launch's
       tile marks are in its block order: an unmarked tile leaves after one load) */
    if (tflag && tflag[bt] != tmark) return;
    const QsMol &M = B.mols[batch_mol(B, bt, lb)];
    const QsGrid G = M.G;
    const int g = G.g, W = (int)M.W;
    const uint64_t *__restrict__ outm = outm_all + M.word_off;
    const uint64_t *__restrict__ shellm = shellm_all + M.word_off;
    const uint32_t *__restrict__ spre = spre_all + M.word_off;
    uint32_t *__restrict__ negm = negm_all + 2 * M.word_off;
    uint32_t *__restrict__ zerom = zerom_all + 2 * M.word_off;
    const int x1 = M.p2_hi;
    const uint32_t nbx = (g + P2_TZ - 1) / P2_TZ, nby = (g + P2_TY - 1) / P2_TY;
    const int tz0 = (int)(lb % nbx) * P2_TZ, ty0 = (int)((lb / nbx) % nby) * P2_TY;
    const int tx0 = M.p2_lo + (int)(lb / (nbx * nby)) * P2R_TX;
    const int ty = threadIdx.x >> 5, tz = threadIdx.x & 31;
    const int gy = ty0 + ty, gz = tz0 + tz;
    const bool vyz = gy < g && gz < g;
    const float dx = G.dx;

    /* the outside rows of the tile and its R halo into LDS; fl: some row has an outside voxel */
    int fl = 0;
    auto load_rows = [&]() {
#pragma unroll
        for (int row = threadIdx.x; row < LX * LY; row += 256) {
            int lx = row / LY, ly = row - lx * LY;
            int gx = tx0 - R + lx, gy_ = ty0 - R + ly;
            uint64_t m = 0ull;
            /* planes outside [p1_lo, p1_hi) are not stored (and are read by no output voxel) */
            if (gx >= M.p1_lo && gx < M.p1_hi && gy_ >= 0 && gy_ < g)
                m = row_bits(outm + ((size_t)gx * g + gy_) * W, W, tz0 - R, LZ);
            rows[row] = m;
            fl |= (m != 0ull);
        }
    };
    /* a tile marked by field pass 1 has a shell voxel: its halo rows are loaded together with its
       own rows, and the first barrier reduces fl */
    const bool marked = tflag != nullptr;

    /* the tile's own shell bits (threads 0-127) and outside bits (threads 128-255) first: an
       unmarked tile without a shell voxel (outside the molecule or deep inside it) is finished
       from them alone, without its halo */
    static_assert(P2R_TX * P2_TY == 128, "one thread per own row of each mask");
    static_assert(P2_TZ == 32, "a tile's z range is one half of a row word");
    if (threadIdx.x < 8) {
        const int t = threadIdx.x;
        const uint64_t w = t < 2 ? nf.w[0] : (t < 4 ? nf.w[1] : (t < 6 ? nf.w[2] : nf.w[3]));
        nfw[t] = (uint32_t)(w >> (32 * (t & 1)));
    }
    int has_shell;
    uint32_t sb = 0u;
    {
        const int t = threadIdx.x & 127, tx = t / P2_TY, ty_ = t - tx * P2_TY;
        const int gx = tx0 + tx, gy_ = ty0 + ty_;
        const uint64_t *__restrict__ src = threadIdx.x < 128 ? shellm : outm;
        uint32_t m = 0u;
        if (gx < x1 && gx < g && gy_ < g) {
            /* the tile's 32 z of the row are one half of a row word (tz0 is a multiple of 32); the
               shell threads also take the row's shell rank there, loaded alongside */
            const size_t wi = ((size_t)gx * g + gy_) * W + (tz0 >> 6);
            const uint64_t wv = src[wi];
            m = (uint32_t)(wv >> (tz0 & 63));
            if (threadIdx.x < 128) sb = spre[wi] + (uint32_t)__popcll(wv & bits_below(tz0 & 63, 0));
        }
        if (marked) load_rows();
        if (threadIdx.x < 128) {
            own[t] = m;
            /* shell bits of each own x plane, over its P2_TY rows (consecutive lanes) */
            uint32_t sx = m;
#pragma unroll
            for (int k = 1; k < P2_TY; k *= 2) sx |= (uint32_t)__shfl_xor((int)sx, k);
            if (ty_ == 0) shx[tx] = sx;
        } else {
            ownout[t] = m;
        }
        /* marked: the tile has a shell voxel, and the barrier tells whether an outside voxel is
           within reach instead */
        has_shell = __syncthreads_or(marked ? fl : (threadIdx.x < 128 && m != 0u));
    }
    if (!marked && !has_shell) {
        /* no field values to store (only shell voxels hold one; the others are the probe radius
           outside and -dx inside): the sign words are the complemented outside bits */
        const uint32_t zvalid = (tz0 + P2_TZ <= g) ? ~0u : ((1u << (g - tz0)) - 1u);
        if (threadIdx.x < P2R_TX * P2_TY) {
            const int tx = threadIdx.x / P2_TY, ty_ = threadIdx.x - tx * P2_TY;
            const int gx = tx0 + tx, gy_ = ty0 + ty_;
            if (gx < x1 && gx < g && gy_ < g) {
                const size_t w32 = ((size_t)gx * g + gy_) * (2 * (size_t)W) + (tz0 >> 5);
                negm[w32] = ~ownout[threadIdx.x] & zvalid;
                zerom[w32] = 0u;
            }
        }
        return;
    }
    /* shell rank of each own row's first tile voxel: the row's shell voxels are stored from there
       (ownout was last read before the barrier above) */
    if (threadIdx.x < 128) sbase[threadIdx.x] = sb;
    /* index of own voxel (tx, ty, tz) in the stored shell field */
    auto fidx = [&](int tx, int ty_, int tz_) {
        const int t = tx * P2_TY + ty_;
        return sbase[t] + (uint32_t)__popc(own[t] & ((1u << tz_) - 1u));
    };

    if (!marked) load_rows();
    /* every y minimum starts unreachable; only listed columns are computed */
    for (int i = threadIdx.x; i < LX * P2_TY * P2_TZ / 4; i += 256) DYw[i] = 0x01010101u * INF;
    /* a column (lx, z) is read by the x pass only for shell voxels of own planes lx - 2R .. lx */
    uint32_t need = 0;
    if (threadIdx.x < LX) {
        const int lx = threadIdx.x;
#pragma unroll
        for (int tx = 0; tx < P2R_TX; tx++)
            if (tx >= lx - 2 * R && tx <= lx) need |= shx[tx];
    }
    if (!(marked ? has_shell : __syncthreads_or(fl))) {
        /* no outside voxel within reach: every voxel of the tile is -dx (a marked tile has passed
           no barrier since its shell ranks were written) */
        if (marked) __syncthreads();
        for (int tx = 0; tx < P2R_TX; tx++) {
            int gx = tx0 + tx;
            if (gx >= x1 || gx >= g) break;
            if (vyz && ((own[tx * P2_TY + ty] >> tz) & 1u)) field[fidx(tx, ty, tz)] = -dx;
            emit_sign_bits(negm, zerom, W, g, gx, gy, tz0, vyz, -dx);
        }
        return;
    }

    /* the work list: columns needed by a shell voxel with an outside voxel within R of their z
       in some row (else every minimum is INF, as filled) */
    static_assert(LX <= 64, "one lane per halo x plane");
    if (threadIdx.x < 64) {
        const int lx = threadIdx.x;
        uint32_t c = 0;
        if (lx < LX) {
            uint64_t o = 0;
#pragma unroll
            for (int ly = 0; ly < LY; ly++) o |= rows[lx * LY + ly];
            /* bit z: any of bits z .. z + 2R (bit k of a row is z = k - R) */
            uint64_t near = 0;
#pragma unroll
            for (int k = 0; k <= 2 * R; k++) near |= o >> k;
            c = (uint32_t)near & need;
            cmask[lx] = c;
        }
        const uint32_t cnt = (uint32_t)__builtin_popcount(c);
        uint32_t inc = cnt;
#pragma unroll
        for (int k = 1; k < 64; k *= 2) {
            const uint32_t t = (uint32_t)__shfl_up((int)inc, k);
            if (lx >= k) inc += t;
        }
        if (lx < LX) cpre[lx] = inc - cnt;
        if (lx == LX - 1) cpre[LX] = inc;
    }
    __syncthreads();
    const uint32_t ntodo = cpre[LX];

    /* z then y, one listed (lx, z) column per thread: the squared z distance of each row to its
       nearest outside voxel within R, straight from the row bits, then the y minima over 2R+1 rows */
    if constexpr (2 * R + 1 <= 16) {
        /* two y outputs per 32-bit register (16-bit halves): the tap sums, minima and equality
           masks in packed 16-bit integer arithmetic */
        typedef uint16_t u16x2 __attribute__((ext_vector_type(2)));
        /* 2^1, 2^2, 2^4, 2^8 per lane, opaque so that the mask bits stay packed multiply-adds */
        u16x2 pw[4] = {{2, 2}, {4, 4}, {16, 16}, {256, 256}};
#pragma unroll
        for (int i = 0; i < 4; i++) asm volatile("" : "+v"(pw[i]));
        for (uint32_t it = threadIdx.x; it < ntodo; it += 256) {
            int lx, z;
            list_column<LX>(cpre, cmask, it, lx, z);
            uint32_t v[LY];
#pragma unroll
            for (int ly = 0; ly < LY; ly++) v[ly] = zdist2<R>(rows[lx * LY + ly], z);
            u16x2 P[LY - 1];
#pragma unroll
            for (int k = 0; k < LY - 1; k++) P[k] = __builtin_bit_cast(u16x2, v[k] | (v[k + 1] << 16));
#pragma unroll
            for (int yy = 0; yy < P2_TY; yy += 2) {
                /* taps below 2^11 (z squares below 1024 plus 49): no 16-bit lane overflows */
                constexpr int NT = 2 * R + 1;
                u16x2 sb[NT], t[NT];
#pragma unroll
                for (int b = 0; b < NT; b++) {
                    const uint16_t c = (uint16_t)((b - R) * (b - R));
                    sb[b] = P[yy + b] + (u16x2){c, c};
                    t[b] = sb[b];
                }
                /* minimum as a tree (no chain of dependent packed operations) */
#pragma unroll
                for (int s = 1; s < NT; s *= 2)
#pragma unroll
                    for (int b = 0; b + s < NT; b += 2 * s) t[b] = __builtin_elementwise_min(t[b], t[b + s]);
                const u16x2 best = t[0];
                /* tie mask: best + 1 - tap saturates to 1 at a minimum and to 0 elsewhere; the
                   bits are combined as a tree as well, pairs of taps first */
                const u16x2 best1 = best + (u16x2){1, 1};
#pragma unroll
                for (int b = 0; b < NT; b++) t[b] = __builtin_elementwise_sub_sat(best1, sb[b]);
#pragma unroll
                for (int s = 1, l = 0; s < NT; s *= 2, l++)
#pragma unroll
                    for (int b = 0; b + s < NT; b += 2 * s) t[b] = t[b + s] * pw[l] + t[b];
                const u16x2 mm = t[0];
                const uint32_t bb = __builtin_bit_cast(uint32_t, __builtin_elementwise_min(best, (u16x2){(uint16_t)INF, (uint16_t)INF}));
                const uint32_t mw = __builtin_bit_cast(uint32_t, mm);
                DY[(lx * P2_TY + yy) * P2_TZ + z] = (uint8_t)bb;
                DY[(lx * P2_TY + yy + 1) * P2_TZ + z] = (uint8_t)(bb >> 16);
                BM[(lx * P2_TY + yy) * P2_TZ + z] = (BMask)(mw & 0xFFFFu);
                BM[(lx * P2_TY + yy + 1) * P2_TZ + z] = (BMask)(mw >> 16);
            }
        }
    } else
    for (uint32_t it = threadIdx.x; it < ntodo; it += 256) {
        int lx, z;
        list_column<LX>(cpre, cmask, it, lx, z);
        int v[LY];
#pragma unroll
        for (int ly = 0; ly < LY; ly++) v[ly] = (int)zdist2<R>(rows[lx * LY + ly], z);
#pragma unroll
        for (int yy = 0; yy < P2_TY; yy++) {
            int best = INF;
#pragma unroll
            for (int b = 0; b <= 2 * R; b++) best = min(best, v[yy + b] + (b - R) * (b - R));
            uint32_t bm = 0;
#pragma unroll
            for (int b = 0; b <= 2 * R; b++) bm |= (uint32_t)(v[yy + b] + (b - R) * (b - R) == best) << b;
            DY[(lx * P2_TY + yy) * P2_TZ + z] = (uint8_t)best;
            BM[(lx * P2_TY + yy) * P2_TZ + z] = (BMask)bm;
        }
    }
    __syncthreads();

    /* x, one (ty, tz) column per thread, in two phases.  Phase 1: each voxel's minimal squared
       norm over its x window; a voxel with nothing to trace back gets its field value and sign
       bits at once, the others join the wave's queue.  Phase 2: the wave's 64 lanes take queued
       voxels 64 at a time (so lanes whose own voxels need nothing do not idle), find the x
       offsets attaining the minimum again, trace the minimisers back through the y masks and
       take the exact float distance.  The sign and zero words gather in LDS and are stored once
       per row at the end. */
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    uint16_t *__restrict__ queue = tq[wid];
    int qn = 0;
    auto trace_items = [&](int count) {
        if (lane >= count) return;
        const uint32_t it = queue[lane];
        const int tx = (int)(it & 15u), ty_ = 2 * wid + (int)((it >> 4) & 1u), tz_ = (int)(it >> 5);
        const int gx = tx0 + tx, gy_ = ty0 + ty_, gz_ = tz0 + tz_;
        /* the x offsets (bit a + R) attaining the minimal norm */
        int nmin = INF;
        uint32_t amask = 0;
        if constexpr (2 * R + 1 <= 16) {
            /* taps a = 2k, 2k + 1 in the 16-bit halves of t[k] (an odd last tap is padded with the
               largest value, which neither lowers the minimum nor equals it); the tie bits are
               taken as in the y pass, then combined by packed multiply-adds: the low halves end up
               holding the even taps' bits at their places, the high halves the odd ones' one place
               lower */
            typedef uint16_t u16x2 __attribute__((ext_vector_type(2)));
            constexpr int NT = 2 * R + 1, NP = (NT + 1) / 2;
            u16x2 t[NP];
#pragma unroll
            for (int k = 0; k < NP; k++) {
                const int a0 = 2 * k, a1 = 2 * k + 1;
                const uint32_t lo = DY[((tx + a0) * P2_TY + ty_) * P2_TZ + tz_];
                const uint32_t hi = (a1 < NT) ? (uint32_t)DY[((tx + a1) * P2_TY + ty_) * P2_TZ + tz_] : 0xFFFFu;
                const uint16_t c0 = (uint16_t)((a0 - R) * (a0 - R)), c1 = (a1 < NT) ? (uint16_t)((a1 - R) * (a1 - R)) : (uint16_t)0;
                t[k] = __builtin_bit_cast(u16x2, lo | (hi << 16)) + (u16x2){c0, c1};
            }
            u16x2 m = t[0];
#pragma unroll
            for (int k = 1; k < NP; k++) m = __builtin_elementwise_min(m, t[k]);
            nmin = (int)min(m.x, m.y);
            const u16x2 n1 = (u16x2){(uint16_t)(nmin + 1), (uint16_t)(nmin + 1)};
#pragma unroll
            for (int k = 0; k < NP; k++) t[k] = __builtin_elementwise_sub_sat(n1, t[k]);
            /* 4, 16, 256 per lane, opaque so that the combination stays packed multiply-adds */
            u16x2 pw[3] = {{4, 4}, {16, 16}, {256, 256}};
#pragma unroll
            for (int i = 0; i < 3; i++) asm volatile("" : "+v"(pw[i]));
#pragma unroll
            for (int st = 1, l = 0; st < NP; st *= 2, l++)
#pragma unroll
                for (int k = 0; k + st < NP; k += 2 * st) t[k] = t[k + st] * pw[l] + t[k];
            amask = (uint32_t)t[0].x | ((uint32_t)t[0].y << 1);
        } else {
#pragma unroll
            for (int a = 0; a <= 2 * R; a++) nmin = min(nmin, (int)DY[((tx + a) * P2_TY + ty_) * P2_TZ + tz_] + (a - R) * (a - R));
#pragma unroll
            for (int a = 0; a <= 2 * R; a++)
                amask |= (uint32_t)((int)DY[((tx + a) * P2_TY + ty_) * P2_TZ + tz_] + (a - R) * (a - R) == nmin) << a;
        }
        const float vx = grid_pos(G.ox, gx, dx), vy = grid_pos(G.oy, gy_, dx), vz = grid_pos(G.oz, gz_, dx);
        /* the smallest squared distance, rooted once at the end: sqrtf rounds correctly and is
           monotone, so the root of the minimum is the minimum of the roots */
        float min_sq = __builtin_inff();
        if ((nfw[nmin >> 5] >> (nmin & 31)) & 1u) {
            /* trace the minimisers back through the separable passes: the x offsets, then their y
               offsets from the recorded masks */
            while (amask) {
                const int a = __builtin_ctz(amask) - R;
                amask &= amask - 1u;
                const int dy = nmin - a * a;
                uint32_t bmask = BM[((tx + R + a) * P2_TY + ty_) * P2_TZ + tz_];
                while (bmask) {
                    const int b = __builtin_ctz(bmask) - R;
                    bmask &= bmask - 1u;
                    /* dz = c^2 exactly, c <= R: the hardware root is within an ulp of c */
                    const int dz = dy - b * b;
                    const int c = (int)(__builtin_amdgcn_sqrtf((float)dz) + 0.5f);
                    const uint64_t m = rows[(tx + R + a) * LY + (ty_ + R + b)];
                    const float qx = grid_pos(G.ox, gx + a, dx), qy = grid_pos(G.oy, gy_ + b, dx);
                    if ((m >> (tz_ + R + c)) & 1ull)
                        min_sq = fminf(dist3_sq(qx, qy, grid_pos(G.oz, gz_ + c, dx), vx, vy, vz), min_sq);
                    if (c && ((m >> (tz_ + R - c)) & 1ull))
                        min_sq = fminf(dist3_sq(qx, qy, grid_pos(G.oz, gz_ - c, dx), vx, vy, vz), min_sq);
                }
            }
        } else {
#pragma unroll 1
            for (int k = norm_start[nmin]; k < enum_end[nmin]; k++) {
                int o = offs[k];
                int a = (int)(int8_t)(o & 0xFF), b = (int)(int8_t)((o >> 8) & 0xFF), c = (int)(int8_t)((o >> 16) & 0xFF);
                uint64_t m = rows[(tx + R + a) * LY + (ty_ + R + b)];
                if (!((m >> (tz_ + R + c)) & 1ull)) continue;
                float d = dist3_sq(grid_pos(G.ox, gx + a, dx), grid_pos(G.oy, gy_ + b, dx), grid_pos(G.oz, gz_ + c, dx), vx, vy,
                                   vz);
                min_sq = fminf(d, min_sq);
            }
        }
        const float min_dist = sqrt_rn(min_sq);
        const float res = (min_dist < 999.0f) ? (QS_PROBE - min_dist) : -dx;
        field[fidx(tx, ty_, tz_)] = res;
        if (res < 0.0f) atomicOr(&sneg[tx * P2_TY + ty_], 1u << tz_);
        if (res == 0.0f) atomicOr(&szero[tx * P2_TY + ty_], 1u << tz_);
    };
    /* voxel tx of the thread's column, whose minimal norm is nmin (INF: no outside voxel within
       the window): settled now, or queued */
    auto handle = [&](int tx, int nmin) {
        const bool out = (rows[(tx + R) * LY + (ty + R)] >> (tz + R)) & 1ull;
        const bool shell = (own[tx * P2_TY + ty] >> tz) & 1u; /* never outside */
        const bool traced = vyz && shell && nmin < INF;
        /* outside: the probe radius; inside, or shell without an outside voxel in reach: -dx */
        if (vyz && shell && !traced) field[fidx(tx, ty, tz)] = -dx;
        const uint64_t nb = __ballot(vyz && !traced && !out);
        if ((lane & 31) == 0) {
            sneg[tx * P2_TY + ty] = (uint32_t)(lane ? (nb >> 32) : nb);
            szero[tx * P2_TY + ty] = 0u;
        }
        const uint64_t tm = __ballot(traced);
        if (traced) queue[qn + __popcll(tm & ((1ull << lane) - 1ull))] = (uint16_t)(tx | ((lane >> 5) << 4) | (tz << 5));
        qn += __popcll(tm);
        if (qn >= 64) {
            wave_lds_sync();
            trace_items(64);
            wave_lds_sync();
            const int rest = qn - 64;
            const uint16_t mv = (lane < rest) ? queue[64 + lane] : (uint16_t)0;
            wave_lds_sync();
            if (lane < rest) queue[lane] = mv;
            wave_lds_sync();
            qn = rest;
        }
    };

    if constexpr (2 * R + 1 <= 16) {
        /* two x outputs per 32-bit register: the minimum over the 2R+1 taps (16-bit lanes) */
        typedef uint16_t u16x2 __attribute__((ext_vector_type(2)));
        u16x2 U[LX - 1];
        {
            uint32_t u[LX];
#pragma unroll
            for (int lx = 0; lx < LX; lx++) u[lx] = DY[(lx * P2_TY + ty) * P2_TZ + tz];
#pragma unroll
            for (int k = 0; k < LX - 1; k++) U[k] = __builtin_bit_cast(u16x2, u[k] | (u[k + 1] << 16));
        }
        static_assert(P2R_TX % 2 == 0, "x outputs in pairs");
#pragma unroll
        for (int tx = 0; tx < P2R_TX; tx += 2) {
            const int gx = tx0 + tx;
            if (gx >= x1 || gx >= g) break;
            const bool two = gx + 1 < x1 && gx + 1 < g;
            const uint32_t sh = ((own[tx * P2_TY + ty] >> tz) & 1u) | (two ? ((own[(tx + 1) * P2_TY + ty] >> tz) & 1u) << 1 : 0u);
            uint32_t nm = (uint32_t)INF * 0x10001u;
            if (vyz && sh) {
                constexpr int NT = 2 * R + 1;
                u16x2 t[NT];
#pragma unroll
                for (int a = 0; a < NT; a++) {
                    const uint16_t c = (uint16_t)((a - R) * (a - R));
                    t[a] = U[tx + a] + (u16x2){c, c};
                }
#pragma unroll
                for (int s2 = 1; s2 < NT; s2 *= 2)
#pragma unroll
                    for (int a = 0; a + s2 < NT; a += 2 * s2) t[a] = __builtin_elementwise_min(t[a], t[a + s2]);
                nm = __builtin_bit_cast(uint32_t, __builtin_elementwise_min(t[0], (u16x2){(uint16_t)INF, (uint16_t)INF}));
            }
            handle(tx, (int)(nm & 0xFFFFu));
            if (two) handle(tx + 1, (int)(nm >> 16));
        }
    } else {
        int u[LX];
#pragma unroll
        for (int lx = 0; lx < LX; lx++) u[lx] = DY[(lx * P2_TY + ty) * P2_TZ + tz];
#pragma unroll
        for (int tx = 0; tx < P2R_TX; tx++) {
            const int gx = tx0 + tx;
            if (gx >= x1 || gx >= g) break;
            int nmin = INF;
            if (vyz && ((own[tx * P2_TY + ty] >> tz) & 1u)) {
#pragma unroll
                for (int a = 0; a <= 2 * R; a++) nmin = min(nmin, u[tx + a] + (a - R) * (a - R));
            }
            handle(tx, nmin);
        }
    }
    if (qn > 0) {
        wave_lds_sync();
        trace_items(qn);
    }
    wave_lds_sync();
    /* the wave's sign and zero words: lane = x (0-15) | row bit << 4 | zero word << 5 */
    {
        const int tx = lane & 15, ty_ = 2 * wid + ((lane >> 4) & 1);
        const int gx = tx0 + tx, gy_ = ty0 + ty_;
        if (gx < x1 && gx < g && gy_ < g) {
            const size_t w32 = ((size_t)gx * g + gy_) * (2 * (size_t)W) + (tz0 >> 5);
            if (lane < 32) negm[w32] = sneg[tx * P2_TY + ty_];
            else zerom[w32] = szero[tx * P2_TY + ty_];
        }
    }
}

__device__ __forceinline__ bool mask_bit(const uint64_t *__restrict__ m, int g, int W, int x, int y, int z)
{
    return (m[((size_t)x * g + y) * W + (z >> 6)] >> (z & 63)) & 1ull;
}

/* direct (2r+1)^3 window, one thread per voxel: any r; also a cross-check of the tile kernel */
__global__ void __launch_bounds__(256) k_pass2_direct(QsBatch B, const uint64_t *__restrict__ outm_all,
                                                      const uint64_t *__restrict__ shellm_all,
                                                      const uint32_t *__restrict__ spre_all, float *__restrict__ field,
                                                      uint32_t *__restrict__ negm_all, uint32_t *__restrict__ zerom_all)
{
    uint32_t lb;
    const QsMol &M = B.mols[batch_mol(B, blockIdx.x, lb)];
    const QsGrid G = M.G;
    const int g = G.g, r = G.r, W = (int)M.W;
    const uint64_t *__restrict__ outm = outm_all + M.word_off;
    const uint64_t *__restrict__ shellm = shellm_all + M.word_off;
    const uint32_t *__restrict__ spre = spre_all + M.word_off;
    uint32_t *__restrict__ negm = negm_all + 2 * M.word_off;
    uint32_t *__restrict__ zerom = zerom_all + 2 * M.word_off;
    const uint32_t nbx = (g + 63) / 64, nby = (g + 3) / 4;
    const int bidx = (int)(lb % nbx);
    int z = bidx * 64 + (threadIdx.x & 63);
    int y = (int)((lb / nbx) % nby) * 4 + (threadIdx.x >> 6);
    int x = M.p2_lo + (int)(lb / (nbx * nby));
    const int x1 = M.p2_hi;
    if (x >= x1 || y >= g) return;
    const bool valid = z < g;
    const float dx = G.dx;
    float res;
    bool shell = false;
    if (!valid || mask_bit(outm, g, W, x, y, z)) {
        res = QS_PROBE;
    } else if (!(shell = mask_bit(shellm, g, W, x, y, z))) {
        res = -dx;
    } else {
        float vx = grid_pos(G.ox, x, dx), vy = grid_pos(G.oy, y, dx), vz = grid_pos(G.oz, z, dx);
        float min_dist = 100000.0f;
        for (int a = -r; a <= r; a++) {
            int qx = iclamp(x + a, 0, g - 1);
            for (int b = -r; b <= r; b++) {
                int qy = iclamp(y + b, 0, g - 1);
                for (int c = -r; c <= r; c++) {
                    int qz = iclamp(z + c, 0, g - 1);
                    if (!mask_bit(outm, g, W, qx, qy, qz)) continue;
                    float d = dist3(grid_pos(G.ox, qx, dx), grid_pos(G.oy, qy, dx), grid_pos(G.oz, qz, dx), vx, vy, vz);
                    min_dist = fminf(d, min_dist);
                }
            }
        }
        res = (min_dist < 999.0f) ? (QS_PROBE - min_dist) : -dx;
    }
    if (shell) {
        const size_t wi = ((size_t)x * g + y) * W + (z >> 6);
        field[spre[wi] + (uint32_t)__popcll(shellm[wi] & bits_below(z & 63, 0))] = res;
    }
    uint64_t n = __ballot(valid && res < 0.0f), zr = __ballot(valid && res == 0.0f);
    if ((threadIdx.x & 63) == 0) {
        size_t w32 = ((size_t)x * g + y) * (2 * (size_t)W) + 2 * (size_t)bidx;
        negm[w32] = (uint32_t)n;
        negm[w32 + 1] = (uint32_t)(n >> 32);
        zerom[w32] = (uint32_t)zr;
        zerom[w32 + 1] = (uint32_t)(zr >> 32);
    }
}

/* sign and zero bits of a caller-given field (stage-wise entry point) */
__global__ void __launch_bounds__(256) k_field_signs(const float *__restrict__ F, int g, int W, uint32_t *__restrict__ negm,
                                                     uint32_t *__restrict__ zerom)
{
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    const int w = blockIdx.x, y = blockIdx.y * 4 + wid, x = blockIdx.z;
    if (y >= g) return;
    const int z = w * 64 + lane;
    const float f = (z < g) ? F[((size_t)x * g + y) * g + z] : 1.0f;
    uint64_t n = __ballot(f < 0.0f), zr = __ballot(f == 0.0f);
    if (lane == 0) {
        size_t w32 = ((size_t)x * g + y) * (2 * (size_t)W) + 2 * (size_t)w;
        negm[w32] = (uint32_t)n;
        negm[w32 + 1] = (uint32_t)(n >> 32);
        zerom[w32] = (uint32_t)zr;
        zerom[w32 + 1] = (uint32_t)(zr >> 32);
    }
}

void qs_launch_field_signs(const float *F, int g, int W, uint32_t *negm, uint32_t *zerom, hipStream_t s)
{
    dim3 grid((unsigned)W, (unsigned)((g + 3) / 4), (unsigned)g);
    hipLaunchKernelGGL(k_field_signs, grid, dim3(256), 0, s, F, g, W, negm, zerom);
}

/* labels (u8, 0 outside 1 shell 2 inside) -> masks, for stage-wise entry points */
__global__ void __launch_bounds__(256) k_labels_to_masks(const uint8_t *__restrict__ labels, int g, int W,
                                                         uint64_t *__restrict__ outm, uint64_t *__restrict__ shellm)
{
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    const int w = blockIdx.x, y = blockIdx.y * 4 + wid, x = blockIdx.z;
    if (y >= g) return;
    int z = w * 64 + lane;
    uint8_t l = (z < g) ? labels[((size_t)x * g + y) * g + z] : 0xFF;
    uint64_t mo = __ballot(l == QS_OUT), ms = __ballot(l == QS_SHELL);
    if (lane == 0) {
        outm[((size_t)x * g + y) * W + w] = mo;
        shellm[((size_t)x * g + y) * W + w] = ms;
    }
}

/* ------------------------------------------------------------------------------------------ */
/* marching cubes with edge-id welding                                                         */
/*
 * Every grid point p owns four vertex slots: the edges from p along +x, +y, +z (slots 0..2) and
 * p itself (slot 3).  An edge carries a vertex when its end points differ in sign (field < 0)
 * and it lies on a polygonised cube; when the non-negative end is exactly zero the vertex is
 * that grid point instead (all such edges then share one vertex).  Vertex ids are the ranks of
 * the occupied slots in (point, slot) order, so they come from one prefix sum and welding is
 * exact.  Triangles follow cube order and the case table's order.
 */

/*
 * Marching cubes on the sign bits of the field.  Field pass 2 writes, per 64-voxel row word,
 * N (field < 0) and Z (field == 0).  For every row word k_mc_count forms with 64-bit logic
 *   the four vertex-slot masks of its points (edge +x, +y, +z with a sign change inside the
 *   emission window and a non-zero non-negative end; the point itself when it is zero and an
 *   incident in-window edge ends negative), and
 *   the mask of polygonised cubes (cube origin in [0, lim)^3, corners not all of one sign),
 * and counts vertices and triangles.  An exclusive scan over row words in linear (x, y, w) order
 * gives each word's first vertex and triangle id; inside a word, a point's vertices come after
 * those of the points below it (popcounts of the slot masks), in slot order x, y, z, point --
 * the (point, slot) order.  Emission and smoothing run only on words that own something.
 */

struct MCView {
    const uint64_t *__restrict__ N;
    const uint64_t *__restrict__ Zm;
    const uint64_t *__restrict__ S4; /* 4 slot masks per row word */
    const uint32_t *__restrict__ vbw; /* first vertex id per row word */
    int g, W, lim;

    __device__ __forceinline__ size_t widx(int x, int y, int z) const { return ((size_t)x * g + y) * W + (z >> 6); }
    __device__ __forceinline__ bool nbit(int x, int y, int z) const { return (N[widx(x, y, z)] >> (z & 63)) & 1ull; }
    __device__ __forceinline__ bool zbit(int x, int y, int z) const { return (Zm[widx(x, y, z)] >> (z & 63)) & 1ull; }
    __device__ __forceinline__ int cube_index(int x, int y, int z) const
    {
        int ci = (int)nbit(x, y, z);
        ci |= (int)nbit(x + 1, y, z) << 1;
        ci |= (int)nbit(x + 1, y + 1, z) << 2;
        ci |= (int)nbit(x, y + 1, z) << 3;
        ci |= (int)nbit(x, y, z + 1) << 4;
        ci |= (int)nbit(x + 1, y, z + 1) << 5;
        ci |= (int)nbit(x + 1, y + 1, z + 1) << 6;
        ci |= (int)nbit(x, y + 1, z + 1) << 7;
        return ci;
    }
    /* the same case from the four row words of the cube's x, y corners (and the next words when
       the cube's top corners lie in them) */
    __device__ __forceinline__ int cube_case(int x, int y, int z) const
    {
        const int b = z & 63;
        const size_t w00 = widx(x, y, z), w10 = w00 + (size_t)g * W, w01 = w00 + W, w11 = w10 + W;
        const uint64_t n00 = N[w00], n10 = N[w10], n11 = N[w11], n01 = N[w01];
        int ci = (int)((n00 >> b) & 1ull) | (int)((n10 >> b) & 1ull) << 1 | (int)((n11 >> b) & 1ull) << 2 |
                 (int)((n01 >> b) & 1ull) << 3;
        if (b < 63) {
            ci |= (int)((n00 >> (b + 1)) & 1ull) << 4 | (int)((n10 >> (b + 1)) & 1ull) << 5 |
                  (int)((n11 >> (b + 1)) & 1ull) << 6 | (int)((n01 >> (b + 1)) & 1ull) << 7;
        } else {
            ci |= (int)(N[w00 + 1] & 1ull) << 4 | (int)(N[w10 + 1] & 1ull) << 5 | (int)(N[w11 + 1] & 1ull) << 6 |
                  (int)(N[w01 + 1] & 1ull) << 7;
        }
        return ci;
    }
    /* vertex id of slot s (0,1,2 edge x,y,z; 3 point) of point (x, y, z) */
    __device__ __forceinline__ uint32_t slot_id(int x, int y, int z, int sl) const
    {
        size_t wi = widx(x, y, z);
        int b = z & 63;
        uint64_t below = (1ull << b) - 1ull;
        const uint64_t *m = S4 + 4 * wi;
        uint64_t m0 = m[0], m1 = m[1], m2 = m[2], m3 = m[3];
        uint32_t id = vbw[wi] + (uint32_t)(__popcll(m0 & below) + __popcll(m1 & below) + __popcll(m2 & below) +
                                           __popcll(m3 & below));
        if (sl > 0) id += (uint32_t)((m0 >> b) & 1ull);
        if (sl > 1) id += (uint32_t)((m1 >> b) & 1ull);
        if (sl > 2) id += (uint32_t)((m2 >> b) & 1ull);
        return id;
    }
    /* vertex of cube edge e of cube (cx, cy, cz), whose case is ci: the sign of the edge's low
       corner is case bit (z offset) * 4 + (0, 1, 2, 3 for xy offsets 00, 10, 11, 01) */
    __device__ __forceinline__ uint32_t corner_vid(int cx, int cy, int cz, int e, int ci) const
    {
        const int lx = (int)((kEdgeLX >> e) & 1u), ly = (int)((kEdgeLY >> e) & 1u), lz = (int)((kEdgeLZ >> e) & 1u);
        int qx = cx + lx, qy = cy + ly, qz = cz + lz;
        int a = (int)((kEdgeAX >> (2 * e)) & 3u);
        int rx = qx + (a == 0), ry = qy + (a == 1), rz = qz + (a == 2);
        const bool qneg = (ci >> (4 * lz + (ly ? (lx ? 2 : 3) : lx))) & 1;
        int nx = qneg ? rx : qx, ny = qneg ? ry : qy, nz = qneg ? rz : qz;
        if (zbit(nx, ny, nz)) return slot_id(nx, ny, nz, 3);
        return slot_id(qx, qy, qz, a);
    }
};

/* countVertexPerCell MarchingCubes.cu:571-658 in word form: slot masks and per-word counts */
__global__ void __launch_bounds__(256) k_mc_count(QsBatch B, const uint64_t *__restrict__ N_all,
                                                  const uint64_t *__restrict__ Zm_all, uint64_t *__restrict__ S4_all,
                                                  uint32_t *__restrict__ cnt_all, uint64_t *__restrict__ NT_all)
{
    uint32_t lb;
    const QsMol &M = B.mols[batch_mol(B, xcd_block(blockIdx.x, gridDim.x), lb)];
    const QsGrid G = M.G;
    const int g = G.g, lim = G.lim, W = (int)M.W;
    const uint64_t *__restrict__ N = N_all + M.word_off;
    const uint64_t *__restrict__ Zm = Zm_all + M.word_off;
    uint64_t *__restrict__ S4 = S4_all + 4 * M.word_off;
    uint32_t *__restrict__ cnt = cnt_all + M.word_off;
    const uint32_t nbx = ((uint32_t)g * W + 255) / 256;
    const int idx = (int)(lb % nbx) * 256 + threadIdx.x;
    const int x = M.mc_lo + (int)(lb / nbx);
    if (idx >= g * W || x >= M.mc_hi) return;
    const int y = idx / W, w = idx - y * W;
    const int z0 = w * 64;
    const size_t wi = ((size_t)x * g + y) * W + w;
    const uint64_t valid = bits_below(g, z0);
    const uint64_t wlt = bits_below(lim, z0) & valid;     /* z < lim */
    const uint64_t wle = bits_below(lim + 1, z0) & valid; /* z <= lim */
    const uint64_t zm1 = (w == 0) ? ~1ull : ~0ull;         /* z >= 1 */

    /* the word and its neighbours one step along x, y and w, loaded together without branches: a
       neighbour outside the grid reads the word itself and is masked to zero (the point itself
       is inside the grid) */
    const size_t sx = (size_t)g * W;
    const uint64_t *__restrict__ np = N + wi;
    const uint64_t *__restrict__ zp = Zm + wi;
    const bool hx = x + 1 < g, hy = y + 1 < g, hw = w + 1 < W;
    const size_t ox = hx ? sx : 0, oy = hy ? (size_t)W : 0, ow = hw ? 1 : 0;
    const uint64_t mx = hx ? ~0ull : 0ull, my = hy ? ~0ull : 0ull, mw = hw ? ~0ull : 0ull;
    const uint64_t n00 = np[0], n00p = np[ow] & mw, n00m = np[w > 0 ? -1 : 0] & (w > 0 ? ~0ull : 0ull);
    const uint64_t z00 = zp[0], z00p = zp[ow] & mw;
    const uint64_t n10 = np[ox] & mx, z10 = zp[ox] & mx;
    const uint64_t n01 = np[oy] & my, z01 = zp[oy] & my;
    /* the cube's far x, y corners and the next words of its top corners */
    const uint64_t n11 = np[ox + oy] & mx & my;
    const uint64_t n10p = np[ox + ow] & mx & mw, n11p = np[ox + oy + ow] & mx & my & mw, n01p = np[oy + ow] & my & mw;
    const uint64_t n00_up = (n00 >> 1) | (n00p << 63); /* bit z = N[z+1] */
    const uint64_t n00_dn = (n00 << 1) | (n00m >> 63); /* bit z = N[z-1] */
    const uint64_t z00_up = (z00 >> 1) | (z00p << 63);

    /* edge masks under condition masks rather than branches (a branch would take its word's
       load with it and wait on it alone) */
    const uint64_t kx = (x + 1 < g && x < lim && y <= lim) ? ~0ull : 0ull;
    const uint64_t ky = (y + 1 < g && x <= lim && y < lim) ? ~0ull : 0ull;
    const uint64_t kz = (x <= lim && y <= lim) ? ~0ull : 0ull;
    const uint64_t ex = (n00 ^ n10) & ~((n00 & z10) | (~n00 & z00)) & wle & kx;
    const uint64_t ey = (n00 ^ n01) & ~((n00 & z01) | (~n00 & z00)) & wle & ky;
    const uint64_t ez = (n00 ^ n00_up) & ~((n00 & z00_up) | (~n00 & z00)) & wlt & kz;
    uint64_t pt = 0;
    if (z00) {
        uint64_t nb = 0;
        if (x + 1 < g && x < lim && y <= lim) nb |= n10 & wle;
        if (x >= 1 && x - 1 < lim && y <= lim) nb |= *(np - sx) & wle;
        if (y + 1 < g && x <= lim && y < lim) nb |= n01 & wle;
        if (y >= 1 && x <= lim && y - 1 < lim) nb |= *(np - W) & wle;
        if (x <= lim && y <= lim) nb |= (n00_up & wlt) | (n00_dn & zm1 & wle);
        pt = z00 & nb;
    }
    uint32_t nv = (uint32_t)(__popcll(ex) + __popcll(ey) + __popcll(ez) + __popcll(pt));
    /* slot masks are read only for the words that hold vertices */
    if (nv) {
        S4[4 * wi + 0] = ex;
        S4[4 * wi + 1] = ey;
        S4[4 * wi + 2] = ez;
        S4[4 * wi + 3] = pt;
    }
    uint32_t nt = 0;
    if (x < lim && x < M.cube_hi && y < lim) {
        const uint64_t c0 = n00, c1 = n10, c2 = n11, c3 = n01;
        const uint64_t c4 = n00_up;
        const uint64_t c5 = (n10 >> 1) | (n10p << 63);
        const uint64_t c6 = (n11 >> 1) | (n11p << 63);
        const uint64_t c7 = (n01 >> 1) | (n01p << 63);
        const uint64_t all = c0 & c1 & c2 & c3 & c4 & c5 & c6 & c7;
        const uint64_t any = c0 | c1 | c2 | c3 | c4 | c5 | c6 | c7;
        const uint64_t act = any & ~all & wlt;
        /* triangles per cube as three bit planes (counts < 8), for the word's 64 cubes at once.
           For a cube with negative corners S (neither none nor all), the table's count is
           |S| - 2 f - 4 h: f is the number of faces whose four corners are in S, and h marks the
           cases whose two corners outside S are opposite ends of a body diagonal.  (The table
           gives k - 2 triangles per component of S with k cut edges; the cut edges number
           3 |S| - 2 e(S) for e(S) edges inside S, and the components |S| - e(S) + f, except for
           the ring of six corners around a body diagonal: one component, which the table
           triangulates as two, one triangle at each corner outside S.)  Checked against every
           case of the table. */
        const auto fadd = [](uint64_t a, uint64_t b, uint64_t c, uint64_t &carry) {
            const uint64_t ab = a ^ b;
            carry = (a & b) | (c & ab);
            return ab ^ c;
        };
        uint64_t a1, b1, e1, t1;
        const uint64_t a0 = fadd(c0, c1, c2, a1), b0 = fadd(c3, c4, c5, b1);
        const uint64_t d0 = c6 ^ c7, d1 = c6 & c7;
        const uint64_t s0 = fadd(a0, b0, d0, e1);    /* |S| bit 0 */
        const uint64_t t0 = fadd(a1, b1, d1, t1);    /* weight-2 sum of the carries */
        const uint64_t s1 = t0 ^ e1, s2 = t1 ^ (t0 & e1); /* |S| bits 1 and 2 (bit 3: all set) */
        const uint64_t fz0 = c0 & c1 & c2 & c3, fz1 = c4 & c5 & c6 & c7, fy0 = c0 & c1 & c4 & c5;
        const uint64_t fy1 = c2 & c3 & c6 & c7, fx0 = c0 & c3 & c4 & c7, fx1 = c1 & c2 & c5 & c6;
        uint64_t x1, y1;
        const uint64_t x0 = fadd(fz0, fz1, fy0, x1), y0 = fadd(fy1, fx0, fx1, y1);
        const uint64_t f0 = x0 ^ y0, f1 = x1 ^ y1 ^ (x0 & y0); /* f mod 4 (f <= 3 unless all set) */
        /* |S| = 6 with no full face: the two corners outside S are a body diagonal */
        const uint64_t h = ~s0 & s1 & s2 & ~(fz0 | fz1 | fy0 | fy1 | fx0 | fx1);
        /* k = |S| - 2 f - 4 h mod 8, bit by bit */
        const uint64_t p0 = s0 & act, p1 = (s1 ^ f0) & act, p2 = (s2 ^ f1 ^ (~s1 & f0) ^ h) & act;
        nt = (uint32_t)(__popcll(p0) + 2 * __popcll(p1) + 4 * __popcll(p2));
        if (nt) {
            uint64_t *__restrict__ NT = NT_all + 3 * (M.word_off + wi);
            NT[0] = p0;
            NT[1] = p1;
            NT[2] = p2;
        }
    }
    /* the word's vertex count (at most 4 x 64) and triangle count (at most 5 x 64) in one word */
    cnt[wi] = nv | (nt << 16);
}

/* row word wi of a grid -> (x, y, word) */
__device__ __forceinline__ void decode_word(uint32_t wi, int g, int W, int &x, int &y, int &w)
{
    const int row = (int)(wi / (uint32_t)W);
    w = (int)(wi - (uint32_t)row * W);
    x = row / g;
    y = row - x * g;
}

/* last index k in [lo, hi] with a[k] <= v, given a[lo] <= v; a non-decreasing */
__device__ __forceinline__ uint32_t last_le(const uint32_t *__restrict__ a, uint32_t lo, uint32_t hi, uint32_t v)
{
    while (lo < hi) {
        const uint32_t mid = (lo + hi + 1) >> 1;
        if (a[mid] <= v) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

/* first active word of every 256-item block of a launch over items [0, n): active word k holds
   items [base[k], base[k + 1]) (the last one up to n), so block b starts in the word whose range
   holds item 256 b.  Words may hold items beyond n (a slab's scan counts the vertices of the next
   slab's first plane, which that slab emits): blocks stop at n.  One launch serves the vertex
   launch (blockIdx.y 0) and the triangle launch (1). */
struct QsStarts {
    const uint32_t *base; /* first item of each active word */
    uint32_t n;           /* items the launch covers */
    const uint32_t *d_n;  /* the item count on the device (within n), or null */
    uint32_t *start;      /* first active word of each 256-item block */
    uint32_t *zero;       /* nzero counters the item kernel adds to, cleared here */
    uint32_t nzero;
};

__global__ void __launch_bounds__(256) k_block_starts(QsStarts sv, QsStarts st, uint32_t nact,
                                                      const uint32_t *__restrict__ d_nact)
{
    const QsStarts A = blockIdx.y ? st : sv;
    /* counts on the device (when given) within the launch's bounds nact, n */
    if (d_nact) nact = min(*d_nact, nact);
    uint32_t n = A.n;
    if (A.d_n) n = min(*A.d_n, n);
    const uint32_t k0 = blockIdx.x * 256u + threadIdx.x;
    for (uint32_t i = k0; i < A.nzero; i += gridDim.x * 256u) A.zero[i] = 0u;
    for (uint32_t k = k0; k < nact; k += gridDim.x * 256u) {
        const uint32_t b0 = (A.base[k] + 255u) >> 8;
        const uint32_t e = (k + 1 < nact) ? min(A.base[k + 1], n) : n;
        for (uint32_t b = b0; (b << 8) < e; b++) A.start[b] = k;
    }
}

/* The active words that hold a block's items, staged in LDS: their first items and word
   indices, and the block's molecule when one molecule holds them all (else -1). */
struct ItemWords {
    uint32_t k0, nk;
    int mol;
    uint32_t base[257], word[257];
};

__device__ __forceinline__ void stage_item_words(ItemWords &S, const uint32_t *__restrict__ base,
                                                 const uint32_t *__restrict__ list, const uint32_t *__restrict__ start,
                                                 uint32_t nact, uint32_t nblocks, uint32_t blk, const uint64_t *__restrict__ wpre,
                                                 const uint32_t *__restrict__ wcoarse)
{
    const uint32_t k0 = start[blk];
    const uint32_t k1 = (blk + 1 < nblocks) ? start[blk + 1] : nact - 1u;
    const uint32_t nk = k1 - k0 + 1u;
    if (threadIdx.x == 0) {
        S.k0 = k0;
        S.nk = nk;
        const int m0 = seg_coarse<QS_COARSE_ITEMS>(wpre, wcoarse, (uint64_t)list[k0]);
        const int m1 = seg_coarse<QS_COARSE_ITEMS>(wpre, wcoarse, (uint64_t)list[k1]);
        S.mol = (m0 == m1) ? m0 : -1;
    }
    if (nk <= 257u)
        for (uint32_t t = threadIdx.x; t < nk; t += 256) {
            S.base[t] = base[k0 + t];
            S.word[t] = list[k0 + t];
        }
    __syncthreads();
}

/* active word (index into the active list, and its row word) of item i of the block */
__device__ __forceinline__ void item_word(const ItemWords &S, const uint32_t *__restrict__ base,
                                          const uint32_t *__restrict__ list, uint32_t i, uint32_t &k, uint32_t &wi,
                                          uint32_t &first)
{
    if (S.nk <= 257u) {
        uint32_t lo = 0, hi = S.nk - 1u;
        while (lo < hi) {
            const uint32_t mid = (lo + hi + 1) >> 1;
            if (S.base[mid] <= i) lo = mid;
            else hi = mid - 1;
        }
        k = S.k0 + lo;
        wi = S.word[lo];
        first = S.base[lo];
    } else {
        k = last_le(base, S.k0, S.k0 + S.nk - 1u, i);
        wi = list[k];
        first = base[k];
    }
}

/* Field value of voxel z of grid row `row` (= x * g + y).  Field pass 2 stores the shell voxels'
   values only, densely in shell-rank order (F[spre[word] + shell voxels of the word below z]);
   every other voxel is the probe radius outside and -dx inside, read from the outside and shell
   masks (W words per row).  shellm == nullptr: F holds every voxel, in grid order. */
__device__ __forceinline__ float field_at(const float *__restrict__ F, const uint32_t *__restrict__ spre,
                                          const uint64_t *__restrict__ outm, const uint64_t *__restrict__ shellm, size_t row,
                                          int g, int W, int z, float dx)
{
    if (!shellm) return F[row * g + z];
    const size_t wi = row * W + (z >> 6);
    const uint64_t sh = shellm[wi], ou = outm[wi];
    const uint32_t base = spre[wi];
    const int b = z & 63;
    if ((sh >> b) & 1ull) return F[base + (uint32_t)__popcll(sh & ((1ull << b) - 1ull))];
    return ((ou >> b) & 1ull) ? QS_PROBE : -dx;
}

/* every voxel of one molecule's field in grid order (out) from the stored shell values (F, global
   shell ranks spre) and the masks, one thread per mask word */
__global__ void __launch_bounds__(256) k_field_expand(const float *__restrict__ F, const uint32_t *__restrict__ spre,
                                                      const uint64_t *__restrict__ outm, const uint64_t *__restrict__ shellm,
                                                      float *__restrict__ out, int g, int W, float dx)
{
    const size_t wi = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (wi >= (size_t)g * g * W) return;
    const size_t row = wi / W;
    const int z0 = (int)(wi - row * W) * 64;
    const uint64_t o = outm[wi], sh = shellm[wi];
    uint32_t k = spre[wi];
    for (int b = 0; b < 64 && z0 + b < g; b++) {
        if ((sh >> b) & 1ull) out[row * g + z0 + b] = F[k++];
        else out[row * g + z0 + b] = ((o >> b) & 1ull) ? QS_PROBE : -dx;
    }
}

void qs_launch_field_expand(const float *F, const uint32_t *spre, const uint64_t *outm, const uint64_t *shellm, float *out, int g,
                            int W, float dx, hipStream_t s)
{
    const size_t n = (size_t)g * g * W;
    if (!n) return;
    k_field_expand<<<(unsigned)((n + 255) / 256), 256, 0, s>>>(F, spre, outm, shellm, out, g, W, dx);
}

/* vertices (vertexInterp/lerp MarchingCubes.cu:715-734, edge oriented low->high), one thread per
   vertex: its word from the active words' first vertex ids, its point and slot from the word's
   four slot masks (vertex ids run over the word's points in z order, slots x, y, z, point) */
__global__ void __launch_bounds__(256) k_mc_vertices(const QsMol *__restrict__ mols, const uint64_t *__restrict__ wpre,
                                                     const uint32_t *__restrict__ wcoarse, const uint32_t *__restrict__ list,
                                                     const uint32_t *__restrict__ avb, const uint32_t *__restrict__ vstart,
                                                     uint32_t nact, uint32_t nv, const uint64_t *__restrict__ S4_all,
                                                     const float *__restrict__ F_all, const uint32_t *__restrict__ spre_all,
                                                     const uint64_t *__restrict__ outm_all,
                                                     const uint64_t *__restrict__ shellm_all, float *__restrict__ vert,
                                                     uint64_t *__restrict__ vkey, const uint32_t *__restrict__ d_counts)
{
    __shared__ ItemWords S;
    /* with d_counts (triangles, vertices, active words) the counts come from the device, the
       launch covering the pool of nv vertices: blocks past the count leave at once */
    if (d_counts) {
        nact = d_counts[2];
        nv = min(d_counts[1], nv);
    }
    const uint32_t nb = (nv + 255u) >> 8;
    /* persistent blocks over the 256-vertex blocks; the grid is a multiple of 8, so block vb runs
       on XCD vb % 8 and the renumbering keeps neighbouring rows' words on one L2 */
    for (uint32_t vb = blockIdx.x; vb < nb; vb += gridDim.x) {
    const uint32_t blk = xcd_group<64>(vb, nb);
    stage_item_words(S, avb, list, vstart, nact, nb, blk, wpre, wcoarse);
    const uint32_t v = blk * 256u + threadIdx.x;
    if (v < nv) {
    uint32_t k, wi, first;
    item_word(S, avb, list, v, k, wi, first);
    const uint32_t local = v - first;
    const QsMol &M = mols[S.mol >= 0 ? S.mol : seg_coarse<QS_COARSE_ITEMS>(wpre, wcoarse, (uint64_t)wi)];
    const QsGrid G = M.G;
    const int g = G.g;
    int x, y, w;
    decode_word((uint32_t)(wi - M.word_off), g, (int)M.W, x, y, w);
    const uint64_t *__restrict__ m = S4_all + 4 * (size_t)wi;
    const uint64_t m0 = m[0], m1 = m[1], m2 = m[2], m3 = m[3];
    /* the point: smallest b whose slots below b + 1 exceed local */
    int lo = 0, hi = 63;
    while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        const uint64_t below = (2ull << mid) - 1ull;
        const uint32_t c = (uint32_t)(__popcll(m0 & below) + __popcll(m1 & below) + __popcll(m2 & below) + __popcll(m3 & below));
        if (c > local) hi = mid;
        else lo = mid + 1;
    }
    const int b = lo;
    const uint64_t below = (1ull << b) - 1ull;
    uint32_t r = local - (uint32_t)(__popcll(m0 & below) + __popcll(m1 & below) + __popcll(m2 & below) + __popcll(m3 & below));
    const uint32_t bits = (uint32_t)((m0 >> b) & 1ull) | (uint32_t)((m1 >> b) & 1ull) << 1 |
                          (uint32_t)((m2 >> b) & 1ull) << 2 | (uint32_t)((m3 >> b) & 1ull) << 3;
    uint32_t bb = bits;
    while (r) {
        bb &= bb - 1u;
        r--;
    }
    const int sl = bb ? __builtin_ctz(bb) : 3;
    const int z = w * 64 + b;
    const float dx = G.dx;
    const size_t pidx = ((size_t)x * g + y) * g + z;
    const float A[3] = {grid_pos(G.ox, x, dx), grid_pos(G.oy, y, dx), grid_pos(G.oz, z, dx)};
    float P[3] = {A[0], A[1], A[2]};
    if (sl < 3) {
        /* stored shell values are addressed by global shell rank; a full field by the molecule's
           voxel offset */
        const float *__restrict__ F = shellm_all ? F_all : F_all + M.vox_off;
        const uint32_t *__restrict__ sp = shellm_all ? spre_all + M.word_off : nullptr;
        const uint64_t *__restrict__ om = outm_all + M.word_off;
        const uint64_t *__restrict__ sm = shellm_all ? shellm_all + M.word_off : nullptr;
        const size_t row = (size_t)x * g + y;
        const int W = (int)M.W;
        const float f0 = field_at(F, sp, om, sm, row, g, W, z, dx);
        const float f1 = field_at(F, sp, om, sm, row + (sl == 0 ? (size_t)g : (sl == 1 ? 1 : 0)), g, W, z + (sl == 2), dx);
        float B[3] = {A[0], A[1], A[2]};
        B[sl] = A[sl] + dx;
        const float t = (0.0f - f0) / (f1 - f0);
#pragma unroll
        for (int c3 = 0; c3 < 3; c3++) {
            float dd = B[c3] - A[c3];
            float s = t * dd;
            P[c3] = A[c3] + s;
        }
    }
    vert[3 * (size_t)v + 0] = P[0];
    vert[3 * (size_t)v + 1] = P[1];
    vert[3 * (size_t)v + 2] = P[2];
    vkey[v] = ((uint64_t)pidx << 2) | (uint64_t)sl;
    }
    __syncthreads(); /* S is staged again for the next block */
    }
}

/* triangles (generateTriangleVerticesSMEM MarchingCubes.cu:736-851 order), one thread per
   triangle: its word from the active words' first triangle ids, its cube from the word's
   triangle-count bit planes, the cube's case from the sign words of its corners */
__global__ void __launch_bounds__(256) k_mc_triangles(const QsMol *__restrict__ mols, const uint64_t *__restrict__ wpre,
                                                      const uint32_t *__restrict__ wcoarse, const uint32_t *__restrict__ list,
                                                      const uint32_t *__restrict__ atb, const uint32_t *__restrict__ tstart,
                                                      uint32_t nact, uint32_t ntri,
                                                      const uint64_t *__restrict__ N_all, const uint64_t *__restrict__ Zm_all,
                                                      const uint64_t *__restrict__ S4_all, const uint32_t *__restrict__ vbw_all,
                                                      const uint64_t *__restrict__ NT_all, int *__restrict__ tri, uint32_t vbase, uint32_t *__restrict__ deg_blk,
                                                      uint32_t *__restrict__ deg_total, const uint32_t *__restrict__ d_counts)
{
    __shared__ ItemWords S;
    /* counts from the device as in k_mc_vertices */
    if (d_counts) {
        nact = d_counts[2];
        ntri = min(d_counts[0], ntri);
    }
    const uint32_t nb = (ntri + 255u) >> 8;
    /* persistent blocks over the 256-triangle blocks, as in k_mc_vertices */
    for (uint32_t vb = blockIdx.x; vb < nb; vb += gridDim.x) {
    const uint32_t blk = xcd_group<64>(vb, nb);
    stage_item_words(S, atb, list, tstart, nact, nb, blk, wpre, wcoarse);
    const uint32_t t = blk * 256u + threadIdx.x;
    int c[3] = {0, 1, 2};
    if (t < ntri) {
    uint32_t k, wi, first;
    item_word(S, atb, list, t, k, wi, first);
    const uint32_t local = t - first;
    const QsMol &M = mols[S.mol >= 0 ? S.mol : seg_coarse<QS_COARSE_ITEMS>(wpre, wcoarse, (uint64_t)wi)];
    const int g = M.G.g;
    int x, y, w;
    decode_word((uint32_t)(wi - M.word_off), g, (int)M.W, x, y, w);
    const uint64_t *__restrict__ np = NT_all + 3 * (size_t)wi;
    const uint64_t p0 = np[0], p1 = np[1], p2 = np[2];
    int lo = 0, hi = 63;
    while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        const uint64_t below = (2ull << mid) - 1ull;
        const uint32_t c = (uint32_t)(__popcll(p0 & below) + 2 * __popcll(p1 & below) + 4 * __popcll(p2 & below));
        if (c > local) hi = mid;
        else lo = mid + 1;
    }
    const int b = lo;
    const uint64_t below = (1ull << b) - 1ull;
    const int j = (int)(local - (uint32_t)(__popcll(p0 & below) + 2 * __popcll(p1 & below) + 4 * __popcll(p2 & below)));
    const int z = w * 64 + b;
    MCView V;
    V.N = N_all + M.word_off;
    V.Zm = Zm_all + M.word_off;
    V.S4 = S4_all + 4 * M.word_off;
    V.vbw = vbw_all + M.word_off;
    V.g = g;
    V.W = (int)M.W;
    V.lim = M.G.lim;
    const int ci = V.cube_case(x, y, z);
    const uint64_t packed = c_tri[ci];
#pragma unroll
    for (int cc = 0; cc < 3; cc++) c[cc] = (int)(vbase + V.corner_vid(x, y, z, tri_entry(packed, 3 * j + cc), ci));
    tri[3 * (size_t)t + 0] = c[0];
    tri[3 * (size_t)t + 1] = c[1];
    tri[3 * (size_t)t + 2] = c[2];
    }
    /* triangles with a repeated corner (zero grid points) are dropped from the output
       (CudaSurf.cu:820-829): counted here per block and in total (lanes past the count hold the
       distinct placeholder corners 0, 1, 2) */
    const uint64_t dm = __ballot(c[0] == c[1] || c[1] == c[2] || c[0] == c[2]);
    if (dm && (threadIdx.x & 63) == (uint32_t)__builtin_ctzll(dm)) {
        atomicAdd(&deg_blk[blk], (uint32_t)__popcll(dm));
        atomicAdd(deg_total, (uint32_t)__popcll(dm));
    }
    __syncthreads(); /* S is staged again for the next block */
    }
}

/* ------------------------------------------------------------------------------------------ */
/* closest atom per vertex (computeClosestAtom Kernels.cu:547-604)                              */
/* closest atom of vertex v (position p) of molecule M, relative to its first atom (-1: none) */
__device__ __forceinline__ int closest_one(const QsMol &M, float px, float py, float pz, const uint32_t *__restrict__ cstart_all,
                                           const float4 *__restrict__ sx, const int *__restrict__ perm, int sorted_ids)
{
    const QsGrid G = M.G;
    const uint32_t *__restrict__ cstart = cstart_all + M.cell_off;
    int cx = cell_of(px, G.ox, G.dxn), cy = cell_of(py, G.oy, G.dxn), cz = cell_of(pz, G.oz, G.dxn);
    const int gn = G.gn;
    /* The reference visits the 27 clamped cells x-major, z fastest, and keeps the first atom whose
       rounded distance is strictly below the best so far.  The three z cells of an (x, y) column
       are consecutive cells, so their atoms are one contiguous range in that same order; a cell
       the clamp visits twice cannot improve on its first visit.  sqrtf is monotonic, so an atom
       can only win if its squared sum is below the best one's; the root is taken only then. */
    const int z0 = iclamp(cz - 1, 0, gn - 1), z1 = iclamp(cz + 1, 0, gn - 1);
    int best = -1;
    float min_d = 999999.0f, min_s = INFINITY;
    /* the nine columns' ranges are loaded before any is walked */
    int lo9[9], hi9[9];
#pragma unroll
    for (int k = 0; k < 9; k++) {
        const int ix = iclamp(cx + k / 3 - 1, 0, gn - 1), iy = iclamp(cy + k % 3 - 1, 0, gn - 1);
        const int h = (gn * gn * ix) + (gn * iy);
        lo9[k] = (int)cstart[h + z0];
        hi9[k] = (int)cstart[h + z1 + 1];
    }
#pragma unroll
    for (int k = 0; k < 9; k++) {
        const int lo = lo9[k], hi = hi9[k];
        for (int id = lo; id < hi; id++) {
            const float4 at = sx[id];
            const float sq = dist3_sq(at.x, at.y, at.z, px, py, pz);
            if (sq < min_s) {
                const float d = sqrt_rn(sq);
                if (d < min_d) {
                    min_d = d;
                    min_s = sq;
                    best = id;
                }
            }
        }
    }
    /* atom indices are returned relative to the molecule's first atom */
    return (best < 0) ? -1 : ((sorted_ids ? best : perm[best]) - (int)M.atom_off);
}

__global__ void __launch_bounds__(256) k_closest_atom(const QsMol *__restrict__ mols, const uint32_t *__restrict__ vpre,
                                                      const uint32_t *__restrict__ vcoarse, const uint32_t *__restrict__ cstart_all,
                                                      const float4 *__restrict__ sx, const int *__restrict__ perm,
                                                      const float *__restrict__ vert, uint32_t nv, int sorted_ids,
                                                      int *__restrict__ out, const uint32_t *__restrict__ d_nv)
{
    if (d_nv) nv = min(*d_nv, nv); /* the device count within the launch's pool */
    const uint32_t nb = (nv + 255u) >> 8;
    /* persistent blocks over the 256-vertex blocks (grid a multiple of 8: see k_mc_vertices) */
    for (uint32_t vb = blockIdx.x; vb < nb; vb += gridDim.x) {
        const uint32_t v = xcd_group<64>(vb, nb) * 256u + threadIdx.x;
        if (v >= nv) continue;
        const QsMol &M = mols[seg_of_wave(vpre, vcoarse, v)];
        out[v] = closest_one(M, vert[3 * (size_t)v], vert[3 * (size_t)v + 1], vert[3 * (size_t)v + 2], cstart_all, sx, perm,
                             sorted_ids);
    }
}

/* ------------------------------------------------------------------------------------------ */
/* Laplacian smoothing (SmoothMesh.cpp:45-63, 96-124) in Jacobi form on the GPU                 */
/*
 * A vertex's neighbour list is, triangle by triangle in triangle order, the two other corners
 * in the order the reference pushes them (corner 0: 1,2; corner 1: 0,2; corner 2: 0,1).  Every
 * triangle that uses an edge vertex lies in one of the 4 cubes around that edge, every triangle
 * that uses a point vertex in one of the 8 cubes around that point, and triangle order is cube
 * order then case-table order.  So the list is walked from the case table, cube by cube in
 * order, without being stored; the sums run in that order, reading the previous pass.
 */
__device__ __forceinline__ void push_pair(const float *__restrict__ src, uint32_t a, uint32_t b, float &sx, float &sy,
                                          float &sz)
{
    const float *q0 = src + 3 * (size_t)a;
    sx = sx + q0[0];
    sy = sy + q0[1];
    sz = sz + q0[2];
    const float *q1 = src + 3 * (size_t)b;
    sx = sx + q1[0];
    sy = sy + q1[1];
    sz = sz + q1[2];
}

/* first triangle id of polygonised cube (x, y, z) (bit b of word wi): the word's first triangle
   tbw[wi] plus the triangles of the word's cubes below b, from its triangle-count bit planes */
__device__ __forceinline__ uint32_t cube_first_tri(const uint64_t *__restrict__ NT, const uint32_t *__restrict__ tbw, size_t wi,
                                                   int b)
{
    const uint64_t *np = NT + 3 * wi;
    const uint64_t below = (1ull << b) - 1ull;
    return tbw[wi] + (uint32_t)(__popcll(np[0] & below) + 2 * __popcll(np[1] & below) + 4 * __popcll(np[2] & below));
}

/* The grid records the smoothing walk reads (one molecule's view of the batch arrays) */
struct SmoothView {
    MCView V;
    const uint64_t *__restrict__ NT;
    const uint32_t *__restrict__ tbw;
};

__device__ __forceinline__ SmoothView smooth_view(const QsMol &M, const uint64_t *__restrict__ N_all,
                                                  const uint64_t *__restrict__ Zm_all, const uint64_t *__restrict__ S4_all,
                                                  const uint32_t *__restrict__ vbw_all, const uint64_t *__restrict__ NT_all,
                                                  const uint32_t *__restrict__ tbw_all)
{
    SmoothView S;
    S.V.N = N_all + M.word_off;
    S.V.Zm = Zm_all + M.word_off;
    S.V.S4 = S4_all + 4 * M.word_off;
    S.V.vbw = vbw_all + M.word_off;
    S.V.g = M.G.g;
    S.V.W = (int)M.W;
    S.V.lim = M.G.lim;
    S.NT = NT_all + 3 * M.word_off;
    S.tbw = tbw_all + M.word_off;
    return S;
}

/* The incident cubes of an edge vertex: the 4 cubes around the edge in cube order (offsets -1 / 0
   of the two other axes, major first); every one inside the grid contains the sign change, so it
   is polygonised and its word's triangle records exist.  Their first triangles and counts (0 for
   a cube outside the grid) are loaded here, all four before any is used. */
__device__ __forceinline__ void smooth_edge_cubes(const SmoothView &S, int x, int y, int z, int sl, int (&nt4)[4],
                                                  uint32_t (&tb4)[4])
{
    const int g = S.V.g, W = S.V.W, lim = S.V.lim;
#pragma unroll
    for (int k = 0; k < 4; k++) {
        const int hi = k >> 1, lo = k & 1;
        const int cx = x - (sl != 0 ? 1 - hi : 0);
        const int cy = y - (sl == 0 ? 1 - hi : (sl == 2 ? 1 - lo : 0));
        const int cz = z - (sl != 2 ? 1 - lo : 0);
        nt4[k] = 0;
        tb4[k] = 0;
        if (cx < 0 || cy < 0 || cz < 0 || cx >= lim || cy >= lim || cz >= lim) continue;
        const size_t wi = ((size_t)cx * g + cy) * W + (cz >> 6);
        const int b = cz & 63;
        const uint64_t *np = S.NT + 3 * wi;
        const uint64_t p0 = np[0], p1 = np[1], p2 = np[2], below = (1ull << b) - 1ull;
        tb4[k] = S.tbw[wi] + (uint32_t)(__popcll(p0 & below) + 2 * __popcll(p1 & below) + 4 * __popcll(p2 & below));
        nt4[k] = (int)((p0 >> b) & 1ull) + 2 * (int)((p1 >> b) & 1ull) + 4 * (int)((p2 >> b) & 1ull);
    }
}

/* vertex v's Jacobi sum over the neighbour list walked from its cubes (edge vertex: the cubes of
   smooth_edge_cubes; zero point: the 8 cubes around it), written to dst */
__device__ __forceinline__ void smooth_sum(const SmoothView &S, uint32_t v, int x, int y, int z, int sl, const int (&nt4)[4],
                                           const uint32_t (&tb4)[4], const int *__restrict__ tri,
                                           const float *__restrict__ src, float *__restrict__ dst)
{
    const int g = S.V.g, W = S.V.W, lim = S.V.lim;
    float sx = 0.0f, sy = 0.0f, sz = 0.0f;
    uint32_t cnt = 0;
    if (sl < 3) {
#pragma unroll
        for (int k = 0; k < 4; k++) {
            for (int t = 0; t < nt4[k]; t++) {
                const int *tt = tri + 3 * (size_t)(tb4[k] + t);
                const uint32_t a = (uint32_t)tt[0], b = (uint32_t)tt[1], c = (uint32_t)tt[2];
                if (a == v) push_pair(src, b, c, sx, sy, sz);
                else if (b == v) push_pair(src, a, c, sx, sy, sz);
                else if (c == v) push_pair(src, a, b, sx, sy, sz);
                else continue;
                cnt += 2;
            }
        }
    } else {
        /* a zero grid point: the 8 cubes around it, every corner that is this vertex */
        for (int k = 0; k < 8; k++) {
            const int cx = x - 1 + (k >> 2), cy = y - 1 + ((k >> 1) & 1), cz = z - 1 + (k & 1);
            if (cx < 0 || cy < 0 || cz < 0 || cx >= lim || cy >= lim || cz >= lim) continue;
            const int ci = S.V.cube_index(cx, cy, cz);
            const int nt = tri_count(ci);
            if (!nt) continue;
            const uint32_t tb = cube_first_tri(S.NT, S.tbw, ((size_t)cx * g + cy) * W + (cz >> 6), cz & 63);
            for (int t = 0; t < nt; t++) {
                const int *tt = tri + 3 * (size_t)(tb + t);
                const uint32_t a = (uint32_t)tt[0], b = (uint32_t)tt[1], c = (uint32_t)tt[2];
                if (a == v) { push_pair(src, b, c, sx, sy, sz); cnt += 2; }
                if (b == v) { push_pair(src, a, c, sx, sy, sz); cnt += 2; }
                if (c == v) { push_pair(src, a, b, sx, sy, sz); cnt += 2; }
            }
        }
    }
    const float fc = (float)cnt;
    dst[3 * (size_t)v + 0] = sx / fc;
    dst[3 * (size_t)v + 1] = sy / fc;
    dst[3 * (size_t)v + 2] = sz / fc;
}

/* grid point (x, y, z) and slot of a vertex key (grid point index << 2 | slot; a molecule's grid
   has fewer than 2^32 points) */
__device__ __forceinline__ void vertex_point(uint64_t key, int g, int &x, int &y, int &z, int &sl)
{
    sl = (int)(key & 3ull);
    const uint32_t p = (uint32_t)(key >> 2), ug = (uint32_t)g;
    const uint32_t xy = p / ug;
    z = (int)(p - xy * ug);
    x = (int)(xy / ug);
    y = (int)(xy - (uint32_t)x * ug);
}

/* one Jacobi pass, one thread per vertex.  A polygonised cube's triangles are consecutive from
   its first triangle id (its word's triangle base plus the triangle-count planes below it), and
   its count is its bit of those planes, so each incident triangle's other corners are read from
   the triangle array.  An edge vertex is a corner of exactly the triangles whose case-table entry
   is its cube edge (no other edge of the cube shares its id), so it finds them by id. */
__global__ void __launch_bounds__(256) k_smooth(const QsMol *__restrict__ mols, const uint32_t *__restrict__ vpre,
                                                const uint32_t *__restrict__ vcoarse,
                                                const uint64_t *__restrict__ N_all, const uint64_t *__restrict__ Zm_all,
                                                const uint64_t *__restrict__ S4_all, const uint32_t *__restrict__ vbw_all,
                                                const uint64_t *__restrict__ vkey, uint32_t nv,
                                                const uint64_t *__restrict__ NT_all,
                                                const uint32_t *__restrict__ tbw_all, const int *__restrict__ tri,
                                                const float *__restrict__ src,
                                                float *__restrict__ dst, const uint32_t *__restrict__ d_nv)
{
    /* a mesh larger than the pool (the run is repeated with a larger one) has corner ids past
       the pool: nothing to smooth then */
    if (d_nv && *d_nv > nv) return;
    if (d_nv) nv = *d_nv;
    const uint32_t nb = (nv + 255u) >> 8;
    /* persistent blocks over the 256-vertex blocks (grid a multiple of 8: see k_mc_vertices) */
    for (uint32_t vb = blockIdx.x; vb < nb; vb += gridDim.x) {
        const uint32_t v = xcd_group<64>(vb, nb) * 256u + threadIdx.x;
        if (v >= nv) continue;
        const QsMol &M = mols[seg_of_wave(vpre, vcoarse, v)];
        const SmoothView S = smooth_view(M, N_all, Zm_all, S4_all, vbw_all, NT_all, tbw_all);
        int x, y, z, sl;
        vertex_point(vkey[v], S.V.g, x, y, z, sl);
        int nt4[4] = {0, 0, 0, 0};
        uint32_t tb4[4] = {0u, 0u, 0u, 0u};
        if (sl < 3) smooth_edge_cubes(S, x, y, z, sl, nt4, tb4);
        smooth_sum(S, v, x, y, z, sl, nt4, tb4, tri, src, dst);
    }
}

/* ------------------------------------------------------------------------------------------ */
/* Laplacian smoothing over a triangle array (SmoothMesh.cpp:45-63, 96-124), for meshes built in
 * x-slabs, whose grid records do not outlive their slab.  Each vertex's incidences (corner index
 * 3 t + j of triangle t) are listed per vertex and sorted, so the Jacobi sums run over the
 * reference's neighbour list in its push order -- triangle by triangle, corner 0: 1,2; corner 1:
 * 0,2; corner 2: 0,1 -- as k_smooth walks it from the grid records: the results are the same bits. */
/* (corner ids outside [0, nv) are skipped: they cannot occur, and a mistake shows as a mesh
   difference rather than a stray write) */
__global__ void __launch_bounds__(256) k_incid_count(const int *__restrict__ tri, uint64_t ncorner, uint32_t nv,
                                                     uint32_t *__restrict__ cnt)
{
    const uint64_t i = (uint64_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= ncorner) return;
    const uint32_t v = (uint32_t)tri[i];
    if (v < nv) atomicAdd(&cnt[v], 1u);
}

__global__ void __launch_bounds__(256) k_incid_fill(const int *__restrict__ tri, uint64_t ncorner, uint32_t nv,
                                                    const uint32_t *__restrict__ start, uint32_t *__restrict__ cursor,
                                                    uint32_t *__restrict__ ent)
{
    const uint64_t i = (uint64_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= ncorner) return;
    const uint32_t v = (uint32_t)tri[i];
    if (v < nv) ent[start[v] + atomicAdd(&cursor[v], 1u)] = (uint32_t)i;
}

/* each vertex's incidences in increasing corner index (insertion sort; a few per vertex) */
__global__ void __launch_bounds__(256) k_incid_sort(const uint32_t *__restrict__ start, uint32_t nv, uint32_t *__restrict__ ent)
{
    const uint32_t v = blockIdx.x * 256 + threadIdx.x;
    if (v >= nv) return;
    const uint32_t a = start[v], b = start[v + 1];
    for (uint32_t i = a + 1; i < b; i++) {
        const uint32_t x = ent[i];
        uint32_t j = i;
        while (j > a && ent[j - 1] > x) {
            ent[j] = ent[j - 1];
            j--;
        }
        ent[j] = x;
    }
}

/* one Jacobi pass over the listed incidences */
__global__ void __launch_bounds__(256) k_smooth_incid(const uint32_t *__restrict__ start, const uint32_t *__restrict__ ent,
                                                      const int *__restrict__ tri, uint32_t nv, const float *__restrict__ src,
                                                      float *__restrict__ dst)
{
    const uint32_t v = xcd_group<64>(blockIdx.x, gridDim.x) * 256 + threadIdx.x;
    if (v >= nv) return;
    float sx = 0.0f, sy = 0.0f, sz = 0.0f;
    uint32_t cnt = 0;
    for (uint32_t k = start[v], e = start[v + 1]; k < e; k++) {
        const uint32_t c = ent[k];
        const uint32_t t = c / 3u, j = c - 3u * t;
        const int *tt = tri + 3 * (size_t)t;
        const uint32_t a = min((uint32_t)tt[j == 0 ? 1 : 0], nv - 1u), b = min((uint32_t)tt[j == 2 ? 1 : 2], nv - 1u);
        push_pair(src, a, b, sx, sy, sz);
        cnt += 2;
    }
    const float fc = (float)cnt;
    dst[3 * (size_t)v + 0] = sx / fc;
    dst[3 * (size_t)v + 1] = sy / fc;
    dst[3 * (size_t)v + 2] = sz / fc;
}

/* ------------------------------------------------------------------------------------------ */
/* Per-vertex normals (the normals field that MeshData leaves commented out, SmoothMesh.h:34):
 * the unit field gradient at the vertex's marching-cubes position.  The gradient at a grid point
 * is the central difference (one-sided at the grid faces) of the field; an edge vertex blends
 * its two end points' gradients with the vertex's edge parameter t, a zero-point vertex takes
 * its point's.  The field grows outwards (negative inside), so the normals point outwards. */
/* one molecule's field, addressed by grid row (x * g + y) and z */
struct FieldView {
    const float *F;
    const uint32_t *spre;
    const uint64_t *outm, *shellm;
    int g, W;
    float dx;
    __device__ __forceinline__ float at(size_t row, int z) const { return field_at(F, spre, outm, shellm, row, g, W, z, dx); }
};

/* derivative along axis a (0 x, 1 y, 2 z) at grid point (row, z), whose coordinate on that axis is i */
__device__ __forceinline__ float grad_axis(const FieldView &F, size_t row, int z, int a, int i, int g, float dx)
{
    const size_t dr = a == 0 ? (size_t)g : (a == 1 ? 1 : 0);
    const int dz = a == 2;
    if (i > 0 && i < g - 1) return (F.at(row + dr, z + dz) - F.at(row - dr, z - dz)) / (2.0f * dx);
    if (i == 0) return (F.at(row + dr, z + dz) - F.at(row, z)) / dx;
    return (F.at(row, z) - F.at(row - dr, z - dz)) / dx;
}

__global__ void __launch_bounds__(256) k_vertex_normals(const QsMol *__restrict__ mols, const uint32_t *__restrict__ vpre,
                                                        const uint32_t *__restrict__ vcoarse, const uint64_t *__restrict__ vkey,
                                                        uint32_t nv, const float *__restrict__ F_all,
                                                        const uint32_t *__restrict__ spre_all, const uint64_t *__restrict__ outm_all,
                                                        const uint64_t *__restrict__ shellm_all, float *__restrict__ out)
{
    const uint32_t v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= nv) return;
    const QsMol &M = mols[seg_of_wave(vpre, vcoarse, v)];
    const int g = M.G.g;
    const float dx = M.G.dx;
    const FieldView F{shellm_all ? F_all : F_all + M.vox_off, shellm_all ? spre_all + M.word_off : nullptr, outm_all + M.word_off,
                      shellm_all ? shellm_all + M.word_off : nullptr, g, (int)M.W, dx};
    const uint64_t key = vkey[v];
    const int sl = (int)(key & 3ull);
    const uint32_t p = (uint32_t)(key >> 2), ug = (uint32_t)g;
    const uint32_t xy = p / ug;
    int c[3];
    c[2] = (int)(p - xy * ug);
    c[0] = (int)(xy / ug);
    c[1] = (int)(xy - (uint32_t)c[0] * ug);
    const size_t row = (size_t)c[0] * g + c[1];
    const int z = c[2];
    float n[3];
#pragma unroll
    for (int a = 0; a < 3; a++) n[a] = grad_axis(F, row, z, a, c[a], g, dx);
    if (sl < 3) {
        const size_t rq = row + (sl == 0 ? (size_t)g : (sl == 1 ? 1 : 0));
        const int zq = z + (sl == 2);
        const float f0 = F.at(row, z), f1 = F.at(rq, zq);
        const float t = (0.0f - f0) / (f1 - f0);
        const float s = 1.0f - t;
#pragma unroll
        for (int a = 0; a < 3; a++) {
            const float g1 = grad_axis(F, rq, zq, a, c[a] + (a == sl), g, dx);
            n[a] = s * n[a] + t * g1;
        }
    }
    const float len = sqrtf(n[0] * n[0] + n[1] * n[1] + n[2] * n[2]);
    const float inv = len > 0.0f ? 1.0f / len : 0.0f;
    out[3 * (size_t)v + 0] = n[0] * inv;
    out[3 * (size_t)v + 1] = n[1] * inv;
    out[3 * (size_t)v + 2] = n[2] * inv;
}

/* Per-vertex colours (the colors field that MeshData leaves commented out, SmoothMesh.h:35): the
   colour of the vertex's closest atom, from a per-atom RGBA table laid out like the atom input
   (a batch's molecules concatenated); a vertex without a closest atom gets (0, 0, 0, 0).  When the
   ids are in the reference's cell-sorted convention (perm != nullptr), the id is mapped back to
   the input order through perm (sorted position -> input index, batch-global) first. */
__global__ void __launch_bounds__(256) k_vertex_colors(const QsMol *__restrict__ mols, const uint32_t *__restrict__ vpre,
                                                       const uint32_t *__restrict__ vcoarse, const int *__restrict__ atom,
                                                       const int *__restrict__ perm, uint32_t nv,
                                                       const float4 *__restrict__ table, float4 *__restrict__ out)
{
    const uint32_t v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= nv) return;
    const QsMol &M = mols[seg_of_wave(vpre, vcoarse, v)];
    const int a = atom[v];
    if (a < 0) {
        out[v] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        return;
    }
    const uint32_t in = perm ? (uint32_t)perm[M.atom_off + (uint32_t)a] : M.atom_off + (uint32_t)a;
    out[v] = table[in];
}

/* ------------------------------------------------------------------------------------------ */
/* degenerate-triangle filter (CudaSurf.cu:820-829), order preserving: block b of 256 triangles
   writes its kept ones from position 256 b - dpre[b] on, dpre being the exclusive prefix of the
   per-block degenerate counts */
__device__ __forceinline__ bool tri_degenerate(const int *__restrict__ tri, uint32_t t)
{
    const int x = tri[3 * (size_t)t], y = tri[3 * (size_t)t + 1], z = tri[3 * (size_t)t + 2];
    return !(x != y && y != z && x != z);
}

/* Per-molecule first vertex and first triangle of a batch, from the word scans: molecule m's are
   the exclusive scans at its first row word (wpre[m]; at or past the scanned words nw for a
   molecule without words, whose mesh is empty and starts at the totals).  vpre / tpre get nmol + 1 entries, the last one the
   totals (counts[1] vertices, counts[0] triangles).  In the same launch, the coarse index of the
   vertex prefix: entry k (of ncoarse, enough for a pool of vertices) = the last molecule whose
   first vertex is at or before vertex k << QS_COARSE_ITEMS, searched over the same first
   vertices read from the word scan. */
__device__ __forceinline__ uint32_t mol_first_vertex(const uint32_t *__restrict__ vbw, const uint64_t *__restrict__ wpre,
                                                     uint64_t nw, int nmol, const uint32_t *__restrict__ counts, int m)
{
    const uint64_t w = (m < nmol) ? wpre[m] : nw;
    return (w < nw) ? vbw[w] : counts[1];
}

__global__ void __launch_bounds__(256) k_mesh_bases(const uint32_t *__restrict__ vbw, const uint32_t *__restrict__ tbw,
                                                    const uint64_t *__restrict__ wpre, uint64_t nw, int nmol,
                                                    const uint32_t *__restrict__ counts, uint32_t *__restrict__ vpre,
                                                    uint64_t *__restrict__ tpre, uint32_t *__restrict__ coarse,
                                                    uint32_t ncoarse)
{
    const uint32_t i = blockIdx.x * 256 + threadIdx.x;
    if (i <= (uint32_t)nmol) {
        const int m = (int)i;
        const uint64_t w = (m < nmol) ? wpre[m] : nw;
        vpre[m] = (w < nw) ? vbw[w] : counts[1];
        tpre[m] = (w < nw) ? tbw[w] : counts[0];
    }
    if (i < ncoarse) {
        const uint64_t t = (uint64_t)i << QS_COARSE_ITEMS;
        int lo = 0, hi = nmol - 1;
        while (lo < hi) {
            const int mid = (lo + hi + 1) >> 1;
            if ((uint64_t)mol_first_vertex(vbw, wpre, nw, nmol, counts, mid) <= t) lo = mid;
            else hi = mid - 1;
        }
        coarse[i] = (uint32_t)lo;
    }
}

/* degenerate triangles per 256-triangle block (and their total, added to *deg_total), as the
   triangle kernel counts them: for triangle arrays assembled from several launches */
__global__ void __launch_bounds__(256) k_deg_count(const int *__restrict__ tri, uint32_t nt, uint32_t *__restrict__ deg_blk,
                                                   uint32_t *__restrict__ deg_total)
{
    __shared__ uint32_t lds[256 / 64 + 1];
    const uint32_t t = blockIdx.x * 256 + threadIdx.x;
    const uint32_t d = (t < nt && tri_degenerate(tri, t)) ? 1u : 0u;
    const uint32_t sum = block_sum<256>(d, lds);
    if (threadIdx.x == 0) {
        deg_blk[blockIdx.x] = sum;
        if (sum) atomicAdd(deg_total, sum);
    }
}

__global__ void __launch_bounds__(256) k_tri_compact(const int *__restrict__ tri, uint32_t nt,
                                                     const uint32_t *__restrict__ dpre, const uint32_t *__restrict__ ndeg,
                                                     int *__restrict__ out, const uint32_t *__restrict__ d_nt)
{
    __shared__ uint32_t wkeep[4];
    if (*ndeg == 0) return; /* nothing to drop: the caller keeps the input array */
    if (d_nt) nt = min(*d_nt, nt); /* the device count within the launch's pool */
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    /* persistent blocks over the 256-triangle blocks */
    for (uint32_t b = blockIdx.x; b * 256u < nt; b += gridDim.x) {
        const uint32_t t = b * 256u + threadIdx.x;
        const bool keep = t < nt && !tri_degenerate(tri, t);
        const uint64_t km = __ballot(keep);
        if (lane == 0) wkeep[wid] = (uint32_t)__popcll(km);
        __syncthreads();
        uint32_t o = b * 256u - dpre[b] + (uint32_t)__popcll(km & ((1ull << lane) - 1ull));
        for (int k = 0; k < wid; k++) o += wkeep[k];
        if (keep) {
            out[3 * (size_t)o] = tri[3 * (size_t)t];
            out[3 * (size_t)o + 1] = tri[3 * (size_t)t + 1];
            out[3 * (size_t)o + 2] = tri[3 * (size_t)t + 2];
        }
        __syncthreads(); /* wkeep is written again for the next block */
    }
}

/* kept triangles before each of n boundaries idx[i] (molecule starts and the end) */
__global__ void __launch_bounds__(256) k_tri_kept_before(const int *__restrict__ tri, const uint32_t *__restrict__ dpre,
                                                         const uint64_t *__restrict__ idx, int n, uint32_t tcap,
                                                         uint32_t *__restrict__ out, const uint32_t *__restrict__ totals,
                                                         const uint32_t *__restrict__ vpre, uint32_t *__restrict__ rb)
{
    /* with rb: the run's read-back in one array, {totals[0..3], vpre[0..n), out[0..n)} */
    if (rb && blockIdx.x == 0)
        for (int j = (int)threadIdx.x; j < 4 + n; j += 256) rb[j] = (j < 4) ? totals[j] : vpre[j - 4];
    /* one wave per boundary: the degenerate triangles of the boundary's 256-triangle block below
       it, four per lane, counted by ballot */
    const int i = blockIdx.x * 4 + (int)(threadIdx.x >> 6), lane = threadIdx.x & 63;
    if (i >= n) return;
    /* (boundaries past the triangle pool occur only when the mesh overflowed it) */
    const uint32_t t = (uint32_t)min(idx[i], (uint64_t)tcap);
    const uint32_t b = t >> 8;
    uint32_t d = 0;
#pragma unroll
    for (int k = 0; k < 4; k++) {
        const uint32_t u = (b << 8) + (uint32_t)(64 * k + lane);
        d += (uint32_t)__popcll(__ballot(u < t && tri_degenerate(tri, u)));
    }
    if (lane == 0) {
        out[i] = t - dpre[b] - d;
        if (rb) rb[4 + n + i] = t - dpre[b] - d;
    }
}

/* plain device copy (16-byte vectors, grid-stride): the reference point for traffic counters */
__global__ void __launch_bounds__(256) k_copy(const uint4 *__restrict__ src, uint4 *__restrict__ dst, uint64_t n16)
{
    for (uint64_t i = (uint64_t)blockIdx.x * 256 + threadIdx.x; i < n16; i += (uint64_t)gridDim.x * 256) dst[i] = src[i];
}

/* triangles of a batch with vertex indices local to their molecule */
__global__ void __launch_bounds__(256) k_tri_localize(const int *__restrict__ tri, uint32_t nt, const uint32_t *__restrict__ tpre,
                                                      const uint32_t *__restrict__ vpre, int nmol, int *__restrict__ out)
{
    uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nt) return;
    const int v0 = (int)vpre[seg_of(tpre, nmol, t)];
    for (int c = 0; c < 3; c++) out[3 * (size_t)t + c] = tri[3 * (size_t)t + c] - v0;
}

/* ------------------------------------------------------------------------------------------ */
/* launchers                                                                                    */

static inline unsigned nblk(uint64_t n, unsigned b) { return (unsigned)((n + b - 1) / b); }

/* Launches over a pool of items whose count lives on the device cover the pool with blocks that
   loop over 256-item blocks up to the count.  They use at most g_persist_blocks blocks, a multiple
   of 8 so that a block keeps its XCD from one item block to the next.  The default is one block
   per 256-item block of the pool (blocks past the count leave at once): fewer, longer-lived blocks
   were slower on these gather-bound kernels (smoothing +15 % and closest atom +11 % with 2048
   blocks, +19 % and +50 % with 1024, on the 50,000-atom 0.2 A mesh). */
static uint32_t g_persist_blocks = 1u << 30;

void qs_set_persistent_blocks(uint32_t n) { g_persist_blocks = std::max<uint32_t>(8u, n & ~7u); }
uint32_t qs_persistent_blocks(void) { return g_persist_blocks; }

static inline unsigned pgrid(uint64_t nb)
{
    return nb <= g_persist_blocks ? (unsigned)std::max<uint64_t>(nb, 1) : g_persist_blocks;
}

/* partial boxes per molecule: about 8 atoms per thread of the largest molecule, 1 to BB_PMAX */
int qs_bbox_parts(uint32_t max_atoms)
{
    const uint32_t p = (max_atoms + 8 * BB_NT - 1) / (8 * BB_NT);
    return (int)std::min<uint32_t>(std::max<uint32_t>(p, 1u), BB_PMAX);
}

size_t qs_bbox_part_floats(int nmol, int parts) { return (size_t)nmol * parts * BB_NV; }

void qs_launch_bbox(const float *pos, const float *rad, const uint32_t *off, uint32_t o0, uint32_t o1, int nmol, int parts,
                    float *part, float *out, hipStream_t s)
{
    /* one partial box per molecule is its box (part[m][0] and out[m] are at the same offset) */
    if (parts == 1) {
        hipLaunchKernelGGL(k_bbox_part, dim3(1, nmol), dim3(BB_NT), 0, s, pos, rad, off, o0, o1, out);
        return;
    }
    hipLaunchKernelGGL(k_bbox_part, dim3(parts, nmol), dim3(BB_NT), 0, s, pos, rad, off, o0, o1, part);
    hipLaunchKernelGGL(k_bbox_final, dim3(nmol), dim3(64), 0, s, part, parts, out);
}

void qs_launch_hash_count(QsBatch B, uint32_t nblocks, const float *pos, int2 *key, uint32_t *count, hipStream_t s)
{
    hipLaunchKernelGGL(k_hash_count, dim3(nblocks), dim3(256), 0, s, B, pos, key, count);
}

void qs_launch_scatter(const int2 *key, int n, const uint32_t *start, int *perm, hipStream_t s)
{
    hipLaunchKernelGGL(k_scatter, dim3(nblk(n, 256)), dim3(256), 0, s, key, n, start, perm);
}

void qs_launch_cell_finalize(const uint32_t *start, const uint32_t *count, int ncell, int *perm, int2 *cse, hipStream_t s)
{
    hipLaunchKernelGGL(k_cell_finalize, dim3(nblk(ncell, 256)), dim3(256), 0, s, start, count, ncell, perm, cse);
}

void qs_launch_gather_aux(QsBatch B, uint32_t nblocks, const int *perm, const float *pos, const float *rad, float4 *sx,
                          float4 *aux, hipStream_t s)
{
    hipLaunchKernelGGL(k_gather_aux, dim3(nblocks), dim3(256), 0, s, B, perm, pos, rad, sx, aux);
}

int qs_pass1_rows_per_wave(QsGrid G)
{
    /* R rows per wavefront with R * dx <= neighbour-cell size */
    int R = (int)(G.dxn / G.dx);
    return R < 1 ? 1 : (R > P1_RMAX ? P1_RMAX : R);
}

/* tuned rows per pass-1 block by voxel size (qs_set_tuned_rows): a short list, looked up by the
   nearest voxel size within 1 % */
static float g_rows_dx[16];
static int g_rows_v[16];
static int g_rows_n = 0;

int qs_set_tuned_rows(float dx, int rows)
{
    if (!(dx > 0.0f)) {
        g_rows_n = 0; /* dx <= 0 clears the table */
        return 0;
    }
    for (int i = 0; i < g_rows_n; i++)
        if (std::fabs(g_rows_dx[i] - dx) <= 0.01f * dx) {
            g_rows_v[i] = rows;
            return 0;
        }
    if (g_rows_n == 16) return -1;
    g_rows_dx[g_rows_n] = dx;
    g_rows_v[g_rows_n++] = rows;
    return 0;
}

int qs_pass1_slab_rows(QsGrid G, int W)
{
    /* 64 rows (or the tuned count for this voxel size, up to 256); both masks in at most 12 KB of LDS
       (a tuned count from a workload of small grids would otherwise give a large grid slabs that
       are too tall: pass 1 of the million-atom shell at 0.5 A, 12 words per row, took 1.22 ms with
       256-row slabs and 0.86 ms with the 64 rows of this cap (0.81 ms with 85), of a 200,000-atom
       molecule at 0.3 A, 11 words per row, 0.64 ms with 93 rows and 0.48 ms with 69, at 0.5 A, 7
       words per row, about 0.30 ms with 146 rows and 0.25 ms with 109); at most 80 neighbour
       cells of y per slab, so the candidate columns (3 x cells times the y cells + 2) stay
       within 256 */
    int yb = P1S_DEFROWS;
    for (int i = 0; i < g_rows_n; i++)
        if (std::fabs(g_rows_dx[i] - G.dx) <= 0.01f * G.dx && g_rows_v[i] >= 1) yb = std::min(P1S_MAXROWS, g_rows_v[i]);
    yb = std::min(yb, std::max(1, 768 / std::max(W, 1)));
    yb = std::min(yb, std::max(1, (int)(80.0f * G.dxn / G.dx)));
    return yb;
}

int qs_pass1_mode(bool direct, bool rows) { return direct ? 2 : (rows ? 1 : 0); }

uint32_t qs_pass1_blocks(const QsMol &M, int mode)
{
    const QsGrid &G = M.G;
    const uint32_t nx = (uint32_t)std::max(M.p1_hi - M.p1_lo, 0);
    if (mode == 2) return M.W * nblk(G.g, 4) * nx;
    if (mode == 1) return nblk(G.g, 4 * M.rows_per_wave) * nx;
    return nblk(G.g, M.slab_rows) * nx;
}

size_t qs_pass1_lds_bytes(const QsMol &M, int mode)
{
    if (mode != 0) return 0;
    return (size_t)16 * M.slab_rows * M.W + (size_t)8 * (M.G.gn + 2);
}

void qs_launch_pass1(QsBatch B, uint32_t nblocks, int mode, size_t lds_bytes, const uint32_t *cstart, const float4 *sx,
                     const float4 *aux, uint64_t *outm, uint64_t *shellm, uint64_t *negm, uint64_t *zerom, uint8_t *tflag,
                     uint8_t tmark, hipStream_t s)
{
    if (mode == 1)
        hipLaunchKernelGGL(k_pass1_rows, dim3(nblocks), dim3(256), 0, s, B, cstart, sx, aux, outm, shellm);
    else
        hipLaunchKernelGGL(k_pass1_slab, dim3(nblocks), dim3(256), lds_bytes, s, B, cstart, sx, aux, outm, shellm, negm, zerom,
                           tflag, tmark);
}

void qs_launch_pass1_direct(QsBatch B, uint32_t nblocks, const int2 *cse, const float4 *sx, uint64_t *outm, uint64_t *shellm,
                            hipStream_t s)
{
    hipLaunchKernelGGL(k_pass1_direct, dim3(nblocks), dim3(256), 0, s, B, cse, sx, outm, shellm);
}

/* field pass 2 kernel choice for half-window r: 0 fixed-R kernel (r <= 9), 1 runtime-r tile
   kernel, 2 direct window */
int qs_pass2_path(int r, bool direct)
{
    if (direct) return 2;
    if (r >= 1 && r <= 9) return 0;
    if (qs_pass2_lds_bytes(r) <= 65536 && 2 * r + P2_TZ <= 64) return 1;
    return 2;
}

uint32_t qs_pass2_blocks(const QsMol &M, int path)
{
    const QsGrid &G = M.G;
    const uint32_t nx = (uint32_t)std::max(M.p2_hi - M.p2_lo, 0);
    if (path == 2) return nblk(G.g, 64) * nblk(G.g, 4) * nx;
    return nblk(G.g, P2_TZ) * nblk(G.g, P2_TY) * nblk(nx, path == 0 ? P2R_TX : P2_TX);
}

/* one launch for a group of molecules sharing r (and so the path) */
void qs_launch_pass2(QsBatch B, uint32_t nblocks, int r, int path, const uint64_t *outm, const uint64_t *shellm,
                     const uint32_t *spre, float *field, uint32_t *negm, uint32_t *zerom, const int *offs, const int *norm_start, const int *enum_end,
                     QsNormFlags nf, const uint8_t *tflag, uint8_t tmark, uint32_t xchunk, hipStream_t s)
{
    if (path == 0) {
        xchunk = std::max(xchunk, 1u);
#define QS_P2R(RR)                                                                                                    \
    case RR:                                                                                                          \
        hipLaunchKernelGGL(k_pass2_r<RR>, dim3(nblocks), dim3(256), 0, s, B, outm, shellm, spre, field, negm, zerom, \
                           offs, norm_start, enum_end, nf, tflag, tmark, xchunk);                                            \
        return;
        switch (r) {
            QS_P2R(1)
            QS_P2R(2)
            QS_P2R(3)
            QS_P2R(4)
            QS_P2R(5)
            QS_P2R(6)
            QS_P2R(7)
            QS_P2R(8)
            QS_P2R(9)
        default:
            break;
        }
#undef QS_P2R
    }
    if (path <= 1) {
        hipLaunchKernelGGL(k_pass2_tile<uint16_t>, dim3(nblocks), dim3(256), qs_pass2_lds_bytes(r), s, B, outm, shellm, spre,
                           field, negm, zerom, offs, norm_start, enum_end);
        return;
    }
    hipLaunchKernelGGL(k_pass2_direct, dim3(nblocks), dim3(256), 0, s, B, outm, shellm, spre, field, negm, zerom);
}

void qs_launch_shell_scan(const uint64_t *shellm, uint64_t nw, uint32_t *sums, uint32_t *totals, uint32_t *spre, hipStream_t s)
{
    if (!nw) return;
    if (nw <= QS_SCAN1_MAX) {
        hipLaunchKernelGGL(k_shell_scan1, dim3(1), dim3(SCAN1_NT), 0, s, shellm, nw, totals, spre);
        return;
    }
    const uint64_t ntiles = qs_scan_tiles(nw);
    const int self = ntiles <= QS_SCAN_SELF_TILES;
    hipLaunchKernelGGL(k_shell_reduce, dim3((unsigned)ntiles), dim3(SCAN_NT), 0, s, shellm, nw, sums);
    if (!self) hipLaunchKernelGGL(k_scan_planes, dim3(1), dim3(SCAN_NT), 0, s, sums, ntiles, totals);
    hipLaunchKernelGGL(k_shell_scan, dim3((unsigned)ntiles), dim3(SCAN_NT), 0, s, shellm, nw, sums, spre, self, totals);
}

void qs_launch_labels_to_masks(const uint8_t *labels, int g, int W, uint64_t *outm, uint64_t *shellm, hipStream_t s)
{
    dim3 grid((unsigned)W, nblk(g, 4), (unsigned)g);
    hipLaunchKernelGGL(k_labels_to_masks, grid, dim3(256), 0, s, labels, g, W, outm, shellm);
}

uint32_t qs_mc_count_blocks(const QsMol &M)
{
    return nblk((uint64_t)M.G.g * M.W, 256) * (uint32_t)std::max(M.mc_hi - M.mc_lo, 0);
}

void qs_launch_mc_count(QsBatch B, uint32_t nblocks, const uint64_t *N, const uint64_t *Zm, uint64_t *S4, uint32_t *cnt,
                        uint64_t *NT, hipStream_t s)
{
    hipLaunchKernelGGL(k_mc_count, dim3(nblocks), dim3(256), 0, s, B, N, Zm, S4, cnt, NT);
}

void qs_launch_mc_emit(const QsMol *mols, const uint64_t *wpre, const uint32_t *wcoarse, const uint32_t *list, uint32_t nact,
                       const uint32_t *avb, uint32_t *vstart, uint32_t nv, const uint64_t *S4, const float *F,
                       const uint32_t *spre, const uint64_t *outm, const uint64_t *shellm, float *vert, uint64_t *vkey,
                       const uint32_t *atb, uint32_t *tstart, uint32_t ntri, const uint64_t *N, const uint64_t *Zm,
                       const uint32_t *vbw, const uint64_t *NT, int *tri, uint32_t vbase, uint32_t *deg_blk,
                       uint32_t *deg_total, const uint32_t *d_counts, hipStream_t s, hipStream_t s_vert, hipEvent_t ev_fork,
                       hipEvent_t ev_vert, bool starts_ready)
{
    if (!s_vert) s_vert = s;
    if (!nact || (!nv && !ntri)) {
        /* nothing to emit: s_vert still follows s up to here */
        if (s_vert != s) {
            (void)hipEventRecord(ev_fork, s);
            (void)hipStreamWaitEvent(s_vert, ev_fork, 0);
            (void)hipEventRecord(ev_vert, s_vert);
        }
        return;
    }
    /* the per-block degenerate counts start at zero (the triangle kernel adds to them), over the
       whole pool of ntri triangles */
    const uint32_t nzero = ntri ? ntri / 256 + 1 : 0u;
    const QsStarts sv = {avb, nv, d_counts ? d_counts + 1 : nullptr, vstart, nullptr, 0u};
    const QsStarts st = {atb, ntri, d_counts ? d_counts : nullptr, tstart, deg_blk, nzero};
    if (!starts_ready)
        hipLaunchKernelGGL(k_block_starts, dim3(pgrid(nblk(std::max<uint32_t>(nact, nzero), 256)), 2), dim3(256), 0, s, sv, st,
                           nact, d_counts ? d_counts + 2 : nullptr);
    /* the vertex kernel on s_vert (after the block starts), beside the triangle kernel on s */
    if (s_vert != s) {
        (void)hipEventRecord(ev_fork, s);
        (void)hipStreamWaitEvent(s_vert, ev_fork, 0);
    }
    if (nv)
        hipLaunchKernelGGL(k_mc_vertices, dim3(pgrid(nblk(nv, 256))), dim3(256), 0, s_vert, mols, wpre, wcoarse, list, avb, vstart, nact,
                           nv, S4, F, spre, outm, shellm, vert, vkey, d_counts);
    if (s_vert != s) (void)hipEventRecord(ev_vert, s_vert);
    if (ntri)
        hipLaunchKernelGGL(k_mc_triangles, dim3(pgrid(nblk(ntri, 256))), dim3(256), 0, s, mols, wpre, wcoarse, list, atb, tstart,
                           nact, ntri, N, Zm, S4, vbw, NT, tri, vbase, deg_blk, deg_total, d_counts);
}

void qs_launch_smooth(const QsMol *mols, const uint32_t *vpre, const uint32_t *vcoarse, const uint64_t *N, const uint64_t *Zm,
                      const uint64_t *S4, const uint32_t *vbw, const uint64_t *vkey, uint32_t nv,
                      const uint64_t *NT, const uint32_t *tbw, const int *tri, const float *src, float *dst,
                      const uint32_t *d_nv, hipStream_t s)
{
    if (!nv) return;
    hipLaunchKernelGGL(k_smooth, dim3(pgrid(nblk(nv, 256))), dim3(256), 0, s, mols, vpre, vcoarse, N, Zm, S4, vbw, vkey, nv,
                       NT, tbw, tri, src, dst, d_nv);
}

void qs_launch_vertex_normals(const QsMol *mols, const uint32_t *vpre, const uint32_t *vcoarse, const uint64_t *vkey, uint32_t nv,
                              const float *F, const uint32_t *spre, const uint64_t *outm, const uint64_t *shellm, float *out,
                              hipStream_t s)
{
    if (!nv) return;
    hipLaunchKernelGGL(k_vertex_normals, dim3(nblk(nv, 256)), dim3(256), 0, s, mols, vpre, vcoarse, vkey, nv, F, spre, outm, shellm,
                       out);
}

void qs_launch_vertex_colors(const QsMol *mols, const uint32_t *vpre, const uint32_t *vcoarse, const int *atom,
                             const int *perm, uint32_t nv, const float4 *table, float4 *out, hipStream_t s)
{
    if (!nv) return;
    hipLaunchKernelGGL(k_vertex_colors, dim3(nblk(nv, 256)), dim3(256), 0, s, mols, vpre, vcoarse, atom, perm, nv, table,
                       out);
}

void qs_launch_closest_atom(const QsMol *mols, const uint32_t *vpre, const uint32_t *vcoarse, const uint32_t *cstart,
                            const float4 *sx,
                            const int *perm, const float *vert, uint32_t nv, int sorted_ids, int *out, const uint32_t *d_nv,
                            hipStream_t s)
{
    if (!nv) return;
    hipLaunchKernelGGL(k_closest_atom, dim3(pgrid(nblk(nv, 256))), dim3(256), 0, s, mols, vpre, vcoarse, cstart, sx, perm, vert,
                       nv, sorted_ids, out, d_nv);
}

void qs_launch_copy(const void *src, void *dst, uint64_t bytes, hipStream_t s)
{
    const uint64_t n16 = bytes / 16;
    if (!n16) return;
    const uint64_t blocks = std::min<uint64_t>((n16 + 255) / 256, 8192);
    hipLaunchKernelGGL(k_copy, dim3((unsigned)blocks), dim3(256), 0, s, (const uint4 *)src, (uint4 *)dst, n16);
}

void qs_launch_tri_localize(const int *tri, uint32_t nt, const uint32_t *tpre, const uint32_t *vpre, int nmol, int *out,
                            hipStream_t s)
{
    if (nt == 0) return;
    hipLaunchKernelGGL(k_tri_localize, dim3(nblk(nt, 256)), dim3(256), 0, s, tri, nt, tpre, vpre, nmol, out);
}

void qs_launch_tri_compact(const int *tri, uint32_t nt, const uint32_t *dpre, const uint32_t *ndeg, int *out,
                           const uint32_t *d_nt, hipStream_t s)
{
    if (!nt) return;
    hipLaunchKernelGGL(k_tri_compact, dim3(pgrid(nblk(nt, 256))), dim3(256), 0, s, tri, nt, dpre, ndeg, out, d_nt);
}

void qs_launch_tri_kept_before(const int *tri, const uint32_t *dpre, const uint64_t *idx, int n, uint32_t tcap, uint32_t *out,
                               const uint32_t *totals, const uint32_t *vpre, uint32_t *rb, hipStream_t s)
{
    if (n <= 0) return;
    hipLaunchKernelGGL(k_tri_kept_before, dim3(nblk(n, 4)), dim3(256), 0, s, tri, dpre, idx, n, tcap, out, totals, vpre, rb);
}

/* scan helpers */
void qs_launch_mc_scans(const uint32_t *cnt, uint64_t nw, uint32_t *sums, uint32_t *vbw,
                        uint32_t *tbw, uint32_t *list, uint32_t *avb, uint32_t *atb, uint32_t *totals, uint32_t list_base,
                        hipStream_t s, QsBlockStarts bs)
{
    if (!nw) return;
    if (nw <= QS_SCAN1_MAX) {
        hipLaunchKernelGGL(k_mc_scan1, dim3(1), dim3(SCAN1_NT), 0, s, cnt, nw, vbw, tbw, list, avb, atb, list_base,
                           totals, bs);
        return;
    }
    const uint64_t ntiles = qs_scan_tiles(nw);
    const int self = ntiles <= QS_SCAN_SELF_TILES;
    hipLaunchKernelGGL(k_mc_reduce3, dim3((unsigned)ntiles), dim3(SCAN_NT), 0, s, cnt, nw, sums, ntiles);
    if (!self) hipLaunchKernelGGL(k_scan_planes, dim3(3), dim3(SCAN_NT), 0, s, sums, ntiles, totals);
    hipLaunchKernelGGL(k_mc_scan3, dim3((unsigned)ntiles), dim3(SCAN_NT), 0, s, cnt, nw, sums, ntiles, vbw, tbw, list,
                       avb, atb, list_base, self, totals, bs);
}

uint64_t qs_scan_tiles(uint64_t n) { return (n + SCAN_TILE - 1) / SCAN_TILE; }

void qs_launch_tile_reduce(const uint32_t *in, uint64_t n, uint32_t *sums, hipStream_t s)
{
    hipLaunchKernelGGL(k_tile_reduce, dim3((unsigned)qs_scan_tiles(n)), dim3(SCAN_NT), 0, s, in, n, sums);
}

void qs_launch_scan1(const uint32_t *in, uint32_t *out, uint64_t n, uint32_t *total, hipStream_t s)
{
    hipLaunchKernelGGL(k_scan1, dim3(1), dim3(SCAN1_NT), 0, s, in, out, n, total);
}

void qs_launch_tile_scan(const uint32_t *in, uint32_t *out, uint64_t n, const uint32_t *offsets, uint32_t *total,
                         hipStream_t s, const uint32_t *self_sums)
{
    hipLaunchKernelGGL(k_tile_scan, dim3((unsigned)qs_scan_tiles(n)), dim3(SCAN_NT), 0, s, in, out, n, offsets, total,
                       self_sums);
}

/* value of a uint32 array at given indices (per-molecule boundaries of scans) */
__global__ void k_gather_u32(const uint32_t *__restrict__ src, const uint64_t *__restrict__ idx, int n,
                             uint32_t *__restrict__ out)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = src[idx[i]];
}

void qs_launch_gather_u32(const uint32_t *src, const uint64_t *idx, int n, uint32_t *out, hipStream_t s)
{
    if (n <= 0) return;
    hipLaunchKernelGGL(k_gather_u32, dim3(nblk(n, 256)), dim3(256), 0, s, src, idx, n, out);
}

void qs_launch_smooth_incid_build(const int *tri, uint32_t nt, uint32_t nv, uint32_t *cnt, uint32_t *start, uint32_t *ent,
                                  hipStream_t s)
{
    const uint64_t nc = 3 * (uint64_t)nt;
    if (!nc || !nv) return;
    /* start[] (nv + 1 entries) holds the exclusive prefix of cnt[]; cnt[] is reused as the fill
       cursors, so the caller clears it between the scan and this fill */
    hipLaunchKernelGGL(k_incid_fill, dim3(nblk(nc, 256)), dim3(256), 0, s, tri, nc, nv, start, cnt, ent);
    hipLaunchKernelGGL(k_incid_sort, dim3(nblk(nv, 256)), dim3(256), 0, s, start, nv, ent);
}

void qs_launch_incid_count(const int *tri, uint32_t nt, uint32_t nv, uint32_t *cnt, hipStream_t s)
{
    const uint64_t nc = 3 * (uint64_t)nt;
    if (!nc || !nv) return;
    hipLaunchKernelGGL(k_incid_count, dim3(nblk(nc, 256)), dim3(256), 0, s, tri, nc, nv, cnt);
}

void qs_launch_smooth_incid(const uint32_t *start, const uint32_t *ent, const int *tri, uint32_t nv, const float *src,
                            float *dst, hipStream_t s)
{
    if (!nv) return;
    hipLaunchKernelGGL(k_smooth_incid, dim3(nblk(nv, 256)), dim3(256), 0, s, start, ent, tri, nv, src, dst);
}

void qs_launch_deg_count(const int *tri, uint32_t nt, uint32_t *deg_blk, uint32_t *deg_total, hipStream_t s)
{
    if (!nt) return;
    hipLaunchKernelGGL(k_deg_count, dim3(nblk(nt, 256)), dim3(256), 0, s, tri, nt, deg_blk, deg_total);
}

void qs_launch_mesh_bases(const uint32_t *vbw, const uint32_t *tbw, const uint64_t *wpre, uint64_t nw, int nmol,
                          const uint32_t *counts, uint32_t *vpre, uint64_t *tpre, uint32_t *vcoarse, uint32_t ncoarse, hipStream_t s)
{
    hipLaunchKernelGGL(k_mesh_bases, dim3(nblk(std::max<uint64_t>((uint64_t)nmol + 1, ncoarse), 256)), dim3(256), 0, s, vbw, tbw,
                       wpre, nw, nmol, counts, vpre, tpre, vcoarse, ncoarse);
}

uint32_t qs_pass2_tiles(const QsMol &M) { return qs_pass2_blocks(M, 0); }
uint32_t qs_pass2_ztiles(const QsMol &M) { return (uint32_t)((M.G.g + P2_TZ - 1) / P2_TZ); }
