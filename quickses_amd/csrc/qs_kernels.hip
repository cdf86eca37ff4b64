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
#include "qs_mc_table.h"

__constant__ uint64_t c_tri[256] = QS_TRI_PACKED_DATA;

/* cube corners and edges (see qs_mc_table.h) */
__constant__ int8_t c_edge_low[12][3] = {{0, 0, 0}, {1, 0, 0}, {0, 1, 0}, {0, 0, 0}, {0, 0, 1}, {1, 0, 1},
                                         {0, 1, 1}, {0, 0, 1}, {0, 0, 0}, {1, 0, 0}, {1, 1, 0}, {0, 1, 0}};
__constant__ int8_t c_edge_axis[12] = {0, 1, 0, 1, 0, 1, 0, 1, 2, 2, 2, 2};

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

/* number of triangles of case ci: the first nibble equal to 0xF, divided by 3 */
__device__ __forceinline__ int tri_count(int ci)
{
    uint64_t p = c_tri[ci];
    /* nibble k is 0xF for k >= 3*ntri; count leading non-0xF nibbles */
    int k = 0;
    while (k < 15 && tri_entry(p, k) != 0xF) k += 3;
    return k / 3;
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

/* out[i] = offsets[tile] + exclusive prefix inside the tile; in may alias out */
__global__ void __launch_bounds__(SCAN_NT) k_tile_scan(const uint32_t *in, uint32_t *out, uint64_t n,
                                                       const uint32_t *__restrict__ offsets,
                                                       uint32_t *__restrict__ total)
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
    uint32_t off = offsets ? offsets[blockIdx.x] : 0u;
    uint32_t run = off + pre;
#pragma unroll
    for (int k = 0; k < SCAN_IPT; k++) {
        uint64_t i = base + k;
        if (i < n) out[i] = run;
        run += v[k];
    }
    if (total && blockIdx.x == gridDim.x - 1 && threadIdx.x == 0) *total = off + tot;
}

/* ------------------------------------------------------------------------------------------ */
/* bounding box (getMinMax CudaSurf.cu:149-174): partial per block, then one block             */

#define BB_NT 256
#define BB_NV 8 /* min x,y,z, max x,y,z, max radius, count of non-finite inputs */
__global__ void __launch_bounds__(BB_NT) k_bbox_partial(const float *__restrict__ pos, const float *__restrict__ rad,
                                                        int n, float *__restrict__ part)
{
    __shared__ float s[BB_NV][BB_NT];
    float v[BB_NV] = {100000.0f, 100000.0f, 100000.0f, -100000.0f, -100000.0f, -100000.0f, 0.0f, 0.0f};
    for (int i = blockIdx.x * BB_NT + threadIdx.x; i < n; i += gridDim.x * BB_NT) {
        for (int c = 0; c < 3; c++) {
            float p = pos[3 * i + c];
            v[c] = fminf(v[c], p);
            v[3 + c] = fmaxf(v[3 + c], p);
            if (!isfinite(p)) v[7] += 1.0f;
        }
        float ra = rad[i];
        v[6] = fmaxf(v[6], ra);
        if (!isfinite(ra)) v[7] += 1.0f;
    }
    for (int c = 0; c < BB_NV; c++) s[c][threadIdx.x] = v[c];
    __syncthreads();
    for (int st = BB_NT / 2; st > 0; st >>= 1) {
        if (threadIdx.x < st)
            for (int c = 0; c < BB_NV; c++) {
                float a = s[c][threadIdx.x], b = s[c][threadIdx.x + st];
                s[c][threadIdx.x] = (c < 3) ? fminf(a, b) : (c < 7 ? fmaxf(a, b) : a + b);
            }
        __syncthreads();
    }
    if (threadIdx.x < BB_NV) part[BB_NV * blockIdx.x + threadIdx.x] = s[threadIdx.x][0];
}

__global__ void __launch_bounds__(BB_NT) k_bbox_final(const float *__restrict__ part, int nparts, float *__restrict__ out)
{
    if (threadIdx.x != 0) return;
    float v[BB_NV] = {100000.0f, 100000.0f, 100000.0f, -100000.0f, -100000.0f, -100000.0f, 0.0f, 0.0f};
    for (int b = 0; b < nparts; b++)
        for (int c = 0; c < BB_NV; c++) {
            float p = part[BB_NV * b + c];
            v[c] = (c < 3) ? fminf(v[c], p) : (c < 7 ? fmaxf(v[c], p) : v[c] + p);
        }
    for (int c = 0; c < BB_NV; c++) out[c] = v[c];
}

/* ------------------------------------------------------------------------------------------ */
/* atom hashing and cell ordering                                                              */

/* hashAtoms Kernels.cu:185-216: cell id x-major; count atoms per cell */
__global__ void __launch_bounds__(256) k_hash_count(const float *__restrict__ pos, int n, QsGrid G,
                                                    int *__restrict__ key, uint32_t *__restrict__ count)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    /* in range for every atom by construction of the grid; clamped only as a memory guard */
    int cx = iclamp(cell_of(pos[3 * i + 0], G.ox, G.dxn), 0, G.gn - 1);
    int cy = iclamp(cell_of(pos[3 * i + 1], G.oy, G.dxn), 0, G.gn - 1);
    int cz = iclamp(cell_of(pos[3 * i + 2], G.oz, G.dxn), 0, G.gn - 1);
    int h = (G.gn * G.gn * cx) + (G.gn * cy) + cz;
    key[i] = h;
    atomicAdd(&count[h], 1u);
}

__global__ void __launch_bounds__(256) k_scatter(const int *__restrict__ key, int n, const uint32_t *__restrict__ start,
                                                 uint32_t *__restrict__ cursor, int *__restrict__ perm)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    int h = key[i];
    uint32_t slot = start[h] + atomicAdd(&cursor[h], 1u);
    perm[slot] = i;
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

__global__ void __launch_bounds__(256) k_gather_atoms(const int *__restrict__ perm, const float *__restrict__ pos,
                                                      const float *__restrict__ rad, int n, float4 *__restrict__ sx)
{
    int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    int id = perm[k];
    sx[k] = make_float4(pos[3 * id], pos[3 * id + 1], pos[3 * id + 2], rad[id]);
}

/* ------------------------------------------------------------------------------------------ */
/* field pass 1: probeIntersection + computeInOrOut Kernels.cu:344-475                         */

/* block = 4 (x) * 8 (y) * 8 (z) voxels */
__global__ void __launch_bounds__(256) k_pass1(QsGrid G, const int2 *__restrict__ cse, const float4 *__restrict__ sx,
                                               uint8_t *__restrict__ labels, int x0, int x1)
{
    const int g = G.g;
    int z = blockIdx.x * 8 + (threadIdx.x & 7);
    int y = blockIdx.y * 8 + ((threadIdx.x >> 3) & 7);
    int x = x0 + blockIdx.z * 4 + (threadIdx.x >> 6);
    if (x >= x1 || y >= g || z >= g) return;
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
    labels[((size_t)x * g + y) * g + z] = res;
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
 * of a ballot mask, then y, then x) and evaluates the reference's float distance only for those
 * candidates.  The result is the reference's value bit for bit.
 */

#define P2_TX 8
#define P2_TY 8
#define P2_TZ 32
#define P2_INF 0xFFFF

size_t qs_pass2_lds_bytes(int r)
{
    size_t LX = P2_TX + 2 * r, LY = P2_TY + 2 * r;
    return LX * LY * 8 + LX * LY * P2_TZ * 2 + LX * P2_TY * P2_TZ * 2;
}

__global__ void __launch_bounds__(256) k_pass2_tile(QsGrid G, const uint8_t *__restrict__ labels,
                                                    float *__restrict__ field, const int *__restrict__ offs,
                                                    const int *__restrict__ norm_start,
                                                    const int *__restrict__ enum_end, int x0, int x1)
{
    extern __shared__ __align__(16) uint8_t smem[];
    const int r = G.r, g = G.g;
    const int LX = P2_TX + 2 * r, LY = P2_TY + 2 * r, LZ = P2_TZ + 2 * r;
    uint64_t *rows = (uint64_t *)smem;
    uint16_t *DZ = (uint16_t *)(smem + (size_t)LX * LY * 8);
    uint16_t *DY = DZ + (size_t)LX * LY * P2_TZ;
    const int tz0 = blockIdx.x * P2_TZ, ty0 = blockIdx.y * P2_TY, tx0 = x0 + blockIdx.z * P2_TX;
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;

    /* 1. outside mask of the haloed tile, one 64-bit ballot per (x,y) row */
    for (int row = wid; row < LX * LY; row += 4) {
        int lx = row / LY, ly = row - lx * LY;
        int gx = tx0 - r + lx, gy = ty0 - r + ly, gz = tz0 - r + lane;
        bool out = false;
        if (lane < LZ && gx >= 0 && gx < g && gy >= 0 && gy < g && gz >= 0 && gz < g)
            out = labels[((size_t)gx * g + gy) * g + gz] == QS_OUT;
        uint64_t m = __ballot(out);
        if (lane == 0) rows[row] = m;
    }
    __syncthreads();

    /* 2. z: squared distance to the nearest outside voxel within |c| <= r along the row */
    for (int w = threadIdx.x; w < LX * LY * P2_TZ; w += 256) {
        int row = w / P2_TZ, z = w - row * P2_TZ;
        uint64_t m = rows[row];
        int c = z + r;
        int best = P2_INF;
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
        DZ[w] = (uint16_t)best;
    }
    __syncthreads();

    /* 3. y */
    for (int w = threadIdx.x; w < LX * P2_TY * P2_TZ; w += 256) {
        int lx = w / (P2_TY * P2_TZ);
        int rem = w - lx * (P2_TY * P2_TZ);
        int ty = rem / P2_TZ, z = rem - ty * P2_TZ;
        int best = P2_INF;
        for (int b = -r; b <= r; b++) {
            int v = DZ[((size_t)lx * LY + (ty + r + b)) * P2_TZ + z];
            if (v != P2_INF) best = min(best, v + b * b);
        }
        DY[w] = (uint16_t)best;
    }
    __syncthreads();

    /* 4. x, then the field value of each output voxel */
    const int ty = threadIdx.x / P2_TZ, tz = threadIdx.x - ty * P2_TZ;
    const int gy = ty0 + ty, gz = tz0 + tz;
    if (gy >= g || gz >= g) return;
    const float dx = G.dx;
    const float vy = grid_pos(G.oy, gy, dx), vz = grid_pos(G.oz, gz, dx);
    for (int tx = 0; tx < P2_TX; tx++) {
        int gx = tx0 + tx;
        if (gx >= x1 || gx >= g) break;
        size_t vid = ((size_t)gx * g + gy) * g + gz;
        uint8_t lab = labels[vid];
        float out;
        if (lab == QS_OUT) {
            out = QS_PROBE;
        } else if (lab == QS_IN) {
            out = -dx;
        } else {
            int nmin = P2_INF;
            for (int a = -r; a <= r; a++) {
                int v = DY[((size_t)(tx + r + a) * P2_TY + ty) * P2_TZ + tz];
                if (v != P2_INF) nmin = min(nmin, v + a * a);
            }
            if (nmin == P2_INF) {
                out = -dx;
            } else {
                const float vx = grid_pos(G.ox, gx, dx);
                float min_dist = 100000.0f;
                for (int k = norm_start[nmin]; k < enum_end[nmin]; k++) {
                    int o = offs[k];
                    int a = (int)(int8_t)(o & 0xFF), b = (int)(int8_t)((o >> 8) & 0xFF), c = (int)(int8_t)((o >> 16) & 0xFF);
                    uint64_t m = rows[(size_t)(tx + r + a) * LY + (ty + r + b)];
                    if (!((m >> (tz + r + c)) & 1ull)) continue;
                    float d = dist3(grid_pos(G.ox, gx + a, dx), grid_pos(G.oy, gy + b, dx), grid_pos(G.oz, gz + c, dx),
                                    vx, vy, vz);
                    min_dist = fminf(d, min_dist);
                }
                out = (min_dist < 999.0f) ? (QS_PROBE - min_dist) : -dx;
            }
        }
        field[vid] = out;
    }
}

/* direct (2r+1)^3 window, one thread per voxel: any r; also a cross-check of the tile kernel */
__global__ void __launch_bounds__(256) k_pass2_direct(QsGrid G, const uint8_t *__restrict__ labels,
                                                      float *__restrict__ field, int x0, int x1)
{
    const int g = G.g, r = G.r;
    int z = blockIdx.x * 64 + (threadIdx.x & 63);
    int y = blockIdx.y * 4 + (threadIdx.x >> 6);
    int x = x0 + blockIdx.z;
    if (x >= x1 || y >= g || z >= g) return;
    size_t vid = ((size_t)x * g + y) * g + z;
    uint8_t lab = labels[vid];
    const float dx = G.dx;
    float out;
    if (lab == QS_OUT) {
        out = QS_PROBE;
    } else if (lab == QS_IN) {
        out = -dx;
    } else {
        float vx = grid_pos(G.ox, x, dx), vy = grid_pos(G.oy, y, dx), vz = grid_pos(G.oz, z, dx);
        float min_dist = 100000.0f;
        for (int a = -r; a <= r; a++) {
            int qx = iclamp(x + a, 0, g - 1);
            for (int b = -r; b <= r; b++) {
                int qy = iclamp(y + b, 0, g - 1);
                for (int c = -r; c <= r; c++) {
                    int qz = iclamp(z + c, 0, g - 1);
                    if (labels[((size_t)qx * g + qy) * g + qz] != QS_OUT) continue;
                    float d = dist3(grid_pos(G.ox, qx, dx), grid_pos(G.oy, qy, dx), grid_pos(G.oz, qz, dx), vx, vy, vz);
                    min_dist = fminf(d, min_dist);
                }
            }
        }
        out = (min_dist < 999.0f) ? (QS_PROBE - min_dist) : -dx;
    }
    field[vid] = out;
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

#define MC_NT 256
#define MC_IPT 4
#define MC_B (MC_NT * MC_IPT)

struct P3 {
    int x, y, z;
};

__device__ __forceinline__ P3 unflatten(uint64_t p, int g)
{
    uint32_t gg = (uint32_t)g;
    uint64_t xy = p / gg;
    P3 r;
    r.z = (int)(p - xy * gg);
    r.x = (int)(xy / gg);
    r.y = (int)(xy - (uint64_t)r.x * gg);
    return r;
}

__device__ __forceinline__ float fat(const float *__restrict__ f, int g, int x, int y, int z)
{
    return f[((size_t)x * g + y) * g + z];
}

/* edge (p, axis) lies on a polygonised cube iff p[axis] < lim and the other two <= lim */
__device__ __forceinline__ bool edge_in_window(int x, int y, int z, int axis, int lim)
{
    bool ok = (axis == 0 ? x < lim : x <= lim);
    ok = ok && (axis == 1 ? y < lim : y <= lim);
    ok = ok && (axis == 2 ? z < lim : z <= lim);
    return ok;
}

__device__ __forceinline__ unsigned slot_mask(const float *__restrict__ F, int g, int lim, int x, int y, int z)
{
    unsigned m = 0;
    float f = fat(F, g, x, y, z);
    bool fn = f < 0.0f;
#pragma unroll
    for (int a = 0; a < 3; a++) {
        int qx = x + (a == 0), qy = y + (a == 1), qz = z + (a == 2);
        if ((a == 0 ? qx : (a == 1 ? qy : qz)) < g && edge_in_window(x, y, z, a, lim)) {
            float fq = fat(F, g, qx, qy, qz);
            if (fn != (fq < 0.0f)) {
                float nonneg = fn ? fq : f;
                if (nonneg != 0.0f) m |= 1u << a;
            }
        }
    }
    if (f == 0.0f) {
        for (int a = 0; a < 3 && !(m & 8u); a++) {
            for (int s = -1; s <= 1; s += 2) {
                int qx = x + (a == 0 ? s : 0), qy = y + (a == 1 ? s : 0), qz = z + (a == 2 ? s : 0);
                int qa = (a == 0 ? qx : (a == 1 ? qy : qz));
                if (qa < 0 || qa >= g) continue;
                int lx = (s < 0) ? qx : x, ly = (s < 0) ? qy : y, lz = (s < 0) ? qz : z;
                if (!edge_in_window(lx, ly, lz, a, lim)) continue;
                if (fat(F, g, qx, qy, qz) < 0.0f) {
                    m |= 8u;
                    break;
                }
            }
        }
    }
    return m;
}

/* countVertexPerCell MarchingCubes.cu:571-658 (cube index) plus vertex slots; per-block sums */
__global__ void __launch_bounds__(MC_NT) k_mc_count(QsGrid G, const float *__restrict__ F, uint16_t *__restrict__ pinfo,
                                                    uint32_t *__restrict__ bs_tri, uint32_t *__restrict__ bs_vert,
                                                    uint64_t p0, uint64_t p1)
{
    __shared__ uint32_t lds[MC_NT / 64 + 1];
    const int g = G.g, lim = G.lim;
    uint64_t base = p0 + (uint64_t)blockIdx.x * MC_B;
    uint32_t st = 0, sv = 0;
#pragma unroll
    for (int k = 0; k < MC_IPT; k++) {
        uint64_t p = base + (uint64_t)k * MC_NT + threadIdx.x;
        if (p >= p1) continue;
        P3 q = unflatten(p, g);
        unsigned ci = 0;
        if (q.x < lim && q.y < lim && q.z < lim) {
            ci = (unsigned)(fat(F, g, q.x, q.y, q.z) < 0.0f);
            ci += (unsigned)(fat(F, g, q.x + 1, q.y, q.z) < 0.0f) * 2u;
            ci += (unsigned)(fat(F, g, q.x + 1, q.y + 1, q.z) < 0.0f) * 4u;
            ci += (unsigned)(fat(F, g, q.x, q.y + 1, q.z) < 0.0f) * 8u;
            ci += (unsigned)(fat(F, g, q.x, q.y, q.z + 1) < 0.0f) * 16u;
            ci += (unsigned)(fat(F, g, q.x + 1, q.y, q.z + 1) < 0.0f) * 32u;
            ci += (unsigned)(fat(F, g, q.x + 1, q.y + 1, q.z + 1) < 0.0f) * 64u;
            ci += (unsigned)(fat(F, g, q.x, q.y + 1, q.z + 1) < 0.0f) * 128u;
        }
        unsigned m = slot_mask(F, g, lim, q.x, q.y, q.z);
        pinfo[p] = (uint16_t)(ci | (m << 8));
        st += tri_count(ci);
        sv += __popc(m);
    }
    uint32_t t = block_sum<MC_NT>(st, lds);
    uint32_t v = block_sum<MC_NT>(sv, lds);
    if (threadIdx.x == 0) {
        bs_tri[blockIdx.x] = t;
        bs_vert[blockIdx.x] = v;
    }
}

/* vertex ids and positions (vertexInterp/lerp MarchingCubes.cu:715-734, edge oriented low->high) */
__global__ void __launch_bounds__(MC_NT) k_mc_verts(QsGrid G, const float *__restrict__ F,
                                                    const uint16_t *__restrict__ pinfo,
                                                    const uint32_t *__restrict__ bo_vert, uint32_t *__restrict__ vbase,
                                                    float *__restrict__ vert, uint64_t p0, uint64_t p1)
{
    __shared__ uint32_t lds[MC_NT / 64 + 1];
    const int g = G.g;
    const float dx = G.dx;
    uint64_t base = p0 + (uint64_t)blockIdx.x * MC_B;
    uint32_t run = bo_vert[blockIdx.x];
    for (int k = 0; k < MC_IPT; k++) {
        uint64_t p = base + (uint64_t)k * MC_NT + threadIdx.x;
        unsigned m = (p < p1) ? (unsigned)(pinfo[p] >> 8) : 0u;
        uint32_t tot;
        uint32_t pre = block_excl_scan<MC_NT>((uint32_t)__popc(m), &tot, lds);
        uint32_t v = run + pre;
        run += tot;
        if (p >= p1 || !m) continue;
        vbase[p] = v;
        P3 q = unflatten(p, g);
        float A[3] = {grid_pos(G.ox, q.x, dx), grid_pos(G.oy, q.y, dx), grid_pos(G.oz, q.z, dx)};
        float f0 = fat(F, g, q.x, q.y, q.z);
#pragma unroll
        for (int a = 0; a < 3; a++) {
            if (!(m & (1u << a))) continue;
            float f1 = fat(F, g, q.x + (a == 0), q.y + (a == 1), q.z + (a == 2));
            float B[3] = {A[0], A[1], A[2]};
            B[a] = A[a] + dx;
            float t = (0.0f - f0) / (f1 - f0);
#pragma unroll
            for (int c = 0; c < 3; c++) {
                float dd = B[c] - A[c];
                float s = t * dd;
                vert[3 * (size_t)v + c] = A[c] + s;
            }
            v++;
        }
        if (m & 8u) {
            vert[3 * (size_t)v + 0] = A[0];
            vert[3 * (size_t)v + 1] = A[1];
            vert[3 * (size_t)v + 2] = A[2];
        }
    }
}

/* triangles: generateTriangleVerticesSMEM MarchingCubes.cu:736-851 order, vertex ids by edge */
__global__ void __launch_bounds__(MC_NT) k_mc_tris(QsGrid G, const float *__restrict__ F,
                                                   const uint16_t *__restrict__ pinfo, const uint32_t *__restrict__ vbase,
                                                   const uint32_t *__restrict__ bo_tri, int *__restrict__ tri, uint64_t p0,
                                                   uint64_t p1)
{
    __shared__ uint32_t lds[MC_NT / 64 + 1];
    const int g = G.g;
    uint64_t base = p0 + (uint64_t)blockIdx.x * MC_B;
    uint32_t run = bo_tri[blockIdx.x];
    for (int k = 0; k < MC_IPT; k++) {
        uint64_t p = base + (uint64_t)k * MC_NT + threadIdx.x;
        unsigned ci = (p < p1) ? (unsigned)(pinfo[p] & 0xFF) : 0u;
        int nt = tri_count(ci);
        uint32_t tot;
        uint32_t pre = block_excl_scan<MC_NT>((uint32_t)nt, &tot, lds);
        uint32_t t0 = run + pre;
        run += tot;
        if (nt == 0) continue;
        P3 q = unflatten(p, g);
        uint64_t packed = c_tri[ci];
        for (int e3 = 0; e3 < 3 * nt; e3++) {
            int e = tri_entry(packed, e3);
            int ax = q.x + c_edge_low[e][0], ay = q.y + c_edge_low[e][1], az = q.z + c_edge_low[e][2];
            int axis = c_edge_axis[e];
            int bx = ax + (axis == 0), by = ay + (axis == 1), bz = az + (axis == 2);
            float fa = fat(F, g, ax, ay, az), fb = fat(F, g, bx, by, bz);
            bool a_neg = fa < 0.0f;
            float fnn = a_neg ? fb : fa;
            uint32_t vid;
            if (fnn == 0.0f) {
                size_t pn = a_neg ? (((size_t)bx * g + by) * g + bz) : (((size_t)ax * g + ay) * g + az);
                vid = vbase[pn] + (uint32_t)__popc((pinfo[pn] >> 8) & 7u);
            } else {
                size_t pa = ((size_t)ax * g + ay) * g + az;
                vid = vbase[pa] + (uint32_t)__popc((pinfo[pa] >> 8) & ((1u << axis) - 1u));
            }
            tri[3 * (size_t)t0 + e3] = (int)vid;
        }
    }
}

/* ------------------------------------------------------------------------------------------ */
/* closest atom per vertex (computeClosestAtom Kernels.cu:547-604)                              */
__global__ void __launch_bounds__(256) k_closest_atom(QsGrid G, const int2 *__restrict__ cse, const float4 *__restrict__ sx,
                                                      const int *__restrict__ perm, const float *__restrict__ vert,
                                                      uint32_t nv, int sorted_ids, int *__restrict__ out)
{
    uint32_t v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= nv) return;
    float px = vert[3 * (size_t)v], py = vert[3 * (size_t)v + 1], pz = vert[3 * (size_t)v + 2];
    int cx = cell_of(px, G.ox, G.dxn), cy = cell_of(py, G.oy, G.dxn), cz = cell_of(pz, G.oz, G.dxn);
    const int gn = G.gn;
    int best = -1;
    float min_d = 999999.0f;
    for (int a = -1; a <= 1; a++) {
        int ix = iclamp(cx + a, 0, gn - 1);
        for (int b = -1; b <= 1; b++) {
            int iy = iclamp(cy + b, 0, gn - 1);
            for (int c = -1; c <= 1; c++) {
                int iz = iclamp(cz + c, 0, gn - 1);
                int2 se = cse[(gn * gn * ix) + (gn * iy) + iz];
                if (se.x >= QS_EMPTY) continue;
                for (int id = se.x; id < se.y; id++) {
                    float4 at = sx[id];
                    float d = dist3(at.x, at.y, at.z, px, py, pz);
                    if (d < min_d) {
                        min_d = d;
                        best = id;
                    }
                }
            }
        }
    }
    out[v] = (best >= 0 && !sorted_ids) ? perm[best] : best;
}

/* ------------------------------------------------------------------------------------------ */
/* Laplacian smoothing (SmoothMesh.cpp:45-63, 96-124) in Jacobi form on the GPU                 */
/*
 * A vertex's neighbour list is, triangle by triangle in triangle order, the two other corners
 * in the order the reference pushes them (corner 0: 1,2; corner 1: 0,2; corner 2: 0,1).  The
 * incidence list (3*t + corner) of each vertex is sorted ascending, which is that order, and
 * the sums run in it, so a pass equals a host pass over the same list that reads the previous
 * pass's positions.
 */
__global__ void __launch_bounds__(256) k_inc_count(const int *__restrict__ tri, uint32_t ncorner, uint32_t *__restrict__ deg)
{
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= ncorner) return;
    atomicAdd(&deg[tri[i]], 1u);
}

__global__ void __launch_bounds__(256) k_inc_fill(const int *__restrict__ tri, uint32_t ncorner,
                                                  const uint32_t *__restrict__ start, uint32_t *__restrict__ cursor,
                                                  uint32_t *__restrict__ inc)
{
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= ncorner) return;
    int v = tri[i];
    uint32_t slot = start[v] + atomicAdd(&cursor[v], 1u);
    inc[slot] = i;
}

__global__ void __launch_bounds__(256) k_inc_sort(const uint32_t *__restrict__ start, const uint32_t *__restrict__ deg,
                                                  uint32_t nv, uint32_t *__restrict__ inc)
{
    uint32_t v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= nv) return;
    uint32_t s = start[v], k = deg[v];
    for (uint32_t a = s + 1; a < s + k; a++) {
        uint32_t x = inc[a];
        uint32_t b = a;
        while (b > s && inc[b - 1] > x) {
            inc[b] = inc[b - 1];
            b--;
        }
        inc[b] = x;
    }
}

__global__ void __launch_bounds__(256) k_jacobi(const int *__restrict__ tri, const uint32_t *__restrict__ start,
                                                const uint32_t *__restrict__ deg, const uint32_t *__restrict__ inc,
                                                uint32_t nv, const float *__restrict__ src, float *__restrict__ dst)
{
    uint32_t v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= nv) return;
    uint32_t s = start[v], k = deg[v];
    float sx = 0.0f, sy = 0.0f, sz = 0.0f;
    for (uint32_t a = s; a < s + k; a++) {
        uint32_t e = inc[a];
        uint32_t t = e / 3u, c = e - 3u * t;
        int n0 = tri[3 * (size_t)t + (c == 0 ? 1 : 0)];
        int n1 = tri[3 * (size_t)t + (c == 2 ? 1 : 2)];
        const float *q0 = src + 3 * (size_t)n0;
        sx = sx + q0[0];
        sy = sy + q0[1];
        sz = sz + q0[2];
        const float *q1 = src + 3 * (size_t)n1;
        sx = sx + q1[0];
        sy = sy + q1[1];
        sz = sz + q1[2];
    }
    float cnt = (float)(2u * k);
    dst[3 * (size_t)v + 0] = sx / cnt;
    dst[3 * (size_t)v + 1] = sy / cnt;
    dst[3 * (size_t)v + 2] = sz / cnt;
}

/* ------------------------------------------------------------------------------------------ */
/* degenerate-triangle filter (CudaSurf.cu:820-829), order preserving                           */
__global__ void __launch_bounds__(256) k_tri_keep(const int *__restrict__ tri, uint32_t nt, uint32_t *__restrict__ keep)
{
    uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nt) return;
    int x = tri[3 * (size_t)t], y = tri[3 * (size_t)t + 1], z = tri[3 * (size_t)t + 2];
    keep[t] = (x != y && y != z && x != z) ? 1u : 0u;
}

__global__ void __launch_bounds__(256) k_tri_compact(const int *__restrict__ tri, uint32_t nt, const uint32_t *__restrict__ pos,
                                                     int *__restrict__ out)
{
    uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nt) return;
    int x = tri[3 * (size_t)t], y = tri[3 * (size_t)t + 1], z = tri[3 * (size_t)t + 2];
    if (x != y && y != z && x != z) {
        uint32_t o = pos[t];
        out[3 * (size_t)o] = x;
        out[3 * (size_t)o + 1] = y;
        out[3 * (size_t)o + 2] = z;
    }
}

/* ------------------------------------------------------------------------------------------ */
/* launchers                                                                                    */

static inline unsigned nblk(uint64_t n, unsigned b) { return (unsigned)((n + b - 1) / b); }

void qs_launch_bbox(const float *pos, const float *rad, int n, float *part, int nparts, float *out, hipStream_t s)
{
    hipLaunchKernelGGL(k_bbox_partial, dim3(nparts), dim3(BB_NT), 0, s, pos, rad, n, part);
    hipLaunchKernelGGL(k_bbox_final, dim3(1), dim3(BB_NT), 0, s, part, nparts, out);
}

void qs_launch_hash_count(const float *pos, int n, QsGrid G, int *key, uint32_t *count, hipStream_t s)
{
    hipLaunchKernelGGL(k_hash_count, dim3(nblk(n, 256)), dim3(256), 0, s, pos, n, G, key, count);
}

void qs_launch_scatter(const int *key, int n, const uint32_t *start, uint32_t *cursor, int *perm, hipStream_t s)
{
    hipLaunchKernelGGL(k_scatter, dim3(nblk(n, 256)), dim3(256), 0, s, key, n, start, cursor, perm);
}

void qs_launch_cell_finalize(const uint32_t *start, const uint32_t *count, int ncell, int *perm, int2 *cse, hipStream_t s)
{
    hipLaunchKernelGGL(k_cell_finalize, dim3(nblk(ncell, 256)), dim3(256), 0, s, start, count, ncell, perm, cse);
}

void qs_launch_gather_atoms(const int *perm, const float *pos, const float *rad, int n, float4 *sx, hipStream_t s)
{
    hipLaunchKernelGGL(k_gather_atoms, dim3(nblk(n, 256)), dim3(256), 0, s, perm, pos, rad, n, sx);
}

void qs_launch_pass1(QsGrid G, const int2 *cse, const float4 *sx, uint8_t *labels, int x0, int x1, hipStream_t s)
{
    dim3 grid(nblk(G.g, 8), nblk(G.g, 8), nblk(x1 - x0, 4));
    hipLaunchKernelGGL(k_pass1, grid, dim3(256), 0, s, G, cse, sx, labels, x0, x1);
}

void qs_launch_pass2_tile(QsGrid G, const uint8_t *labels, float *field, const int *offs, const int *norm_start,
                          const int *enum_end, int x0, int x1, hipStream_t s)
{
    dim3 grid(nblk(G.g, P2_TZ), nblk(G.g, P2_TY), nblk(x1 - x0, P2_TX));
    size_t lds = qs_pass2_lds_bytes(G.r);
    hipLaunchKernelGGL(k_pass2_tile, grid, dim3(256), lds, s, G, labels, field, offs, norm_start, enum_end, x0, x1);
}

void qs_launch_pass2_direct(QsGrid G, const uint8_t *labels, float *field, int x0, int x1, hipStream_t s)
{
    dim3 grid(nblk(G.g, 64), nblk(G.g, 4), (unsigned)(x1 - x0));
    hipLaunchKernelGGL(k_pass2_direct, grid, dim3(256), 0, s, G, labels, field, x0, x1);
}

unsigned qs_mc_blocks(uint64_t npoints) { return nblk(npoints, MC_B); }

void qs_launch_mc_count(QsGrid G, const float *F, uint16_t *pinfo, uint32_t *bs_tri, uint32_t *bs_vert, uint64_t p0,
                        uint64_t p1, hipStream_t s)
{
    hipLaunchKernelGGL(k_mc_count, dim3(qs_mc_blocks(p1 - p0)), dim3(MC_NT), 0, s, G, F, pinfo, bs_tri, bs_vert, p0, p1);
}

void qs_launch_mc_verts(QsGrid G, const float *F, const uint16_t *pinfo, const uint32_t *bo_vert, uint32_t *vbase,
                        float *vert, uint64_t p0, uint64_t p1, hipStream_t s)
{
    hipLaunchKernelGGL(k_mc_verts, dim3(qs_mc_blocks(p1 - p0)), dim3(MC_NT), 0, s, G, F, pinfo, bo_vert, vbase, vert, p0, p1);
}

void qs_launch_mc_tris(QsGrid G, const float *F, const uint16_t *pinfo, const uint32_t *vbase, const uint32_t *bo_tri,
                       int *tri, uint64_t p0, uint64_t p1, hipStream_t s)
{
    hipLaunchKernelGGL(k_mc_tris, dim3(qs_mc_blocks(p1 - p0)), dim3(MC_NT), 0, s, G, F, pinfo, vbase, bo_tri, tri, p0, p1);
}

void qs_launch_closest_atom(QsGrid G, const int2 *cse, const float4 *sx, const int *perm, const float *vert, uint32_t nv,
                            int sorted_ids, int *out, hipStream_t s)
{
    if (!nv) return;
    hipLaunchKernelGGL(k_closest_atom, dim3(nblk(nv, 256)), dim3(256), 0, s, G, cse, sx, perm, vert, nv, sorted_ids, out);
}

void qs_launch_inc_count(const int *tri, uint32_t ncorner, uint32_t *deg, hipStream_t s)
{
    if (!ncorner) return;
    hipLaunchKernelGGL(k_inc_count, dim3(nblk(ncorner, 256)), dim3(256), 0, s, tri, ncorner, deg);
}

void qs_launch_inc_fill(const int *tri, uint32_t ncorner, const uint32_t *start, uint32_t *cursor, uint32_t *inc,
                        hipStream_t s)
{
    if (!ncorner) return;
    hipLaunchKernelGGL(k_inc_fill, dim3(nblk(ncorner, 256)), dim3(256), 0, s, tri, ncorner, start, cursor, inc);
}

void qs_launch_inc_sort(const uint32_t *start, const uint32_t *deg, uint32_t nv, uint32_t *inc, hipStream_t s)
{
    if (!nv) return;
    hipLaunchKernelGGL(k_inc_sort, dim3(nblk(nv, 256)), dim3(256), 0, s, start, deg, nv, inc);
}

void qs_launch_jacobi(const int *tri, const uint32_t *start, const uint32_t *deg, const uint32_t *inc, uint32_t nv,
                      const float *src, float *dst, hipStream_t s)
{
    if (!nv) return;
    hipLaunchKernelGGL(k_jacobi, dim3(nblk(nv, 256)), dim3(256), 0, s, tri, start, deg, inc, nv, src, dst);
}

void qs_launch_tri_keep(const int *tri, uint32_t nt, uint32_t *keep, hipStream_t s)
{
    if (!nt) return;
    hipLaunchKernelGGL(k_tri_keep, dim3(nblk(nt, 256)), dim3(256), 0, s, tri, nt, keep);
}

void qs_launch_tri_compact(const int *tri, uint32_t nt, const uint32_t *pos, int *out, hipStream_t s)
{
    if (!nt) return;
    hipLaunchKernelGGL(k_tri_compact, dim3(nblk(nt, 256)), dim3(256), 0, s, tri, nt, pos, out);
}

/* scan helpers */
uint64_t qs_scan_tiles(uint64_t n) { return (n + SCAN_TILE - 1) / SCAN_TILE; }

void qs_launch_tile_reduce(const uint32_t *in, uint64_t n, uint32_t *sums, hipStream_t s)
{
    hipLaunchKernelGGL(k_tile_reduce, dim3((unsigned)qs_scan_tiles(n)), dim3(SCAN_NT), 0, s, in, n, sums);
}

void qs_launch_tile_scan(const uint32_t *in, uint32_t *out, uint64_t n, const uint32_t *offsets, uint32_t *total,
                         hipStream_t s)
{
    hipLaunchKernelGGL(k_tile_scan, dim3((unsigned)qs_scan_tiles(n)), dim3(SCAN_NT), 0, s, in, out, n, offsets, total);
}
