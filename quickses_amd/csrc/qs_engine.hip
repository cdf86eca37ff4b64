/* Host side of the SES engine: grid setup, stage sequencing on one HIP stream, device
 * workspace, and the C ABI (include/quickses.h).
 *
 * One call runs, on the current device:
 *   bbox -> grid (CudaSurf.cu:149-174, 250-253, 548-576, 629)
 *   hash + count, scan, scatter, per-cell order, gather (CudaSurf.cu:594-622)
 *   field pass 1 -> labels, field pass 2 -> field (CudaSurf.cu:653-729)
 *   MC count, scans, vertices, triangles (CudaSurf.cu:373-458)
 *   closest atom (CudaSurf.cu:460-467), Laplacian passes (CudaSurf.cu:743)
 *   degenerate filter (CudaSurf.cu:820-829)
 * The whole grid is one slice; x-slabs ("slice" option) bound the per-slab field memory and
 * give the same mesh.
 */
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <string>
#include <vector>

#include "../../include/quickses.h"
#include "qs_common.h"
#include "qs_kernels.h"

namespace {

/* ------------------------------------------------------------------------------------------ */
/* errors                                                                                       */

int g_err = 0;
char g_errmsg[512] = "";

int set_error(int code, const char *fmt, ...)
{
    g_err = code;
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_errmsg, sizeof(g_errmsg), fmt, ap);
    va_end(ap);
    return code;
}

#define QS_HIP(call)                                                                                     \
    do {                                                                                                 \
        hipError_t e_ = (call);                                                                          \
        if (e_ != hipSuccess)                                                                            \
            return set_error(2, "%s failed: %s (%s:%d)", #call, hipGetErrorString(e_), __FILE__, __LINE__); \
    } while (0)

#define QS_TRY(expr)          \
    do {                      \
        int rc_ = (expr);     \
        if (rc_) return rc_;  \
    } while (0)

/* ------------------------------------------------------------------------------------------ */
/* options                                                                                      */

struct Options {
    int slice = 0;
    int lim_mode = 0;
    int atom_ids_sorted = 0;
    int pass1_direct = 0;
    int pass2_direct = 0;
    int trace = 0;
    int timing = 0;
};
Options g_opt;

/* ------------------------------------------------------------------------------------------ */
/* grow-only device buffers                                                                     */

struct DevBuf {
    void *p = nullptr;
    size_t cap = 0;
};

int ensure(DevBuf &b, size_t bytes)
{
    if (bytes <= b.cap && b.p) return 0;
    if (b.p) {
        QS_HIP(hipFree(b.p));
        b.p = nullptr;
        b.cap = 0;
    }
    size_t want = std::max<size_t>(bytes + bytes / 8, 256);
    QS_HIP(hipMalloc(&b.p, want));
    b.cap = want;
    return 0;
}

template <class T>
T *as(DevBuf &b) { return (T *)b.p; }

/* ------------------------------------------------------------------------------------------ */
/* grid setup: the reference's host formulas, evaluated in single precision                    */

struct BBox {
    float vmin[3], vmax[3], max_rad;
};

void bbox_host(const float *pos, const float *rad, unsigned n, BBox &b, int &nonfinite)
{
    /* getMinMax CudaSurf.cu:149-174 with std::min / std::max */
    for (int c = 0; c < 3; c++) {
        b.vmin[c] = 100000.0f;
        b.vmax[c] = -100000.0f;
    }
    b.max_rad = 0.0f;
    nonfinite = 0;
    for (unsigned a = 0; a < n; a++) {
        for (int c = 0; c < 3; c++) {
            float p = pos[3 * (size_t)a + c];
            if (!std::isfinite(p)) nonfinite++;
            b.vmin[c] = std::min(b.vmin[c], p);
            b.vmax[c] = std::max(b.vmax[c], p);
        }
        if (!std::isfinite(rad[a])) nonfinite++;
        b.max_rad = std::max(b.max_rad, rad[a]);
    }
}

void grid_from_bbox(const BBox &b, float reso, QsGrid &G)
{
    /* computeMaxDist CudaSurf.cu:250-253 */
    float ext = std::max(b.vmax[0] - b.vmin[0], std::max(b.vmax[1] - b.vmin[1], b.vmax[2] - b.vmin[2]));
    float max_dist = ext + (2 * b.max_rad);
    max_dist = max_dist + (4 * QS_PROBE);
    /* CudaSurf.cu:548-576, 629 */
    G.dxn = QS_PROBE + b.max_rad;
    float o[3];
    for (int c = 0; c < 3; c++) {
        float t = b.vmin[c] - b.max_rad;
        o[c] = t - QS_PROBE;
    }
    G.ox = o[0];
    G.oy = o[1];
    G.oz = o[2];
    G.dx = reso;
    G.gn = (int)std::ceil(max_dist / G.dxn);
    G.g = (int)std::ceil(max_dist / reso);
    G.r = (int)std::ceil(QS_PROBE / reso);
    G.lim = 0;
}

/* ------------------------------------------------------------------------------------------ */
/* pass-2 candidate tables: window offsets sorted by squared norm, and for each minimal norm the
   end of the range of norms whose float distance can still be the minimum                      */

struct OffsetTables {
    int r = -1;
    float dx = 0, cmax = 0, bound = 0;
    std::vector<int> offs, norm_start, enum_end;
};

void build_offset_tables(int r, float dx, float coord_max, OffsetTables &T)
{
    T.r = r;
    T.dx = dx;
    T.cmax = coord_max;
    const int nmax = 3 * r * r;
    std::vector<std::vector<int>> by_norm(nmax + 1);
    for (int a = -r; a <= r; a++)
        for (int b = -r; b <= r; b++)
            for (int c = -r; c <= r; c++) {
                int n = a * a + b * b + c * c;
                int packed = (a & 0xFF) | ((b & 0xFF) << 8) | ((c & 0xFF) << 16);
                by_norm[n].push_back(packed);
            }
    T.offs.clear();
    T.norm_start.assign(nmax + 2, 0);
    for (int n = 0; n <= nmax; n++) {
        T.norm_start[n] = (int)T.offs.size();
        T.offs.insert(T.offs.end(), by_norm[n].begin(), by_norm[n].end());
    }
    T.norm_start[nmax + 1] = (int)T.offs.size();
    /* A single-precision distance between two grid points differs from dx*sqrt(n) by less than
       delta = 16 u (M + 2 r dx), u = 2^-24 and M the largest coordinate magnitude (each coordinate
       carries two roundings, the difference, squares, sums and the root one each).  Candidates
       with sqrt(n) > sqrt(nmin) + 2 delta / dx cannot attain the minimum. */
    double u = 1.0 / 16777216.0;
    double delta = 16.0 * u * ((double)coord_max + 2.0 * r * (double)dx);
    T.bound = (float)delta;
    T.enum_end.assign(nmax + 1, 0);
    for (int n = 0; n <= nmax; n++) {
        double lim = std::sqrt((double)n) + 2.0 * delta / (double)dx;
        int hi = n;
        while (hi + 1 <= nmax && std::sqrt((double)(hi + 1)) <= lim) hi++;
        T.enum_end[n] = T.norm_start[hi + 1];
    }
}

/* ------------------------------------------------------------------------------------------ */
/* engine                                                                                       */

std::vector<std::pair<std::string, float>> g_stage_ms;

struct StageTimer {
    bool on = false;
    hipStream_t s = nullptr;
    std::vector<std::pair<std::string, hipEvent_t>> ev;
    void mark(const char *name)
    {
        if (!on) return;
        hipEvent_t e;
        if (hipEventCreate(&e) != hipSuccess) return;
        hipEventRecord(e, s);
        ev.emplace_back(name, e);
    }
    void report(const QsGrid &G, unsigned n)
    {
        if (!on || ev.size() < 2) return;
        hipEventSynchronize(ev.back().second);
        g_stage_ms.clear();
        for (size_t i = 1; i < ev.size(); i++) {
            float ms = 0;
            hipEventElapsedTime(&ms, ev[i - 1].second, ev[i].second);
            g_stage_ms.emplace_back(ev[i].first, ms);
        }
        if (!g_opt.trace) {
            for (auto &p : ev) hipEventDestroy(p.second);
            ev.clear();
            return;
        }
        std::string line = "{\"qs_trace\": {\"n_atoms\": " + std::to_string(n) + ", \"g\": " + std::to_string(G.g);
        for (size_t i = 1; i < ev.size(); i++) {
            float ms = 0;
            hipEventElapsedTime(&ms, ev[i - 1].second, ev[i].second);
            char buf[96];
            snprintf(buf, sizeof(buf), ", \"%s_ms\": %.4f", ev[i].first.c_str(), ms);
            line += buf;
        }
        line += "}}";
        fprintf(stderr, "%s\n", line.c_str());
        for (auto &p : ev) hipEventDestroy(p.second);
        ev.clear();
    }
};

struct Engine {
    int device = -1;
    hipStream_t stream = nullptr;

    DevBuf in_pos, in_rad, bb_part, bb_out;
    DevBuf key, count, start, cursor, perm, cse, sx;
    DevBuf labels, outm, shellm, field, negm, zerom;
    DevBuf s4, vcnt, tcnt, vbw, tbw, apos, alist, totals;
    DevBuf vert, vert2, vert3, vkey, cidx, tbase, tri, tri_out, atom;
    DevBuf keep, keeppos;
    DevBuf scan_tmp;
    DevBuf t_offs, t_norm, t_end;
    OffsetTables tables;

    /* last result */
    QsGrid G{};
    unsigned n_atoms = 0;
    uint32_t nv = 0, nt_all = 0, nt = 0, n_active = 0;
    float *res_vert = nullptr;
    int *res_tri = nullptr, *res_atom = nullptr;

    int init()
    {
        int dev = 0;
        QS_HIP(hipGetDevice(&dev));
        if (device == dev && stream) return 0;
        device = dev;
        QS_HIP(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking));
        return 0;
    }

    /* exclusive scan of n uint32 (in may equal out); total written to d_total if given */
    int scan(const uint32_t *in, uint32_t *out, uint64_t n, uint32_t *d_total)
    {
        if (n == 0) {
            if (d_total) QS_HIP(hipMemsetAsync(d_total, 0, sizeof(uint32_t), stream));
            return 0;
        }
        /* level sizes */
        std::vector<uint64_t> lv;
        uint64_t m = n;
        while (qs_scan_tiles(m) > 1) {
            m = qs_scan_tiles(m);
            lv.push_back(m);
        }
        size_t need = 0;
        for (uint64_t x : lv) need += (x + 63) / 64 * 64;
        QS_TRY(ensure(scan_tmp, std::max<size_t>(need, 64) * sizeof(uint32_t)));
        std::vector<uint32_t *> lvl_ptr;
        uint32_t *p = as<uint32_t>(scan_tmp);
        for (uint64_t x : lv) {
            lvl_ptr.push_back(p);
            p += (x + 63) / 64 * 64;
        }
        /* reduce down the levels */
        const uint32_t *src = in;
        uint64_t len = n;
        for (size_t l = 0; l < lv.size(); l++) {
            qs_launch_tile_reduce(src, len, lvl_ptr[l], stream);
            src = lvl_ptr[l];
            len = lv[l];
        }
        /* scan back up */
        for (int l = (int)lv.size() - 1; l >= 0; l--) {
            const uint32_t *offs = (l + 1 < (int)lv.size()) ? lvl_ptr[l + 1] : nullptr;
            qs_launch_tile_scan(lvl_ptr[l], lvl_ptr[l], lv[l], offs, nullptr, stream);
        }
        qs_launch_tile_scan(in, out, n, lv.empty() ? nullptr : lvl_ptr[0], d_total, stream);
        QS_HIP(hipGetLastError());
        return 0;
    }

    int run(const float *d_pos, const float *d_rad, unsigned n, float reso, int smooth, const BBox *host_bbox);
    int mesh_from_field(StageTimer &tm, bool with_atoms);
    int pass2_tables();
};

Engine g_eng;

int Engine::run(const float *d_pos, const float *d_rad, unsigned n, float reso, int smooth, const BBox *host_bbox)
{
    nv = nt = nt_all = 0;
    n_atoms = n;
    res_vert = nullptr;
    res_tri = nullptr;
    res_atom = nullptr;
    if (!(reso > 0.0f) || !std::isfinite(reso)) return set_error(3, "voxel size must be a positive finite number");
    if (n <= 1) return 0; /* CudaSurf.cu:539-543 */
    if (n > (1u << 30)) return set_error(3, "too many atoms (%u)", n);

    StageTimer tm;
    tm.on = g_opt.trace != 0 || g_opt.timing != 0;
    tm.s = stream;
    tm.mark("start");

    /* ---- bbox and grid ---- */
    BBox bb;
    if (host_bbox) {
        bb = *host_bbox;
    } else {
        int nparts = (int)std::min<unsigned>((n + 255) / 256, 1024);
        QS_TRY(ensure(bb_part, (size_t)nparts * 8 * sizeof(float)));
        QS_TRY(ensure(bb_out, 8 * sizeof(float)));
        qs_launch_bbox(d_pos, d_rad, (int)n, as<float>(bb_part), nparts, as<float>(bb_out), stream);
        float h[8];
        QS_HIP(hipMemcpyAsync(h, bb_out.p, sizeof(h), hipMemcpyDeviceToHost, stream));
        QS_HIP(hipStreamSynchronize(stream));
        if (h[7] != 0.0f) return set_error(3, "atom positions or radii contain NaN or Inf");
        for (int c = 0; c < 3; c++) {
            bb.vmin[c] = h[c];
            bb.vmax[c] = h[3 + c];
        }
        bb.max_rad = h[6];
    }
    tm.mark("bbox");
    for (int c = 0; c < 3; c++)
        if (!(bb.vmin[c] > -100000.0f && bb.vmax[c] < 100000.0f && bb.vmin[c] <= bb.vmax[c]))
            return set_error(3, "atom coordinates must lie strictly within (-1e5, 1e5) Angstrom");
    grid_from_bbox(bb, reso, G);
    const int g = G.g, gn = G.gn;
    if (g < 2 || gn < 1) return set_error(3, "degenerate grid (g=%d)", g);
    const uint64_t V = (uint64_t)g * g * g;
    const uint64_t ncell = (uint64_t)gn * gn * gn;
    if (V >= (1ull << 32)) return set_error(4, "grid of %d^3 points exceeds 2^32 points", g);
    if (ncell >= (1ull << 31)) return set_error(4, "neighbour grid too large");
    int lim = g_opt.lim_mode ? (g - 2 * G.r - 1) : (g - 1);
    G.lim = std::max(lim, 0);

    /* ---- atoms into cells ---- */
    QS_TRY(ensure(key, (size_t)n * 4));
    QS_TRY(ensure(count, ncell * 4));
    QS_TRY(ensure(start, (ncell + 1) * 4));
    QS_TRY(ensure(cursor, ncell * 4));
    QS_TRY(ensure(perm, (size_t)n * 4));
    QS_TRY(ensure(cse, ncell * 8));
    QS_TRY(ensure(sx, (size_t)n * 16));
    QS_HIP(hipMemsetAsync(count.p, 0, ncell * 4, stream));
    QS_HIP(hipMemsetAsync(cursor.p, 0, ncell * 4, stream));
    qs_launch_hash_count(d_pos, (int)n, G, as<int>(key), as<uint32_t>(count), stream);
    /* start[c] = atoms in cells < c; start[ncell] = n */
    QS_TRY(scan(as<uint32_t>(count), as<uint32_t>(start), ncell, as<uint32_t>(start) + ncell));
    qs_launch_scatter(as<int>(key), (int)n, as<uint32_t>(start), as<uint32_t>(cursor), as<int>(perm), stream);
    qs_launch_cell_finalize(as<uint32_t>(start), as<uint32_t>(count), (int)ncell, as<int>(perm), as<int2>(cse), stream);
    qs_launch_gather_atoms(as<int>(perm), d_pos, d_rad, (int)n, as<float4>(sx), stream);
    QS_HIP(hipGetLastError());
    tm.mark("cells");

    /* ---- field ---- */
    const int W = (g + 63) / 64;
    const size_t nwords = (size_t)g * g * W;
    QS_TRY(ensure(outm, nwords * 8));
    QS_TRY(ensure(shellm, nwords * 8));
    QS_TRY(ensure(field, V * 4));
    if (g_opt.pass1_direct)
        qs_launch_pass1_direct(G, as<int2>(cse), as<float4>(sx), as<uint64_t>(outm), as<uint64_t>(shellm), W, 0, g, stream);
    else
        qs_launch_pass1(G, as<uint32_t>(start), as<float4>(sx), as<uint64_t>(outm), as<uint64_t>(shellm), W, 0, g, stream);
    QS_HIP(hipGetLastError());
    tm.mark("pass1");
    QS_TRY(ensure(negm, nwords * 8));
    QS_TRY(ensure(zerom, nwords * 8));
    QS_TRY(pass2_tables());
    qs_launch_pass2(G, as<uint64_t>(outm), as<uint64_t>(shellm), W, as<float>(field), as<uint32_t>(negm), as<uint32_t>(zerom),
                    as<int>(t_offs), as<int>(t_norm), as<int>(t_end), 0, g, g_opt.pass2_direct != 0, stream);
    QS_HIP(hipGetLastError());
    tm.mark("pass2");

    QS_TRY(mesh_from_field(tm, true));

    /* ---- Laplacian passes (Jacobi) ---- */
    res_vert = as<float>(vert);
    if (smooth > 0 && nv > 0 && nt_all > 0) {
        const int W = (g + 63) / 64;
        QS_TRY(ensure(vert2, (size_t)nv * 12));
        /* `vert` keeps the unsmoothed positions; passes alternate between vert2 and vert3 */
        if (smooth > 1) QS_TRY(ensure(vert3, (size_t)nv * 12));
        const float *src = as<float>(vert);
        float *dst = as<float>(vert2);
        for (int it = 0; it < smooth; it++) {
            qs_launch_smooth(G, as<uint64_t>(negm), as<uint64_t>(zerom), as<uint64_t>(s4), as<uint32_t>(vbw), W,
                             as<uint64_t>(vkey), nv, as<uint8_t>(cidx), as<uint32_t>(tbase), as<int>(tri), src, dst, stream);
            src = dst;
            dst = (dst == as<float>(vert2)) ? as<float>(vert3) : as<float>(vert2);
        }
        QS_HIP(hipGetLastError());
        res_vert = (float *)src;
        tm.mark("smooth");
    }

    /* ---- degenerate filter ---- */
    QS_TRY(ensure(keep, (size_t)std::max<uint32_t>(nt_all, 1) * 4));
    QS_TRY(ensure(keeppos, (size_t)std::max<uint32_t>(nt_all, 1) * 4));
    QS_TRY(ensure(tri_out, (size_t)std::max<uint32_t>(nt_all, 1) * 12));
    QS_TRY(ensure(totals, 16));
    qs_launch_tri_keep(as<int>(tri), nt_all, as<uint32_t>(keep), stream);
    QS_TRY(scan(as<uint32_t>(keep), as<uint32_t>(keeppos), nt_all, as<uint32_t>(totals)));
    qs_launch_tri_compact(as<int>(tri), nt_all, as<uint32_t>(keeppos), as<int>(tri_out), stream);
    uint32_t kept = 0;
    QS_HIP(hipMemcpyAsync(&kept, totals.p, 4, hipMemcpyDeviceToHost, stream));
    QS_HIP(hipStreamSynchronize(stream));
    tm.mark("filter");
    nt = kept;
    res_tri = as<int>(tri_out);
    res_atom = as<int>(atom);
    tm.report(G, n);
    return 0;
}

/* marching cubes on the sign bits of `field` (grid G), results into vert / tri / atom */
int Engine::mesh_from_field(StageTimer &tm, bool with_atoms)
{
    const int g = G.g, W = (g + 63) / 64;
    const uint64_t nw = (uint64_t)g * g * W;
    QS_TRY(ensure(s4, nw * 32));
    QS_TRY(ensure(vcnt, nw * 4));
    QS_TRY(ensure(tcnt, nw * 4));
    QS_TRY(ensure(vbw, nw * 4));
    QS_TRY(ensure(tbw, nw * 4));
    QS_TRY(ensure(apos, nw * 4));
    QS_TRY(ensure(alist, nw * 4));
    QS_TRY(ensure(totals, 16));
    qs_launch_mc_count(G, as<uint64_t>(negm), as<uint64_t>(zerom), W, as<uint64_t>(s4), as<uint32_t>(vcnt), as<uint32_t>(tcnt),
                       0, g, stream);
    QS_TRY(scan(as<uint32_t>(tcnt), as<uint32_t>(tbw), nw, as<uint32_t>(totals)));
    QS_TRY(scan(as<uint32_t>(vcnt), as<uint32_t>(vbw), nw, as<uint32_t>(totals) + 1));
    qs_launch_mc_flag(as<uint32_t>(vcnt), as<uint32_t>(tcnt), nw, as<uint32_t>(apos), stream);
    QS_TRY(scan(as<uint32_t>(apos), as<uint32_t>(apos), nw, as<uint32_t>(totals) + 2));
    qs_launch_mc_active(as<uint32_t>(vcnt), as<uint32_t>(tcnt), as<uint32_t>(apos), nw, as<uint32_t>(alist), stream);
    uint32_t tot[3];
    QS_HIP(hipMemcpyAsync(tot, totals.p, 12, hipMemcpyDeviceToHost, stream));
    QS_HIP(hipStreamSynchronize(stream));
    tm.mark("mc_count");
    nt_all = tot[0];
    nv = tot[1];
    const uint32_t nact = tot[2];
    if ((uint64_t)nt_all * 3 >= (1ull << 31) || nv >= (1u << 31)) return set_error(4, "mesh too large");
    QS_TRY(ensure(vert, (size_t)std::max<uint32_t>(nv, 1) * 12));
    QS_TRY(ensure(vkey, (size_t)std::max<uint32_t>(nv, 1) * 8));
    QS_TRY(ensure(tri, (size_t)std::max<uint32_t>(nt_all, 1) * 12));
    const uint64_t V = (uint64_t)g * g * g;
    QS_TRY(ensure(cidx, V));
    QS_TRY(ensure(tbase, V * 4));
    n_active = nact;
    qs_launch_mc_emit(G, as<uint64_t>(negm), as<uint64_t>(zerom), as<uint64_t>(s4), as<uint32_t>(vbw), W, as<float>(field),
                      as<uint32_t>(alist), nact, as<uint32_t>(tbw), as<float>(vert), as<uint64_t>(vkey), as<int>(tri),
                      as<uint8_t>(cidx), as<uint32_t>(tbase), stream);
    QS_HIP(hipGetLastError());
    tm.mark("mc_emit");
    if (with_atoms) {
        QS_TRY(ensure(atom, (size_t)std::max<uint32_t>(nv, 1) * 4));
        qs_launch_closest_atom(G, as<int2>(cse), as<float4>(sx), as<int>(perm), as<float>(vert), nv, g_opt.atom_ids_sorted,
                               as<int>(atom), stream);
        QS_HIP(hipGetLastError());
        tm.mark("closest_atom");
    }
    return 0;
}

/* candidate tables of field pass 2 for the current grid (built when r, dx or the coordinate
   range change) */
int Engine::pass2_tables()
{
    float cmax = 0.0f;
    for (float o : {G.ox, G.oy, G.oz}) cmax = std::max(cmax, std::max(std::fabs(o), std::fabs(o + (float)G.g * G.dx)));
    cmax = std::max(cmax, (float)G.g * G.dx);
    float cmax2 = 1.0f;
    while (cmax2 < cmax) cmax2 *= 2.0f;
    if (tables.r == G.r && tables.dx == G.dx && tables.cmax == cmax2) return 0;
    build_offset_tables(G.r, G.dx, cmax2, tables);
    QS_TRY(ensure(t_offs, tables.offs.size() * 4));
    QS_TRY(ensure(t_norm, tables.norm_start.size() * 4));
    QS_TRY(ensure(t_end, tables.enum_end.size() * 4));
    QS_HIP(hipMemcpyAsync(t_offs.p, tables.offs.data(), tables.offs.size() * 4, hipMemcpyHostToDevice, stream));
    QS_HIP(hipMemcpyAsync(t_norm.p, tables.norm_start.data(), tables.norm_start.size() * 4, hipMemcpyHostToDevice, stream));
    QS_HIP(hipMemcpyAsync(t_end.p, tables.enum_end.data(), tables.enum_end.size() * 4, hipMemcpyHostToDevice, stream));
    QS_HIP(hipStreamSynchronize(stream));
    return 0;
}

/* host-side result of API_computeSES (CudaSurf.cu:778-785) */
float *g_vertices = nullptr;
int *g_triangles = nullptr;
int *g_atom_ids = nullptr;
unsigned g_ntri_idx = 0, g_nvert = 0;

}  // namespace

/* ========================================================================================== */
/* C ABI                                                                                        */

extern "C" {

int API_lastError(void) { return g_err; }
const char *qs_last_error_message(void) { return g_errmsg; }

int qs_set_option(const char *key, int value)
{
    std::string k = key ? key : "";
    if (k == "slice") g_opt.slice = value;
    else if (k == "lim_mode") g_opt.lim_mode = value;
    else if (k == "atom_ids_sorted") g_opt.atom_ids_sorted = value;
    else if (k == "pass1_direct") g_opt.pass1_direct = value;
    else if (k == "pass2_direct") g_opt.pass2_direct = value;
    else if (k == "trace") g_opt.trace = value;
    else if (k == "timing") g_opt.timing = value;
    else return -1;
    return 0;
}

int qs_get_option(const char *key)
{
    std::string k = key ? key : "";
    if (k == "slice") return g_opt.slice;
    if (k == "lim_mode") return g_opt.lim_mode;
    if (k == "atom_ids_sorted") return g_opt.atom_ids_sorted;
    if (k == "pass1_direct") return g_opt.pass1_direct;
    if (k == "pass2_direct") return g_opt.pass2_direct;
    if (k == "trace") return g_opt.trace;
    if (k == "timing") return g_opt.timing;
    return -1;
}

void API_computeSES(float resoSES, qs_float3 *atomPos, float *atomRad, unsigned int N, qs_float3 *out_vertices,
                    unsigned int *NVert, int *out_triangles, unsigned int *NTri, int doSmoothing)
{
    (void)out_vertices;
    (void)out_triangles;
    g_err = 0;
    g_errmsg[0] = 0;
    if (NVert) *NVert = 0;
    if (NTri) *NTri = 0;
    g_nvert = g_ntri_idx = 0;
    /* results of a previous call that were not freed are released (the reference leaks them) */
    free(g_vertices);
    free(g_triangles);
    free(g_atom_ids);
    g_vertices = nullptr;
    g_triangles = nullptr;
    g_atom_ids = nullptr;
    auto fail = [&]() {
        g_vertices = (float *)malloc(12);
        g_triangles = (int *)malloc(12);
        g_atom_ids = (int *)malloc(4);
    };
    if (N > 1 && (!atomPos || !atomRad)) {
        set_error(3, "null atom arrays");
        fail();
        return;
    }
    if (N <= 1) { /* empty mesh without touching the device (CudaSurf.cu:539-543) */
        fail();
        return;
    }
    if (g_eng.init()) {
        fail();
        return;
    }
    BBox bb{};
    int nonfinite = 0;
    const float *hp = (const float *)atomPos;
    if (N > 1) {
        bbox_host(hp, atomRad, N, bb, nonfinite);
        if (nonfinite) {
            set_error(3, "atom positions or radii contain NaN or Inf");
            fail();
            return;
        }
    }
    int rc = 0;
    if (N > 1) {
        Engine &E = g_eng;
        rc = ensure(E.in_pos, (size_t)N * 12);
        if (!rc) rc = ensure(E.in_rad, (size_t)N * 4);
        if (!rc && hipMemcpyAsync(E.in_pos.p, hp, (size_t)N * 12, hipMemcpyHostToDevice, E.stream) != hipSuccess)
            rc = set_error(2, "H2D copy failed");
        if (!rc && hipMemcpyAsync(E.in_rad.p, atomRad, (size_t)N * 4, hipMemcpyHostToDevice, E.stream) != hipSuccess)
            rc = set_error(2, "H2D copy failed");
        if (!rc) rc = E.run(as<float>(E.in_pos), as<float>(E.in_rad), N, resoSES, doSmoothing, &bb);
    } else {
        g_eng.nv = g_eng.nt = 0;
    }
    if (rc) {
        fail();
        return;
    }
    Engine &E = g_eng;
    g_nvert = E.nv;
    g_ntri_idx = 3 * E.nt;
    g_vertices = (float *)malloc((size_t)std::max<uint32_t>(E.nv, 1) * 12);
    g_triangles = (int *)malloc((size_t)std::max<uint32_t>(E.nt, 1) * 12);
    g_atom_ids = (int *)malloc((size_t)std::max<uint32_t>(E.nv, 1) * 4);
    if (E.nv) {
        if (hipMemcpyAsync(g_vertices, E.res_vert, (size_t)E.nv * 12, hipMemcpyDeviceToHost, E.stream) != hipSuccess ||
            hipMemcpyAsync(g_atom_ids, E.res_atom, (size_t)E.nv * 4, hipMemcpyDeviceToHost, E.stream) != hipSuccess) {
            set_error(2, "D2H copy failed");
            g_nvert = g_ntri_idx = 0;
            return;
        }
    }
    if (E.nt && hipMemcpyAsync(g_triangles, E.res_tri, (size_t)E.nt * 12, hipMemcpyDeviceToHost, E.stream) != hipSuccess) {
        set_error(2, "D2H copy failed");
        g_nvert = g_ntri_idx = 0;
        return;
    }
    if (hipStreamSynchronize(E.stream) != hipSuccess) {
        set_error(2, "stream synchronisation failed");
        g_nvert = g_ntri_idx = 0;
        return;
    }
    if (NVert) *NVert = g_nvert;
    if (NTri) *NTri = g_ntri_idx;
}

int *API_getTriangles(bool invertTriangles)
{
    if (invertTriangles && g_triangles) {
        for (unsigned t = 0; t < g_ntri_idx / 3; t++) std::swap(g_triangles[3 * t], g_triangles[3 * t + 1]);
    }
    return g_triangles;
}

qs_float3 *API_getVertices(void) { return (qs_float3 *)g_vertices; }

int *API_getAtomIdPerVert(void) { return g_atom_ids; }

void API_freeMesh(void)
{
    free(g_vertices);
    free(g_triangles);
    free(g_atom_ids);
    g_vertices = nullptr;
    g_triangles = nullptr;
    g_atom_ids = nullptr;
    g_nvert = g_ntri_idx = 0;
}

int qs_compute_device(const float *d_pos, const float *d_rad, unsigned int N, float resoSES, int doSmoothing,
                      unsigned int *nvert, unsigned int *nidx, float **d_vert, int **d_tri, int **d_atom_id)
{
    g_err = 0;
    g_errmsg[0] = 0;
    if (nvert) *nvert = 0;
    if (nidx) *nidx = 0;
    QS_TRY(g_eng.init());
    QS_TRY(g_eng.run(d_pos, d_rad, N, resoSES, doSmoothing, nullptr));
    if (nvert) *nvert = g_eng.nv;
    if (nidx) *nidx = 3 * g_eng.nt;
    if (d_vert) *d_vert = g_eng.res_vert;
    if (d_tri) *d_tri = g_eng.res_tri;
    if (d_atom_id) *d_atom_id = g_eng.res_atom;
    return 0;
}

int qs_stage_times(char *names, int names_cap, float *ms, int max_stages)
{
    int k = 0;
    std::string all;
    for (auto &p : g_stage_ms) {
        if (k >= max_stages) break;
        if (ms) ms[k] = p.second;
        all += p.first;
        all += ',';
        k++;
    }
    if (names && names_cap > 0) {
        strncpy(names, all.c_str(), (size_t)names_cap - 1);
        names[names_cap - 1] = 0;
    }
    return k;
}

int qs_last_grid(float *fp, int *ip)
{
    const QsGrid &G = g_eng.G;
    if (fp) {
        fp[0] = G.ox;
        fp[1] = G.oy;
        fp[2] = G.oz;
        fp[3] = G.dx;
        fp[4] = G.dxn;
    }
    if (ip) {
        ip[0] = G.g;
        ip[1] = G.gn;
        ip[2] = G.r;
        ip[3] = G.lim;
    }
    return 0;
}

int64_t qs_fetch(const char *what, void *host, int64_t bytes)
{
    Engine &E = g_eng;
    std::string w = what ? what : "";
    const QsGrid &G = E.G;
    uint64_t V = (uint64_t)G.g * G.g * G.g;
    uint64_t nc = (uint64_t)G.gn * G.gn * G.gn;
    const void *src = nullptr;
    int64_t need = 0;
    if (w == "labels") {
        if (bytes < (int64_t)V || !E.outm.p) return -1;
        const int g = G.g, W = (g + 63) / 64;
        size_t nw = (size_t)g * g * W;
        std::vector<uint64_t> mo(nw), ms(nw);
        if (hipMemcpy(mo.data(), E.outm.p, nw * 8, hipMemcpyDeviceToHost) != hipSuccess) return -1;
        if (hipMemcpy(ms.data(), E.shellm.p, nw * 8, hipMemcpyDeviceToHost) != hipSuccess) return -1;
        uint8_t *out = (uint8_t *)host;
        for (size_t row = 0; row < (size_t)g * g; row++)
            for (int z = 0; z < g; z++) {
                uint64_t bit = 1ull << (z & 63);
                size_t wi = row * W + (z >> 6);
                out[row * g + z] = (mo[wi] & bit) ? QS_OUT : ((ms[wi] & bit) ? QS_SHELL : QS_IN);
            }
        return (int64_t)V;
    }
    else if (w == "field") { src = E.field.p; need = (int64_t)V * 4; }
    else if (w == "perm") { src = E.perm.p; need = (int64_t)E.n_atoms * 4; }
    else if (w == "cells") { src = E.cse.p; need = (int64_t)nc * 8; }
    else if (w == "vert_unsmoothed") { src = E.vert.p; need = (int64_t)E.nv * 12; }
    else if (w == "tri_all") { src = E.tri.p; need = (int64_t)E.nt_all * 12; }
    else return -1;
    if (!src || bytes < need) return -1;
    if (need && hipMemcpy(host, src, (size_t)need, hipMemcpyDeviceToHost) != hipSuccess) return -1;
    return need;
}

static void grid_from_params(const float *fp, const int *ip, QsGrid &G)
{
    G.ox = fp[0];
    G.oy = fp[1];
    G.oz = fp[2];
    G.dx = fp[3];
    G.dxn = fp[4];
    G.g = ip[0];
    G.gn = ip[1];
    G.r = ip[2];
    G.lim = ip[3];
}

int qs_stage_pass2(const float *fp, const int *ip, const uint8_t *labels, float *field)
{
    g_err = 0;
    Engine &E = g_eng;
    QS_TRY(E.init());
    QsGrid G;
    grid_from_params(fp, ip, G);
    uint64_t V = (uint64_t)G.g * G.g * G.g;
    const int W = (G.g + 63) / 64;
    const size_t nwords = (size_t)G.g * G.g * W;
    QS_TRY(ensure(E.labels, V));
    QS_TRY(ensure(E.outm, nwords * 8));
    QS_TRY(ensure(E.shellm, nwords * 8));
    QS_TRY(ensure(E.field, V * 4));
    QS_HIP(hipMemcpy(E.labels.p, labels, V, hipMemcpyHostToDevice));
    qs_launch_labels_to_masks(as<uint8_t>(E.labels), G.g, W, as<uint64_t>(E.outm), as<uint64_t>(E.shellm), E.stream);
    QS_TRY(ensure(E.negm, nwords * 8));
    QS_TRY(ensure(E.zerom, nwords * 8));
    E.G = G;
    QS_TRY(E.pass2_tables());
    qs_launch_pass2(G, as<uint64_t>(E.outm), as<uint64_t>(E.shellm), W, as<float>(E.field), as<uint32_t>(E.negm),
                    as<uint32_t>(E.zerom), as<int>(E.t_offs), as<int>(E.t_norm), as<int>(E.t_end), 0, G.g,
                    g_opt.pass2_direct != 0, E.stream);
    QS_HIP(hipGetLastError());
    QS_HIP(hipStreamSynchronize(E.stream));
    QS_HIP(hipMemcpy(field, E.field.p, V * 4, hipMemcpyDeviceToHost));
    return 0;
}

int qs_stage_mesh(const float *fp, const int *ip, const float *field, unsigned int *nvert, unsigned int *ntri, float *vert,
                  int64_t vert_cap, int *tri, int64_t tri_cap)
{
    g_err = 0;
    Engine &E = g_eng;
    QS_TRY(E.init());
    grid_from_params(fp, ip, E.G);
    uint64_t V = (uint64_t)E.G.g * E.G.g * E.G.g;
    QS_TRY(ensure(E.field, V * 4));
    QS_HIP(hipMemcpy(E.field.p, field, V * 4, hipMemcpyHostToDevice));
    const int W = (E.G.g + 63) / 64;
    const size_t nwords = (size_t)E.G.g * E.G.g * W;
    QS_TRY(ensure(E.negm, nwords * 8));
    QS_TRY(ensure(E.zerom, nwords * 8));
    qs_launch_field_signs(as<float>(E.field), E.G.g, W, as<uint32_t>(E.negm), as<uint32_t>(E.zerom), E.stream);
    StageTimer tm;
    QS_TRY(E.mesh_from_field(tm, false));
    QS_HIP(hipStreamSynchronize(E.stream));
    if (nvert) *nvert = E.nv;
    if (ntri) *ntri = E.nt_all;
    if ((int64_t)E.nv * 3 > vert_cap || (int64_t)E.nt_all * 3 > tri_cap) return set_error(5, "output capacity too small");
    if (E.nv) QS_HIP(hipMemcpy(vert, E.vert.p, (size_t)E.nv * 12, hipMemcpyDeviceToHost));
    if (E.nt_all) QS_HIP(hipMemcpy(tri, E.tri.p, (size_t)E.nt_all * 12, hipMemcpyDeviceToHost));
    return 0;
}

}  // extern "C"
