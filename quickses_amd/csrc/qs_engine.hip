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
 * The whole grid is one slice.  A batch of molecules runs as one sequence of launches over
 * concatenated arrays: each molecule owns a range of neighbour cells, row words, voxels,
 * vertices and triangles (QsMol), and each launch maps its blocks to molecules by a prefix of
 * per-molecule block counts (QsBatch).
 */
#include <dlfcn.h>
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
#include "qs_engine.h"

namespace qs {

/* field pass 1: a slab launch of fewer blocks than this (four per CU) splits its slabs, down to
   slabs of this many rows */
constexpr uint64_t P1_FILL_BLOCKS = 1024;
constexpr int P1_FILL_MINROWS = 8;

thread_local int g_err = 0;
thread_local char g_errmsg[512] = "";
thread_local std::vector<std::pair<std::string, float>> g_stage_ms;
thread_local int g_grow_shift = 3;
thread_local Arena *g_arena = nullptr;
Options g_opt;

int set_error(int code, const char *fmt, ...)
{
    g_err = code;
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_errmsg, sizeof(g_errmsg), fmt, ap);
    va_end(ap);
    return code;
}

int ensure(DevBuf &b, size_t bytes)
{
    if (bytes <= b.cap && b.p) return 0;
    if (g_arena) {
        /* carve from the fixed region (the buffer's earlier piece, if any, stays unused) */
        const size_t off = (g_arena->used + 255) & ~(size_t)255;
        const size_t want = std::max<size_t>(bytes, 256);
        if (off + want > g_arena->cap)
            return set_error(6, "the fixed scratch region (%zu MiB) is too small for this molecule (%zu MiB and more needed)",
                             g_arena->cap >> 20, (off + want + (1u << 20) - 1) >> 20);
        b.p = g_arena->base + off;
        b.cap = want;
        g_arena->used = off + want;
        return 0;
    }
    if (b.p) {
        QS_HIP(hipFree(b.p));
        b.p = nullptr;
        b.cap = 0;
    }
    size_t want = std::max<size_t>(bytes + (bytes >> g_grow_shift), 256);
    QS_HIP(hipMalloc(&b.p, want));
    b.cap = want;
    return 0;
}

void release(DevBuf &b)
{
    if (b.p && !g_arena) (void)hipFree(b.p);
    b.p = nullptr;
    b.cap = 0;
}

/* The tuning cache (tuning.json, written by quickses_amd/autotune.py): its "pass1_slab_rows" object,
   voxel size -> rows per field-pass-1 block, e.g. {"0.2": 64, "0.5": 32}.  Returns the number of
   entries read, or -1 when the file cannot be read. */
int load_tuning_file(const char *path)
{
    FILE *f = fopen(path, "rb");
    if (!f) return -1;
    std::string text;
    char buf[4096];
    size_t k;
    while ((k = fread(buf, 1, sizeof(buf), f)) > 0) text.append(buf, k);
    fclose(f);
    const size_t key = text.find("\"pass1_slab_rows\"");
    if (key == std::string::npos) return 0;
    const size_t open = text.find('{', key), close = text.find('}', key);
    if (open == std::string::npos || close == std::string::npos || close < open) return 0;
    int n = 0;
    size_t at = open + 1;
    while (at < close) {
        const size_t q0 = text.find('"', at);
        if (q0 == std::string::npos || q0 >= close) break;
        const size_t q1 = text.find('"', q0 + 1);
        const size_t colon = text.find(':', q1);
        if (q1 == std::string::npos || colon == std::string::npos || colon >= close) break;
        const float dx = strtof(text.c_str() + q0 + 1, nullptr);
        const int rows = (int)strtol(text.c_str() + colon + 1, nullptr, 10);
        if (dx > 0.0f && rows > 0 && qs_set_tuned_rows(dx, rows) == 0) n++;
        at = text.find(',', colon);
        if (at == std::string::npos) break;
        at++;
    }
    return n;
}

/* once per process: the tuning cache beside the library (QS_TUNING_FILE names another file;
   "none" reads none) */
void load_default_tuning()
{
    static bool done = false;
    if (done) return;
    done = true;
    const char *env = getenv("QS_TUNING_FILE");
    if (env && !strcmp(env, "none")) return;
    if (env && *env) {
        (void)load_tuning_file(env);
        return;
    }
    Dl_info info;
    if (!dladdr((void *)&load_default_tuning, &info) || !info.dli_fname) return;
    std::string p = info.dli_fname;
    const size_t slash = p.rfind('/');
    p = (slash == std::string::npos ? std::string(".") : p.substr(0, slash)) + "/tuning.json";
    (void)load_tuning_file(p.c_str());
}

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


int Engine::reserve(uint64_t vox, uint64_t atoms, uint64_t cells, uint64_t verts, int nmol)
{
    const uint64_t nw = vox / 32 + 1, tris = 2 * verts + 1;
    vox += 1;
    atoms += 1;
    cells += 1;
    verts += 1;
    for (DevBuf *b : {&outm, &shellm, &negm, &zerom}) QS_TRY(ensure(*b, nw * 8));
    for (DevBuf *b : {&wcnt, &vbw, &tbw, &alist, &avb, &atb}) QS_TRY(ensure(*b, nw * 4));
    QS_TRY(ensure(s4, nw * 32));
    QS_TRY(ensure(ntb, nw * 24));
    QS_TRY(ensure(field, vox * 4));
    QS_TRY(ensure(spre, nw * 4));
    QS_TRY(ensure(stot, 8));
    QS_TRY(ensure(in_pos, atoms * 12));
    for (DevBuf *b : {&in_rad, &perm}) QS_TRY(ensure(*b, atoms * 4));
    QS_TRY(ensure(key, atoms * 8));
    for (DevBuf *b : {&sx, &aux}) QS_TRY(ensure(*b, atoms * 16));
    QS_TRY(ensure(count, (cells + 1) * 4));
    QS_TRY(ensure(start, (cells + 1) * 4));
    QS_TRY(ensure(cse, cells * 8));
    for (DevBuf *b : {&vert, &vert2, &vert3}) QS_TRY(ensure(*b, verts * 12));
    QS_TRY(ensure(vkey, verts * 8));
    QS_TRY(ensure(atom, verts * 4));
    for (DevBuf *b : {&tri, &tri_out, &tri_loc}) QS_TRY(ensure(*b, tris * 12));
    for (DevBuf *b : {&vstart, &deg_blk, &keeppos, &tstart}) QS_TRY(ensure(*b, (tris / 256 + 2) * 4));
    const size_t m = (size_t)nmol + 1;
    QS_TRY(ensure(d_off, m * 8 + 64));
    QS_TRY(ensure(bb_out, m * 8 * sizeof(float)));
    QS_TRY(ensure(bb_part, qs_bbox_part_floats((int)m, qs_bbox_parts((uint32_t)std::min<uint64_t>(atoms, UINT32_MAX))) *
                               sizeof(float)));
    QS_TRY(ensure(bounds, m * 8));
    QS_TRY(ensure(rback, (4 + 2 * m) * 4));
    QS_TRY(ensure(d_mols, m * sizeof(QsMol)));
    return 0;
}

/* per-molecule grids and offsets into the concatenated arrays */
int Engine::setup(const std::vector<MolIn> &in, const std::vector<BBox> &bb)
{
    const int nmol = (int)in.size();
    mols.assign(nmol, QsMol{});
    uint64_t cell = 0, word = 0, vox = 0;
    n_atoms_total = 0;
    for (int m = 0; m < nmol; m++) {
        QsMol &M = mols[m];
        M.atom_off = in[m].atom_off;
        M.n_atoms = in[m].n;
        n_atoms_total += in[m].n;
        if (!(in[m].reso > 0.0f) || !std::isfinite(in[m].reso))
            return set_error(3, "molecule %d: voxel size must be a positive finite number", m);
        if (in[m].n > 1) {
            for (int c = 0; c < 3; c++)
                if (!(bb[m].vmin[c] > -100000.0f && bb[m].vmax[c] < 100000.0f && bb[m].vmin[c] <= bb[m].vmax[c]))
                    return set_error(3, "molecule %d: atom coordinates must lie strictly within (-1e5, 1e5) Angstrom", m);
            grid_from_bbox(bb[m], in[m].reso, M.G);
            QsGrid &G = M.G;
            if (G.g < 2 || G.gn < 1) return set_error(3, "molecule %d: degenerate grid (g=%d)", m, G.g);
            /* one molecule of any size runs in x-slabs; a batch's grids run whole */
            if (nmol > 1 && (uint64_t)G.g * G.g * G.g >= (1ull << 32))
                return set_error(4, "molecule %d: a grid of %d^3 points is meshed alone, not in a batch", m, G.g);
            if ((uint64_t)G.g * G.g * ((G.g + 63) / 64) >= (1ull << 32))
                return set_error(4, "molecule %d: grid of %d^3 points", m, G.g);
            int lim = g_opt.lim_mode ? (G.g - 2 * G.r - 1) : (G.g - 1);
            G.lim = std::max(lim, 0);
            M.W = (uint32_t)((G.g + 63) / 64);
            M.rows_per_wave = qs_pass1_rows_per_wave(G);
            M.slab_rows = qs_pass1_slab_rows(G, (int)M.W);
        } else {
            /* N <= 1: empty mesh (CudaSurf.cu:539-543); its atoms sit in one cell and no grid is run */
            M.G = QsGrid{0.f, 0.f, 0.f, in[m].reso, 1.0f, 0, 1, 1, 0};
            M.W = 0;
            M.rows_per_wave = 1;
            M.slab_rows = 1;
        }
        full_planes(M);
        const uint64_t gg = (uint64_t)M.G.g;
        M.cell_off = (uint32_t)cell;
        M.word_off = word;
        M.vox_off = vox;
        cell += (uint64_t)M.G.gn * M.G.gn * M.G.gn;
        word += gg * gg * M.W;
        vox += gg * gg * gg;
    }
    if (cell >= (1ull << 31)) return set_error(4, "neighbour grids too large for one batch");
    if (word >= (1ull << 32)) return set_error(4, "grids too large for one batch");
    ncell_total = cell;
    nw_total = word;
    vox_total = vox;
    scan_w0 = 0;
    scan_nw = word;
    return 0;
}

/* coarse index of a prefix array pre[0..n] (pre[n] = total): entry k = the segment holding item
   k << shift */
template <class T>
void append_coarse(std::vector<uint32_t> &out, const T *pre, int n, uint64_t total, int shift)
{
    const uint64_t K = (total >> shift) + 1;
    int j = 0;
    for (uint64_t k = 0; k < K; k++) {
        const uint64_t t = k << shift;
        while (j + 1 < n && (uint64_t)pre[j + 1] <= t) j++;
        out.push_back((uint32_t)j);
    }
}

/* block prefixes of the launches (one uint32 array, uploaded once) */
int Engine::build_meta(size_t &at_atoms, size_t &at_p1, size_t &at_mc, size_t &at_wpre,
                       uint32_t &nb_atoms, uint32_t &nb_p1, uint32_t &nb_mc)
{
    const int nmol = (int)mols.size();
    std::vector<uint32_t> &meta = h_meta;
    meta.assign(1, 0u); /* index 0 is never a prefix start (mid_at == 0 means none) */
    /* each block prefix (nmol + 1 entries) is followed by its coarse index */
    auto prefix = [&](auto blocks_of, uint32_t &total) {
        size_t at = meta.size();
        uint64_t acc = 0;
        for (int m = 0; m < nmol; m++) {
            meta.push_back((uint32_t)acc);
            acc += blocks_of(mols[m]);
        }
        meta.push_back((uint32_t)acc);
        total = (uint32_t)acc;
        std::vector<uint32_t> pre(meta.begin() + at, meta.end());
        append_coarse(meta, pre.data(), nmol, acc, QS_COARSE_BLOCKS);
        return at;
    };
    at_atoms = prefix([](const QsMol &M) { return (uint64_t)((M.n_atoms + 255) / 256); }, nb_atoms);
    const int p1mode = qs_pass1_mode(g_opt.pass1_direct != 0, g_opt.pass1_rows != 0);
    /* A run whose slab launch would leave the device under-filled (a small molecule whose planes
       are one slab each) takes slabs of fewer rows, down to P1_FILL_MINROWS: the masks do not
       depend on the slab height, only the number of blocks does. */
    if (p1mode == 0) {
        for (;;) {
            uint64_t total = 0;
            int rows = 0;
            for (const QsMol &M : mols)
                if (M.G.g) {
                    total += qs_pass1_blocks(M, 0);
                    rows = std::max(rows, M.slab_rows);
                }
            if (total >= P1_FILL_BLOCKS || rows <= P1_FILL_MINROWS) break;
            for (QsMol &M : mols)
                if (M.G.g) M.slab_rows = std::max(P1_FILL_MINROWS, std::min(M.slab_rows, (rows + 1) / 2));
        }
    }
    at_p1 = prefix([&](const QsMol &M) { return (uint64_t)(M.G.g ? qs_pass1_blocks(M, p1mode) : 0); }, nb_p1);
    p1_lds = 0;
    for (const QsMol &M : mols)
        if (M.G.g) p1_lds = std::max(p1_lds, qs_pass1_lds_bytes(M, p1mode));
    at_mc = prefix([](const QsMol &M) { return (uint64_t)(M.G.g ? qs_mc_count_blocks(M) : 0); }, nb_mc);
    /* word prefix (64-bit) stored as pairs, then its coarse index */
    if (meta.size() & 1) meta.push_back(0u);
    at_wpre = meta.size();
    std::vector<uint64_t> wp(nmol + 1);
    for (int m = 0; m <= nmol; m++) {
        uint64_t w = (m < nmol) ? phys_word_off(mols[m]) : nw_total;
        wp[m] = w;
        meta.push_back((uint32_t)w);
        meta.push_back((uint32_t)(w >> 32));
    }
    append_coarse(meta, wp.data(), nmol, nw_total, QS_COARSE_ITEMS);
    /* field pass 2 groups: same kernel path, r, dx and coordinate range */
    groups.resize(0);
    std::vector<P2Group> gs;
    for (int m = 0; m < nmol; m++) {
        const QsGrid &G = mols[m].G;
        if (!G.g) continue;
        float cmax = 0.0f;
        for (float o : {G.ox, G.oy, G.oz}) cmax = std::max(cmax, std::max(std::fabs(o), std::fabs(o + (float)G.g * G.dx)));
        cmax = std::max(cmax, (float)G.g * G.dx);
        float cmax2 = 1.0f;
        while (cmax2 < cmax) cmax2 *= 2.0f;
        const int path = qs_pass2_path(G.r, g_opt.pass2_direct != 0);
        size_t k = 0;
        for (; k < gs.size(); k++)
            if (gs[k].path == path && gs[k].r == G.r && gs[k].dx == G.dx && gs[k].cmax == cmax2) break;
        if (k == gs.size()) {
            gs.emplace_back();
            gs.back().path = path;
            gs.back().r = G.r;
            gs.back().dx = G.dx;
            gs.back().cmax = cmax2;
        }
        gs[k].mols.push_back(m);
    }
    /* tile marks of the molecules that the fixed-R pass-2 kernel runs on, in its tile order */
    ntflags = 0;
    for (QsMol &M : mols) M.tile_off = ~0u;
    for (auto &grp : gs)
        if (grp.path == 0)
            for (int m : grp.mols) {
                mols[m].tile_off = (uint32_t)ntflags;
                ntflags += qs_pass2_tiles(mols[m]);
            }
    if (ntflags >= (1ull << 32)) return set_error(4, "too many field-pass-2 tiles for one batch");
    for (auto &grp : gs) {
        grp.bpre_at = meta.size();
        uint64_t acc = 0;
        grp.xchunk = 1;
        for (int m : grp.mols) {
            meta.push_back((uint32_t)acc);
            acc += qs_pass2_blocks(mols[m], grp.path);
            grp.xchunk = std::max(grp.xchunk, 4u * (uint32_t)qs_pass2_ztiles(mols[m]));
        }
        meta.push_back((uint32_t)acc);
        grp.nblocks = (uint32_t)acc;
        std::vector<uint32_t> pre(meta.begin() + grp.bpre_at, meta.end());
        append_coarse(meta, pre.data(), (int)grp.mols.size(), acc, QS_COARSE_BLOCKS);
        grp.mid_at = meta.size();
        for (int m : grp.mols) meta.push_back((uint32_t)m);
    }
    groups.swap(gs);
    return 0;
}

/* candidate tables of field pass 2 for one group (built once per (r, dx, coordinate range)) */
int Engine::group_tables(P2Group &grp)
{
    for (P2Tables *t : tables)
        if (t->r == grp.r && t->dx == grp.dx && t->cmax == grp.cmax) {
            grp.tab = t;
            return 0;
        }
    OffsetTables T;
    build_offset_tables(grp.r, grp.dx, grp.cmax, T);
    P2Tables *t = new P2Tables{grp.r, grp.dx, grp.cmax, DevBuf{}, DevBuf{}, DevBuf{}, QsNormFlags{{0, 0, 0, 0}}};
    for (int n = 0; n < 256 && n + 1 < (int)T.norm_start.size(); n++)
        if (T.enum_end[n] == T.norm_start[n + 1]) t->flags.w[n >> 6] |= 1ull << (n & 63);
    tables.push_back(t);
    QS_TRY(ensure(t->offs, T.offs.size() * 4));
    QS_TRY(ensure(t->norm, T.norm_start.size() * 4));
    QS_TRY(ensure(t->end, T.enum_end.size() * 4));
    /* synchronous copies: T is freed on return */
    QS_HIP(hipMemcpy(t->offs.p, T.offs.data(), T.offs.size() * 4, hipMemcpyHostToDevice));
    QS_HIP(hipMemcpy(t->norm.p, T.norm_start.data(), T.norm_start.size() * 4, hipMemcpyHostToDevice));
    QS_HIP(hipMemcpy(t->end.p, T.enum_end.data(), T.enum_end.size() * 4, hipMemcpyHostToDevice));
    grp.tab = t;
    return 0;
}

/* field passes 1 and 2 over every molecule of the batch (cells must be ready) */
int Engine::field_stage(uint32_t nb_p1, size_t at_p1, StageTimer &tm)
{
    const int nmol = (int)mols.size();
    QS_TRY(ensure(outm, std::max<uint64_t>(nw_total, 1) * 8));
    QS_TRY(ensure(shellm, std::max<uint64_t>(nw_total, 1) * 8));
    QS_TRY(ensure(field, std::max<uint64_t>(vox_total, 1) * 4));
    QS_TRY(ensure(negm, std::max<uint64_t>(nw_total, 1) * 8));
    QS_TRY(ensure(zerom, std::max<uint64_t>(nw_total, 1) * 8));
    QS_TRY(ensure(spre, std::max<uint64_t>(nw_total, 1) * 4));
    QS_TRY(ensure(stot, 8));
    QS_TRY(ensure(scan_tmp, qs_scan_tiles(std::max<uint64_t>(nw_total, 1)) * sizeof(uint32_t)));
    /* the slab kernel of field pass 1 writes every voxel's default sign words and marks the
       pass-2 tiles that hold a shell voxel; pass 2 then runs on the marked tiles only */
    const int p1mode = qs_pass1_mode(g_opt.pass1_direct != 0, g_opt.pass1_rows != 0);
    uint8_t *tf = nullptr;
    if (p1mode == 0 && ntflags) {
        QS_TRY(ensure(tflag, ntflags));
        tf = as<uint8_t>(tflag);
    }
    /* this run's tile mark (the flags are not cleared: see qs_launch_pass1) */
    tmark = (uint8_t)(tmark == 255 ? 1 : tmark + 1);
    if (nb_p1) {
        if (p1mode == 2)
            qs_launch_pass1_direct(batch(at_p1, 0, nmol), nb_p1, as<int2>(cse), as<float4>(sx), as<uint64_t>(outm),
                                   as<uint64_t>(shellm), stream);
        else
            qs_launch_pass1(batch(at_p1, 0, nmol), nb_p1, p1mode, p1_lds, as<uint32_t>(start), as<float4>(sx),
                            as<float4>(aux), as<uint64_t>(outm), as<uint64_t>(shellm),
                            p1mode == 0 ? as<uint64_t>(negm) : nullptr, p1mode == 0 ? as<uint64_t>(zerom) : nullptr, tf,
                            tmark, stream);
        QS_HIP(hipGetLastError());
    }
    tm.mark("pass1");
    /* every row word's shell rank: pass 2 stores the shell voxels' values densely in that order
       (the shell voxels are at most the voxels, which the field is sized for) */
    qs_launch_shell_scan(as<uint64_t>(shellm), nw_total, as<uint32_t>(scan_tmp), as<uint32_t>(stot), as<uint32_t>(spre), stream);
    tm.mark("shell_scan");
    for (auto &grp : groups) {
        QS_TRY(group_tables(grp));
        qs_launch_pass2(batch(grp.bpre_at, grp.mid_at, (int)grp.mols.size()), grp.nblocks, grp.r, grp.path, as<uint64_t>(outm),
                        as<uint64_t>(shellm), as<uint32_t>(spre), as<float>(field), as<uint32_t>(negm), as<uint32_t>(zerom), as<int>(grp.tab->offs),
                        as<int>(grp.tab->norm), as<int>(grp.tab->end), grp.tab->flags,
                        (tf && grp.path == 0) ? tf + mols[grp.mols[0]].tile_off : nullptr, tmark, grp.xchunk, stream);
        QS_HIP(hipGetLastError());
    }
    field_sparse = true;
    tm.mark("pass2");
    return 0;
}

/* every voxel of every molecule's field in grid order (pass 2 stores the shell voxels only):
   *dense receives the device array, molecule m at its voxel offset */
int Engine::fill_field(const void **dense)
{
    *dense = field.p;
    if (!field_sparse) return 0;
    QS_TRY(ensure(field_dense, std::max<uint64_t>(vox_total, 1) * 4));
    for (const QsMol &M : mols)
        qs_launch_field_expand(as<float>(field), as<uint32_t>(spre) + M.word_off, as<uint64_t>(outm) + M.word_off,
                               as<uint64_t>(shellm) + M.word_off, as<float>(field_dense) + M.vox_off, M.G.g, (int)M.W,
                               M.G.dx, stream);
    QS_HIP(hipGetLastError());
    *dense = field_dense.p;
    return 0;
}

/* marching cubes over every molecule (sign masks must be ready) */
int Engine::mesh_stage(size_t at_mc, uint32_t nb_mc, size_t at_wpre, StageTimer &tm, bool with_atoms)
{
    const int nmol = (int)mols.size();
    const uint64_t nw = std::max<uint64_t>(nw_total, 1);
    QS_TRY(ensure(s4, nw * 32));
    QS_TRY(ensure(wcnt, nw * 4));
    QS_TRY(ensure(vbw, nw * 4));
    QS_TRY(ensure(tbw, nw * 4));
    QS_TRY(ensure(alist, nw * 4));
    QS_TRY(ensure(avb, nw * 4));
    QS_TRY(ensure(atb, nw * 4));
    QS_TRY(ensure(ntb, nw * 24));
    QS_TRY(ensure(totals, 16));
    /* Vertex and triangle pools.  The totals stay on the device: the emission, closest-atom,
       smoothing and filter kernels read them there and run as persistent blocks over the pools,
       so nothing waits for the host until the run's results are read back at its end, which is
       also where a pool that was too small shows (the run is then repeated with larger pools).
       The pools are planned before the word scans, which write the emission launches' block
       starts for them. */
    plan_pools();
    QS_TRY(ensure(vstart, ((size_t)pool_v / 256 + 1) * 4));
    QS_TRY(ensure(tstart, ((size_t)pool_t / 256 + 1) * 4));
    /* degenerate triangles per 256-triangle block and in total (totals[3]), counted by the
       triangle kernel (the block counts zeroed by the word scans, totals[3] by the scans) */
    QS_TRY(ensure(deg_blk, ((size_t)pool_t / 256 + 1) * 4));
    if (nb_mc)
        qs_launch_mc_count(batch(at_mc, 0, nmol), nb_mc, as<uint64_t>(negm), as<uint64_t>(zerom), as<uint64_t>(s4),
                           as<uint32_t>(wcnt), as<uint64_t>(ntb), stream);
    if (scan_nw) {
        QS_TRY(ensure(scan_tmp, 3 * qs_scan_tiles(scan_nw) * sizeof(uint32_t)));
        QsBlockStarts bs;
        bs.vstart = as<uint32_t>(vstart);
        bs.nv = pool_v / 256 + 1;
        bs.tstart = as<uint32_t>(tstart);
        bs.nt = pool_t / 256 + 1;
        bs.zero = as<uint32_t>(deg_blk);
        bs.nzero = pool_t / 256 + 1;
        qs_launch_mc_scans(as<uint32_t>(wcnt) + scan_w0, scan_nw, as<uint32_t>(scan_tmp),
                           as<uint32_t>(vbw) + scan_w0, as<uint32_t>(tbw) + scan_w0, as<uint32_t>(alist), as<uint32_t>(avb),
                           as<uint32_t>(atb), as<uint32_t>(totals), (uint32_t)scan_w0, stream, bs);
        QS_HIP(hipGetLastError());
    } else {
        QS_HIP(hipMemsetAsync(totals.p, 0, 16, stream));
    }
    tm.mark("mc_count");
    /* per-molecule first vertex and triangle (the scans at each molecule's first word, from the
       word prefix) and the vertex coarse index, computed on the device */
    const uint64_t *wpre = (const uint64_t *)(as<uint32_t>(d_meta) + at_wpre);
    const uint32_t *wcoarse = as<uint32_t>(d_meta) + at_wpre + 2 * ((size_t)nmol + 1);
    QS_TRY(ensure(bounds, ((size_t)nmol + 1) * 8));
    QS_TRY(ensure(tpre_d, ((size_t)nmol + 1) * 8));
    const uint32_t ncoarse = (pool_v >> QS_COARSE_ITEMS) + 2;
    QS_TRY(ensure(d_off, ((size_t)nmol + 1) * 8 + (size_t)ncoarse * 4 + 64));
    uint32_t *d_vpre = as<uint32_t>(d_off) + (nmol + 1);
    qs_launch_mesh_bases(as<uint32_t>(vbw), as<uint32_t>(tbw), wpre, nw_total, nmol, as<uint32_t>(totals), d_vpre,
                         as<uint64_t>(tpre_d), d_vpre + nmol + 1, ncoarse, stream);
    QS_TRY(ensure(vert, (size_t)pool_v * 12));
    QS_TRY(ensure(vkey, (size_t)pool_v * 8));
    QS_TRY(ensure(tri, (size_t)pool_t * 12));
    /* every scanned word may be active */
    const uint32_t nact_max = (uint32_t)std::max<uint64_t>(scan_nw, 1);
    const uint32_t *counts = as<uint32_t>(totals);
    /* Without stage timing the closest-atom pass (which reads only the unsmoothed vertices) runs
       on the side stream beside the smoothing passes, and is joined before the filter (ev_side).
       For one molecule the vertex kernel runs there too, before it, beside the triangle kernel
       (the smoothing passes wait for the vertices, ev_vert): the 50,000-atom 0.2 A mesh 1.972 ->
       1.956 ms, while a batch of 256 molecules was 0.4 % slower that way. */
    const bool overlap = with_atoms && !tm.on;
    const bool fork_vertices = overlap && nmol == 1;
    qs_launch_mc_emit(as<QsMol>(d_mols), wpre, wcoarse, as<uint32_t>(alist), nact_max, as<uint32_t>(avb), as<uint32_t>(vstart),
                      pool_v, as<uint64_t>(s4), as<float>(field), field_sparse ? as<uint32_t>(spre) : nullptr, as<uint64_t>(outm),
                      field_sparse ? as<uint64_t>(shellm) : nullptr, as<float>(vert), as<uint64_t>(vkey), as<uint32_t>(atb),
                      as<uint32_t>(tstart), pool_t, as<uint64_t>(negm), as<uint64_t>(zerom), as<uint32_t>(vbw),
                      as<uint64_t>(ntb), as<int>(tri), 0u, as<uint32_t>(deg_blk), as<uint32_t>(totals) + 3, counts, stream,
                      fork_vertices ? side : stream, ev_emit, ev_vert, scan_nw != 0);
    QS_HIP(hipGetLastError());
    vert_pending = fork_vertices;
    tm.mark("mc_emit");
    if (with_atoms) {
        QS_TRY(ensure(atom, (size_t)pool_v * 4));
        hipStream_t cs = overlap ? side : stream;
        if (overlap && !fork_vertices) {
            QS_HIP(hipEventRecord(ev_emit, stream));
            QS_HIP(hipStreamWaitEvent(side, ev_emit, 0));
        }
        ids_sorted = g_opt.atom_ids_sorted != 0;
        qs_launch_closest_atom(as<QsMol>(d_mols), d_vpre, d_vpre + nmol + 1, as<uint32_t>(start), as<float4>(sx), as<int>(perm),
                               as<float>(vert), pool_v, ids_sorted ? 1 : 0, as<int>(atom), counts + 1, cs);
        QS_HIP(hipGetLastError());
        if (overlap) {
            QS_HIP(hipEventRecord(ev_side, side));
            side_pending = true;
        }
        tm.mark("closest_atom");
    }
    return 0;
}

/* Vertex and triangle pools of a run: about 8 A^2 of surface per atom (measured 2-6 A^2 on the
   benchmark molecules, cavities included), at least what earlier runs needed (cap_v, cap_t) */
void Engine::plan_pools()
{
    uint64_t ev = 4096;
    for (const QsMol &M : mols)
        if (M.G.g) ev += 256 + (uint64_t)(8.0 * M.n_atoms / ((double)M.G.dx * M.G.dx));
    ev = std::min<uint64_t>(ev, 4 * vox_total + 4096);
    ev = std::max<uint64_t>(ev, cap_v);
    uint64_t et = std::max<uint64_t>(2 * ev + 4096, cap_t);
    pool_v = (uint32_t)std::min<uint64_t>(ev, (1ull << 31) - 1);
    pool_t = (uint32_t)std::min<uint64_t>(et, (1ull << 31) / 3 - 1);
}

/* after a run's results are read: did its pools hold the mesh?  If not, they grow to fit */
bool Engine::pools_overflowed(uint32_t nv_, uint32_t nt_)
{
    if (nv_ <= pool_v && nt_ <= pool_t) return false;
    cap_v = std::max<uint64_t>(cap_v, (uint64_t)nv_ + nv_ / 4 + 4096);
    cap_t = std::max<uint64_t>(cap_t, (uint64_t)nt_ + nt_ / 4 + 4096);
    return true;
}

/* grow a buffer to at least `bytes`, keeping its first `keep` bytes */
static int grow_keep(DevBuf &b, size_t bytes, size_t keep, hipStream_t s)
{
    if (bytes <= b.cap && b.p) return 0;
    DevBuf nb;
    const size_t want = std::max<size_t>(bytes + (bytes >> 2), 256);
    QS_HIP(hipMalloc(&nb.p, want));
    nb.cap = want;
    if (keep && b.p) {
        QS_HIP(hipMemcpyAsync(nb.p, b.p, keep, hipMemcpyDeviceToDevice, s));
        QS_HIP(hipStreamSynchronize(s));
    }
    release(b);
    b = nb;
    return 0;
}

/* device bytes of the grid arrays per x plane: per row word the outside, shell, sign and zero
   masks (32 B), slot masks (32 B), triangle-count planes (24 B) and six counts and scans
   (24 B); per voxel the field (4 B) */
static uint64_t plane_bytes(const QsMol &M)
{
    const uint64_t g = (uint64_t)M.G.g;
    return g * M.W * 112 + g * g * 4;
}

/* Planes each slab owns (CudaSurf.cu:626-636 sizes its slices from the SLICE setting instead),
   0 when the whole grid runs at once: that is when it fits the memory budget and its grid points
   fit 32-bit indexes (the grid-walk smoothing keys them so).  Option slab_planes forces slabs of
   that many planes; option slab_mb sets the budget (default: half the device's free memory). */
int Engine::slab_plan(const QsMol &M, int &planes)
{
    planes = 0;
    const int g = M.G.g, R = M.G.r;
    if (g_opt.slab_planes > 0 && g_opt.slab_planes < g) {
        planes = g_opt.slab_planes;
        return 0;
    }
    const uint64_t pb = plane_bytes(M);
    uint64_t budget = (uint64_t)g_opt.slab_mb << 20;
    if (!budget) {
        size_t free_b = 0, total_b = 0;
        QS_HIP(hipMemGetInfo(&free_b, &total_b));
        budget = free_b / 2;
    }
    const bool fits = pb * (uint64_t)g <= budget;
    if (fits && (uint64_t)g * g * g < (1ull << 32)) return 0;
    const uint64_t halo = 2 * (uint64_t)R + 3;
    const uint64_t stored = budget / pb;
    if (stored < halo + 8) return set_error(1, "grid of %d^3 points: not enough device memory for an x-slab", g);
    planes = (int)std::min<uint64_t>((uint64_t)g, stored - halo);
    return 0;
}

/* Marching cubes of one slab (CudaSurf.cu:373-458 for one slice): count the slab's words, scan
   the owned grid points' words and those of the next slab's first plane (so the vertex ids the
   triangles take there are the ones that slab emits), emit the owned vertices at global id
   vbase + local id and the slab's triangles at tbase. */
int Engine::slab_mesh(size_t at_mc, uint32_t nb_mc, size_t at_wpre, uint64_t owned_word, uint32_t vbase, uint32_t tbase,
                      uint32_t &n_owned, uint32_t &n_tri)
{
    const uint64_t nw = std::max<uint64_t>(nw_total, 1);
    QS_TRY(ensure(s4, nw * 32));
    for (DevBuf *b : {&wcnt, &vbw, &tbw, &alist, &avb, &atb}) QS_TRY(ensure(*b, nw * 4));
    QS_TRY(ensure(ntb, nw * 24));
    QS_TRY(ensure(totals, 16));
    if (nb_mc)
        qs_launch_mc_count(batch(at_mc, 0, 1), nb_mc, as<uint64_t>(negm), as<uint64_t>(zerom), as<uint64_t>(s4),
                           as<uint32_t>(wcnt), as<uint64_t>(ntb), stream);
    QS_TRY(ensure(scan_tmp, 3 * qs_scan_tiles(std::max<uint64_t>(scan_nw, 1)) * sizeof(uint32_t)));
    qs_launch_mc_scans(as<uint32_t>(wcnt) + scan_w0, scan_nw, as<uint32_t>(scan_tmp),
                       as<uint32_t>(vbw) + scan_w0, as<uint32_t>(tbw) + scan_w0, as<uint32_t>(alist), as<uint32_t>(avb),
                       as<uint32_t>(atb), as<uint32_t>(totals), (uint32_t)scan_w0, stream);
    QS_HIP(hipGetLastError());
    uint32_t hb[4] = {0, 0, 0, 0};
    uint32_t *hs = (uint32_t *)stage(sizeof(hb));
    uint32_t *hd = hs ? hs : hb;
    QS_HIP(hipMemcpyAsync(hd, totals.p, 12, hipMemcpyDeviceToHost, stream));
    if (owned_word != ~0ull)
        QS_HIP(hipMemcpyAsync(hd + 3, as<uint32_t>(vbw) + owned_word, 4, hipMemcpyDeviceToHost, stream));
    QS_HIP(hipStreamSynchronize(stream));
    if (hs) memcpy(hb, hs, sizeof(hb));
    n_tri = hb[0];
    const uint32_t nact = hb[2];
    n_owned = (owned_word != ~0ull) ? hb[3] : hb[1];
    if ((uint64_t)vbase + n_owned >= (1ull << 31) || 3 * ((uint64_t)tbase + n_tri) >= (1ull << 31))
        return set_error(4, "mesh too large");
    QS_TRY(grow_keep(vert, ((size_t)vbase + n_owned + 1) * 12, (size_t)vbase * 12, stream));
    QS_TRY(grow_keep(tri, ((size_t)tbase + n_tri + 1) * 12, (size_t)tbase * 12, stream));
    QS_TRY(ensure(vkey, ((size_t)n_owned + 1) * 8));
    QS_TRY(ensure(vstart, ((size_t)n_owned / 256 + 1) * 4));
    QS_TRY(ensure(tstart, ((size_t)n_tri / 256 + 1) * 4));
    QS_TRY(ensure(deg_blk, ((size_t)n_tri / 256 + 1) * 4));
    const uint64_t *wpre = (const uint64_t *)(as<uint32_t>(d_meta) + at_wpre);
    const uint32_t *wcoarse = as<uint32_t>(d_meta) + at_wpre + 4;
    qs_launch_mc_emit(as<QsMol>(d_mols), wpre, wcoarse, as<uint32_t>(alist), nact, as<uint32_t>(avb), as<uint32_t>(vstart),
                      n_owned, as<uint64_t>(s4), as<float>(field), as<uint32_t>(spre), as<uint64_t>(outm), as<uint64_t>(shellm),
                      as<float>(vert) + 3 * (size_t)vbase, as<uint64_t>(vkey), as<uint32_t>(atb), as<uint32_t>(tstart), n_tri,
                      as<uint64_t>(negm), as<uint64_t>(zerom), as<uint32_t>(vbw), as<uint64_t>(ntb),
                      as<int>(tri) + 3 * (size_t)tbase, vbase, as<uint32_t>(deg_blk), as<uint32_t>(totals) + 3, nullptr, stream);
    QS_HIP(hipGetLastError());
    return 0;
}

/* One molecule in x-slabs (the reference's slice loop, CudaSurf.cu:626-754, with disjoint cube
   ownership instead of overlapping slices).  Slab k owns the cubes with origin x in [x0, x1).
   Its arrays hold planes [x0 - 1 - R, x1 + 2 + R): field pass 1 labels all of them, field pass 2
   evaluates [x0 - 1, x1 + 2) (its window reads R planes either side), the marching-cubes count
   covers grid points [x0, x1 + 1) and polygonises cubes below x1.  Every stage therefore sees
   exactly the values the whole-grid run computes there.  Vertex ids are ranks in global
   (point, slot) order and triangles follow global cube order, so the slabs' meshes concatenate
   into the whole-grid mesh: the seams weld exactly.  The per-vertex stages then run on the whole
   mesh: closest atom, Laplacian passes over the triangle array, the degenerate filter. */
int Engine::run_slabs(int planes, int smooth, StageTimer &tm)
{
    QsMol &M = mols[0];
    const int g = M.G.g, R = M.G.r;
    const uint64_t plane_w = (uint64_t)g * M.W, plane_v = (uint64_t)g * g;
    multi_slab = true;
    n_slabs = (g + planes - 1) / planes;
    uint64_t vbase = 0, tbase = 0;
    StageTimer none;
    for (int x0 = 0; x0 < g; x0 += planes) {
        const int x1 = std::min(g, x0 + planes);
        M.p1_lo = std::max(0, x0 - 1 - R);
        M.p1_hi = std::min(g, x1 + 2 + R);
        M.p2_lo = std::max(0, x0 - 1);
        M.p2_hi = std::min(g, x1 + 2);
        M.mc_lo = x0;
        M.mc_hi = std::min(g, x1 + 1);
        M.cube_hi = x1;
        /* global plane x lands at array plane x - p1_lo (unsigned wrap-around) */
        M.word_off = 0ull - (uint64_t)M.p1_lo * plane_w;
        M.vox_off = 0ull - (uint64_t)M.p1_lo * plane_v;
        nw_total = (uint64_t)(M.p1_hi - M.p1_lo) * plane_w;
        vox_total = (uint64_t)(M.p1_hi - M.p1_lo) * plane_v;
        scan_w0 = (uint64_t)(M.mc_lo - M.p1_lo) * plane_w;
        scan_nw = (uint64_t)(M.mc_hi - M.mc_lo) * plane_w;
        /* the first word of the next slab's first plane: the vertices before it are this slab's */
        const uint64_t owned_word = (x1 < g) ? (uint64_t)(x1 - M.p1_lo) * plane_w : ~0ull;
        size_t at_atoms, at_p1, at_mc, at_wpre;
        uint32_t nb_atoms, nb_p1, nb_mc;
        QS_TRY(build_meta(at_atoms, at_p1, at_mc, at_wpre, nb_atoms, nb_p1, nb_mc));
        QS_TRY(ensure(d_meta, h_meta.size() * 4));
        QS_TRY(upload(d_mols.p, mols.data(), sizeof(QsMol)));
        QS_TRY(upload(d_meta.p, h_meta.data(), h_meta.size() * 4));
        QS_TRY(field_stage(nb_p1, at_p1, none));
        uint32_t n_owned = 0, n_tri = 0;
        QS_TRY(slab_mesh(at_mc, nb_mc, at_wpre, owned_word, (uint32_t)vbase, (uint32_t)tbase, n_owned, n_tri));
        vbase += n_owned;
        tbase += n_tri;
    }
    tm.mark("slabs");
    nv = (uint32_t)vbase;
    nt_all = (uint32_t)tbase;
    n_active = 0;
    M.vert_off = 0;
    M.n_vert = nv;
    M.tri_off = 0;
    M.n_tri = nt_all;
    vert_off.assign(2, 0);
    vert_off[1] = nv;
    QS_TRY(upload(d_mols.p, mols.data(), sizeof(QsMol)));
    std::vector<uint32_t> &vpre = h_vpre;
    vpre.assign({0u, nv});
    {
        std::vector<uint32_t> vc;
        append_coarse(vc, vpre.data(), 1, nv, QS_COARSE_ITEMS);
        vpre.insert(vpre.end(), vc.begin(), vc.end());
    }
    QS_TRY(ensure(d_off, 2 * 4 + vpre.size() * 4 + 64));
    uint32_t *d_vpre = as<uint32_t>(d_off) + 2;
    QS_TRY(upload(d_vpre, vpre.data(), vpre.size() * 4));

    /* closest atom per unsmoothed vertex, beside the smoothing passes unless stages are timed */
    /* (vert and tri already hold the whole mesh: grow_keep sized them) */
    QS_TRY(ensure(atom, (size_t)std::max<uint32_t>(nv, 1) * 4));
    const bool overlap = !tm.on;
    hipStream_t cs = overlap ? side : stream;
    if (overlap) {
        QS_HIP(hipEventRecord(ev_emit, stream));
        QS_HIP(hipStreamWaitEvent(side, ev_emit, 0));
    }
    ids_sorted = g_opt.atom_ids_sorted != 0;
    qs_launch_closest_atom(as<QsMol>(d_mols), d_vpre, d_vpre + 2, as<uint32_t>(start), as<float4>(sx), as<int>(perm),
                           as<float>(vert), nv, ids_sorted ? 1 : 0, as<int>(atom), nullptr, cs);
    QS_HIP(hipGetLastError());
    if (overlap) {
        QS_HIP(hipEventRecord(ev_side, side));
        side_pending = true;
    }
    tm.mark("closest_atom");

    /* Laplacian passes (Jacobi) over the triangle array's incidences */
    res_vert = as<float>(vert);
    if (smooth > 0 && nv > 0 && nt_all > 0) {
        QS_TRY(ensure(vert2, (size_t)nv * 12));
        if (smooth > 1) QS_TRY(ensure(vert3, (size_t)nv * 12));
        QS_TRY(ensure(inc_cnt, (size_t)nv * 4));
        QS_TRY(ensure(inc_start, ((size_t)nv + 1) * 4));
        QS_TRY(ensure(inc_ent, (size_t)nt_all * 12));
        QS_HIP(hipMemsetAsync(inc_cnt.p, 0, (size_t)nv * 4, stream));
        qs_launch_incid_count(as<int>(tri), nt_all, nv, as<uint32_t>(inc_cnt), stream);
        QS_TRY(scan(as<uint32_t>(inc_cnt), as<uint32_t>(inc_start), nv, as<uint32_t>(inc_start) + nv));
        QS_HIP(hipMemsetAsync(inc_cnt.p, 0, (size_t)nv * 4, stream));
        qs_launch_smooth_incid_build(as<int>(tri), nt_all, nv, as<uint32_t>(inc_cnt), as<uint32_t>(inc_start),
                                     as<uint32_t>(inc_ent), stream);
        const float *src = as<float>(vert);
        float *dst = as<float>(vert2);
        for (int it = 0; it < smooth; it++) {
            qs_launch_smooth_incid(as<uint32_t>(inc_start), as<uint32_t>(inc_ent), as<int>(tri), nv, src, dst, stream);
            src = dst;
            dst = (dst == as<float>(vert2)) ? as<float>(vert3) : as<float>(vert2);
        }
        QS_HIP(hipGetLastError());
        res_vert = (float *)src;
        tm.mark("smooth");
    }
    return 0;
}

int Engine::run_batch(const float *d_pos, const float *d_rad, const std::vector<MolIn> &in, int smooth,
                      const std::vector<BBox> *host_bbox)
{
    const int nmol = (int)in.size();
    /* (inputs queued through the staging by upload_begin stay) */
    if (!inputs_staged) QS_TRY(upload_reset());
    inputs_staged = false;
    batch_valid = false;
    nv = nt = nt_all = n_active = 0;
    res_vert = nullptr;
    res_tri = nullptr;
    res_atom = nullptr;
    multi_slab = false;
    n_slabs = 1;
    vert_off.assign(nmol + 1, 0);
    tri_off.assign(nmol + 1, 0);
    mols.clear();
    if (nmol == 0) {
        batch_valid = true;
        return 0;
    }
    StageTimer tm;
    tm.on = (g_opt.trace != 0 || g_opt.timing != 0) && !g_arena; /* (timing events are not created in a fixed region's run) */
    tm.s = stream;
    tm.mark("start");

    /* ---- bounding boxes (getMinMax CudaSurf.cu:149-174) ---- */
    std::vector<BBox> bb(nmol);
    if (host_bbox) {
        bb = *host_bbox;
    } else {
        std::vector<uint32_t> off(nmol + 1);
        for (int m = 0; m < nmol; m++) off[m] = in[m].atom_off;
        off[nmol] = in[nmol - 1].atom_off + in[nmol - 1].n;
        QS_TRY(ensure(d_off, ((size_t)nmol + 1) * 4 * 2 + 64));
        QS_TRY(ensure(bb_out, (size_t)nmol * 8 * sizeof(float)));
        uint32_t max_n = 0;
        for (int m = 0; m < nmol; m++) max_n = std::max(max_n, in[m].n);
        const int parts = qs_bbox_parts(max_n);
        QS_TRY(ensure(bb_part, qs_bbox_part_floats(nmol, parts) * sizeof(float)));
        /* (one molecule's atom range goes in the launch's arguments) */
        if (nmol > 1) QS_TRY(upload(d_off.p, off.data(), (nmol + 1) * 4));
        qs_launch_bbox(d_pos, d_rad, nmol > 1 ? as<uint32_t>(d_off) : nullptr, off[0], off[1], nmol, parts, as<float>(bb_part),
                       as<float>(bb_out), stream);
        std::vector<float> h((size_t)nmol * 8);
        float *hs = (float *)stage(h.size() * 4);
        QS_HIP(hipMemcpyAsync(hs ? (void *)hs : (void *)h.data(), bb_out.p, h.size() * 4, hipMemcpyDeviceToHost, stream));
        QS_HIP(hipStreamSynchronize(stream));
        if (hs) memcpy(h.data(), hs, h.size() * 4);
        for (int m = 0; m < nmol; m++) {
            if (in[m].n > 1 && h[8 * m + 7] != 0.0f)
                return set_error(3, "molecule %d: atom positions or radii contain NaN or Inf", m);
            for (int c = 0; c < 3; c++) {
                bb[m].vmin[c] = h[8 * m + c];
                bb[m].vmax[c] = h[8 * m + 3 + c];
            }
            bb[m].max_rad = h[8 * m + 6];
        }
    }
    tm.mark("bbox");
    QS_TRY(setup(in, bb));
    size_t at_atoms, at_p1, at_mc, at_wpre;
    uint32_t nb_atoms, nb_p1, nb_mc;
    QS_TRY(build_meta(at_atoms, at_p1, at_mc, at_wpre, nb_atoms, nb_p1, nb_mc));
    QS_TRY(ensure(d_mols, mols.size() * sizeof(QsMol)));
    QS_TRY(ensure(d_meta, h_meta.size() * 4));
    QS_TRY(upload(d_mols.p, mols.data(), mols.size() * sizeof(QsMol)));
    QS_TRY(upload(d_meta.p, h_meta.data(), h_meta.size() * 4));

    /* ---- atoms into cells (CudaSurf.cu:594-622), global over the batch ---- */
    const uint32_t n = n_atoms_total;
    const uint64_t ncell = ncell_total;
    /* per atom its cell and rank within the cell; per cell its atom count (cleared by a fill) */
    QS_TRY(ensure(key, (size_t)std::max<uint32_t>(n, 1) * 8));
    QS_TRY(ensure(count, ncell * 4));
    QS_TRY(ensure(start, (ncell + 1) * 4));
    QS_TRY(ensure(perm, (size_t)std::max<uint32_t>(n, 1) * 4));
    QS_TRY(ensure(cse, ncell * 8));
    QS_TRY(ensure(sx, (size_t)std::max<uint32_t>(n, 1) * 16));
    QS_TRY(ensure(aux, (size_t)std::max<uint32_t>(n, 1) * 16));
    QS_HIP(hipMemsetAsync(count.p, 0, ncell * 4, stream));
    if (nb_atoms) qs_launch_hash_count(batch(at_atoms, 0, nmol), nb_atoms, d_pos, as<int2>(key), as<uint32_t>(count), stream);
    /* start[c] = atoms in cells < c (global); start[ncell] = n */
    QS_TRY(scan(as<uint32_t>(count), as<uint32_t>(start), ncell, as<uint32_t>(start) + ncell));
    if (n) {
        qs_launch_scatter(as<int2>(key), (int)n, as<uint32_t>(start), as<int>(perm), stream);
        qs_launch_cell_finalize(as<uint32_t>(start), as<uint32_t>(count), (int)ncell, as<int>(perm), as<int2>(cse), stream);
        qs_launch_gather_aux(batch(at_atoms, 0, nmol), nb_atoms, as<int>(perm), d_pos, d_rad, as<float4>(sx), as<float4>(aux),
                             stream);
    }
    QS_HIP(hipGetLastError());
    tm.mark("cells");

    /* ---- field, mesh: the whole grid at once, or one molecule in x-slabs ---- */
    int slab_planes = 0;
    if (nmol == 1 && mols[0].G.g && !g_arena) QS_TRY(slab_plan(mols[0], slab_planes));
    if (g_arena && nmol == 1 && (uint64_t)mols[0].G.g * mols[0].G.g * mols[0].G.g >= (1ull << 32))
        return set_error(6, "grid of %d^3 points: too large for the fixed scratch region", mols[0].G.g);
    if (slab_planes) {
        QS_TRY(run_slabs(slab_planes, smooth, tm));
    } else {
    QS_TRY(field_stage(nb_p1, at_p1, tm));
    QS_TRY(mesh_stage(at_mc, nb_mc, at_wpre, tm, true));

    /* ---- Laplacian passes (Jacobi), over the vertex pool ---- */
    res_vert = as<float>(vert);
    const uint32_t *d_vpre = as<uint32_t>(d_off) + (nmol + 1);
    if (smooth > 0) {
        QS_TRY(ensure(vert2, (size_t)pool_v * 12));
        /* `vert` keeps the unsmoothed positions; passes alternate between vert2 and vert3 */
        if (smooth > 1) QS_TRY(ensure(vert3, (size_t)pool_v * 12));
        const float *src = as<float>(vert);
        float *dst = as<float>(vert2);
        if (vert_pending) {
            QS_HIP(hipStreamWaitEvent(stream, ev_vert, 0));
            vert_pending = false;
        }
        for (int it = 0; it < smooth; it++) {
            qs_launch_smooth(as<QsMol>(d_mols), d_vpre, d_vpre + nmol + 1, as<uint64_t>(negm), as<uint64_t>(zerom), as<uint64_t>(s4),
                             as<uint32_t>(vbw), as<uint64_t>(vkey), pool_v, as<uint64_t>(ntb), as<uint32_t>(tbw), as<int>(tri),
                             src, dst, as<uint32_t>(totals) + 1, stream);
            src = dst;
            dst = (dst == as<float>(vert2)) ? as<float>(vert3) : as<float>(vert2);
        }
        QS_HIP(hipGetLastError());
        res_vert = (float *)src;
        tm.mark("smooth");
    }
    }

    /* the closest-atom pass (after the vertex kernel on the same stream) joins the main stream */
    if (side_pending) {
        QS_HIP(hipStreamWaitEvent(stream, ev_side, 0));
        side_pending = false;
    }
    if (vert_pending) {
        QS_HIP(hipStreamWaitEvent(stream, ev_vert, 0));
        vert_pending = false;
    }

    /* ---- degenerate filter (CudaSurf.cu:820-829), per-molecule kept counts ---- */
    /* The triangle kernel counted them (totals[3]).  The compaction runs without waiting for that
       count: with none it writes nothing and the input array is the result.  Kept triangles go to
       256 b - (degenerate ones in blocks before b) + rank in block b; the kept counts before each
       molecule's first triangle give the offsets, equal to the input offsets when none drop.  A
       slab run knows its triangle count; otherwise the filter covers the triangle pool and takes
       the count from the device. */
    const uint32_t tcap = multi_slab ? nt_all : pool_t;
    const uint64_t nblk_t = multi_slab ? ((uint64_t)nt_all + 255) / 256 : (uint64_t)pool_t / 256 + 1;
    if (multi_slab) {
        /* the slabs' triangle launches counted per slab: count again over the whole array */
        QS_TRY(ensure(deg_blk, (nblk_t + 1) * 4));
        QS_HIP(hipMemsetAsync(as<uint32_t>(totals) + 3, 0, 4, stream));
        qs_launch_deg_count(as<int>(tri), nt_all, as<uint32_t>(deg_blk), as<uint32_t>(totals) + 3, stream);
        QS_HIP(hipGetLastError());
        /* the molecule's first triangle and the end, for the kept counts */
        h_bidx.assign({0ull, (uint64_t)nt_all});
        QS_TRY(ensure(tpre_d, 2 * 8));
        QS_TRY(upload(tpre_d.p, h_bidx.data(), 2 * 8));
    }
    QS_TRY(ensure(keeppos, (nblk_t + 1) * 4));
    QS_TRY(ensure(tri_out, (size_t)std::max<uint32_t>(tcap, 1) * 12));
    QS_TRY(scan(as<uint32_t>(deg_blk), as<uint32_t>(keeppos), nblk_t, as<uint32_t>(keeppos) + nblk_t));
    qs_launch_tri_compact(as<int>(tri), tcap, as<uint32_t>(keeppos), as<uint32_t>(totals) + 3, as<int>(tri_out),
                          multi_slab ? nullptr : as<uint32_t>(totals), stream);
    QS_TRY(ensure(bounds, ((size_t)nmol + 1) * 8));
    /* the run's one read-back, gathered by the last kernel: totals, per-molecule first vertices
       and kept-triangle offsets */
    std::vector<uint32_t> hb(4 + 2 * ((size_t)nmol + 1));
    QS_TRY(ensure(rback, hb.size() * 4));
    qs_launch_tri_kept_before(as<int>(tri), as<uint32_t>(keeppos), as<uint64_t>(tpre_d), nmol + 1, tcap, as<uint32_t>(bounds),
                              as<uint32_t>(totals), as<uint32_t>(d_off) + (nmol + 1), as<uint32_t>(rback), stream);
    uint32_t *fs = (uint32_t *)stage(hb.size() * 4);
    uint32_t *fd = fs ? fs : hb.data();
    QS_HIP(hipMemcpyAsync(fd, rback.p, hb.size() * 4, hipMemcpyDeviceToHost, stream));
    QS_HIP(hipStreamSynchronize(stream));
    if (fs) memcpy(hb.data(), fs, hb.size() * 4);
    tm.mark("filter");
    if (!multi_slab) {
        nt_all = hb[0];
        nv = hb[1];
        n_active = hb[2];
        if (pools_overflowed(nv, nt_all)) {
            /* the mesh did not fit the pools, which have grown to fit it: run again */
            if (reruns) return set_error(4, "mesh pools overflowed again (%u vertices, %u triangles)", nv, nt_all);
            reruns = 1;
            if (g_arena) {
                g_arena->used = 0;
                forget_buffers();
            }
            const int rc = run_batch(d_pos, d_rad, in, smooth, host_bbox);
            reruns = 0;
            return rc;
        }
        if ((uint64_t)nt_all * 3 >= (1ull << 31) || nv >= (1u << 31)) return set_error(4, "mesh too large");
    }
    const uint32_t ndeg = hb[3];
    vert_off.assign(hb.begin() + 4, hb.begin() + 4 + nmol + 1);
    tri_off.assign(hb.begin() + 4 + nmol + 1, hb.end());
    for (int m = 0; m < nmol; m++) {
        mols[m].vert_off = vert_off[m];
        mols[m].n_vert = vert_off[m + 1] - vert_off[m];
    }
    nt = tri_off[nmol];
    res_tri = ndeg ? as<int>(tri_out) : as<int>(tri);
    res_atom = as<int>(atom);
    batch_valid = true; /* device copies: vertex offsets after the molecule prefixes in d_off, triangle offsets in bounds */
    tm.report(mols[0].G, n);
    return 0;
}

}  // namespace qs

using namespace qs;

namespace {

Engine g_eng;

/* host-side result of API_computeSES (CudaSurf.cu:778-785) */
float *g_vertices = nullptr;
int *g_triangles = nullptr;
int *g_atom_ids = nullptr;
unsigned g_ntri_idx = 0, g_nvert = 0;

/* Host result arrays.  Arrays of PIN_MIN bytes or more are pinned (hipHostMalloc), so the
   device-to-host copies are queued at once and run at the link's speed; smaller ones (and every
   array of an empty mesh, which must not touch the device) are ordinary memory.  API_freeMesh
   keeps the result's arrays (the largest set so far) for the next call: pinning, like first-touch
   page faults in fresh pageable memory, costs more than the copy itself. */
struct HostArray {
    void *p = nullptr;
    size_t cap = 0;
    bool pinned = false;
};
HostArray g_cur[3], g_spare[3]; /* vertices, triangles, atom ids */

constexpr size_t PIN_MIN = 4096;

HostArray host_alloc(size_t bytes)
{
    HostArray a;
    if (bytes >= PIN_MIN && hipHostMalloc(&a.p, bytes, hipHostMallocDefault) == hipSuccess) {
        a.pinned = true;
    } else {
        (void)hipGetLastError();
        a.p = malloc(bytes);
    }
    a.cap = a.p ? bytes : 0;
    return a;
}

void host_free(HostArray &a)
{
    if (a.p) {
        if (a.pinned) (void)hipHostFree(a.p);
        else free(a.p);
    }
    a = HostArray{};
}

void *host_array(int k, size_t bytes)
{
    HostArray &s = g_spare[k];
    if (s.p && s.cap >= bytes) {
        g_cur[k] = s;
        s = HostArray{};
    } else {
        g_cur[k] = host_alloc(bytes);
    }
    return g_cur[k].p;
}

/* the last batch's meshes in library-owned host arrays (qs_batch_fetch_host), grow-only */
HostArray g_batch_host[3];

/* pipelined host fetch (qs_batch_fetch_host_async / qs_batch_wait_host): two slots, each a
   device staging copy of one batch's meshes and their pinned host arrays.  The staging copies run
   on the library stream after the batch; the device-to-host copies run on a copy stream (default
   priority), so that they overlap the next batch's work. */
struct HostSlot {
    DevBuf v, t, a;
    HostArray hv, ht, ha;
    hipEvent_t staged = nullptr, done = nullptr;
    uint32_t nv = 0, nt = 0;
    bool pending = false, ready = false;
};
HostSlot g_slot[2];
hipStream_t g_copy = nullptr;
int g_copy_dev = -1;

/* pinned host array of any size (ordinary memory only if pinning fails): a device-to-host copy
   into pageable memory would run synchronously, and the pipelined fetch must return at once */
HostArray host_alloc_pinned(size_t bytes)
{
    HostArray a;
    if (hipHostMalloc(&a.p, bytes, hipHostMallocDefault) == hipSuccess) {
        a.pinned = true;
    } else {
        (void)hipGetLastError();
        a.p = malloc(bytes);
    }
    a.cap = a.p ? bytes : 0;
    return a;
}

void grow_host(HostArray &a, size_t bytes)
{
    if (a.cap < bytes) {
        host_free(a);
        a = host_alloc_pinned(bytes);
    }
}

/* the pipelined fetch's copy stream, events and slot buffers for the current device: when the
   device changed, the old slots' copies are waited for and their stream, events and device
   buffers released on the device that owns them */
int copy_slots_for(int dev)
{
    if (g_copy && g_copy_dev == dev) return 0;
    if (g_copy) {
        int cur = 0;
        QS_HIP(hipGetDevice(&cur));
        QS_HIP(hipSetDevice(g_copy_dev));
        for (HostSlot &S : g_slot) {
            if (S.pending && S.done) (void)hipEventSynchronize(S.done);
            S.pending = S.ready = false;
            if (S.staged) (void)hipEventDestroy(S.staged);
            if (S.done) (void)hipEventDestroy(S.done);
            S.staged = S.done = nullptr;
            release(S.v);
            release(S.t);
            release(S.a);
        }
        (void)hipStreamDestroy(g_copy);
        g_copy = nullptr;
        g_copy_dev = -1;
        QS_HIP(hipSetDevice(cur));
    }
    QS_HIP(hipStreamCreateWithFlags(&g_copy, hipStreamNonBlocking));
    for (HostSlot &S : g_slot) {
        QS_HIP(hipEventCreateWithFlags(&S.staged, hipEventDisableTiming));
        QS_HIP(hipEventCreateWithFlags(&S.done, hipEventDisableTiming));
    }
    g_copy_dev = dev;
    return 0;
}

void *batch_host_array(int k, size_t bytes)
{
    HostArray &a = g_batch_host[k];
    if (a.cap < bytes) {
        host_free(a);
        a = host_alloc_pinned(bytes);
    }
    return a.p;
}

/* one-molecule batch with a given grid (stage-wise entry points) */
int single_grid(Engine &E, const QsGrid &G, size_t &at_mc, uint32_t &nb_mc, size_t &at_wpre)
{
    E.mols.assign(1, QsMol{});
    QsMol &M = E.mols[0];
    M.G = G;
    M.W = (uint32_t)((G.g + 63) / 64);
    M.rows_per_wave = qs_pass1_rows_per_wave(G);
    M.slab_rows = qs_pass1_slab_rows(G, (int)M.W);
    full_planes(M);
    E.n_atoms_total = 0;
    E.ncell_total = (uint64_t)G.gn * G.gn * G.gn;
    E.nw_total = (uint64_t)G.g * G.g * M.W;
    E.vox_total = (uint64_t)G.g * G.g * G.g;
    E.scan_w0 = 0;
    E.scan_nw = E.nw_total;
    E.multi_slab = false;
    E.batch_valid = false;
    E.vert_off.assign(2, 0);
    E.tri_off.assign(2, 0);
    size_t at_atoms, at_p1;
    uint32_t nb_atoms, nb_p1;
    QS_TRY(E.build_meta(at_atoms, at_p1, at_mc, at_wpre, nb_atoms, nb_p1, nb_mc));
    QS_TRY(ensure(E.d_mols, sizeof(QsMol)));
    QS_TRY(ensure(E.d_meta, E.h_meta.size() * 4));
    QS_HIP(hipMemcpyAsync(E.d_mols.p, E.mols.data(), sizeof(QsMol), hipMemcpyHostToDevice, E.stream));
    QS_HIP(hipMemcpyAsync(E.d_meta.p, E.h_meta.data(), E.h_meta.size() * 4, hipMemcpyHostToDevice, E.stream));
    return 0;
}

int check_batch_args(const unsigned *atom_off, int nmol, const float *reso, std::vector<MolIn> &in)
{
    if (nmol < 0 || (nmol > 0 && (!atom_off || !reso))) return set_error(3, "bad batch description");
    in.resize(nmol);
    for (int m = 0; m < nmol; m++) {
        if (atom_off[m + 1] < atom_off[m]) return set_error(3, "atom offsets must be non-decreasing");
        in[m].atom_off = atom_off[m];
        in[m].n = atom_off[m + 1] - atom_off[m];
        in[m].reso = reso[m];
    }
    if (nmol && atom_off[0] != 0) return set_error(3, "atom offsets must start at 0");
    if (nmol && atom_off[nmol] > (1u << 30)) return set_error(3, "too many atoms (%u)", atom_off[nmol]);
    return 0;
}

void clear_host_result()
{
    for (int k = 0; k < 3; k++) {
        HostArray &c = g_cur[k], &s = g_spare[k];
        if (c.p && c.cap > s.cap) {
            host_free(s);
            s = c;
            c = HostArray{};
        } else {
            host_free(c);
        }
    }
    g_vertices = nullptr;
    g_triangles = nullptr;
    g_atom_ids = nullptr;
    g_nvert = g_ntri_idx = 0;
}

}  // namespace

/* ------------------------------------------------------------------------------------------ */
/* The fixed device region of qs_compute_device_noalloc.  Every device the library runs on holds
 * one until the process ends.  In host code `scratch` only names the symbol: its device address,
 * looked up once per device with hipGetSymbolAddress, is what buffers are carved from. */
#define QS_SCRATCH_MB 256
static __device__ __align__(256) unsigned char scratch[(size_t)QS_SCRATCH_MB << 20];

namespace {

constexpr int QS_MAX_DEV = 16;
Engine g_eng_na[QS_MAX_DEV];        /* the no-allocation entry point's engine per device */
void *g_scratch_ptr[QS_MAX_DEV];    /* device address of `scratch` on each device */

}  // namespace

/* ========================================================================================== */
/* C ABI                                                                                        */

extern "C" {

int API_lastError(void) { return g_err; }
const char *qs_last_error_message(void) { return g_errmsg; }

int qs_load_tuning(const char *path)
{
    g_err = 0;
    if (!path) {
        qs_set_tuned_rows(0.0f, 0);
        return 0;
    }
    const int n = load_tuning_file(path);
    if (n < 0) return -set_error(3, "tuning file %s cannot be read", path);
    return n;
}

int qs_set_tuning(const char *key, float reso, int value)
{
    std::string k = key ? key : "";
    if (k == "pass1_slab_rows") return qs_set_tuned_rows(reso, value);
    return -1;
}

int qs_set_option(const char *key, int value)
{
    std::string k = key ? key : "";
    if (k == "lim_mode") g_opt.lim_mode = value;
    else if (k == "atom_ids_sorted") g_opt.atom_ids_sorted = value;
    else if (k == "pass1_direct") g_opt.pass1_direct = value;
    else if (k == "pass1_rows") g_opt.pass1_rows = value;
    else if (k == "pass2_direct") g_opt.pass2_direct = value;
    else if (k == "trace") g_opt.trace = value;
    else if (k == "timing") g_opt.timing = value;
    else if (k == "slab_planes") g_opt.slab_planes = value < 0 ? 0 : value;
    else if (k == "slab_mb") g_opt.slab_mb = value < 0 ? 0 : value;
    else if (k == "persistent_blocks") qs_set_persistent_blocks((uint32_t)std::max(value, 8));
    else return -1;
    return 0;
}

int qs_get_option(const char *key)
{
    std::string k = key ? key : "";
    if (k == "lim_mode") return g_opt.lim_mode;
    if (k == "atom_ids_sorted") return g_opt.atom_ids_sorted;
    if (k == "pass1_direct") return g_opt.pass1_direct;
    if (k == "pass1_rows") return g_opt.pass1_rows;
    if (k == "pass2_direct") return g_opt.pass2_direct;
    if (k == "trace") return g_opt.trace;
    if (k == "timing") return g_opt.timing;
    if (k == "slab_planes") return g_opt.slab_planes;
    if (k == "slab_mb") return g_opt.slab_mb;
    if (k == "persistent_blocks") return (int)qs_persistent_blocks();
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
    /* results of a previous call that were not freed are released (the reference leaks them) */
    clear_host_result();
    auto fail = [&]() {
        clear_host_result();
        g_vertices = (float *)host_array(0, 12);
        g_triangles = (int *)host_array(1, 12);
        g_atom_ids = (int *)host_array(2, 4);
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
    if (N > (1u << 30)) {
        set_error(3, "too many atoms (%u)", N);
        fail();
        return;
    }
    const float *hp = (const float *)atomPos;
    std::vector<BBox> bb(1);
    int nonfinite = 0;
    bbox_host(hp, atomRad, N, bb[0], nonfinite);
    if (nonfinite) {
        set_error(3, "atom positions or radii contain NaN or Inf");
        fail();
        return;
    }
    Engine &E = g_eng;
    int rc = E.init();
    if (!rc) rc = ensure(E.in_pos, (size_t)N * 12);
    if (!rc) rc = ensure(E.in_rad, (size_t)N * 4);
    if (!rc) rc = E.upload_begin();
    if (!rc) rc = E.upload_input(E.in_pos.p, hp, (size_t)N * 12);
    if (!rc) rc = E.upload_input(E.in_rad.p, atomRad, (size_t)N * 4);
    std::vector<MolIn> in(1);
    in[0].atom_off = 0;
    in[0].n = N;
    in[0].reso = resoSES;
    if (!rc) rc = E.run_batch(as<float>(E.in_pos), as<float>(E.in_rad), in, doSmoothing, &bb);
    if (rc) {
        fail();
        return;
    }
    g_nvert = E.nv;
    g_ntri_idx = 3 * E.nt;
    g_vertices = (float *)host_array(0, (size_t)std::max<uint32_t>(E.nv, 1) * 12);
    g_triangles = (int *)host_array(1, (size_t)std::max<uint32_t>(E.nt, 1) * 12);
    g_atom_ids = (int *)host_array(2, (size_t)std::max<uint32_t>(E.nv, 1) * 4);
    if (!g_vertices || !g_triangles || !g_atom_ids) {
        set_error(1, "out of host memory");
        fail();
        return;
    }
    bool ok = true;
    if (E.nv) {
        ok = hipMemcpyAsync(g_vertices, E.res_vert, (size_t)E.nv * 12, hipMemcpyDeviceToHost, E.stream) == hipSuccess &&
             hipMemcpyAsync(g_atom_ids, E.res_atom, (size_t)E.nv * 4, hipMemcpyDeviceToHost, E.stream) == hipSuccess;
    }
    if (ok && E.nt)
        ok = hipMemcpyAsync(g_triangles, E.res_tri, (size_t)E.nt * 12, hipMemcpyDeviceToHost, E.stream) == hipSuccess;
    if (ok) ok = hipStreamSynchronize(E.stream) == hipSuccess;
    if (!ok) {
        set_error(2, "D2H copy failed");
        fail();
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

void API_freeMesh(void) { clear_host_result(); }

int qs_compute_device(const float *d_pos, const float *d_rad, unsigned int N, float resoSES, int doSmoothing,
                      unsigned int *nvert, unsigned int *nidx, float **d_vert, int **d_tri, int **d_atom_id)
{
    g_err = 0;
    g_errmsg[0] = 0;
    if (nvert) *nvert = 0;
    if (nidx) *nidx = 0;
    if (N > (1u << 30)) return set_error(3, "too many atoms (%u)", N);
    QS_TRY(g_eng.init());
    std::vector<MolIn> in(1);
    in[0].atom_off = 0;
    in[0].n = N;
    in[0].reso = resoSES;
    QS_TRY(g_eng.run_batch(d_pos, d_rad, in, doSmoothing, nullptr));
    if (nvert) *nvert = g_eng.nv;
    if (nidx) *nidx = 3 * g_eng.nt;
    if (d_vert) *d_vert = g_eng.res_vert;
    if (d_tri) *d_tri = g_eng.res_tri;
    if (d_atom_id) *d_atom_id = g_eng.res_atom;
    return 0;
}

int qs_compute_device_noalloc(const float *d_pos, const float *d_rad, unsigned int N, float resoSES, int doSmoothing,
                              float *d_vert, unsigned int vert_cap, int *d_tri, unsigned int tri_cap, int *d_atom_id,
                              unsigned int *nvert, unsigned int *nidx)
{
    g_err = 0;
    g_errmsg[0] = 0;
    if (nvert) *nvert = 0;
    if (nidx) *nidx = 0;
    if (N > (1u << 30)) return set_error(3, "too many atoms (%u)", N);
    int dev = 0;
    QS_HIP(hipGetDevice(&dev));
    if (dev < 0 || dev >= QS_MAX_DEV) return set_error(3, "device %d: the fixed scratch region serves devices 0-%d", dev,
                                                       QS_MAX_DEV - 1);
    Engine &E = g_eng_na[dev];
    /* set-up, once per device with that device current: the device address of its region, the
       engine's streams and events (no memory is allocated) */
    if (!g_scratch_ptr[dev]) {
        void *p = nullptr;
        QS_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        if (!p) return set_error(2, "no device address for the fixed scratch region");
        g_scratch_ptr[dev] = p;
    }
    QS_TRY(E.init());
    Arena arena;
    arena.base = (char *)g_scratch_ptr[dev];
    arena.cap = (size_t)QS_SCRATCH_MB << 20;
    E.forget_buffers();
    struct ArenaScope {
        explicit ArenaScope(Arena *a) { g_arena = a; }
        ~ArenaScope() { g_arena = nullptr; }
    } scope(&arena);
    std::vector<MolIn> in(1);
    in[0].atom_off = 0;
    in[0].n = N;
    in[0].reso = resoSES;
    QS_TRY(E.run_batch(d_pos, d_rad, in, doSmoothing, nullptr));
    if (E.nv > vert_cap || E.nt > tri_cap)
        return set_error(5, "output capacity too small: %u vertices and %u triangles (capacity %u and %u)", E.nv, E.nt,
                         vert_cap, tri_cap);
    if ((E.nv && (!d_vert || !d_atom_id)) || (E.nt && !d_tri)) return set_error(3, "output pointer is NULL");
    if (E.nv) {
        QS_HIP(hipMemcpyAsync(d_vert, E.res_vert, (size_t)E.nv * 12, hipMemcpyDeviceToDevice, E.stream));
        QS_HIP(hipMemcpyAsync(d_atom_id, E.res_atom, (size_t)E.nv * 4, hipMemcpyDeviceToDevice, E.stream));
    }
    if (E.nt) QS_HIP(hipMemcpyAsync(d_tri, E.res_tri, (size_t)E.nt * 12, hipMemcpyDeviceToDevice, E.stream));
    QS_HIP(hipStreamSynchronize(E.stream));
    if (nvert) *nvert = E.nv;
    if (nidx) *nidx = 3 * E.nt;
    return 0;
}

int qs_compute_batch_device(const float *d_pos, const float *d_rad, const unsigned int *atom_off, int nmol,
                            const float *reso, int doSmoothing, unsigned int *vert_off, unsigned int *tri_off,
                            float **d_vert, int **d_tri, int **d_atom_id)
{
    g_err = 0;
    g_errmsg[0] = 0;
    std::vector<MolIn> in;
    QS_TRY(check_batch_args(atom_off, nmol, reso, in));
    QS_TRY(g_eng.init());
    QS_TRY(g_eng.run_batch(d_pos, d_rad, in, doSmoothing, nullptr));
    for (int m = 0; m <= nmol; m++) {
        if (vert_off) vert_off[m] = g_eng.vert_off[m];
        if (tri_off) tri_off[m] = g_eng.tri_off[m];
    }
    if (d_vert) *d_vert = g_eng.res_vert;
    if (d_tri) *d_tri = g_eng.res_tri;
    if (d_atom_id) *d_atom_id = g_eng.res_atom;
    return 0;
}

int qs_compute_batch(const float *pos, const float *rad, const unsigned int *atom_off, int nmol, const float *reso,
                     int doSmoothing, unsigned int *nvert, unsigned int *ntri)
{
    g_err = 0;
    g_errmsg[0] = 0;
    std::vector<MolIn> in;
    QS_TRY(check_batch_args(atom_off, nmol, reso, in));
    const unsigned n = nmol ? atom_off[nmol] : 0;
    if (n && (!pos || !rad)) return set_error(3, "null atom arrays");
    std::vector<BBox> bb(nmol);
    for (int m = 0; m < nmol; m++) {
        int nonfinite = 0;
        bbox_host(pos + 3 * (size_t)in[m].atom_off, rad + in[m].atom_off, in[m].n, bb[m], nonfinite);
        if (in[m].n > 1 && nonfinite) return set_error(3, "molecule %d: atom positions or radii contain NaN or Inf", m);
    }
    Engine &E = g_eng;
    QS_TRY(E.init());
    QS_TRY(ensure(E.in_pos, (size_t)std::max(n, 1u) * 12));
    QS_TRY(ensure(E.in_rad, (size_t)std::max(n, 1u) * 4));
    QS_TRY(E.upload_begin());
    QS_TRY(E.upload_input(E.in_pos.p, pos, (size_t)n * 12));
    QS_TRY(E.upload_input(E.in_rad.p, rad, (size_t)n * 4));
    QS_TRY(E.run_batch(as<float>(E.in_pos), as<float>(E.in_rad), in, doSmoothing, &bb));
    for (int m = 0; m < nmol; m++) {
        if (nvert) nvert[m] = E.vert_off[m + 1] - E.vert_off[m];
        if (ntri) ntri[m] = E.tri_off[m + 1] - E.tri_off[m];
    }
    return 0;
}

int qs_batch_offsets(unsigned int *vert_off, unsigned int *tri_off)
{
    Engine &E = g_eng;
    if (!E.batch_valid) return -set_error(3, "no batch result");
    const int nmol = (int)E.mols.size();
    for (int m = 0; m <= nmol; m++) {
        if (vert_off) vert_off[m] = E.vert_off[m];
        if (tri_off) tri_off[m] = E.tri_off[m];
    }
    return nmol;
}

int qs_batch_fetch(float *vert, int *tri, int *atom_id)
{
    Engine &E = g_eng;
    if (!E.batch_valid) return set_error(3, "no batch result");
    const int nmol = (int)E.mols.size();
    if (E.nv && vert) QS_HIP(hipMemcpy(vert, E.res_vert, (size_t)E.nv * 12, hipMemcpyDeviceToHost));
    if (E.nv && atom_id) QS_HIP(hipMemcpy(atom_id, E.res_atom, (size_t)E.nv * 4, hipMemcpyDeviceToHost));
    if (E.nt && tri) {
        QS_TRY(ensure(E.tri_loc, (size_t)E.nt * 12));
        qs_launch_tri_localize(E.res_tri, E.nt, as<uint32_t>(E.bounds), as<uint32_t>(E.d_off) + (nmol + 1), nmol,
                               as<int>(E.tri_loc), E.stream);
        QS_HIP(hipGetLastError());
        QS_HIP(hipMemcpyAsync(tri, E.tri_loc.p, (size_t)E.nt * 12, hipMemcpyDeviceToHost, E.stream));
        QS_HIP(hipStreamSynchronize(E.stream));
    }
    return 0;
}

int qs_batch_fetch_host(const float **vert, const int **tri, const int **atom_id)
{
    Engine &E = g_eng;
    if (!E.batch_valid) return set_error(3, "no batch result");
    const int nmol = (int)E.mols.size();
    float *hv = (float *)batch_host_array(0, (size_t)std::max<uint32_t>(E.nv, 1) * 12);
    int *ht = (int *)batch_host_array(1, (size_t)std::max<uint32_t>(E.nt, 1) * 12);
    int *ha = (int *)batch_host_array(2, (size_t)std::max<uint32_t>(E.nv, 1) * 4);
    if (!hv || !ht || !ha) return set_error(1, "out of host memory");
    if (E.nv) {
        QS_HIP(hipMemcpyAsync(hv, E.res_vert, (size_t)E.nv * 12, hipMemcpyDeviceToHost, E.stream));
        QS_HIP(hipMemcpyAsync(ha, E.res_atom, (size_t)E.nv * 4, hipMemcpyDeviceToHost, E.stream));
    }
    if (E.nt) {
        QS_TRY(ensure(E.tri_loc, (size_t)E.nt * 12));
        qs_launch_tri_localize(E.res_tri, E.nt, as<uint32_t>(E.bounds), as<uint32_t>(E.d_off) + (nmol + 1), nmol,
                               as<int>(E.tri_loc), E.stream);
        QS_HIP(hipGetLastError());
        QS_HIP(hipMemcpyAsync(ht, E.tri_loc.p, (size_t)E.nt * 12, hipMemcpyDeviceToHost, E.stream));
    }
    QS_HIP(hipStreamSynchronize(E.stream));
    if (vert) *vert = hv;
    if (tri) *tri = ht;
    if (atom_id) *atom_id = ha;
    return 0;
}

int qs_batch_fetch_host_async(int slot)
{
    Engine &E = g_eng;
    if (!E.batch_valid) return set_error(3, "no batch result");
    if (slot < 0 || slot > 1) return set_error(3, "slot must be 0 or 1");
    QS_TRY(copy_slots_for(E.device));
    HostSlot &S = g_slot[slot];
    if (S.pending) QS_HIP(hipEventSynchronize(S.done)); /* its previous copies still read the slot */
    S.pending = S.ready = false;
    const int nmol = (int)E.mols.size();
    const size_t nv = E.nv, nt = E.nt;
    QS_TRY(ensure(S.v, std::max<size_t>(nv, 1) * 12));
    QS_TRY(ensure(S.a, std::max<size_t>(nv, 1) * 4));
    QS_TRY(ensure(S.t, std::max<size_t>(nt, 1) * 12));
    grow_host(S.hv, std::max<size_t>(nv, 1) * 12);
    grow_host(S.ha, std::max<size_t>(nv, 1) * 4);
    grow_host(S.ht, std::max<size_t>(nt, 1) * 12);
    if (!S.hv.p || !S.ha.p || !S.ht.p) return set_error(1, "out of host memory");
    if (nv) {
        QS_HIP(hipMemcpyAsync(S.v.p, E.res_vert, nv * 12, hipMemcpyDeviceToDevice, E.stream));
        QS_HIP(hipMemcpyAsync(S.a.p, E.res_atom, nv * 4, hipMemcpyDeviceToDevice, E.stream));
    }
    if (nt) {
        qs_launch_tri_localize(E.res_tri, (uint32_t)nt, as<uint32_t>(E.bounds), as<uint32_t>(E.d_off) + (nmol + 1), nmol,
                               as<int>(S.t), E.stream);
        QS_HIP(hipGetLastError());
    }
    QS_HIP(hipEventRecord(S.staged, E.stream));
    QS_HIP(hipStreamWaitEvent(g_copy, S.staged, 0));
    if (nv) {
        QS_HIP(hipMemcpyAsync(S.hv.p, S.v.p, nv * 12, hipMemcpyDeviceToHost, g_copy));
        QS_HIP(hipMemcpyAsync(S.ha.p, S.a.p, nv * 4, hipMemcpyDeviceToHost, g_copy));
    }
    if (nt) QS_HIP(hipMemcpyAsync(S.ht.p, S.t.p, nt * 12, hipMemcpyDeviceToHost, g_copy));
    QS_HIP(hipEventRecord(S.done, g_copy));
    S.nv = (uint32_t)nv;
    S.nt = (uint32_t)nt;
    S.pending = true;
    return 0;
}

int qs_batch_wait_host(int slot, const float **vert, const int **tri, const int **atom_id, unsigned int *nvert,
                       unsigned int *ntri)
{
    if (slot < 0 || slot > 1) return set_error(3, "slot must be 0 or 1");
    HostSlot &S = g_slot[slot];
    if (S.pending) {
        QS_HIP(hipEventSynchronize(S.done));
        S.pending = false;
        S.ready = true;
    }
    if (!S.ready) return set_error(3, "no fetch started on slot %d", slot);
    if (vert) *vert = (const float *)S.hv.p;
    if (tri) *tri = (const int *)S.ht.p;
    if (atom_id) *atom_id = (const int *)S.ha.p;
    if (nvert) *nvert = S.nv;
    if (ntri) *ntri = S.nt;
    return 0;
}

int qs_batch_mesh(int m, float *vert, int *tri, int *atom_id)
{
    Engine &E = g_eng;
    if (!E.batch_valid) return set_error(3, "no batch result");
    if (m < 0 || m + 1 >= (int)E.vert_off.size() || m + 1 >= (int)E.tri_off.size()) return set_error(3, "no molecule %d", m);
    const uint32_t v0 = E.vert_off[m], nvm = E.vert_off[m + 1] - v0;
    const uint32_t t0 = E.tri_off[m], ntm = E.tri_off[m + 1] - t0;
    if (nvm && vert) QS_HIP(hipMemcpy(vert, E.res_vert + 3 * (size_t)v0, (size_t)nvm * 12, hipMemcpyDeviceToHost));
    if (nvm && atom_id) QS_HIP(hipMemcpy(atom_id, E.res_atom + v0, (size_t)nvm * 4, hipMemcpyDeviceToHost));
    if (ntm && tri) {
        QS_HIP(hipMemcpy(tri, E.res_tri + 3 * (size_t)t0, (size_t)ntm * 12, hipMemcpyDeviceToHost));
        for (size_t i = 0; i < 3 * (size_t)ntm; i++) tri[i] -= (int)v0;
    }
    return 0;
}

/* normals of the last computation's vertices (batch order), on the library's stream */
static int launch_normals(Engine &E, float *d_out)
{
    if (!E.batch_valid) return set_error(3, "no mesh to take normals of");
    if (E.multi_slab) return set_error(3, "normals: the last mesh was built in x-slabs; its field is not kept");
    const int nmol = (int)E.mols.size();
    const uint32_t *d_vpre = as<uint32_t>(E.d_off) + (nmol + 1);
    qs_launch_vertex_normals(as<QsMol>(E.d_mols), d_vpre, d_vpre + nmol + 1, as<uint64_t>(E.vkey), E.nv, as<float>(E.field),
                             E.field_sparse ? as<uint32_t>(E.spre) : nullptr, as<uint64_t>(E.outm),
                             E.field_sparse ? as<uint64_t>(E.shellm) : nullptr, d_out, E.stream);
    QS_HIP(hipGetLastError());
    return 0;
}

int qs_vertex_normals_device(float *d_out, float *ms)
{
    g_err = 0;
    Engine &E = g_eng;
    if (!E.batch_valid) return set_error(3, "no mesh to take normals of");
    if (!d_out && E.nv) return set_error(3, "normals: output pointer is NULL");
    hipEvent_t a, b;
    QS_HIP(hipEventCreate(&a));
    QS_HIP(hipEventCreate(&b));
    QS_HIP(hipEventRecord(a, E.stream));
    const int rc = launch_normals(E, d_out);
    QS_HIP(hipEventRecord(b, E.stream));
    QS_HIP(hipEventSynchronize(b));
    float t = 0.0f;
    QS_HIP(hipEventElapsedTime(&t, a, b));
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
    if (ms) *ms = t;
    return rc;
}

int qs_vertex_normals(float *out, unsigned int nvert)
{
    g_err = 0;
    Engine &E = g_eng;
    if (!E.batch_valid) return set_error(3, "no mesh to take normals of");
    if (nvert != E.nv) return set_error(3, "normals: %u vertices asked, the last mesh has %u", nvert, E.nv);
    if (!E.nv) return 0;
    if (!out) return set_error(3, "normals: output pointer is NULL");
    QS_TRY(ensure(E.normals, (size_t)E.nv * 12));
    QS_TRY(launch_normals(E, as<float>(E.normals)));
    QS_HIP(hipMemcpyAsync(out, E.normals.p, (size_t)E.nv * 12, hipMemcpyDeviceToHost, E.stream));
    QS_HIP(hipStreamSynchronize(E.stream));
    return 0;
}

int qs_vertex_colors(const float *atom_rgba, unsigned int natoms, float *out, unsigned int nvert)
{
    g_err = 0;
    Engine &E = g_eng;
    if (!E.batch_valid) return set_error(3, "no mesh to colour");
    if (nvert != E.nv) return set_error(3, "colours: %u vertices asked, the last mesh has %u", nvert, E.nv);
    if (natoms != E.n_atoms_total) return set_error(3, "colours: %u atom colours given, the last mesh has %u atoms", natoms,
                                                    E.n_atoms_total);
    if (!E.nv) return 0;
    if (!atom_rgba || !out) return set_error(3, "colours: NULL array");
    QS_TRY(ensure(E.normals, (size_t)E.nv * 16));
    QS_TRY(ensure(E.colors_in, (size_t)std::max(natoms, 1u) * 16));
    QS_HIP(hipMemcpyAsync(E.colors_in.p, atom_rgba, (size_t)natoms * 16, hipMemcpyHostToDevice, E.stream));
    const int nmol = (int)E.mols.size();
    const uint32_t *d_vpre = as<uint32_t>(E.d_off) + (nmol + 1);
    qs_launch_vertex_colors(as<QsMol>(E.d_mols), d_vpre, d_vpre + nmol + 1, E.res_atom, E.ids_sorted ? as<int>(E.perm) : nullptr,
                            E.nv, as<float4>(E.colors_in), as<float4>(E.normals), E.stream);
    QS_HIP(hipGetLastError());
    QS_HIP(hipMemcpyAsync(out, E.normals.p, (size_t)E.nv * 16, hipMemcpyDeviceToHost, E.stream));
    QS_HIP(hipStreamSynchronize(E.stream));
    return 0;
}

int qs_copy_device(void *dst, const void *src, uint64_t bytes, float *ms)
{
    g_err = 0;
    if (bytes % 16 || ((uintptr_t)dst | (uintptr_t)src) % 16) return set_error(3, "copy: 16-byte sizes and addresses");
    QS_TRY(g_eng.init());
    hipEvent_t a, b;
    QS_HIP(hipEventCreate(&a));
    QS_HIP(hipEventCreate(&b));
    QS_HIP(hipEventRecord(a, g_eng.stream));
    qs_launch_copy(src, dst, bytes, g_eng.stream);
    QS_HIP(hipGetLastError());
    QS_HIP(hipEventRecord(b, g_eng.stream));
    QS_HIP(hipEventSynchronize(b));
    float t = 0.0f;
    QS_HIP(hipEventElapsedTime(&t, a, b));
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
    if (ms) *ms = t;
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

int qs_batch_grid(int m, float *fp, int *ip)
{
    if (m < 0 || m >= (int)g_eng.mols.size()) return -1;
    const QsGrid &G = g_eng.mols[m].G;
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

int qs_last_grid(float *fp, int *ip) { return qs_batch_grid(0, fp, ip); }

int qs_last_slab_count(void) { return g_eng.batch_valid ? g_eng.n_slabs : 0; }

int64_t qs_fetch(const char *what, void *host, int64_t bytes)
{
    Engine &E = g_eng;
    std::string w = what ? what : "";
    if (E.mols.empty()) return -1;
    const QsGrid &G = E.mols[0].G;
    const uint64_t V = E.vox_total;
    const void *src = nullptr;
    int64_t need = 0;
    const bool grid_kept = !E.multi_slab; /* a slab run keeps only its last slab's grid arrays */
    if (!grid_kept && (w == "labels" || w == "field" || w == "shellm")) return -1;
    if (w == "labels") {
        /* molecule 0 of the last run, reconstructed from the outside / shell masks */
        const uint64_t V0 = (uint64_t)G.g * G.g * G.g;
        if (bytes < (int64_t)V0 || !E.outm.p) return -1;
        const int g = G.g, W = (g + 63) / 64;
        size_t nw = (size_t)g * g * W;
        std::vector<uint64_t> mo(nw), ms(nw);
        if (nw && hipMemcpy(mo.data(), E.outm.p, nw * 8, hipMemcpyDeviceToHost) != hipSuccess) return -1;
        if (nw && hipMemcpy(ms.data(), E.shellm.p, nw * 8, hipMemcpyDeviceToHost) != hipSuccess) return -1;
        uint8_t *out = (uint8_t *)host;
        for (size_t row = 0; row < (size_t)g * g; row++)
            for (int z = 0; z < g; z++) {
                uint64_t bit = 1ull << (z & 63);
                size_t wi = row * W + (z >> 6);
                out[row * g + z] = (mo[wi] & bit) ? QS_OUT : ((ms[wi] & bit) ? QS_SHELL : QS_IN);
            }
        return (int64_t)V0;
    }
    else if (w == "field") {
        if (E.fill_field(&src) != 0) return -1;
        if (hipStreamSynchronize(E.stream) != hipSuccess) return -1;
        need = (int64_t)V * 4;
    }
    else if (w == "shellm") { src = E.shellm.p; need = (int64_t)E.nw_total * 8; }
    else if (w == "perm") { src = E.perm.p; need = (int64_t)E.n_atoms_total * 4; }
    else if (w == "cells") { src = E.cse.p; need = (int64_t)E.ncell_total * 8; }
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
    if (G.g < 2 || (uint64_t)G.g * G.g * G.g >= (1ull << 32)) return set_error(3, "bad grid");
    size_t at_mc, at_wpre;
    uint32_t nb_mc;
    QS_TRY(single_grid(E, G, at_mc, nb_mc, at_wpre));
    const uint64_t V = E.vox_total;
    const int W = (int)E.mols[0].W;
    QS_TRY(ensure(E.labels, V));
    QS_TRY(ensure(E.outm, E.nw_total * 8));
    QS_TRY(ensure(E.shellm, E.nw_total * 8));
    QS_HIP(hipMemcpyAsync(E.labels.p, labels, V, hipMemcpyHostToDevice, E.stream));
    qs_launch_labels_to_masks(as<uint8_t>(E.labels), G.g, W, as<uint64_t>(E.outm), as<uint64_t>(E.shellm), E.stream);
    QS_HIP(hipGetLastError());
    /* pass 1 is skipped: the labels stand in for it */
    QS_TRY(ensure(E.field, V * 4));
    QS_TRY(ensure(E.negm, E.nw_total * 8));
    QS_TRY(ensure(E.zerom, E.nw_total * 8));
    QS_TRY(ensure(E.spre, E.nw_total * 4));
    QS_TRY(ensure(E.stot, 8));
    QS_TRY(ensure(E.scan_tmp, qs_scan_tiles(E.nw_total) * sizeof(uint32_t)));
    qs_launch_shell_scan(as<uint64_t>(E.shellm), E.nw_total, as<uint32_t>(E.scan_tmp), as<uint32_t>(E.stot), as<uint32_t>(E.spre),
                         E.stream);
    for (auto &grp : E.groups) {
        QS_TRY(E.group_tables(grp));
        qs_launch_pass2(E.batch(grp.bpre_at, grp.mid_at, (int)grp.mols.size()), grp.nblocks, grp.r, grp.path,
                        as<uint64_t>(E.outm), as<uint64_t>(E.shellm), as<uint32_t>(E.spre), as<float>(E.field), as<uint32_t>(E.negm),
                        as<uint32_t>(E.zerom), as<int>(grp.tab->offs), as<int>(grp.tab->norm), as<int>(grp.tab->end),
                        grp.tab->flags, nullptr, 0, grp.xchunk, E.stream);
        QS_HIP(hipGetLastError());
    }
    E.field_sparse = true;
    const void *dense = nullptr;
    QS_TRY(E.fill_field(&dense));
    QS_HIP(hipStreamSynchronize(E.stream));
    QS_HIP(hipMemcpy(field, dense, V * 4, hipMemcpyDeviceToHost));
    return 0;
}

int qs_stage_mesh(const float *fp, const int *ip, const float *field, unsigned int *nvert, unsigned int *ntri, float *vert,
                  int64_t vert_cap, int *tri, int64_t tri_cap)
{
    g_err = 0;
    Engine &E = g_eng;
    QS_TRY(E.init());
    QsGrid G;
    grid_from_params(fp, ip, G);
    if (G.g < 2 || (uint64_t)G.g * G.g * G.g >= (1ull << 32)) return set_error(3, "bad grid");
    size_t at_mc, at_wpre;
    uint32_t nb_mc;
    QS_TRY(single_grid(E, G, at_mc, nb_mc, at_wpre));
    const uint64_t V = E.vox_total;
    QS_TRY(ensure(E.field, V * 4));
    QS_HIP(hipMemcpyAsync(E.field.p, field, V * 4, hipMemcpyHostToDevice, E.stream));
    E.field_sparse = false; /* the caller's field holds every voxel */
    const int W = (int)E.mols[0].W;
    QS_TRY(ensure(E.negm, E.nw_total * 8));
    QS_TRY(ensure(E.zerom, E.nw_total * 8));
    qs_launch_field_signs(as<float>(E.field), G.g, W, as<uint32_t>(E.negm), as<uint32_t>(E.zerom), E.stream);
    StageTimer tm;
    for (int attempt = 0;; attempt++) {
        QS_TRY(E.mesh_stage(at_mc, nb_mc, at_wpre, tm, false));
        uint32_t hb[4];
        QS_HIP(hipMemcpyAsync(hb, E.totals.p, 16, hipMemcpyDeviceToHost, E.stream));
        QS_HIP(hipStreamSynchronize(E.stream));
        E.nt_all = hb[0];
        E.nv = hb[1];
        if (!E.pools_overflowed(E.nv, E.nt_all)) break;
        if (attempt) return set_error(4, "mesh pools overflowed again");
    }
    if (nvert) *nvert = E.nv;
    if (ntri) *ntri = E.nt_all;
    if ((int64_t)E.nv * 3 > vert_cap || (int64_t)E.nt_all * 3 > tri_cap) return set_error(5, "output capacity too small");
    if (E.nv) QS_HIP(hipMemcpy(vert, E.vert.p, (size_t)E.nv * 12, hipMemcpyDeviceToHost));
    if (E.nt_all) QS_HIP(hipMemcpy(tri, E.tri.p, (size_t)E.nt_all * 12, hipMemcpyDeviceToHost));
    return 0;
}

}  // extern "C"
