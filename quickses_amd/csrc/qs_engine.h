/* Host engine of the SES path (internal to libquickses.so): error state, options, device
 * buffers, grid setup and the batched stage sequence (Engine).  Used by the C ABI
 * (qs_engine.hip) and the request server (qs_serve.hip). */
#ifndef QS_ENGINE_H
#define QS_ENGINE_H

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <cstdio>
#include <string>
#include <vector>

#include "qs_common.h"
#include "qs_kernels.h"

namespace qs {

/* ------------------------------------------------------------------------------------------ */
/* errors                                                                                       */

/* error state of the calling thread */
extern thread_local int g_err;
extern thread_local char g_errmsg[512];

int set_error(int code, const char *fmt, ...);

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
    int lim_mode = 0;
    int atom_ids_sorted = 0;
    int pass1_direct = 0;
    int pass1_rows = 0;
    int pass2_direct = 0;
    int trace = 0;
    int timing = 0;
    /* x-slabs: planes owned per slab (0: chosen from the memory budget), and that budget in MiB
       (0: half the device's free memory) */
    int slab_planes = 0;
    int slab_mb = 0;
};
extern Options g_opt;

/* ------------------------------------------------------------------------------------------ */
/* grow-only device buffers                                                                     */

struct DevBuf {
    void *p = nullptr;
    size_t cap = 0;
};

/* grow (never shrink) to at least `bytes`; contents are not kept */
/* headroom of a grown buffer: bytes >> g_grow_shift extra (3 = 1/8; 0 doubles, for threads
   that must rarely reallocate, since freeing device memory waits for the whole device) */
extern thread_local int g_grow_shift;
int ensure(DevBuf &b, size_t bytes);
void release(DevBuf &b);

/* A fixed device region that ensure() carves buffers from, in place of allocating them, while a
   thread's g_arena points at it (the no-allocation entry point).  Buffers are carved in order
   and never freed; the region is reused from its start by the next run. */
struct Arena {
    char *base = nullptr;
    size_t cap = 0, used = 0;
};
extern thread_local Arena *g_arena;

template <class T>
inline T *as(DevBuf &b) { return (T *)b.p; }

/* ------------------------------------------------------------------------------------------ */
/* grid setup: the reference's host formulas, evaluated in single precision                    */

struct BBox {
    float vmin[3], vmax[3], max_rad;
};

void bbox_host(const float *pos, const float *rad, unsigned n, BBox &b, int &nonfinite);

void grid_from_bbox(const BBox &b, float reso, QsGrid &G);

/* the tuning cache: pass-1 rows per voxel size from a JSON file (entries read, -1 unreadable),
   and the one beside the library, read once per process */
int load_tuning_file(const char *path);
void load_default_tuning();

/* ------------------------------------------------------------------------------------------ */
/* pass-2 candidate tables: window offsets sorted by squared norm, and for each minimal norm the
   end of the range of norms whose float distance can still be the minimum                      */

struct OffsetTables {
    int r = -1;
    float dx = 0, cmax = 0, bound = 0;
    std::vector<int> offs, norm_start, enum_end;
};

void build_offset_tables(int r, float dx, float coord_max, OffsetTables &T);

/* ------------------------------------------------------------------------------------------ */
/* engine                                                                                       */

/* per-stage times of the calling thread's last timed run */
extern thread_local std::vector<std::pair<std::string, float>> g_stage_ms;

struct StageTimer {
    bool on = false;
    hipStream_t s = nullptr;
    std::vector<std::pair<std::string, hipEvent_t>> ev;
    void mark(const char *name)
    {
        if (!on) return;
        hipEvent_t e;
        if (hipEventCreate(&e) != hipSuccess) return;
        (void)hipEventRecord(e, s);
        ev.emplace_back(name, e);
    }
    void report(const QsGrid &G, unsigned n)
    {
        if (!on || ev.size() < 2) return;
        (void)hipEventSynchronize(ev.back().second);
        g_stage_ms.clear();
        for (size_t i = 1; i < ev.size(); i++) {
            float ms = 0;
            (void)hipEventElapsedTime(&ms, ev[i - 1].second, ev[i].second);
            g_stage_ms.emplace_back(ev[i].first, ms);
        }
        if (!g_opt.trace) {
            for (auto &p : ev) (void)hipEventDestroy(p.second);
            ev.clear();
            return;
        }
        std::string line = "{\"qs_trace\": {\"n_atoms\": " + std::to_string(n) + ", \"g\": " + std::to_string(G.g);
        for (size_t i = 1; i < ev.size(); i++) {
            float ms = 0;
            (void)hipEventElapsedTime(&ms, ev[i - 1].second, ev[i].second);
            char buf[96];
            snprintf(buf, sizeof(buf), ", \"%s_ms\": %.4f", ev[i].first.c_str(), ms);
            line += buf;
        }
        line += "}}";
        fprintf(stderr, "%s\n", line.c_str());
        for (auto &p : ev) (void)hipEventDestroy(p.second);
        ev.clear();
    }
};

/* the whole grid as one slab: every stage over planes [0, g) */
inline void full_planes(QsMol &M)
{
    const int g = M.G.g;
    M.p1_lo = M.p2_lo = M.mc_lo = 0;
    M.p1_hi = M.p2_hi = M.mc_hi = M.cube_hi = g;
}

/* the array index of a molecule's first stored word (word_off addresses global plane 0) */
inline uint64_t phys_word_off(const QsMol &M) { return M.word_off + (uint64_t)M.p1_lo * M.G.g * M.W; }

/* host description of one molecule of a batch */
struct MolIn {
    uint32_t atom_off = 0, n = 0;
    float reso = 0.5f;
};

/* field-pass-2 candidate tables on the device, kept for reuse by later calls */
struct P2Tables {
    int r;
    float dx, cmax;
    DevBuf offs, norm, end;
    QsNormFlags flags;
};

/* molecules sharing a field-pass-2 kernel (path, r) and candidate tables (dx, coordinate range) */
struct P2Group {
    int path, r;
    float dx, cmax;
    std::vector<int> mols;
    size_t bpre_at = 0, mid_at = 0;
    uint32_t nblocks = 0;
    uint32_t xchunk = 1; /* tiles per XCD turn of the fixed-R kernel: 4 rows of z tiles */
    P2Tables *tab = nullptr;
};

struct Engine {
    int device = -1;
    hipStream_t stream = nullptr, side = nullptr;
    hipEvent_t ev_emit = nullptr, ev_side = nullptr, ev_vert = nullptr;
    bool side_pending = false; /* closest-atom work on `side` not yet joined into `stream` */
    bool vert_pending = false; /* the vertex kernel on `side` not yet joined into `stream` */

    DevBuf in_pos, in_rad, d_off, bb_part, bb_out, d_mols, d_meta;
    DevBuf key /* cell and rank within the cell per atom */, count, start, perm, cse, sx, aux;
    DevBuf labels, outm, shellm, field, negm, zerom;
    DevBuf s4, wcnt /* per row word: vertex count | triangle count << 16 */, vbw, tbw, alist, totals, bounds;
    DevBuf rback; /* the run's read-back: totals, vertex and kept-triangle offsets */
    DevBuf ntb, avb, atb; /* triangle-count bit planes per word; first vertex / triangle of each active word */
    DevBuf vstart, tstart; /* first active word of each 256-vertex / 256-triangle block */
    DevBuf normals;        /* staging for qs_vertex_normals / qs_vertex_colors */
    DevBuf colors_in;      /* per-atom colour table of qs_vertex_colors */
    DevBuf vert, vert2, vert3, vkey, tri, tri_out, atom;
    DevBuf deg_blk, keeppos, tri_loc;
    DevBuf scan_tmp;
    DevBuf inc_cnt, inc_start, inc_ent; /* per-vertex triangle incidences (smoothing of slab meshes) */
    DevBuf tpre_d;                      /* per-molecule first triangle (u64, nmol + 1), on the device */
    DevBuf tflag;                       /* field-pass-2 tiles holding a shell voxel (u8 per tile) */
    DevBuf spre, stot;                  /* shell rank of each row word (u32), shell total (2 x u32) */
    DevBuf field_dense;                 /* the whole field in grid order, expanded on request */
    std::vector<P2Group> groups;
    std::vector<P2Tables *> tables;
    /* host staging of uploads that kernels read later in the same call */
    std::vector<uint32_t> h_meta, h_vpre;
    std::vector<uint64_t> h_bidx; /* first triangle of each molecule, staged for the filter */
    /* pinned staging for a run's small uploads (molecule records, prefixes, indices): copies from
       it are queued at once, where the runtime stages pageable sources on the host first */
    void *up = nullptr;
    size_t up_cap = 0, up_used = 0, up_need = 0;
    int upload(void *dst, const void *src, size_t bytes)
    {
        if (!bytes) return 0;
        const size_t off = (up_used + 255) & ~(size_t)255;
        if (up && off + bytes <= up_cap) {
            memcpy((char *)up + off, src, bytes);
            up_used = off + bytes;
            QS_HIP(hipMemcpyAsync(dst, (char *)up + off, bytes, hipMemcpyHostToDevice, stream));
            return 0;
        }
        up_need = std::max(up_need, off + bytes);
        QS_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, stream));
        return 0;
    }
    /* a pinned landing area for a download that is read after the next stream synchronisation,
       or nullptr when the staging is full (the caller then copies into its own memory) */
    void *stage(size_t bytes)
    {
        const size_t off = (up_used + 255) & ~(size_t)255;
        if (up && off + bytes <= up_cap) {
            up_used = off + bytes;
            return (char *)up + off;
        }
        up_need = std::max(up_need, off + bytes);
        return nullptr;
    }
    /* at the start of a run: the last run's uploads have been read (a run ends with a stream
       synchronisation; one that stopped early is waited for here), and the staging grows to what
       the last run needed */
    int upload_reset()
    {
        if (up_used || up_need) QS_HIP(hipStreamSynchronize(stream));
        if (up_need > up_cap && !g_arena) { /* (a run in a fixed region allocates no staging) */
            if (up) (void)hipHostFree(up);
            up = nullptr;
            up_cap = 0;
            const size_t cap = std::max<size_t>(2 * up_need, 1 << 16);
            if (hipHostMalloc(&up, cap, hipHostMallocDefault) == hipSuccess) up_cap = cap;
            else (void)hipGetLastError();
        }
        up_used = up_need = 0;
        return 0;
    }
    /* A run's atom inputs from host memory, queued before run_batch through the same staging
       (run_batch then keeps them): upload_begin, then upload_input per array.  Arrays above
       STAGE_INPUT_MAX are copied from the caller's memory directly. */
    static constexpr size_t STAGE_INPUT_MAX = (size_t)8 << 20;
    bool inputs_staged = false;
    int upload_begin()
    {
        QS_TRY(upload_reset());
        inputs_staged = true;
        return 0;
    }
    int upload_input(void *dst, const void *src, size_t bytes)
    {
        if (bytes > STAGE_INPUT_MAX) {
            QS_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, stream));
            return 0;
        }
        return upload(dst, src, bytes);
    }

    /* last result */
    std::vector<QsMol> mols;
    std::vector<uint32_t> vert_off, tri_off; /* per molecule, nmol + 1; triangles after the filter */
    uint32_t n_atoms_total = 0;
    uint64_t ncell_total = 0, nw_total = 0, vox_total = 0;
    uint32_t nv = 0, nt_all = 0, nt = 0, n_active = 0;
    size_t p1_lds = 0; /* dynamic LDS of the field pass 1 launch (largest molecule of the batch) */
    bool batch_valid = false; /* result arrays, vert_off / tri_off and their device copies belong to run_batch */
    /* field holds the shell voxels' values only (field pass 2), densely in shell-rank order (spre);
       the others come from outm / shellm.  Else field holds every voxel in grid order. */
    bool field_sparse = false;
    /* the last run's atom ids are cell-sorted positions (option atom_ids_sorted at run time) */
    bool ids_sorted = false;
    /* row words the marching-cubes scans cover: [scan_w0, scan_w0 + scan_nw) */
    uint64_t scan_w0 = 0, scan_nw = 0;
    /* the last run meshed its grid in x-slabs (n_slabs of them): the grid arrays hold only the
       last slab, so the field, labels and normals of the mesh are not available */
    bool multi_slab = false;
    int n_slabs = 1;
    /* vertex and triangle pools of the current run, and the least pools later runs start from
       (grown when a run needed more) */
    uint32_t pool_v = 0, pool_t = 0;
    uint64_t ntflags = 0; /* tile flags of the run's fixed-R pass-2 molecules */
    uint8_t tmark = 0;    /* the field passes' tile mark, 1-255, changed every run */
    uint64_t cap_v = 0, cap_t = 0;
    int reruns = 0;
    float *res_vert = nullptr;
    int *res_tri = nullptr, *res_atom = nullptr;

    /* drops every device buffer record without freeing it (buffers carved from a fixed region,
       which the next run reuses from its start), and the cached pass-2 tables with them */
    void forget_buffers()
    {
        for (DevBuf *b : {&in_pos, &in_rad, &d_off, &bb_part, &bb_out, &d_mols, &d_meta, &key, &count, &start,
                          &perm, &cse, &sx, &aux, &labels, &outm, &shellm, &field, &negm, &zerom, &s4, &wcnt,
                          &vbw, &tbw, &alist, &ntb, &avb, &atb, &vstart, &tstart, &normals, &colors_in, &totals, &bounds, &rback, &vert, &vert2, &vert3, &vkey,
                          &tri, &tri_out, &atom, &deg_blk, &keeppos, &tri_loc, &scan_tmp, &inc_cnt, &inc_start, &inc_ent,
                          &tpre_d, &tflag, &spre, &stot, &field_dense})
            *b = DevBuf{};
        for (P2Tables *t : tables) delete t;
        tables.clear();
        groups.clear();
        batch_valid = false;
    }

    /* frees every device buffer and the stream (the engine can be initialised again) */
    void release_all()
    {
        for (DevBuf *b : {&in_pos, &in_rad, &d_off, &bb_part, &bb_out, &d_mols, &d_meta, &key, &count, &start,
                          &perm, &cse, &sx, &aux, &labels, &outm, &shellm, &field, &negm, &zerom, &s4, &wcnt,
                          &vbw, &tbw, &alist, &ntb, &avb, &atb, &vstart, &tstart, &normals, &colors_in, &totals, &bounds, &rback, &vert, &vert2, &vert3, &vkey,
                          &tri, &tri_out, &atom, &deg_blk, &keeppos, &tri_loc, &scan_tmp, &inc_cnt, &inc_start, &inc_ent,
                          &tpre_d, &tflag, &spre, &stot, &field_dense})
            release(*b);
        for (P2Tables *t : tables) {
            release(t->offs);
            release(t->norm);
            release(t->end);
            delete t;
        }
        tables.clear();
        groups.clear();
        if (stream) (void)hipStreamSynchronize(stream);
        if (up) (void)hipHostFree(up);
        up = nullptr;
        up_cap = up_used = up_need = 0;
        if (stream) (void)hipStreamDestroy(stream);
        stream = nullptr;
        if (side) (void)hipStreamDestroy(side);
        side = nullptr;
        if (ev_emit) (void)hipEventDestroy(ev_emit);
        if (ev_side) (void)hipEventDestroy(ev_side);
        if (ev_vert) (void)hipEventDestroy(ev_vert);
        ev_emit = ev_side = ev_vert = nullptr;
        device = -1;
        batch_valid = false;
    }

    int init()
    {
        load_default_tuning();
        int dev = 0;
        QS_HIP(hipGetDevice(&dev));
        if (device == dev && stream) return 0;
        device = dev;
        QS_HIP(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking));
        /* a second stream (default priority) for the closest-atom pass, which overlaps smoothing */
        QS_HIP(hipStreamCreateWithFlags(&side, hipStreamNonBlocking));
        QS_HIP(hipEventCreateWithFlags(&ev_emit, hipEventDisableTiming));
        QS_HIP(hipEventCreateWithFlags(&ev_side, hipEventDisableTiming));
        QS_HIP(hipEventCreateWithFlags(&ev_vert, hipEventDisableTiming));
        return 0;
    }

    /* exclusive scan of n uint32 (in may equal out); total written to d_total if given */
    int scan(const uint32_t *in, uint32_t *out, uint64_t n, uint32_t *d_total)
    {
        if (n == 0) {
            if (d_total) QS_HIP(hipMemsetAsync(d_total, 0, sizeof(uint32_t), stream));
            return 0;
        }
        if (qs_scan_tiles(n) > 1 && n <= QS_SCAN1_MAX) {
            qs_launch_scan1(in, out, n, d_total, stream);
            QS_HIP(hipGetLastError());
            return 0;
        }
        if (qs_scan_tiles(n) > 1 && qs_scan_tiles(n) <= QS_SCAN_SELF_TILES) {
            /* the tile sums, then the tiles, each adding up the sums before it */
            QS_TRY(ensure(scan_tmp, 64 * sizeof(uint32_t)));
            qs_launch_tile_reduce(in, n, as<uint32_t>(scan_tmp), stream);
            qs_launch_tile_scan(in, out, n, nullptr, d_total, stream, as<uint32_t>(scan_tmp));
            QS_HIP(hipGetLastError());
            return 0;
        }
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
        const uint32_t *src = in;
        uint64_t len = n;
        for (size_t l = 0; l < lv.size(); l++) {
            qs_launch_tile_reduce(src, len, lvl_ptr[l], stream);
            src = lvl_ptr[l];
            len = lv[l];
        }
        for (int l = (int)lv.size() - 1; l >= 0; l--) {
            const uint32_t *offs = (l + 1 < (int)lv.size()) ? lvl_ptr[l + 1] : nullptr;
            qs_launch_tile_scan(lvl_ptr[l], lvl_ptr[l], lv[l], offs, nullptr, stream);
        }
        qs_launch_tile_scan(in, out, n, lv.empty() ? nullptr : lvl_ptr[0], d_total, stream);
        QS_HIP(hipGetLastError());
        return 0;
    }

    QsBatch batch(size_t bpre_at, size_t mid_at, int n) const
    {
        QsBatch B;
        B.mols = (const QsMol *)d_mols.p;
        B.bpre = (const uint32_t *)d_meta.p + bpre_at;
        B.coarse = B.bpre + n + 1; /* the coarse index follows each prefix */
        B.mid = mid_at ? (const int *)((const uint32_t *)d_meta.p + mid_at) : nullptr;
        B.n = n;
        return B;
    }

    int setup(const std::vector<MolIn> &in, const std::vector<BBox> &bb);
    int build_meta(size_t &at_atoms, size_t &at_p1, size_t &at_mc, size_t &at_wpre,
                   uint32_t &nb_atoms, uint32_t &nb_p1, uint32_t &nb_mc);
    /* allocate the workspace for batches up to these totals in one go */
    int reserve(uint64_t vox, uint64_t atoms, uint64_t cells, uint64_t verts, int nmol);
    int run_batch(const float *d_pos, const float *d_rad, const std::vector<MolIn> &in, int smooth,
                  const std::vector<BBox> *host_bbox);
    int field_stage(uint32_t nb_p1, size_t at_p1, StageTimer &tm);
    int fill_field(const void **dense);
    int mesh_stage(size_t at_mc, uint32_t nb_mc, size_t at_wpre, StageTimer &tm, bool with_atoms);
    int group_tables(P2Group &grp);
    void plan_pools();
    bool pools_overflowed(uint32_t nv_, uint32_t nt_);
    /* x-slab form of one molecule's field and mesh stages: owned planes per slab (0: one pass) */
    int slab_plan(const QsMol &M, int &planes);
    int run_slabs(int planes, int smooth, StageTimer &tm);
    int slab_mesh(size_t at_mc, uint32_t nb_mc, size_t at_wpre, uint64_t owned_word, uint32_t vbase, uint32_t tbase,
                  uint32_t &n_owned, uint32_t &n_tri);
};


}  // namespace qs

#endif
