/* Launchers of the HIP kernels (qs_kernels.hip).  Every per-molecule kernel runs over a batch:
 * a QsBatch maps each block to a molecule (QsMol) and to a block index inside it.  Prefix
 * arrays over blocks, row words or vertices come with coarse indexes (qs_common.h). */
#ifndef QS_KERNELS_H
#define QS_KERNELS_H

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "qs_common.h"

/* atoms into cells */
/* bounding boxes: `parts` partial boxes per molecule (qs_bbox_parts of the largest atom count) */
int qs_bbox_parts(uint32_t max_atoms);
size_t qs_bbox_part_floats(int nmol, int parts);
/* off: per-molecule atom offsets (nmol + 1); or nullptr for one molecule of atoms [o0, o1) */
void qs_launch_bbox(const float *pos, const float *rad, const uint32_t *off, uint32_t o0, uint32_t o1, int nmol, int parts,
                    float *part, float *out, hipStream_t s);
/* key: (cell, rank within the cell) per atom; count: atoms per cell (cleared before) */
void qs_launch_hash_count(QsBatch B, uint32_t nblocks, const float *pos, int2 *key, uint32_t *count, hipStream_t s);
void qs_launch_scatter(const int2 *key, int n, const uint32_t *start, int *perm, hipStream_t s);
void qs_launch_cell_finalize(const uint32_t *start, const uint32_t *count, int ncell, int *perm, int2 *cse, hipStream_t s);
/* atoms in cell order (sx) with their exact thresholds and neighbour cells (aux) */
void qs_launch_gather_aux(QsBatch B, uint32_t nblocks, const int *perm, const float *pos, const float *rad, float4 *sx,
                          float4 *aux, hipStream_t s);

/* field pass 1 */
int qs_pass1_rows_per_wave(QsGrid G);
int qs_pass1_slab_rows(QsGrid G, int W);
/* rows per pass-1 slab block at voxel size dx (an upper bound under the LDS and candidate-column
   limits); dx <= 0 clears the table; -1 when it is full */
int qs_set_tuned_rows(float dx, int rows);
/* field pass 1 kernel: 0 slab scatter (default), 1 row gather, 2 per-voxel */
int qs_pass1_mode(bool direct, bool rows);
/* blocks of one molecule's launch over its planes [p1_lo, p1_hi) */
uint32_t qs_pass1_blocks(const QsMol &M, int mode);
size_t qs_pass1_lds_bytes(const QsMol &M, int mode);
/* The slab kernel (mode 0) also writes the sign and zero words every voxel has when its pass-2
   tile holds no shell voxel (negm, zerom), and marks the tiles that do (tflag, at each molecule's
   tile_off, set to tmark); both may be NULL.  The flags are not cleared between runs: a run uses a
   mark of its own, and a stale flag that equals it only makes pass 2 rewrite the same sign words
   for a tile without shell voxels. */
void qs_launch_pass1(QsBatch B, uint32_t nblocks, int mode, size_t lds_bytes, const uint32_t *cstart, const float4 *sx,
                     const float4 *aux, uint64_t *outm, uint64_t *shellm, uint64_t *negm, uint64_t *zerom, uint8_t *tflag,
                     uint8_t tmark, hipStream_t s);
void qs_launch_pass1_direct(QsBatch B, uint32_t nblocks, const int2 *cse, const float4 *sx, uint64_t *outm, uint64_t *shellm,
                            hipStream_t s);

/* field pass 2 */
size_t qs_pass2_lds_bytes(int r);
int qs_pass2_path(int r, bool direct);
/* blocks of one molecule's launch over its planes [p2_lo, p2_hi) */
uint32_t qs_pass2_blocks(const QsMol &M, int path);
/* the shell voxels' values go to field at their shell ranks (spre, from qs_launch_shell_scan);
   tflag (path 0 only, may be NULL): the marks of the launch's tiles in its block order (from the
   group's first molecule's tile_off on); tiles field pass 1 did not mark with tmark are skipped; xchunk
   (path 0): consecutive tiles each XCD takes in turn */
void qs_launch_pass2(QsBatch B, uint32_t nblocks, int r, int path, const uint64_t *outm, const uint64_t *shellm,
                     const uint32_t *spre, float *field, uint32_t *negm, uint32_t *zerom, const int *offs, const int *norm_start, const int *enum_end,
                     QsNormFlags nf, const uint8_t *tflag, uint8_t tmark, uint32_t xchunk, hipStream_t s);
/* pass-2 tiles of the fixed-R kernel of one molecule (its tile-flag count) */
uint32_t qs_pass2_tiles(const QsMol &M);
/* z tiles per (x, y) tile row of the fixed-R pass-2 kernel */
uint32_t qs_pass2_ztiles(const QsMol &M);

/* shell ranks (spre, one per word) of the words [0, nw) of shellm: the place of each word's first
   shell voxel in the stored field (sums: qs_scan_tiles(nw) scratch entries; totals: 2 entries,
   the total shell voxels first) */
void qs_launch_shell_scan(const uint64_t *shellm, uint64_t nw, uint32_t *sums, uint32_t *totals, uint32_t *spre, hipStream_t s);
/* field pass 2 stores the shell voxels' values only, in shell-rank order: one molecule's whole
   field in grid order (out), from them and its outside / shell masks (spre, outm, shellm at the
   molecule's first word; F the stored values of the whole launch) */
void qs_launch_field_expand(const float *F, const uint32_t *spre, const uint64_t *outm, const uint64_t *shellm, float *out, int g,
                            int W, float dx, hipStream_t s);

/* stage-wise helpers (one molecule) */
void qs_launch_field_signs(const float *F, int g, int W, uint32_t *negm, uint32_t *zerom, hipStream_t s);
void qs_launch_labels_to_masks(const uint8_t *labels, int g, int W, uint64_t *outm, uint64_t *shellm, hipStream_t s);

/* marching cubes */
/* blocks of one molecule's launch over its grid points of planes [mc_lo, mc_hi) */
uint32_t qs_mc_count_blocks(const QsMol &M);
/* per row word: slot masks (words with vertices only), vertex and triangle counts, triangle-count
   bit planes (words with triangles only) */
/* cnt: per row word its vertex count | triangle count << 16 */
void qs_launch_mc_count(QsBatch B, uint32_t nblocks, const uint64_t *N, const uint64_t *Zm, uint64_t *S4, uint32_t *cnt,
                        uint64_t *NT, hipStream_t s);
/* one thread per vertex / per triangle over the active words (list, with their first vertex ids
   avb and first triangle ids atb), after one launch that gives vstart / tstart each 256-item
   block's first active word (nv / 256 + 1 and ntri / 256 + 1 entries) and clears deg_blk, unless
   starts_ready says that the word scans did (QsBlockStarts).  The
   field F holds the shell voxels' values only, at their shell ranks (spre), the others come from
   the outside / shell masks (shellm == nullptr: F holds every voxel in grid order, spre is
   unused).  Triangle corners are vertex ids plus vbase; degenerate triangles are counted per
   256-triangle block (deg_blk) and in total.  The block starts and the triangle kernel run on s;
   the vertex kernel on s_vert: when that is another stream it waits there for the block starts
   (event ev_fork), and ev_vert is recorded on s_vert after it. */
void qs_launch_mc_emit(const QsMol *mols, const uint64_t *wpre, const uint32_t *wcoarse, const uint32_t *list, uint32_t nact,
                       const uint32_t *avb, uint32_t *vstart, uint32_t nv, const uint64_t *S4, const float *F,
                       const uint32_t *spre, const uint64_t *outm, const uint64_t *shellm, float *vert, uint64_t *vkey,
                       const uint32_t *atb, uint32_t *tstart, uint32_t ntri, const uint64_t *N, const uint64_t *Zm,
                       const uint32_t *vbw, const uint64_t *NT, int *tri, uint32_t vbase, uint32_t *deg_blk,
                       uint32_t *deg_total, const uint32_t *d_counts, hipStream_t s, hipStream_t s_vert = nullptr,
                       hipEvent_t ev_fork = nullptr, hipEvent_t ev_vert = nullptr, bool starts_ready = false);

/* Launches whose item counts live on the device take them from the d_* pointers (when given),
   the host counts being the pools the launch covers: marching-cubes totals are counts[0]
   triangles, counts[1] vertices, counts[2] active words. */

/* per-molecule first vertex (vpre, u32) and first triangle (tpre, u64) from the word scans at each
   molecule's first word (wpre[m]; at or past the nw scanned words for none), nmol + 1 entries
   each ending with the totals, and the vertex coarse index (ncoarse entries), in one launch */
void qs_launch_mesh_bases(const uint32_t *vbw, const uint32_t *tbw, const uint64_t *wpre, uint64_t nw, int nmol,
                          const uint32_t *counts, uint32_t *vpre, uint64_t *tpre, uint32_t *vcoarse, uint32_t ncoarse, hipStream_t s);

/* per-vertex stages */
void qs_launch_closest_atom(const QsMol *mols, const uint32_t *vpre, const uint32_t *vcoarse, const uint32_t *cstart,
                            const float4 *sx,
                            const int *perm, const float *vert, uint32_t nv, int sorted_ids, int *out, const uint32_t *d_nv,
                            hipStream_t s);
void qs_launch_smooth(const QsMol *mols, const uint32_t *vpre, const uint32_t *vcoarse, const uint64_t *N, const uint64_t *Zm,
                      const uint64_t *S4, const uint32_t *vbw, const uint64_t *vkey, uint32_t nv,
                      const uint64_t *NT, const uint32_t *tbw, const int *tri, const float *src, float *dst,
                      const uint32_t *d_nv, hipStream_t s);

/* Laplacian passes over a triangle array (meshes built in x-slabs): per-vertex incidence counts
   (cnt, nv entries, cleared by the caller) from the triangles, then, after an exclusive scan of
   cnt into start (nv + 1) and a second clear of cnt, the incidence lists sorted by corner index
   (ent, 3 nt entries); each pass reads src and writes dst */
void qs_launch_incid_count(const int *tri, uint32_t nt, uint32_t nv, uint32_t *cnt, hipStream_t s);
void qs_launch_smooth_incid_build(const int *tri, uint32_t nt, uint32_t nv, uint32_t *cnt, uint32_t *start, uint32_t *ent,
                                  hipStream_t s);
void qs_launch_smooth_incid(const uint32_t *start, const uint32_t *ent, const int *tri, uint32_t nv, const float *src,
                            float *dst, hipStream_t s);
/* degenerate triangles per 256-triangle block into deg_blk and their total added to *deg_total */
void qs_launch_deg_count(const int *tri, uint32_t nt, uint32_t *deg_blk, uint32_t *deg_total, hipStream_t s);

/* unit field gradient at each vertex's marching-cubes position, 3 floats per vertex (field as for
   the vertex kernel) */
void qs_launch_vertex_normals(const QsMol *mols, const uint32_t *vpre, const uint32_t *vcoarse, const uint64_t *vkey, uint32_t nv,
                              const float *F, const uint32_t *spre, const uint64_t *outm, const uint64_t *shellm, float *out,
                              hipStream_t s);
/* colour of each vertex's closest atom from a per-atom RGBA table (batch atom layout); perm maps
   cell-sorted atom ids to input indices (nullptr: the ids are input indices) */
void qs_launch_vertex_colors(const QsMol *mols, const uint32_t *vpre, const uint32_t *vcoarse, const int *atom,
                             const int *perm, uint32_t nv, const float4 *table, float4 *out, hipStream_t s);
/* degenerate-triangle filter: dpre = exclusive prefix of the per-256-triangle-block degenerate
   counts that the triangle kernel leaves, ndeg their total (device); with ndeg == 0 the compaction
   writes nothing */
void qs_launch_tri_compact(const int *tri, uint32_t nt, const uint32_t *dpre, const uint32_t *ndeg, int *out,
                           const uint32_t *d_nt, hipStream_t s);
/* kept triangles before each of n boundaries (idx, triangle ids) into out; with rb also the run's
   read-back gathered in one array: {totals[0..3], vpre[0..n), out[0..n)} */
void qs_launch_tri_kept_before(const int *tri, const uint32_t *dpre, const uint64_t *idx, int n, uint32_t tcap, uint32_t *out,
                               const uint32_t *totals, const uint32_t *vpre, uint32_t *rb, hipStream_t s);
void qs_launch_tri_localize(const int *tri, uint32_t nt, const uint32_t *tpre, const uint32_t *vpre, int nmol, int *out,
                            hipStream_t s);

/* blocks of the persistent launches (rounded down to a multiple of 8, at least 8) */
void qs_set_persistent_blocks(uint32_t n);
uint32_t qs_persistent_blocks(void);

/* device-wide exclusive scan pieces */
uint64_t qs_scan_tiles(uint64_t n);
/* inputs of at most this many entries are scanned by one block in one launch (the generic scan,
   the shell ranks and the marching-cubes word scans; one block's memory throughput makes more
   entries slower than several blocks); inputs of at most QS_SCAN_SELF_TILES tiles of 2,048 are
   scanned in two launches, each tile adding up the sums of the tiles before it itself */
#define QS_SCAN1_MAX 8192u
#define QS_SCAN_SELF_TILES 64u
void qs_launch_scan1(const uint32_t *in, uint32_t *out, uint64_t n, uint32_t *total, hipStream_t s);
void qs_launch_tile_reduce(const uint32_t *in, uint64_t n, uint32_t *sums, hipStream_t s);
/* (self_sums: the raw sums of at most QS_SCAN_SELF_TILES tiles, offsets then nullptr) */
void qs_launch_tile_scan(const uint32_t *in, uint32_t *out, uint64_t n, const uint32_t *offsets, uint32_t *total,
                         hipStream_t s, const uint32_t *self_sums = nullptr);
void qs_launch_copy(const void *src, void *dst, uint64_t bytes, hipStream_t s);
/* The emission launches' block starts, written by the word scans when the pools are known
   before them: vstart / tstart hold the first active word of each 256-vertex / 256-triangle block
   (nv / nt entries), and the nzero counters at zero (the triangle kernel's degenerate counts) are
   cleared.  A null vstart writes none of them (qs_launch_mc_emit then computes them). */
struct QsBlockStarts {
    uint32_t *vstart = nullptr;
    uint32_t nv = 0;
    uint32_t *tstart = nullptr;
    uint32_t nt = 0;
    uint32_t *zero = nullptr;
    uint32_t nzero = 0;
};

/* marching-cubes word scans over the packed counts of qs_launch_mc_count: vbw, tbw, the
   active-word list and totals {triangles, vertices, active words}; sums holds
   3 * qs_scan_tiles(nw) entries; list entries are list_base + the word's index in cnt; and the
   block starts of bs */
void qs_launch_mc_scans(const uint32_t *cnt, uint64_t nw, uint32_t *sums, uint32_t *vbw,
                        uint32_t *tbw, uint32_t *list, uint32_t *avb, uint32_t *atb, uint32_t *totals, uint32_t list_base,
                        hipStream_t s, QsBlockStarts bs = QsBlockStarts{});
void qs_launch_gather_u32(const uint32_t *src, const uint64_t *idx, int n, uint32_t *out, hipStream_t s);

#endif
