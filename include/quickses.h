/* quickses: SES (solvent excluded surface) meshes on AMD Instinct MI355X (gfx950).
 *
 * C ABI of libquickses.so.  The first five entry points are drop-in replacements of the
 * reference's exported functions and keep their signatures and semantics; the rest are
 * additions that the reference does not have.
 *
 * Types: qs_float3 is three packed floats (12 bytes, 4-byte aligned), the layout of the
 * float3 the reference's ABI passes (CudaSurf.h:75-83).
 */
#ifndef QUICKSES_H
#define QUICKSES_H

#include <stdbool.h>
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
    float x, y, z;
} qs_float3;

/* ---------------------------------------------------------------------------------------- */
/* Reference-compatible entry points                                                          */

/* Replaces API_computeSES (CudaSurf.h:77-78, CudaSurf.cu:787-841).
 * resoSES: voxel size in Angstrom; atomPos/atomRad: N atom centres and radii (host memory);
 * doSmoothing: number of Laplacian passes (<= 0: none).  out_vertices and out_triangles are
 * ignored, as in the reference; results are read with the getters below.  *NVert = number of
 * vertices, *NTri = number of triangle INDICES (3 per triangle) after dropping triangles with a
 * repeated vertex.  N <= 1 gives an empty mesh.  On error the mesh is empty and
 * API_lastError() is non-zero (the reference exits the process instead). */
void API_computeSES(float resoSES, qs_float3 *atomPos, float *atomRad, unsigned int N, qs_float3 *out_vertices,
                    unsigned int *NVert, int *out_triangles, unsigned int *NTri, int doSmoothing);

/* Replaces API_getTriangles (CudaSurf.h:79, CudaSurf.cu:845-857): flat int32[NTri]; with
 * invertTriangles the first two indices of every triangle are swapped in place. */
int *API_getTriangles(bool invertTriangles);

/* Replaces API_getVertices (CudaSurf.h:80, CudaSurf.cu:858-861): qs_float3[NVert]. */
qs_float3 *API_getVertices(void);

/* Replaces API_freeMesh (CudaSurf.h:81, CudaSurf.cu:867-872): frees the three result arrays. */
void API_freeMesh(void);

/* Replaces API_getAtomIdPerVert (CudaSurf.h:82, CudaSurf.cu:862-865): int32[NVert], the
 * atom nearest to each vertex (before smoothing) among the 27 neighbour cells; -1 if none.
 * By default the index into the caller's atom arrays; with qs_set_option("atom_ids_sorted", 1)
 * the index into the cell-sorted atom order, which is what the reference returns. */
int *API_getAtomIdPerVert(void);

/* ---------------------------------------------------------------------------------------- */
/* Additions                                                                                   */

/* Error code of the last call (0 = success) and its message. */
int API_lastError(void);
const char *qs_last_error_message(void);

/* Options (process-wide, like the reference's globals):
 *   "lim_mode"        0 = polygonise every cube whose corners are valid (default),
 *                     1 = stop at cube g-2r-2 as the reference's single slice does
 *   "atom_ids_sorted" 0 = input atom index (default), 1 = cell-sorted index (reference)
 *   "pass1_direct"    1 = evaluate field pass 1 per voxel over its 27 cells (cross-check)
 *   "pass1_rows"      1 = evaluate field pass 1 by row gather, one wavefront per group of rows,
 *                     instead of the default slab scatter (cross-check)
 *   "pass2_direct"    1 = evaluate field pass 2 with the (2r+1)^3 window kernel (cross-check)
 *   "trace"           1 = print per-stage GPU times as JSON lines on stderr
 *   "timing"          1 = record per-stage GPU times (hipEvents on the library's stream)
 *   "slab_planes"     N > 0 = mesh a single molecule in x-slabs of N planes (the reference's
 *                     slices, CudaSurf.cu:626-754, with disjoint cube ownership: the mesh is the
 *                     whole-grid mesh bit for bit); 0 = slabs only when the grid does not fit the
 *                     memory budget or has 2^32 points or more (default)
 *   "slab_mb"         device-memory budget of one slab's grid arrays in MiB; 0 = half the
 *                     device's free memory (default)
 * Returns 0, or -1 for an unknown key. */
int qs_set_option(const char *key, int value);
int qs_get_option(const char *key);

/* Tuning (no reference counterpart; the reference has no tunables): per voxel size, rows per
 * field-pass-1 block ("pass1_slab_rows"; an upper bound under the kernel's memory limits).
 * qs_set_tuning sets one entry (returns 0, -1 for an unknown key or a full table);
 * qs_load_tuning reads the "pass1_slab_rows" object of a JSON file such as the tuning.json that
 * quickses_amd/autotune.py writes (returns the entries read, a negative error code when the file
 * cannot be read; NULL clears the table).  The
 * library reads tuning.json beside itself once per process on first use (environment variable
 * QS_TUNING_FILE names another file, "none" none). */
int qs_set_tuning(const char *key, float reso, int value);
int qs_load_tuning(const char *path);

/* Device-resident form of API_computeSES: d_pos (3*N floats) and d_rad (N floats) are device
 * pointers on the current HIP device; the mesh stays on the device in buffers owned by the
 * library, valid until the next call.  *nvert, *nidx as for API_computeSES.  Returns 0 or an
 * error code.  The call is synchronous. */
int qs_compute_device(const float *d_pos, const float *d_rad, unsigned int N, float resoSES, int doSmoothing,
                      unsigned int *nvert, unsigned int *nidx, float **d_vert, int **d_tri, int **d_atom_id);

/* qs_compute_device without allocating: no counterpart in the reference; for callers that cannot
 * allocate device memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid allocation).  Every working buffer is carved from a fixed region of 256 MiB per device,
 * a file-scope device array that each device the library runs on holds until the process ends;
 * a molecule whose buffers do not fit (about 20 M grid points, a 20,000-atom molecule at 0.5 A
 * takes about 100 MiB) is rejected with error 6 and a message.  The mesh is written to the
 * caller's device buffers: d_vert (3 * vert_cap floats), d_atom_id (vert_cap ints), d_tri
 * (3 * tri_cap ints); when it has more vertices or triangles than that, error 5 and nothing is
 * written.  *nvert, *nidx as for qs_compute_device (nidx = 3 * triangles).  The same mesh as
 * qs_compute_device, bit for bit; the call is synchronous.  Calls share the device's region
 * (it keeps the last call's contents, and every run clears what it needs), so calls that
 * overlap on one device race.  The first call on a device creates the library's streams and
 * events for it. */
int qs_compute_device_noalloc(const float *d_pos, const float *d_rad, unsigned int N, float resoSES, int doSmoothing,
                              float *d_vert, unsigned int vert_cap, int *d_tri, unsigned int tri_cap, int *d_atom_id,
                              unsigned int *nvert, unsigned int *nidx);

/* Batched form: nmol molecules in one sequence of launches.  Molecule m owns atoms
 * [atom_off[m], atom_off[m+1]) of the concatenated pos (3 floats per atom) / rad arrays and is
 * meshed at voxel size reso[m]; its mesh equals what qs_compute_device gives for it alone.
 * atom_off has nmol+1 entries (host memory), starting at 0.  Molecules with fewer than 2 atoms
 * get an empty mesh.  Results stay on the device (library buffers, valid until the next call):
 * vertices of molecule m are [vert_off[m], vert_off[m+1]) of *d_vert, its triangles
 * [tri_off[m], tri_off[m+1]) of *d_tri (3 ints each, holding indices into *d_vert), atom ids
 * local to the molecule.  vert_off / tri_off (host, nmol+1 entries each) may be NULL.
 * Synchronous.  Returns 0 or an error code. */
int qs_compute_batch_device(const float *d_pos, const float *d_rad, const unsigned int *atom_off, int nmol,
                            const float *reso, int doSmoothing, unsigned int *vert_off, unsigned int *tri_off,
                            float **d_vert, int **d_tri, int **d_atom_id);

/* Batched form on host arrays: per-molecule vertex and triangle counts into nvert[m], ntri[m]
 * (host, nmol entries each, may be NULL); qs_batch_mesh then copies molecule m's mesh out with
 * vertex indices local to the molecule (vert: 3*nvert[m] floats, tri: 3*ntri[m] ints, atom_id:
 * nvert[m] ints; any may be NULL). */
int qs_compute_batch(const float *pos, const float *rad, const unsigned int *atom_off, int nmol, const float *reso,
                     int doSmoothing, unsigned int *nvert, unsigned int *ntri);
int qs_batch_mesh(int m, float *vert, int *tri, int *atom_id);
/* Offsets of the last batch (host arrays of nmol+1 entries, may be NULL); returns nmol, or a
 * negative error code when there is no batch result. */
int qs_batch_offsets(unsigned int *vert_off, unsigned int *tri_off);
/* All meshes of the last batch at once: vertices (3 * vert_off[nmol] floats), triangles
 * (3 * tri_off[nmol] ints, vertex indices local to each molecule), atom ids (vert_off[nmol]). */
int qs_batch_fetch(float *vert, int *tri, int *atom_id);
/* The same arrays in host memory owned by the library (pinned, so the copies run at the link's
 * speed), valid until the next qs_batch_fetch_host call; any pointer argument may be NULL. */
int qs_batch_fetch_host(const float **vert, const int **tri, const int **atom_id);
/* Pipelined form: qs_batch_fetch_host_async(slot) (slot 0 or 1) starts copying the last batch's
 * meshes to the slot's pinned host arrays and returns at once; the next qs_compute_batch may
 * start while the copies run.  qs_batch_wait_host(slot, ...) waits for them and gives the arrays
 * (layout as qs_batch_fetch_host) and the vertex and triangle totals, valid until the slot's next
 * fetch.  A slot's next fetch first waits for its previous copies. */
int qs_batch_fetch_host_async(int slot);
int qs_batch_wait_host(int slot, const float **vert, const int **tri, const int **atom_id, unsigned int *nvert,
                       unsigned int *ntri);

/* Per-vertex normals of the last computation (plain or batch; vertices in the same order as its
 * vertex arrays): the unit gradient of the SES field at each vertex's marching-cubes position
 * (before smoothing), pointing outwards.  The reference's MeshData declares such a normals array
 * but leaves it commented out (SmoothMesh.h:34); there is no reference routine to follow.  The
 * gradient at a grid point is the central difference of the field (one-sided on the grid's
 * faces); an edge vertex blends its edge's two end points with its edge parameter.
 * qs_vertex_normals copies 3 * nvert floats to host memory (nvert must equal the last mesh's
 * vertex count; after API_computeSES that is *NVert).  qs_vertex_normals_device writes them to
 * a device array and reports the kernel time in *ms (may be NULL).  Returns 0 or an error code. */
int qs_vertex_normals(float *out, unsigned int nvert);
int qs_vertex_normals_device(float *d_out, float *ms);

/* Per-vertex colours of the last computation: vertex v gets the RGBA (4 floats) of its closest
 * atom (the atom id of API_getAtomIdPerVert / the batch getters) from atom_rgba, natoms * 4 floats
 * laid out like the atom input (a batch's molecules concatenated); a vertex without a closest
 * atom gets zeros.  MeshData declares a colors array but leaves it commented out
 * (SmoothMesh.h:35).  out receives nvert * 4 floats.  Returns 0 or an error code. */
int qs_vertex_colors(const float *atom_rgba, unsigned int natoms, float *out, unsigned int nvert);

/* Device-to-device copy of `bytes` (multiple of 16, 16-byte aligned device pointers) by a
 * plain vector-copy kernel on the library's stream; *ms = its time.  A measurement reference:
 * hardware traffic counters of this kernel calibrate those of the SES kernels. */
int qs_copy_device(void *dst, const void *src, uint64_t bytes, float *ms);

/* Per-stage GPU times of the last computation (needs option "timing" or "trace"): stage names
 * comma-separated into names, milliseconds into ms; returns the number of stages. */
int qs_stage_times(char *names, int names_cap, float *ms, int max_stages);

/* Grid of the last computation (its first molecule, or molecule m of the last batch):
 * fparams = {origin x,y,z, dx, dxn}, iparams = {g, gn, r, lim}.  Returns 0, or -1 for no such
 * molecule. */
int qs_last_grid(float *fparams, int *iparams);
/* Number of x-slabs the last computation's grid ran in (1: the whole grid at once; 0: no
 * result).  After a run in several slabs the field, labels and normals of the mesh are not
 * available (qs_vertex_normals returns an error). */
int qs_last_slab_count(void);
int qs_batch_grid(int m, float *fparams, int *iparams);

/* Copies of the last computation's intermediate arrays to host memory (stage-wise checks; for
 * a batch, the concatenated arrays, and "labels" of its first molecule):
 *   "labels" uint8[g^3] (0 outside, 1 shell, 2 inside), "field" float[g^3] (every voxel: field
 *   pass 2 stores the shell voxels only, the others are filled in for the copy),
 *   "shellm" uint64[g*g*ceil(g/64)] (shell mask, bit z%64 of word (x*g+y)*ceil(g/64)+z/64),
 *   "perm" int32[N] (cell-sorted order), "cells" int32[2*gn^3] (start,end pairs),
 *   "vert_unsmoothed" float[3*NVert], "tri_all" int32[3*T] (before the degenerate filter).
 * Returns the number of bytes copied, or -1. */
int64_t qs_fetch(const char *what, void *host, int64_t bytes);

/* Stage entry points on caller-given inputs (host memory):
 * field pass 2 from labels; marching cubes + weld from a field. */
int qs_stage_pass2(const float *fparams, const int *iparams, const uint8_t *labels, float *field);
int qs_stage_mesh(const float *fparams, const int *iparams, const float *field, unsigned int *nvert,
                  unsigned int *ntri, float *vert, int64_t vert_cap, int *tri, int64_t tri_cap);

/* Request server (no reference counterpart; the reference meshes one molecule per synchronous
 * call).  Requests of any size and voxel size are submitted at any time; `workers` threads,
 * each with its own engine and HIP stream, take pending requests in arrival order while their
 * estimated device memory fits hbm_budget_bytes / workers, mesh them as one batch and keep
 * each mesh in a device-resident result arena of result_bytes until it is released.  Batches
 * of different workers overlap on the device.  A result equals API_computeSES on the same
 * request alone, with triangle indices local to the request.
 *   qs_serve_open    starts the workers on the current device; one server per process
 *   qs_serve_submit  copies the request and returns its id (>= 0) or -error
 *   qs_serve_poll    0 = done (device pointers into the arena, valid until release;
 *                    latency_ms[3] = {submit -> resident, submit -> batch start, batch size}),
 *                    1 = not finished after wait_ms (-1 waits), < 0 = error
 *   qs_serve_fetch   waits, then copies the mesh to host arrays (any may be NULL)
 *   qs_serve_release frees the request's result
 *   qs_serve_stats   {batches, requests batched, largest batch, most workers busy at once,
 *                    arena bytes in use, arena bytes, requests pending, express batches};
 *                    returns 8
 *   qs_serve_set     "max_batch_voxels": cap on the grid points of one batch (0 = none, the
 *                    default; a request larger than the cap runs alone); "max_batch": cap on
 *                    the requests of one batch (1..256, default 256); "express_voxels": with 2
 *                    workers or more, the last one takes only requests of at most that many
 *                    grid points, oldest first, so that small requests need not wait behind
 *                    large batches (0 = off, the default)
 *   qs_serve_close   finishes pending work, stops the workers, frees everything */
int qs_serve_open(uint64_t hbm_budget_bytes, uint64_t result_bytes, int workers);
int64_t qs_serve_submit(const float *pos, const float *rad, unsigned int N, float reso, int doSmoothing);
int qs_serve_poll(int64_t id, int wait_ms, unsigned int *nvert, unsigned int *ntri, float **d_vert, int **d_tri,
                  int **d_atom_id, double *latency_ms);
int qs_serve_fetch(int64_t id, float *vert, int *tri, int *atom_id);
int qs_serve_release(int64_t id);
int qs_serve_set(const char *key, uint64_t value);
int qs_serve_stats(uint64_t *out, int n);
int qs_serve_close(void);

/* File formats of the reference's command-line tool (host only, no GPU).
 * qs_pdb_read: atoms of a PDB file as the reference's reader takes them (cpdb/cpdb.cpp:34-158,
 *   211-275; radii by element, CudaSurf.h:64-73): "ATOM  " records up to the first "END*"
 *   record, alternate locations other than 'A' / ' ' dropped, hydrogens kept, file order.
 *   *pos (3 floats per atom) and *rad are malloc'ed (release with qs_free).  Returns 0, or 2
 *   if the file cannot be read.
 * qs_obj_write: the reference's OBJ layout (writeToObj CudaSurf.cu:329-371): "v %.3f %.3f %.3f"
 *   lines, an empty line, then "f b+1 a+1 c+1" for each triangle (a, b, c).  Returns 0 or 2.
 * qs_io_error: message of the last failure of either. */
int qs_pdb_read(const char *path, float **pos, float **rad, unsigned int *n);
int qs_obj_write(const char *path, const float *vert, unsigned int nvert, const int *tri, unsigned int ntri);
const char *qs_io_error(void);
void qs_free(void *p);

/* Seeded synthetic molecules (test and benchmark inputs). */
int qs_synth_ball(uint64_t seed, int n, double density, double min_dist, float *pos, float *rad);
int qs_synth_shell(uint64_t seed, int n, double r_in, double r_out, double min_dist, float *pos, float *rad);
int qs_synth_jitter(uint64_t seed, int n, double sigma, const float *base, float *out);
int qs_synth_sizes(uint64_t seed, int count, double lo, double hi, int *out);

#ifdef __cplusplus
}
#endif
#endif
