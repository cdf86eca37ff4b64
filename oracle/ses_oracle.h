/* CPU oracle for the SES mesh path -- TEST INFRASTRUCTURE ONLY.
 *
 * A host restatement, in plain C with IEEE single precision (no FMA contraction, correctly
 * rounded sqrtf and division), of the reference's SES hot path:
 *   grid setup      CudaSurf.cu:149-174, 250-253, 548-576, 629
 *   atom hashing    Kernels.cu:119-131, 185-216
 *   stable sort     CudaSurf.cu:610-612 (thrust::sort by cell id; canonical order (cell, index))
 *   cell ranges     Kernels.cu:306-341
 *   field pass 1    Kernels.cu:344-475
 *   field pass 2    Kernels.cu:478-545
 *   MC classify     MarchingCubes.cu:571-658
 *   MC vertices     MarchingCubes.cu:715-851
 *   quantised weld  MarchingCubes.cu:853-862, CudaSurf.cu:428-458
 *   closest atom    Kernels.cu:547-604
 *   smoothing       SmoothMesh.cpp:45-63, 96-124 (Gauss-Seidel) and its Jacobi form
 *   API concat      CudaSurf.cu:787-841
 * plus a double-precision (FP64) evaluation of the same mathematics.
 *
 * Only tests, the smoke check and the CPU-baseline leg of the benchmark may load this library.
 * The shipped path (libquickses.so) never links or calls it.
 *
 * Pins (what checks this restatement independently of itself; see tests/test_oracle.py):
 *   grid setup, hash, sort, cell ranges: numpy restatement + Python sorted() on the same inputs;
 *   pass 1: brute force over all atoms (no cell grid) in FP64, equal except at ulp ties;
 *   pass 2: closed form for an isolated atom, FP64 brute force, separable-min identity;
 *   MC table: every case's edge set equals its sign-change set; 820 triangles;
 *   mesh: single atom -> sphere area/volume (closed form), watertight, Euler characteristic 2;
 *   weld: edge-id weld vs quantised weld partitions agree except at coincident vertices;
 *   smoothing: Jacobi == Gauss-Seidel on a mesh with an independent-set vertex order,
 *              umbrella operator preserves a constant field / plane.
 */
#ifndef SES_ORACLE_H
#define SES_ORACLE_H

#include <stdint.h>
#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#define OR_PROBE 1.4f
#define OR_EPS 0.001f
#define OR_EMPTY (2147483647 - 1)

/* voxel labels produced by field pass 1 */
enum { OR_OUT = 0, OR_SHELL = 1, OR_IN = 2 };

typedef struct {
    int n_atoms;
    float reso;
    float minv[3], maxv[3];
    float max_rad;
    float max_dist;
    float dx_n;       /* neighbour-grid cell size = probe + max radius */
    float origin[3];  /* shared by the neighbour grid and the SES grid */
    int gn;           /* neighbour grid cells per axis */
    int g;            /* SES grid points per axis */
    int r;            /* pass-2 window half-width in voxels = ceil(probe / reso) */
} or_grid;

int or_grid_setup(const float *pos, const float *rad, int n, float reso, or_grid *out);

/* hash[i] = cell id of atom i */
void or_hash(const or_grid *G, const float *pos, int n, int *hash);
/* perm[k] = original index of the k-th atom in (cell, index) order;
   cell_start/end[gn^3] (OR_EMPTY when empty); sorted_xyzr[4n] */
void or_sort_cells(const or_grid *G, const float *pos, const float *rad, int n, const int *hash,
                   int *perm, int *cell_start, int *cell_end, float *sorted_xyzr);

/* labels[g^3], x-major linear order (x slowest, z fastest) */
void or_pass1(const or_grid *G, const int *cell_start, const int *cell_end, const float *sorted_xyzr,
              uint8_t *labels);
/* field[g^3] from labels, brute-force (2r+1)^3 window exactly as Kernels.cu:478-545 */
void or_pass2(const or_grid *G, const uint8_t *labels, float *field);
/* the same on x-plane ranges (slab-local arrays) */
void or_pass1_range(const or_grid *G, const int *cs, const int *ce, const float *sx, int x0, int x1, uint8_t *labels);
void or_pass2_range(const or_grid *G, const uint8_t *labels, int lx0, int lx1, int x0, int x1, float *field);
/* CPU timing sample: the whole path (field, soup, quantised weld, closest atom, one in-place
   Laplacian pass, filter) on a slice of `planes` middle x-planes; returns seconds */
double or_bench_slab(const float *pos, const float *rad, int n, float reso, int planes, int *g_out, int64_t *voxels,
                     int64_t *n_tri);

/* cube index per cube origin (g^3 entries, 0 outside [0,lim)^3) */
void or_mc_classify(int g, int lim, const float *field, uint8_t *cubeidx);

typedef struct {
    int64_t n_vert;
    int64_t n_tri;
    float *vert;      /* 3*n_vert */
    int *tri;         /* 3*n_tri */
    int *atom_id;     /* n_vert */
} or_mesh;

void or_mesh_free(or_mesh *m);

/* Edge-id welded mesh: vertex = (grid point, axis) with a sign change inside the emission
   window [0,lim)^3, or a zero-valued grid point (snap). Vertices in key order, triangles in
   cube order / table order. */
int or_mc_edge_weld(const or_grid *G, int lim, const float *field, or_mesh *m);
/* vertices of or_mc_edge_weld(G, lim, field) (same order, same edges) interpolated in double
   from field64 (parameter clamped to the edge); returns their count (out holds 3 doubles per
   vertex; point, if not NULL, the grid point of each vertex) */
int64_t or_mc_vertices_fp64(const or_grid *G, int lim, const float *field, const double *field64, double *out,
                            int64_t *point);

/* per-vertex unit field gradients in FP64 from the FP32 field, in edge_weld vertex order
   (the library's normals definition; the reference declares but never fills such an array) */
int64_t or_vertex_normals_fp64(const or_grid *G, int lim, const float *field, double *out);

/* Reference-form soup (one vertex per triangle corner, per-cube interpolation) and the
   reference's quantised weld: vertices sorted by quantised position. */
int or_mc_soup(const or_grid *G, int lim, const float *field, float **soup, int64_t *n_soup);
int or_quantised_weld(const float *soup, int64_t n_soup, or_mesh *m);

/* closest atom per vertex (sorted-array index, as Kernels.cu:547-604) */
void or_closest_atom(const or_grid *G, const int *cell_start, const int *cell_end,
                     const float *sorted_xyzr, const float *vert, int64_t n_vert, int *out);

/* Laplacian smoothing with the reference's neighbour multiset. gauss_seidel=1: in place in
   vertex order (SmoothMesh.cpp:96-124); 0: Jacobi (every vertex reads the previous pass). */
void or_smooth(float *vert, int64_t n_vert, const int *tri, int64_t n_tri, int times, int gauss_seidel);

/* drop triangles with a repeated index (CudaSurf.cu:820-829); returns the kept count */
int64_t or_filter_degenerate(int *tri, int64_t n_tri);

/* Whole pipeline on one grid (one slice covering the whole grid).
   weld: 0 = edge-id weld, 1 = quantised weld.  smooth_gs: 0 Jacobi, 1 Gauss-Seidel.
   lim_mode: 0 = every cube with valid corners (lim = g-1), 1 = reference window (g-2r-1).
   atom_ids_sorted: 1 = sorted-array indices (reference), 0 = input indices. */
typedef struct {
    int weld;
    int smooth_gs;
    int lim_mode;
    int atom_ids_sorted;
} or_options;

int or_compute_ses(const float *pos, const float *rad, int n, float reso, int smooth_times,
                   const or_options *opt, or_mesh *out);

/* FP64 evaluation of passes 1+2 on the same grid (positions and radii widened to double). */
void or_field_fp64(const or_grid *G, const float *pos, const float *rad, int n,
                   uint8_t *labels64, double *field64);

/* number of threads OpenMP will use */
int or_num_threads(void);

#ifdef __cplusplus
}
#endif
#endif
