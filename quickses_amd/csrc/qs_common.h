/* Shared host/device definitions for the SES engine. */
#ifndef QS_COMMON_H
#define QS_COMMON_H

#include <stdint.h>

#define QS_PROBE 1.4f         /* probe radius, Angstrom (Kernels.cu:26) */
#define QS_EPS 0.001f         /* weld / outside tolerance (Kernels.cu:27) */
#define QS_EMPTY (2147483647 - 1) /* empty-cell marker in cell ranges (Kernels.cu:25) */

/* voxel labels written by field pass 1 */
#define QS_OUT 0   /* no atom within r + probe: field = probe      */
#define QS_SHELL 1 /* within r + probe of an atom, inside none      */
#define QS_IN 2    /* within r - dx of some atom: field = -dx       */

/* One SES grid: neighbour grid and SES grid share the origin (CudaSurf.cu:564-574). */
struct QsGrid {
    float ox, oy, oz; /* origin */
    float dx;         /* SES voxel size */
    float dxn;        /* neighbour-cell size = probe + max radius */
    int g;            /* SES grid points per axis */
    int gn;           /* neighbour cells per axis */
    int r;            /* pass-2 half window = ceil(probe / dx) */
    int lim;          /* cubes (i,j,k) with i,j,k < lim are polygonised */
};

/* Field pass 2: bit n is set when a window minimum of squared integer norm n can only be
   attained by offsets of norm n (no longer offset is within rounding of it), n < 256. */
struct QsNormFlags {
    uint64_t w[4];
};

/* One molecule of a batch: its grid and where its rows live in the concatenated arrays. */
struct QsMol {
    QsGrid G;
    uint32_t atom_off, n_atoms; /* atoms (input order and cell order) */
    uint32_t cell_off, W;       /* first neighbour cell; 64-voxel words per grid row */
    uint64_t word_off;          /* first row word (bit masks, slot masks, counts) */
    uint64_t vox_off;           /* first voxel (field, cube records) */
    uint32_t vert_off, n_vert;  /* vertices (set after counting) */
    uint32_t tri_off, n_tri;    /* triangles (set after counting) */
    int rows_per_wave;          /* field pass 1, row kernel: rows per wavefront */
    int slab_rows;              /* field pass 1, slab kernel: y rows per block */
    /* x planes each stage produces (the whole grid, [0, g), unless the grid runs in x-slabs):
       field pass 1 labels [p1_lo, p1_hi), field pass 2 evaluates [p2_lo, p2_hi), the
       marching-cubes count covers grid points [mc_lo, mc_hi), and cubes with origin
       x < cube_hi are polygonised.  word_off and vox_off address global plane x, so a slab
       whose arrays start at plane p1_lo has offsets that wrap below zero. */
    int p1_lo, p1_hi, p2_lo, p2_hi, mc_lo, mc_hi, cube_hi;
    /* first of the molecule's field-pass-2 tile flags (in pass-2 tile order; ~0: none kept) */
    uint32_t tile_off;
};

/* Coarse indexes of prefix arrays: coarse[k] = the segment that holds item k << shift, with
   shift QS_COARSE_BLOCKS for block prefixes and QS_COARSE_ITEMS for row-word and vertex
   prefixes. */
#define QS_COARSE_BLOCKS 6
#define QS_COARSE_ITEMS 12

/* A launch over a batch: block b belongs to molecule mid[j] (or j) where
   bpre[j] <= b < bpre[j+1]; coarse indexes bpre. */
struct QsBatch {
    const QsMol *mols;
    const uint32_t *bpre;
    const uint32_t *coarse;
    const int *mid;
    int n;
};

#endif
