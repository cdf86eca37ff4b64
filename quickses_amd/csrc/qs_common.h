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

#endif
