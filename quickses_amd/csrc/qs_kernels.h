/* Launchers of the HIP kernels (qs_kernels.hip). */
#ifndef QS_KERNELS_H
#define QS_KERNELS_H

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "qs_common.h"

void qs_launch_bbox(const float *pos, const float *rad, int n, float *part, int nparts, float *out, hipStream_t s);
void qs_launch_hash_count(const float *pos, int n, QsGrid G, int *key, uint32_t *count, hipStream_t s);
void qs_launch_scatter(const int *key, int n, const uint32_t *start, uint32_t *cursor, int *perm, hipStream_t s);
void qs_launch_cell_finalize(const uint32_t *start, const uint32_t *count, int ncell, int *perm, int2 *cse, hipStream_t s);
void qs_launch_gather_atoms(const int *perm, const float *pos, const float *rad, int n, float4 *sx, hipStream_t s);
void qs_launch_pass1(QsGrid G, const int2 *cse, const float4 *sx, uint8_t *labels, int x0, int x1, hipStream_t s);
size_t qs_pass2_lds_bytes(int r);
void qs_launch_pass2_tile(QsGrid G, const uint8_t *labels, float *field, const int *offs, const int *norm_start,
                          const int *enum_end, int x0, int x1, hipStream_t s);
void qs_launch_pass2_direct(QsGrid G, const uint8_t *labels, float *field, int x0, int x1, hipStream_t s);
unsigned qs_mc_blocks(uint64_t npoints);
void qs_launch_mc_count(QsGrid G, const float *F, uint16_t *pinfo, uint32_t *bs_tri, uint32_t *bs_vert, uint64_t p0,
                        uint64_t p1, hipStream_t s);
void qs_launch_mc_verts(QsGrid G, const float *F, const uint16_t *pinfo, const uint32_t *bo_vert, uint32_t *vbase,
                        float *vert, uint64_t p0, uint64_t p1, hipStream_t s);
void qs_launch_mc_tris(QsGrid G, const float *F, const uint16_t *pinfo, const uint32_t *vbase, const uint32_t *bo_tri,
                       int *tri, uint64_t p0, uint64_t p1, hipStream_t s);
void qs_launch_closest_atom(QsGrid G, const int2 *cse, const float4 *sx, const int *perm, const float *vert, uint32_t nv,
                            int sorted_ids, int *out, hipStream_t s);
void qs_launch_inc_count(const int *tri, uint32_t ncorner, uint32_t *deg, hipStream_t s);
void qs_launch_inc_fill(const int *tri, uint32_t ncorner, const uint32_t *start, uint32_t *cursor, uint32_t *inc,
                        hipStream_t s);
void qs_launch_inc_sort(const uint32_t *start, const uint32_t *deg, uint32_t nv, uint32_t *inc, hipStream_t s);
void qs_launch_jacobi(const int *tri, const uint32_t *start, const uint32_t *deg, const uint32_t *inc, uint32_t nv,
                      const float *src, float *dst, hipStream_t s);
void qs_launch_tri_keep(const int *tri, uint32_t nt, uint32_t *keep, hipStream_t s);
void qs_launch_tri_compact(const int *tri, uint32_t nt, const uint32_t *pos, int *out, hipStream_t s);
uint64_t qs_scan_tiles(uint64_t n);
void qs_launch_tile_reduce(const uint32_t *in, uint64_t n, uint32_t *sums, hipStream_t s);
void qs_launch_tile_scan(const uint32_t *in, uint32_t *out, uint64_t n, const uint32_t *offsets, uint32_t *total,
                         hipStream_t s);

#endif
