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
void qs_launch_pass1(QsGrid G, const uint32_t *cstart, const float4 *sx, uint64_t *outm, uint64_t *shellm, int W, int x0,
                     int x1, hipStream_t s);
void qs_launch_pass1_direct(QsGrid G, const int2 *cse, const float4 *sx, uint64_t *outm, uint64_t *shellm, int W, int x0,
                            int x1, hipStream_t s);
size_t qs_pass2_lds_bytes(int r);
int qs_launch_pass2(QsGrid G, const uint64_t *outm, const uint64_t *shellm, int W, float *field, uint32_t *negm,
                    uint32_t *zerom, const int *offs, const int *norm_start, const int *enum_end, int x0, int x1, bool direct,
                    hipStream_t s);
void qs_launch_field_signs(const float *F, int g, int W, uint32_t *negm, uint32_t *zerom, hipStream_t s);
void qs_launch_labels_to_masks(const uint8_t *labels, int g, int W, uint64_t *outm, uint64_t *shellm, hipStream_t s);
void qs_launch_mc_count(QsGrid G, const uint64_t *N, const uint64_t *Zm, int W, uint64_t *S4, uint32_t *vcnt, uint32_t *tcnt,
                        int x0, int x1, hipStream_t s);
void qs_launch_mc_flag(const uint32_t *vcnt, const uint32_t *tcnt, uint64_t nw, uint32_t *flag, hipStream_t s);
void qs_launch_mc_active(const uint32_t *vcnt, const uint32_t *tcnt, const uint32_t *apos, uint64_t nw, uint32_t *list,
                         hipStream_t s);
void qs_launch_mc_emit(QsGrid G, const uint64_t *N, const uint64_t *Zm, const uint64_t *S4, const uint32_t *vbw, int W,
                       const float *F, const uint32_t *list, uint32_t nact, const uint32_t *tbw, float *vert, uint64_t *vkey,
                       int *tri, uint8_t *cidx, uint32_t *tbase, hipStream_t s);
void qs_launch_smooth(QsGrid G, const uint64_t *N, const uint64_t *Zm, const uint64_t *S4, const uint32_t *vbw, int W,
                      const uint64_t *vkey, uint32_t nv, const uint8_t *cidx, const uint32_t *tbase, const int *tri,
                      const float *src, float *dst, hipStream_t s);
void qs_launch_closest_atom(QsGrid G, const int2 *cse, const float4 *sx, const int *perm, const float *vert, uint32_t nv,
                            int sorted_ids, int *out, hipStream_t s);
void qs_launch_tri_keep(const int *tri, uint32_t nt, uint32_t *keep, hipStream_t s);
void qs_launch_tri_compact(const int *tri, uint32_t nt, const uint32_t *pos, int *out, hipStream_t s);
uint64_t qs_scan_tiles(uint64_t n);
void qs_launch_tile_reduce(const uint32_t *in, uint64_t n, uint32_t *sums, hipStream_t s);
void qs_launch_tile_scan(const uint32_t *in, uint32_t *out, uint64_t n, const uint32_t *offsets, uint32_t *total,
                         hipStream_t s);

#endif
