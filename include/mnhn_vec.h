/*
 * mnhn_vec.h -- C ABI of the gfx950 epsilon-neighbour engine for float vectors
 * (PCA projections, k-mer frequency vectors): the distance scans behind the
 * reference's L1 / L2 DBSCAN (dbscan.c:563-780, built with -D_SCAN_L1 /
 * -D_SCAN_L2).
 *
 * The reference computes, for a query point p and every point x, the float32
 * accumulator over the dimensions k = 0 .. dims-1 in order
 *   L2:  acc = fma(p_k - x_k, p_k - x_k, acc)     (its AVX2 + FMA path)
 *   L1:  acc = |p_k - x_k| + acc
 * and keeps x when acc <= eps * eps (L2, the product taken in float) or
 * acc <= eps (L1).  The engine evaluates the same float32 operations in the
 * same order for every unordered pair (the accumulator is symmetric, since
 * |p - x| = |x - p| exactly), so its lists are bit-for-bit the reference's.
 */
#ifndef MNHN_VEC_H
#define MNHN_VEC_H

#include <stdint.h>

#include "mnhn_sw.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct mnhn_vec_ctx mnhn_vec_ctx;

/* Norms. */
#define MNHN_NORM_L1 1
#define MNHN_NORM_L2 2

/* Epsilon lists of float vectors: row i lists every j != i with acc(i, j) <=
 * thr, columns ascending, with acc itself (float) in val. */
typedef struct {
  int32_t n_rows;
  int32_t norm;
  float thr;
  int32_t reserved;
  int64_t nnz;
  int64_t *row_ptr; /* [n_rows + 1] */
  int32_t *col;     /* [nnz] */
  float *val;       /* [nnz] */
} mnhn_vcsr;

/* Upload n points of dims dimensions given dimension-major, as the
 * reference's dataset.values (values[k][i], dataset.h:27-35). */
int mnhn_vec_create(const float *const *values, int dims, int n, mnhn_vec_ctx **out);
void mnhn_vec_destroy(mnhn_vec_ctx *ctx);

/* All unordered pairs once, both directions listed.  thr is compared with
 * the accumulator as is: pass eps * eps (float) for L2, eps for L1. */
int mnhn_vec_eps_csr(mnhn_vec_ctx *ctx, int norm, float thr, mnhn_vcsr *out);
/* The only way to release a mnhn_vcsr: the arrays of large lists are pinned
 * host memory, which free() must not release. */
void mnhn_vcsr_free(mnhn_vcsr *csr);

/* Kernel milliseconds and pair count of the last mnhn_vec_eps_csr. */
int mnhn_vec_last_timing(const mnhn_vec_ctx *ctx, double *kernel_ms, double *pairs);

#ifdef __cplusplus
}
#endif

#endif /* MNHN_VEC_H */
