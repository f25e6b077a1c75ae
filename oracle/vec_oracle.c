/*
 * vec_oracle.c -- CPU restatement of the reference's L1 / L2 vector scans.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the shipped library links, loads or
 * calls this file; only tests/, __graft_entry__.smoke() and bench.py's CPU
 * baseline leg use it, and only as the checker.
 *
 * Follows region_query of dbscan.c:563-780 (built with -D_SCAN_L2 or
 * -D_SCAN_L1 and the makefile's -march flags, i.e. its AVX2 + FMA branch):
 *   - for the query point p and each candidate x, in float32, over the
 *     dimensions k = 0 .. dims-1 in order:
 *       L2: acc = fma(p_k - x_k, p_k - x_k, acc)   (dbscan.c:648-659)
 *       L1: acc = |p_k - x_k| + acc                (dbscan.c:660-664)
 *     starting from acc = +0 (dbscan.c:640);
 *   - x is a neighbour when acc <= eps * eps (L2, the product in float,
 *     dbscan.c:577, 676-677) or acc <= eps (L1, dbscan.c:678-679).
 * This file computes the accumulator of every ordered pair (query row p,
 * candidate column x) in exactly that order, with fmaf from the C library
 * (correctly rounded) for the L2 step; the threshold test is left to the
 * caller.
 *
 * Pinned by tests/test_oracle.py against: the compiled reference's
 * cluster_dbscan_pca / cluster_dbscan_kmerL1 / cluster_dbscan_kmerL2 /
 * adaptive_clustering_PCA / adaptive_clustering_kmer_L1/L2 outputs
 * (tests/golden/vectors), among them a case whose neighbourhoods differ
 * between the fused and the unfused L2 step; exact integer L1 / L2 distances
 * of integer-valued vectors (numpy int64); float64 distances within rounding.
 */
#include <math.h>
#include <pthread.h>
#include <stdint.h>
#include <stdlib.h>

typedef struct {
  const float *const *values; /* values[k][i] */
  int dims, n, norm;
  const int32_t *rows;
  int n_rows;
  float *out; /* [n_rows][n] */
  int tid, nthreads;
} vec_job;

static void *vec_worker(void *arg) {
  vec_job *jb = (vec_job *)arg;
  for (int r = jb->tid; r < jb->n_rows; r += jb->nthreads) {
    const int p = jb->rows[r];
    float *o = jb->out + (int64_t)r * jb->n;
    for (int x = 0; x < jb->n; x++) {
      float acc = 0.0f;
      for (int k = 0; k < jb->dims; k++) {
        const float d = jb->values[k][p] - jb->values[k][x];
        if (jb->norm == 2)
          acc = fmaf(d, d, acc);
        else
          acc = fabsf(d) + acc;
      }
      o[x] = acc;
    }
  }
  return NULL;
}

/* Accumulators of the query rows rows[0..n_rows) against every point:
 * out[r * n + x].  norm is 1 (L1) or 2 (L2). */
int oracle_vec_rows(const float *const *values, int dims, int n, int norm, const int32_t *rows,
                    int n_rows, float *out, int nthreads) {
  if (norm != 1 && norm != 2) return 1;
  if (nthreads < 1) nthreads = 1;
  pthread_t *th = (pthread_t *)malloc(sizeof(pthread_t) * (size_t)nthreads);
  vec_job *jobs = (vec_job *)malloc(sizeof(vec_job) * (size_t)nthreads);
  for (int t = 0; t < nthreads; t++) {
    jobs[t] = (vec_job){values, dims, n, norm, rows, n_rows, out, t, nthreads};
    pthread_create(&th[t], NULL, vec_worker, &jobs[t]);
  }
  for (int t = 0; t < nthreads; t++) pthread_join(th[t], NULL);
  free(th);
  free(jobs);
  return 0;
}
