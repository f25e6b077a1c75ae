/*
 * sw_oracle.c -- CPU restatement of the reference's Smith-Waterman distance.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the shipped library links, loads or
 * calls this file; only tests/, __graft_entry__.smoke() and bench.py's CPU
 * baseline leg use it, and only as the checker.
 *
 * Follows smith-waterman.c:16-81 of the reference:
 *   - scoring: match +4, mismatch -4, linear gap 3   (smith-waterman.c:4-14)
 *   - DP cells (i, j) for i in [1, la], j in [1, lb]; A[0] and B[0] are never
 *     compared (smith-waterman.c:56-58)
 *   - the reference reads A[la] and B[lb], which are the NUL terminators of
 *     the C strings (dataset.c:332-333): they match each other and mismatch
 *     every base.  Here the terminator is an explicit sentinel, so the inputs
 *     need not be NUL-terminated.
 *   - H = max(0, diag + s, left - 3, up - 3); result = 4*max(la, lb) - max H
 *     (smith-waterman.c:42-46, 62-80)
 * The restatement keeps one row of H (O(la) memory) instead of the full
 * (la+1)(lb+1) matrix; the recurrence and its evaluation order are unchanged.
 *
 * Pinned by tests/test_oracle.py against: golden pair vectors written by the
 * compiled reference (tests/golden/pairs_golden.json), brute force over all
 * {A,C} strings up to length 6 (tests/golden), and the closed-form invariants
 * (symmetry, d(x,x) = 0, 4|la-lb| <= d <= 4 max(la,lb), length-1 pairs).
 */
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include <pthread.h>

#define ORACLE_MATCH 4
#define ORACLE_GAP 3

/* Byte of x at DP index k (1-based), where index len is the terminator. */
static inline int oracle_sym(const char *x, size_t len, size_t k) {
  return k < len ? (unsigned char)x[k] : 0x100; /* 0x100 = terminator */
}

unsigned int oracle_sw_distance(const char *a, const char *b, size_t la,
                                size_t lb) {
  int optimum = (int)(la > lb ? la : lb) * ORACLE_MATCH;
  if (la == 0 || lb == 0) return (unsigned int)optimum;
  int *row = (int *)calloc(la + 1, sizeof(int)); /* row[i] = H[i][j-1] */
  int best = 0;
  for (size_t j = 1; j <= lb; j++) {
    int bj = oracle_sym(b, lb, j);
    int diag = 0;  /* H[0][j-1] */
    int left = 0;  /* H[0][j]   */
    for (size_t i = 1; i <= la; i++) {
      int up_prev = row[i]; /* H[i][j-1] in reference orientation */
      int s = (oracle_sym(a, la, i) == bj) ? ORACLE_MATCH : -ORACLE_MATCH;
      int h = diag + s;
      int g1 = left - ORACLE_GAP;
      int g2 = up_prev - ORACLE_GAP;
      if (h < g1) h = g1;
      if (h < g2) h = g2;
      if (h < 0) h = 0;
      row[i] = h;
      if (best < h) best = h;
      diag = up_prev;
      left = h;
    }
  }
  free(row);
  return (unsigned int)(optimum - best);
}

/* ---- batch helpers (pthreads) used by the tests and the CPU baseline ---- */

typedef struct {
  const char *const *seqs;
  const size_t *lens;
  const int32_t *ij; /* [2m] pairs, or NULL for all unique pairs */
  int64_t m;
  int n;
  int32_t *out;
  int tid, nthreads;
} oracle_job;

static void *oracle_pairs_worker(void *arg) {
  oracle_job *jb = (oracle_job *)arg;
  for (int64_t k = jb->tid; k < jb->m; k += jb->nthreads) {
    int i = jb->ij[2 * k], j = jb->ij[2 * k + 1];
    jb->out[k] = (int32_t)oracle_sw_distance(jb->seqs[i], jb->seqs[j],
                                             jb->lens[i], jb->lens[j]);
  }
  return NULL;
}

/* d(seq[ij[2k]], seq[ij[2k+1]]) for k < m, spread over nthreads threads. */
int oracle_sw_pairs(const char *const *seqs, const size_t *lens,
                    const int32_t *ij, int64_t m, int32_t *out, int nthreads) {
  if (nthreads < 1) nthreads = 1;
  pthread_t *th = (pthread_t *)malloc(sizeof(pthread_t) * nthreads);
  oracle_job *jobs = (oracle_job *)malloc(sizeof(oracle_job) * nthreads);
  for (int t = 0; t < nthreads; t++) {
    jobs[t] = (oracle_job){seqs, lens, ij, m, 0, out, t, nthreads};
    pthread_create(&th[t], NULL, oracle_pairs_worker, &jobs[t]);
  }
  for (int t = 0; t < nthreads; t++) pthread_join(th[t], NULL);
  free(th);
  free(jobs);
  return 0;
}

static void *oracle_matrix_worker(void *arg) {
  oracle_job *jb = (oracle_job *)arg;
  int n = jb->n;
  /* rows are dealt round-robin; row i computes j > i and mirrors */
  for (int i = jb->tid; i < n; i += jb->nthreads) {
    jb->out[(int64_t)i * n + i] = (int32_t)oracle_sw_distance(
        jb->seqs[i], jb->seqs[i], jb->lens[i], jb->lens[i]);
    for (int j = i + 1; j < n; j++) {
      int32_t d = (int32_t)oracle_sw_distance(jb->seqs[i], jb->seqs[j],
                                              jb->lens[i], jb->lens[j]);
      jb->out[(int64_t)i * n + j] = d;
      jb->out[(int64_t)j * n + i] = d;
    }
  }
  return NULL;
}

/* Full n x n distance matrix (row-major int32), unique pairs computed once. */
int oracle_sw_matrix(const char *const *seqs, const size_t *lens, int n,
                     int32_t *out, int nthreads) {
  if (nthreads < 1) nthreads = 1;
  pthread_t *th = (pthread_t *)malloc(sizeof(pthread_t) * nthreads);
  oracle_job *jobs = (oracle_job *)malloc(sizeof(oracle_job) * nthreads);
  for (int t = 0; t < nthreads; t++) {
    jobs[t] = (oracle_job){seqs, lens, NULL, 0, n, out, t, nthreads};
    pthread_create(&th[t], NULL, oracle_matrix_worker, &jobs[t]);
  }
  for (int t = 0; t < nthreads; t++) pthread_join(th[t], NULL);
  free(th);
  free(jobs);
  return 0;
}
