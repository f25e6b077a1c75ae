/*
 * cpu_allpairs.c -- CPU timing harness for the all-pairs SW distance.
 *
 * TEST/BENCH INFRASTRUCTURE ONLY (the CPU baseline leg of bench.py and the
 * oracle tests).  It is compiled twice:
 *   - into oracle/liboracle_sw.so, calling oracle_sw_distance (the C
 *     restatement in sw_oracle.c);
 *   - with -DCPU_ALLPAIRS_REF into oracle/_ref/libref_sw.so, calling the
 *     reference's own score() (smith-waterman.c:16-81) compiled from its
 *     source.
 *
 * Two modes, matching how the reference spends its time:
 *   ordered = 1: every ordered pair (i, j), one call per pair, each call given
 *                a caller-owned (maxlen+1)^2 work matrix, as the CPU
 *                region_query does (dbscan.c:525-561): N^2 calls per pass.
 *   ordered = 0: unique pairs i < j only, rows dealt round-robin to threads.
 */
#include <pthread.h>
#include <stdint.h>
#include <stdlib.h>
#include <time.h>

#ifdef CPU_ALLPAIRS_REF
unsigned int score(char *A, char *B, size_t len_a, size_t len_b, int *work);
static inline unsigned int pair_distance(const char *a, const char *b,
                                         size_t la, size_t lb, int *work) {
  return score((char *)a, (char *)b, la, lb, work);
}
#define CPU_ALLPAIRS_NAME ref_cpu_allpairs
#else
unsigned int oracle_sw_distance(const char *a, const char *b, size_t la,
                                size_t lb);
static inline unsigned int pair_distance(const char *a, const char *b,
                                         size_t la, size_t lb, int *work) {
  (void)work;
  return oracle_sw_distance(a, b, la, lb);
}
#define CPU_ALLPAIRS_NAME oracle_cpu_allpairs
#endif

typedef struct {
  const char *const *seqs;
  const size_t *lens;
  int n, ordered, tid, nthreads;
  size_t maxlen;
  int32_t *out; /* n*n row-major, or NULL */
  uint64_t checksum;
} allpairs_job;

static void *allpairs_worker(void *arg) {
  allpairs_job *jb = (allpairs_job *)arg;
  int n = jb->n;
  int *work = (int *)malloc(sizeof(int) * (jb->maxlen + 1) * (jb->maxlen + 1));
  uint64_t cs = 0;
  for (int i = jb->tid; i < n; i += jb->nthreads) {
    int j0 = jb->ordered ? 0 : i + 1;
    for (int j = j0; j < n; j++) {
      /* argument order as in region_query: score(seq[j], seq[point], ...) */
      unsigned int d = pair_distance(jb->seqs[j], jb->seqs[i], jb->lens[j],
                                     jb->lens[i], work);
      cs += d;
      if (jb->out) jb->out[(int64_t)i * n + j] = (int32_t)d;
    }
  }
  free(work);
  jb->checksum = cs;
  return NULL;
}

/* Returns wall seconds; *checksum = sum of all computed distances. */
double CPU_ALLPAIRS_NAME(const char *const *seqs, const size_t *lens, int n,
                         int nthreads, int ordered, int32_t *out,
                         uint64_t *checksum) {
  if (nthreads < 1) nthreads = 1;
  size_t maxlen = 0;
  for (int i = 0; i < n; i++)
    if (lens[i] > maxlen) maxlen = lens[i];
  pthread_t *th = (pthread_t *)malloc(sizeof(pthread_t) * nthreads);
  allpairs_job *jobs = (allpairs_job *)malloc(sizeof(allpairs_job) * nthreads);
  struct timespec t0, t1;
  clock_gettime(CLOCK_MONOTONIC, &t0);
  for (int t = 0; t < nthreads; t++) {
    jobs[t] = (allpairs_job){seqs, lens, n, ordered, t, nthreads, maxlen, out, 0};
    pthread_create(&th[t], NULL, allpairs_worker, &jobs[t]);
  }
  uint64_t cs = 0;
  for (int t = 0; t < nthreads; t++) {
    pthread_join(th[t], NULL);
    cs += jobs[t].checksum;
  }
  clock_gettime(CLOCK_MONOTONIC, &t1);
  free(th);
  free(jobs);
  if (checksum) *checksum = cs;
  return (double)(t1.tv_sec - t0.tv_sec) + 1e-9 * (double)(t1.tv_nsec - t0.tv_nsec);
}
