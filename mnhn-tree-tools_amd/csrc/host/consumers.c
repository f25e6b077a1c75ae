/*
 * consumers.c -- the reference's other Smith-Waterman consumers on the gfx950
 * engine: the SW filter (filter.c), the comparison statistics (comparison.c)
 * and dataset_to_fasta (dataset.c:85-110).
 *
 * Every distance comes from mnhn_sw_block over an engine built for the two
 * data sets; the host code keeps the reference's arithmetic, including the
 * accumulator types of comparison.c:
 *   mean  (comparison.c:113-150, 152-188): an int sum of the unsigned
 *     distances (wraps modulo 2^32), divided by n1 * n2 as a double;
 *   sigma (comparison.c:190-238, 240-277): an int sum to which each
 *     (d - mean)^2 is added through a double and truncated back to int (on
 *     x86-64 a value outside the int range becomes INT_MIN), then
 *     sqrt(sum) / (n1 + n2);
 *   silhouette (comparison.c:20-75, 279-313): per row of the first set the
 *     exact sums of its distances to the first set (itself included) and to
 *     the second, a = internal / (n1 - 1), b = external / n2, s as in :65-72,
 *     and the mean of s over the rows.
 * The reference hands rows to n_threads threads in a racy order, so with
 * several threads its sums are added in an order that varies from run to run;
 * this code adds them in row order, which is the reference's result for one
 * thread (the distances themselves do not depend on threads).
 */
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <unistd.h>

#include "mnhn_host.h"

static void die(const char *what, int rc) {
  printf("%s failed: %s\n", what, mnhn_sw_strerror(rc));
  fflush(stdout);
  _exit(1);
}

/* Engine over a's sequences followed by b's. */
static mnhn_sw_ctx *engine_for(dataset a, dataset b) {
  const int n = a.n_values + b.n_values;
  const char **seqs = (const char **)malloc(sizeof(char *) * (size_t)(n ? n : 1));
  size_t *lens = (size_t *)malloc(sizeof(size_t) * (size_t)(n ? n : 1));
  for (int i = 0; i < a.n_values; i++) {
    seqs[i] = a.sequences[i];
    lens[i] = a.sequence_lengths[i];
  }
  for (int i = 0; i < b.n_values; i++) {
    seqs[a.n_values + i] = b.sequences[i];
    lens[a.n_values + i] = b.sequence_lengths[i];
  }
  mnhn_sw_ctx *sw = NULL;
  int rc = mnhn_sw_create(seqs, lens, n, NULL, &sw);
  free(seqs);
  free(lens);
  if (rc) die("mnhn_sw_create", rc);
  return sw;
}

/* Calls fn(row, dist[0..nc)) for every row of a against all of b, in row
 * order, computing blocks of rows at a time. */
typedef void (*row_fn)(void *arg, int row, const int32_t *dist, int nc);

static void for_each_row(dataset a, dataset b, row_fn fn, void *arg) {
  const int n1 = a.n_values, n2 = b.n_values;
  if (n1 == 0) return;
  mnhn_sw_ctx *sw = engine_for(a, b);
  int32_t *cols = (int32_t *)malloc(sizeof(int32_t) * (size_t)(n2 ? n2 : 1));
  for (int k = 0; k < n2; k++) cols[k] = n1 + k;
  const size_t padded = (size_t)((n2 + 127) / 128) * 128;
  size_t chunk = padded ? ((size_t)1 << 27) / padded : (size_t)n1;
  if (chunk < 1) chunk = 1;
  if (chunk > (size_t)n1) chunk = (size_t)n1;
  int32_t *rows = (int32_t *)malloc(sizeof(int32_t) * chunk);
  int32_t *dist = (int32_t *)malloc(sizeof(int32_t) * chunk * (size_t)(n2 ? n2 : 1));
  for (int r0 = 0; r0 < n1; r0 += (int)chunk) {
    const int nr = (int)((size_t)(n1 - r0) < chunk ? (size_t)(n1 - r0) : chunk);
    for (int r = 0; r < nr; r++) rows[r] = r0 + r;
    if (n2 > 0) {
      int rc = mnhn_sw_block(sw, rows, nr, cols, n2, dist);
      if (rc) die("mnhn_sw_block", rc);
    }
    for (int r = 0; r < nr; r++) fn(arg, r0 + r, dist + (size_t)r * n2, n2);
  }
  free(rows);
  free(dist);
  free(cols);
  mnhn_sw_destroy(sw);
}

/* ---- filter.c ---- */

static dataset copy_selected(dataset in_set, const char *keep) {
  dataset out;
  memset(&out, 0, sizeof(out));
  const int n = in_set.n_values;
  out.sequences = (char **)malloc(sizeof(char *) * (size_t)(n ? n : 1));
  out.sequence_lengths = (size_t *)malloc(sizeof(size_t) * (size_t)(n ? n : 1));
  int counter = 0;
  for (int i = 0; i < n; i++) {
    if (!keep[i]) continue;
    const size_t L = in_set.sequence_lengths[i];
    out.sequences[counter] = (char *)malloc(L + 1);
    memcpy(out.sequences[counter], in_set.sequences[i], L);
    out.sequences[counter][L] = 0;
    out.sequence_lengths[counter] = L;
    if (L > out.max_sequence_length) out.max_sequence_length = L;
    counter++;
  }
  out.n_values = counter;
  return out;
}

dataset filter_ds_by_size(dataset in_set, size_t size, size_t plus, size_t minus) {
  char *keep = (char *)calloc((size_t)(in_set.n_values ? in_set.n_values : 1), 1);
  for (int i = 0; i < in_set.n_values; i++) {
    const size_t L = in_set.sequence_lengths[i];
    /* size_t arithmetic as in filter.c:31-32 (size - minus wraps when minus > size) */
    keep[i] = L <= size + plus && L >= size - minus;
  }
  dataset out = copy_selected(in_set, keep);
  free(keep);
  return out;
}

dataset filter_by_SW_distance(dataset in_set, char *target, size_t target_length,
                              unsigned int max_distance, int n_threads) {
  (void)n_threads; /* one device pass replaces the threads of filter.c:96-108 */
  const int n = in_set.n_values;
  char *keep = (char *)calloc((size_t)(n ? n : 1), 1);
  if (n > 0) {
    dataset t;
    memset(&t, 0, sizeof(t));
    t.sequences = &target;
    t.sequence_lengths = &target_length;
    t.n_values = 1;
    mnhn_sw_ctx *sw = engine_for(in_set, t);
    int32_t *dist = (int32_t *)malloc(sizeof(int32_t) * (size_t)(n + 1));
    int rc = mnhn_sw_one_vs_all(sw, n, dist);
    if (rc) die("mnhn_sw_one_vs_all", rc);
    for (int i = 0; i < n; i++) keep[i] = max_distance >= (unsigned int)dist[i];
    free(dist);
    mnhn_sw_destroy(sw);
  }
  dataset out = copy_selected(in_set, keep);
  free(keep);
  return out;
}

/* ---- dataset.c:85-110 ---- */

void dataset_to_fasta(FILE *f, dataset ds) {
  for (int j = 0; j < ds.n_values; j++) {
    fprintf(f, ">sequence_%i\n", j);
    const size_t L = ds.sequence_lengths[j];
    for (size_t k = 0; k < L; k++) {
      if (k != 0 && k % 50 == 0) fputc('\n', f);
      fputc(ds.sequences[j][k], f);
    }
    fprintf(f, "\n");
  }
}

/* ---- comparison.c ---- */

typedef struct {
  unsigned long *matrix;
} matrix_arg;

static void matrix_row(void *arg, int row, const int32_t *dist, int nc) {
  unsigned long *m = ((matrix_arg *)arg)->matrix + (size_t)row * (size_t)nc;
  for (int k = 0; k < nc; k++) m[k] = (unsigned long)(uint32_t)dist[k];
}

unsigned long *smith_waterman_distances_matrix(dataset ds_one, dataset ds_two) {
  matrix_arg a;
  a.matrix = (unsigned long *)malloc(sizeof(unsigned long) * (size_t)ds_one.n_values *
                                     (size_t)(ds_two.n_values ? ds_two.n_values : 1));
  for_each_row(ds_one, ds_two, matrix_row, &a);
  return a.matrix;
}

typedef struct {
  uint32_t sum; /* the reference's int accumulator, modulo 2^32 */
} mean_arg;

static void mean_row(void *arg, int row, const int32_t *dist, int nc) {
  (void)row;
  mean_arg *m = (mean_arg *)arg;
  for (int k = 0; k < nc; k++) m->sum += (uint32_t)dist[k];
}

double mean_from_smith_waterman_datasets(dataset ds_one, dataset ds_two, size_t n_threads) {
  (void)n_threads;
  mean_arg m = {0};
  for_each_row(ds_one, ds_two, mean_row, &m);
  const double local = (double)(int32_t)m.sum;
  return local / ((double)ds_one.n_values * (double)ds_two.n_values);
}

typedef struct {
  double mean;
  int32_t sum;
} sigma_arg;

/* x86-64 cvttsd2si: truncation, INT_MIN for NaN or out-of-range values */
static int32_t to_int_x86(double t) {
  if (t > -2147483649.0 && t < 2147483648.0) return (int32_t)t;
  return INT32_MIN;
}

static void sigma_row(void *arg, int row, const int32_t *dist, int nc) {
  (void)row;
  sigma_arg *s = (sigma_arg *)arg;
  for (int k = 0; k < nc; k++) {
    const double v = (double)(uint32_t)dist[k];
    s->sum = to_int_x86((double)s->sum + (v - s->mean) * (v - s->mean));
  }
}

double sigma_from_smith_waterman_datasets(double mean, dataset ds_one, dataset ds_two,
                                          size_t n_threads) {
  (void)n_threads;
  sigma_arg s = {mean, 0};
  for_each_row(ds_one, ds_two, sigma_row, &s);
  const double sigma = (double)s.sum;
  return sqrt(sigma) / (ds_one.n_values + ds_two.n_values);
}

typedef struct {
  int n1;
  int64_t *internal; /* per row of ds_one, the exact sums of distances */
  int64_t *external;
  int phase;
} sil_arg;

static void sil_row(void *arg, int row, const int32_t *dist, int nc) {
  sil_arg *s = (sil_arg *)arg;
  int64_t sum = 0;
  for (int k = 0; k < nc; k++) sum += (uint32_t)dist[k];
  (s->phase == 0 ? s->internal : s->external)[row] = sum;
}

double silhouette_from_smith_waterman_datasets(dataset ds_one, dataset ds_two, size_t n_threads) {
  (void)n_threads;
  const int n1 = ds_one.n_values;
  sil_arg s;
  s.n1 = n1;
  s.internal = (int64_t *)calloc((size_t)(n1 ? n1 : 1), sizeof(int64_t));
  s.external = (int64_t *)calloc((size_t)(n1 ? n1 : 1), sizeof(int64_t));
  s.phase = 0;
  for_each_row(ds_one, ds_one, sil_row, &s);
  s.phase = 1;
  for_each_row(ds_one, ds_two, sil_row, &s);
  double s_mean = 0;
  for (int i = 0; i < n1; i++) {
    /* sums of integers below 2^53: the reference's running double sums are exact */
    const double internal = (double)s.internal[i], external = (double)s.external[i];
    const double a = internal / ((double)ds_one.n_values - 1.);
    const double b = external / ((double)ds_two.n_values);
    double v = 0;
    if (a < b) {
      v = 1 - a / b;
    } else if (a == b) {
      v = 0;
    } else if (a > b) {
      v = b / a - 1;
    }
    s_mean += v;
  }
  free(s.internal);
  free(s.external);
  return s_mean / (double)n1;
}

double mean_from_smith_waterman_distance_matrix(unsigned long *matrix, dataset ds_one,
                                                dataset ds_two) {
  double sum = 0.;
  for (size_t i = 0; i < (size_t)ds_one.n_values * (size_t)ds_two.n_values; i++) sum += (double)matrix[i];
  return sum / ((double)ds_one.n_values * (double)ds_two.n_values);
}

double sigma_from_smith_waterman_distance_matrix(unsigned long *matrix, double mean, dataset ds_one,
                                                 dataset ds_two) {
  double sigma = 0.;
  for (size_t i = 0; i < (size_t)ds_one.n_values * (size_t)ds_two.n_values; i++)
    sigma += ((double)matrix[i] - mean) * ((double)matrix[i] - mean);
  return sqrt(sigma) / ((double)ds_one.n_values + (double)ds_two.n_values);
}

void shortest_longest_distance_in_matrix(unsigned long *matrix, unsigned long *min, unsigned long *max,
                                         dataset ds_one, dataset ds_two) {
  min[0] = (unsigned long)-1;
  max[0] = 0;
  for (size_t i = 0; i < (size_t)ds_one.n_values * (size_t)ds_two.n_values; i++) {
    if (matrix[i] < min[0]) min[0] = matrix[i];
    if (matrix[i] > max[0]) max[0] = matrix[i];
  }
}

/* comparison.c:343-387: every pair j < i with its Smith-Waterman distance and
 * the L2 / L1 distances of the projections, "%i %i %lf %lf %lf\n". */
void print_SW_PCA_L_comparison(FILE *f, dataset ds) {
  fprintf(f, "i j SW L1 L2\n");
  const int n = ds.n_values;
  if (n < 2) return;
  mnhn_sw_ctx *sw = NULL;
  int rc = mnhn_sw_create((const char *const *)ds.sequences, ds.sequence_lengths, n, NULL, &sw);
  if (rc) die("mnhn_sw_create", rc);
  /* rows in blocks against the columns j < the block's last row */
  int32_t *cols = (int32_t *)malloc(sizeof(int32_t) * (size_t)n);
  for (int k = 0; k < n; k++) cols[k] = k;
  const int rows_per_block = 1024;
  int32_t *rows = (int32_t *)malloc(sizeof(int32_t) * rows_per_block);
  int32_t *dist = (int32_t *)malloc(sizeof(int32_t) * (size_t)rows_per_block * (size_t)n);
  for (int r0 = 1; r0 < n; r0 += rows_per_block) {
    const int nr = n - r0 < rows_per_block ? n - r0 : rows_per_block;
    const int nc = r0 + nr - 1; /* columns 0 .. last row - 1 */
    for (int r = 0; r < nr; r++) rows[r] = r0 + r;
    rc = mnhn_sw_block(sw, rows, nr, cols, nc, dist);
    if (rc) die("mnhn_sw_block", rc);
    for (int r = 0; r < nr; r++) {
      const int i = r0 + r;
      for (int j = 0; j < i; j++) {
        const double dist_sw = (double)(uint32_t)dist[(size_t)r * nc + j];
        double l2 = 0, l1 = 0;
        for (int k = 0; k < ds.n_dimensions; k++) {
          const double buf = (double)(ds.values[k][j] - ds.values[k][i]);
          l2 += buf * buf;
          l1 += fabs(buf);
        }
        l2 = sqrt(l2);
        fprintf(f, "%i %i %lf %lf %lf\n", i, j, dist_sw, l2, l1);
      }
    }
  }
  free(rows);
  free(dist);
  free(cols);
  mnhn_sw_destroy(sw);
}

