/*
 * find_closest_sequence_SW -- the database sequence with the smallest
 * Smith-Waterman distance to a query, from one one-vs-all pass of the gfx950
 * engine.  Same arguments, stdout and output file as the reference's
 * (find_closest_sequence_SW.c:65-153): database fasta, query fasta, 1-based
 * query index, thread count (unused: one device pass), output fasta; prints
 * "Smith Waterman distance = %i \n" and writes the closest sequence with
 * create_single_cluster_file.  Ties go to the lowest database index, as the
 * reference's strict "<" over its ordered thread ranges does.
 *
 * The reference scores the chosen query with the length of the query file's
 * FIRST sequence (:52); when the chosen sequence is longer or shorter than
 * that, it compares past its end (undefined) or truncates it.  This tool uses
 * the chosen sequence's own length; results agree whenever the index is 1 or
 * the two lengths are equal.
 */
#include <limits.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <unistd.h>

#include "mnhn_host.h"

static void file_error(char *path) {
  printf("failed to open file %s\n", path);
  _exit(1);
}

int main(int argc, char **argv) {
  unsigned long idx = 0;
  int n_threads = 0;
  char filename[256];
  if (argc < 6) {
    printf("Arguments are: \n"
           "[file] Fasta file of the database to find the closest sequence in\n"
           "[file] Fasta file containing the query sequence \n"
           " [int] use n_th sequence in fasta query file \n"
           " [int] number of threads to use for this search \n"
           "[file] output fasta containing the closest sequence in the"
           " dataset\n");
    return 1;
  }
  sscanf(argv[3], "%lu", &idx);
  sscanf(argv[4], "%i", &n_threads);
  sscanf(argv[5], "%255s", filename);
  FILE *f = fopen(argv[1], "r");
  if (!f) file_error(argv[1]);
  dataset ds = dataset_from_fasta(f);
  fclose(f);
  FILE *q = fopen(argv[2], "r");
  if (!q) file_error(argv[2]);
  dataset query = dataset_from_fasta(q);
  fclose(q);
  if (idx < 1 || idx > (unsigned long)query.n_values) {
    printf("query index %lu out of range (1..%i)\n", idx, query.n_values);
    _exit(1);
  }
  const int n = ds.n_values;
  const char **seqs = (const char **)malloc(sizeof(char *) * (size_t)(n + 1));
  size_t *lens = (size_t *)malloc(sizeof(size_t) * (size_t)(n + 1));
  for (int i = 0; i < n; i++) {
    seqs[i] = ds.sequences[i];
    lens[i] = ds.sequence_lengths[i];
  }
  seqs[n] = query.sequences[idx - 1];
  lens[n] = query.sequence_lengths[idx - 1];
  mnhn_sw_ctx *sw = NULL;
  int rc = mnhn_sw_create(seqs, lens, n + 1, NULL, &sw);
  if (rc) {
    printf("mnhn_sw_create failed: %s\n", mnhn_sw_strerror(rc));
    _exit(1);
  }
  int32_t *dist = (int32_t *)malloc(sizeof(int32_t) * (size_t)(n + 1));
  rc = mnhn_sw_one_vs_all(sw, n, dist);
  if (rc) {
    printf("mnhn_sw_one_vs_all failed: %s\n", mnhn_sw_strerror(rc));
    _exit(1);
  }
  int minimum = INT_MAX, best = -1;
  for (int i = 0; i < n; i++)
    if (dist[i] < minimum) {
      minimum = dist[i];
      best = i;
    }
  mnhn_sw_destroy(sw);
  cluster cl;
  cl.members = &best;
  cl.n_members = 1;
  cl.id = 0;
  printf("Smith Waterman distance = %i \n", minimum);
  fflush(stdout);
  if (best >= 0) create_single_cluster_file(filename, cl, ds);
  free(dist);
  free(seqs);
  free(lens);
  free_sequences_from_dataset(ds);
  free_sequences_from_dataset(query);
  return 0;
}
