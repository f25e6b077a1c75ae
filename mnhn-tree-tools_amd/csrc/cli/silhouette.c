/*
 * silhouette -- silhouette of the first FASTA set against the second under
 * Smith-Waterman distances, on the gfx950 engine.  Same arguments and stdout
 * as the reference's silhouette (silhouette.c:6-34): fasta1 fasta2 n_threads;
 * prints "%lf\n" of silhouette_from_smith_waterman_datasets.
 */
#include <stdio.h>
#include <stdlib.h>
#include <unistd.h>

#include "mnhn_host.h"

int main(int argc, char **argv) {
  size_t n_threads = 1;
  if (argc < 4) {
    printf("Arguments are [fasta1] [fasta2] [n_threads]\n");
    return 1;
  }
  sscanf(argv[3], "%lu", &n_threads);
  FILE *f1 = fopen(argv[1], "r");
  FILE *f2 = fopen(argv[2], "r");
  if (!f1 || !f2) {
    printf("failed to open file %s\n", f1 ? argv[2] : argv[1]);
    _exit(1);
  }
  dataset ds_one = dataset_from_fasta(f1);
  fclose(f1);
  dataset ds_two = dataset_from_fasta(f2);
  fclose(f2);
  printf("%lf\n", silhouette_from_smith_waterman_datasets(ds_one, ds_two, n_threads));
  free_sequences_from_dataset(ds_one);
  free_sequences_from_dataset(ds_two);
  return 0;
}
