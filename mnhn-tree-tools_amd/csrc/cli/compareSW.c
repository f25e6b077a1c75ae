/*
 * compareSW -- mean and sigma of the Smith-Waterman distances between two
 * FASTA sets, on the gfx950 engine.  Same arguments and stdout as the
 * reference's compareSW (compareSW.c:7-50): fasta1 fasta2 n_threads; prints
 * "mean = %lf, sigma = %lf \n" with the reference's accumulator arithmetic
 * (mean_from_smith_waterman_datasets, sigma_from_smith_waterman_datasets).
 */
#include <stdio.h>
#include <stdlib.h>
#include <unistd.h>

#include "mnhn_host.h"

int main(int argc, char **argv) {
  size_t n_threads = 1;
  if (argc < 3) {
    printf("Arguments are [fasta1] [fasta2] [n_threads]\n");
    return 1;
  }
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
  if (argc > 3) sscanf(argv[3], "%lu", &n_threads);
  double mean = mean_from_smith_waterman_datasets(ds_one, ds_two, n_threads);
  double sigma = sigma_from_smith_waterman_datasets(mean, ds_one, ds_two, n_threads);
  printf("mean = %lf, sigma = %lf \n", mean, sigma);
  free_sequences_from_dataset(ds_one);
  free_sequences_from_dataset(ds_two);
  return 0;
}
