/*
 * adaptive_clustering_SW -- adaptive epsilon sweep over a FASTA file with
 * Smith-Waterman distances from the gfx950 engine.  Same arguments, stdout
 * and split-set files as the reference's adaptive_clustering_SW(_GPU)
 * (adaptive_clustering.c:54-200): fasta eps_start eps_inc minpts prefix
 * n_threads.  Epsilon values are scanned in order by one thread, which is the
 * reference's deterministic (n_threads = 1) behaviour; the GPU build of the
 * reference only accepts one thread (dbscan.c:1185-1190).
 */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <unistd.h>

#include "mnhn_host.h"

int main(int argc, char **argv) {
  float epsilon_start = 0, epsilon_inc = 0;
  int minpts = 0, n_threads = 1;
  if (argc < 7) {
    printf("Arguments are: \n"
           "   [file] FASTA file containing sequence dataset \n"
           "   (float) initial epsilon value \n"
           "   (float) epsilon increase between cluster search \n"
           "   (int) minimum numbers of clusters in epsilon neigbourhood \n"
           "   (string) path/andprefix of split_set_files \n"
           "   (int) number of threads to use \n");
    return 1;
  }
  sscanf(argv[2], "%f", &epsilon_start);
  sscanf(argv[3], "%f", &epsilon_inc);
  sscanf(argv[4], "%i", &minpts);
  char *prefix = (char *)malloc(strlen(argv[5]) + 1);
  sscanf(argv[5], "%s", prefix);
  sscanf(argv[6], "%i", &n_threads);
  FILE *f = fopen(argv[1], "r");
  if (!f) {
    printf("failed to open file %s\n", argv[1]);
    _exit(1);
  }
  dataset ds = dataset_from_fasta(f);
  fclose(f);
  printf("Dataset read!\n");
  adaptive_dbscan(dbscan_SW_GPU, ds, epsilon_start, epsilon_inc, minpts, prefix, n_threads);
  free_sequences_from_dataset(ds);
  free(prefix);
  return 0;
}
