/*
 * cluster_dbscan_SW -- one DBSCAN pass over a FASTA file with Smith-Waterman
 * distances from the gfx950 engine.  Same arguments, stdout and output files
 * as the reference's cluster_dbscan_SW / cluster_dbscan_SW_GPU
 * (cluster_dbscan_SW.c:25-71): fasta epsilon minpts prefix; prints
 * "%u clusters obtained \n" and writes prefix-%03d cluster FASTA files when
 * fewer than 500 clusters were found.
 */
#include <stdio.h>
#include <stdlib.h>
#include <unistd.h>

#include "mnhn_host.h"

int main(int argc, char **argv) {
  float epsilon = 0;
  int minpts = 0;
  if (argc < 5) {
    printf("Arguments are: \n"
           " [file] Sequences in FASTA \n"
           " Epsilon \n"
           " minPoints \n"
           " output-files prefix \n");
    return 1;
  }
  sscanf(argv[2], "%f", &epsilon);
  sscanf(argv[3], "%i", &minpts);
  FILE *f = fopen(argv[1], "r");
  if (!f) {
    printf("failed to open file %s\n", argv[1]);
    _exit(1);
  }
  dataset ds = dataset_from_fasta(f);
  fclose(f);
  mnhn_gpu_stuff g = mnhn_gpu_initialization(ds);
  split_set s = dbscan_SW_GPU(ds, epsilon, minpts, g);
  mnhn_gpu_destroy(g);
  printf("%u clusters obtained \n", s.n_clusters);
  if (s.n_clusters < 500) create_cluster_files(argv[4], s, ds);
  free_sequences_from_dataset(ds);
  free_split_set_and_associated_clusters(s);
  return 0;
}
