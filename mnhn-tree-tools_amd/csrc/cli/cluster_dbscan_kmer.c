/*
 * cluster_dbscan_kmerL1 / cluster_dbscan_kmerL2 (built from this file with
 * -DMNHN_KMER_L1 or -DMNHN_KMER_L2) -- one L1 / L2 DBSCAN pass over k-mer
 * count vectors with the gfx950 vector engine.  Same arguments, stdout and
 * output files as the reference's cluster_dbscan_kmer.c:18-77: fasta kmers
 * epsilon minpts prefix; prints "%u clusters obtained \n" and writes
 * prefix-%03d cluster FASTA files when fewer than 500 clusters were found.
 * (The reference accepts four arguments and then reads a missing prefix;
 * here the prefix is required.)
 */
#include <fcntl.h>
#include <stdio.h>
#include <stdlib.h>
#include <sys/stat.h>
#include <sys/types.h>
#include <unistd.h>

#include "mnhn_host.h"

static void file_error(const char *path) {
  printf("failed to open file %s\n", path);
  _exit(1);
}

int main(int argc, char **argv) {
  float epsilon = 0;
  int minpts = 0;
  if (argc < 6) {
    printf("Arguments are: \n"
           " [file] Sequences in FASTA \n"
           " [file] Corresponding kmers \n"
           " Epsilon \n"
           " minPoints \n"
           " output-files prefix \n");
    return 1;
  }
  sscanf(argv[3], "%f", &epsilon);
  sscanf(argv[4], "%i", &minpts);
  int fd = open(argv[2], O_RDONLY);
  if (fd == -1) file_error(argv[2]);
  data_shape shape = shape_from_kmer_file(fd);
  FILE *kf = fdopen(fd, "r");
  dataset ds = load_kmer_from_file_into_dataset(kf, shape);
  fclose(kf);
#if defined(MNHN_KMER_L1)
  split_set s = dbscan_L1(ds, epsilon, minpts);
#else
  split_set s = dbscan_L2(ds, epsilon, minpts);
#endif
  free_values_from_dataset(ds);
  printf("%u clusters obtained \n", s.n_clusters);
  FILE *f = fopen(argv[1], "r");
  if (!f) file_error(argv[1]);
  ds = dataset_from_fasta(f);
  fclose(f);
  if (s.n_clusters < 500) create_cluster_files(argv[5], s, ds);
  free_sequences_from_dataset(ds);
  free_split_set_and_associated_clusters(s);
  return 0;
}
