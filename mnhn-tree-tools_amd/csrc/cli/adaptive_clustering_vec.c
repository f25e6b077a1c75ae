/*
 * adaptive_clustering_PCA / adaptive_clustering_kmer_L1 /
 * adaptive_clustering_kmer_L2 (built from this file with -DMNHN_CLUSTER_PCA,
 * -DMNHN_KMER_L1 or -DMNHN_KMER_L2) -- adaptive epsilon sweep over PCA
 * projections (L2) or k-mer count vectors (L1 / L2) with the gfx950 vector
 * engine.  Same arguments, stdout and split-set files as the reference's
 * adaptive_clustering.c:54-200 built with -D_CLUSTER_PCA / -D_CLUSTER_KMER_L1 /
 * -D_CLUSTER_KMER_L2:
 *   PCA:  fasta eps_start eps_inc minpts prefix n_threads dimensions projections
 *   kmer: kmers eps_start eps_inc minpts prefix n_threads
 * Epsilon values are scanned in order by one thread (the reference's
 * n_threads = 1 behaviour).  (The reference checks for fewer arguments than it
 * reads; here all of them are required.)
 */
#include <fcntl.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <sys/stat.h>
#include <sys/types.h>
#include <unistd.h>

#include "mnhn_host.h"

static void file_error(const char *path) {
  printf("failed to open file %s\n", path);
  _exit(1);
}

int main(int argc, char **argv) {
  float epsilon_start = 0, epsilon_inc = 0;
  int minpts = 0, n_threads = 1;
#if defined(MNHN_CLUSTER_PCA)
  const int need = 9;
#else
  const int need = 7;
#endif
  if (argc < need) {
    printf("Arguments are: \n"
#if defined(MNHN_CLUSTER_PCA)
           "   [file] FASTA file containing sequence dataset \n"
#else
           "   [file] kmer file\n"
#endif
           "   (float) initial epsilon value \n"
           "   (float) epsilon increase between cluster search \n"
           "   (int) minimum numbers of clusters in epsilon neigbourhood \n"
           "   (string) path/andprefix of split_set_files \n"
           "   (int) number of threads to use \n"
#if defined(MNHN_CLUSTER_PCA)
           "   (int) dimensions in projection file (PCA obtained from kmers)\n"
           "   [FILE] PCA projections file obtained i.e. from kers\n"
#endif
    );
    return 1;
  }
  sscanf(argv[2], "%f", &epsilon_start);
  sscanf(argv[3], "%f", &epsilon_inc);
  sscanf(argv[4], "%i", &minpts);
  char *prefix = (char *)malloc(strlen(argv[5]) + 1);
  sscanf(argv[5], "%s", prefix);
  sscanf(argv[6], "%i", &n_threads);
  dataset ds;
#if defined(MNHN_CLUSTER_PCA)
  int dimensions = 0;
  sscanf(argv[7], "%i", &dimensions);
  FILE *f = fopen(argv[1], "r");
  if (!f) file_error(argv[1]);
  ds = dataset_from_fasta(f);
  fclose(f);
  FILE *pf = fopen(argv[8], "r");
  if (!pf) file_error(argv[8]);
  load_projections_from_file_into_dataset(pf, (size_t)dimensions, &ds);
  fclose(pf);
#else
  int fd = open(argv[1], O_RDONLY);
  if (fd == -1) file_error(argv[1]);
  data_shape shape = shape_from_kmer_file(fd);
  FILE *kf = fdopen(fd, "r");
  ds = load_kmer_from_file_into_dataset(kf, shape);
  fclose(kf);
#endif
  printf("Dataset read!\n");
#if defined(MNHN_KMER_L1)
  adaptive_dbscan(dbscan_L1_GPU, ds, epsilon_start, epsilon_inc, minpts, prefix, n_threads);
#else
  adaptive_dbscan(dbscan_L2_GPU, ds, epsilon_start, epsilon_inc, minpts, prefix, n_threads);
#endif
#if defined(MNHN_CLUSTER_PCA)
  free_dataset(ds);
#else
  free_values_from_dataset(ds);
#endif
  free(prefix);
  return 0;
}
