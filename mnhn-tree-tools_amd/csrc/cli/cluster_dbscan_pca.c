/*
 * cluster_dbscan_pca -- one L2 DBSCAN pass over PCA projections with the
 * gfx950 vector engine.  Same arguments, stdout and output files as the
 * reference's cluster_dbscan_pca (cluster_dbscan_pca.c:16-68): fasta
 * projections dimensions epsilon minpts prefix [values-prefix]; prints
 * "%u clusters obtained \n" and, when fewer than 500 clusters were found,
 * writes prefix-%03d cluster FASTA files and, with the seventh argument,
 * values-prefix-%03d files of the members' projections.
 */
#include <stdio.h>
#include <stdlib.h>
#include <unistd.h>

#include "mnhn_host.h"

static void file_error(const char *path) {
  printf("failed to open file %s\n", path);
  _exit(1);
}

int main(int argc, char **argv) {
  size_t dimensions = 0;
  float epsilon = 0;
  int minpts = 0;
  if (argc < 7) {
    printf("Arguments are: \n"
           " [file] Sequences in FASTA \n"
           " [file] Corresponding PCA from kmers \n"
           " dimensions: dimensions in projections file \n"
           " Epsilon \n"
           " minPoints \n"
           " output-files prefix (FASTA) \n"
           " optional: out-files with values (Projections)\n");
    return 1;
  }
  sscanf(argv[3], "%zu", &dimensions);
  sscanf(argv[4], "%f", &epsilon);
  sscanf(argv[5], "%i", &minpts);
  FILE *f = fopen(argv[1], "r");
  if (!f) file_error(argv[1]);
  dataset ds = dataset_from_fasta(f);
  fclose(f);
  FILE *pf = fopen(argv[2], "r");
  if (!pf) file_error(argv[2]);
  load_projections_from_file_into_dataset(pf, dimensions, &ds);
  fclose(pf);
  split_set s = dbscan_L2(ds, epsilon, minpts);
  printf("%u clusters obtained \n", s.n_clusters);
  if (s.n_clusters < 500) {
    create_cluster_files(argv[6], s, ds);
    if (argc == 8) create_cluster_files_with_values(argv[7], s, ds);
  }
  free_split_set_and_associated_clusters(s);
  free_dataset(ds);
  return 0;
}
