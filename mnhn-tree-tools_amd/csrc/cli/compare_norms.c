/*
 * compare_norms -- every pair's Smith-Waterman distance next to the L2 and L1
 * distances of its PCA projections, on the gfx950 engine.  Same arguments and
 * stdout as the reference's compare_norms (compare_norms.c:12-39): fasta,
 * projections file, number of projections per sequence.
 */
#include <stdio.h>
#include <stdlib.h>
#include <unistd.h>

#include "mnhn_host.h"

static void file_error(char *path) {
  printf("failed to open file %s\n", path);
  _exit(1);
}

int main(int argc, char **argv) {
  int dimensions = 0;
  if (argc < 4) {
    printf("Arguments are: \n");
    printf("  [FILE-in] fasta file to print distance in different norms from");
    printf("  [FILE-in] PCA projections file\n");
    printf("  (int) number of projections in projections file\n");
    _exit(1);
  }
  FILE *fasta_f = fopen(argv[1], "r");
  if (!fasta_f) file_error(argv[1]);
  dataset ds = dataset_from_fasta(fasta_f);
  fclose(fasta_f);
  sscanf(argv[3], "%i", &dimensions);
  FILE *projection_f = fopen(argv[2], "r");
  if (!projection_f) file_error(argv[2]);
  load_projections_from_file_into_dataset(projection_f, (size_t)dimensions, &ds);
  fclose(projection_f);
  print_SW_PCA_L_comparison(stdout, ds);
  free_values_from_dataset(ds);
  free_sequences_from_dataset(ds);
  return 0;
}
