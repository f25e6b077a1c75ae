/*
 * find_satellite -- length window, then Smith-Waterman distance to a target
 * sequence, on the gfx950 engine.  Same arguments and stdout as the
 * reference's find_satellite (find_satellite.c:12-74): dataset fasta, target
 * fasta (its first sequence is the target), size, plus, minus, maximum
 * distance, thread count (unused); the kept sequences go to stdout through
 * dataset_to_fasta.
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
  size_t size = 0, plus = 0, minus = 0;
  unsigned int max_distance = 0;
  int n_threads = 0;
  if (argc < 7) {
    printf("Arguments are: \n"
           " [file] Fasta dataset to filter \n"
           " [file] Fasta target sequence for Smith Waterman filter \n"
           " (int) size filter size \n"
           " (int) allow sequences in the range [size size+this_value] \n"
           " (int) allow sequences in the tange [size-this_value size] \n"
           " (int) maximum allowed Smith Waterman distance from target \n"
           " (int) number of threads to use for this computation \n");
    return 1;
  }
  sscanf(argv[3], "%lu", &size);
  sscanf(argv[4], "%lu", &plus);
  sscanf(argv[5], "%lu", &minus);
  sscanf(argv[6], "%u", &max_distance);
  if (argc > 7) sscanf(argv[7], "%i", &n_threads);
  FILE *f = fopen(argv[1], "r");
  if (!f) file_error(argv[1]);
  dataset ds = dataset_from_fasta(f);
  fclose(f);
  FILE *t = fopen(argv[2], "r");
  if (!t) file_error(argv[2]);
  dataset target_set = dataset_from_fasta(t);
  fclose(t);
  if (target_set.n_values < 1) {
    printf("no target sequence in %s\n", argv[2]);
    _exit(1);
  }
  dataset sized = filter_ds_by_size(ds, size, plus, minus);
  free_sequences_from_dataset(ds);
  dataset kept = filter_by_SW_distance(sized, target_set.sequences[0], target_set.sequence_lengths[0],
                                       max_distance, n_threads);
  free_sequences_from_dataset(sized);
  dataset_to_fasta(stdout, kept);
  free_sequences_from_dataset(kept);
  free_sequences_from_dataset(target_set);
  return 0;
}
