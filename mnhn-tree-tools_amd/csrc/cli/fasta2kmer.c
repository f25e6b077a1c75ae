/*
 * fasta2kmer -- k-mer frequency table of a FASTA file, counted on the GPU.
 * Same arguments and stdout as the reference's fasta2kmer (fasta2kmer.c:7-45):
 * fasta kmer_length n_threads protein(1/0); one line per sequence,
 * "sequence_<i>" and the tab-separated counts.  (The reference opens the file
 * before checking the arguments and reads all four of them after checking
 * for two; here all four are required first.)
 */
#include <stdio.h>
#include <stdlib.h>

#include "mnhn_host.h"

int main(int argc, char **argv) {
  size_t length = 0;
  int n_threads = 1, protein = 0;
  if (argc < 5) {
    printf("Arguments are: \n"
           "  [file] fasta file \n"
           "  [int] kmer length \n"
           "  [int] number of threads \n"
           "  [int] protein ? if yes 1 otherwise 0) \n");
    return 1;
  }
  FILE *in = fopen(argv[1], "r");
  if (in == NULL) {
    printf("Could not open fasta file %s \n", argv[1]);
    return 1;
  }
  sscanf(argv[2], "%zu", &length);
  sscanf(argv[3], "%d", &n_threads);
  sscanf(argv[4], "%i", &protein);
  dataset ds = dataset_from_fasta(in);
  fclose(in);
  kmer_frequencies freq = frequencies_from_dataset(ds, length, n_threads, protein);
  write_kmer_base(stdout, freq);
  free_kmer_frequencies(freq);
  free_sequences_from_dataset(ds);
  return 0;
}
