/*
 * kmers.c -- the reference's k-mer frequency interface (kmers.h of the
 * reference) over the gfx950 counter (mnhn_kmer.h):
 *   number_of_kmers_from_length / number_of_protein_kmers_from_length
 *     (kmers.c:67-84)
 *   frequencies_from_dataset (kmers.c:288-376): one device pass over all
 *     sequences instead of one scan per k-mer and sequence;
 *   write_kmer_base (kmers.c:378-389): the same text as fprintf's
 *     "sequence_%i\t" / "%hi\t" / "%hi\n", formatted by hand into a row
 *     buffer and written with one fwrite per row;
 *   free_kmer_frequencies (kmers.c:391-398).
 */
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <unistd.h>

#include "mnhn_host.h"

size_t number_of_kmers_from_length(size_t kmer_length) {
  if (kmer_length == 0) return 0;
  size_t n = 1;
  for (size_t i = 0; i < kmer_length; i++) n *= 4;
  return n;
}

size_t number_of_protein_kmers_from_length(size_t protein_kmer_length) {
  if (protein_kmer_length == 0) return 0;
  size_t n = 1;
  for (size_t i = 0; i < protein_kmer_length; i++) n *= 20;
  return n;
}

kmer_frequencies frequencies_from_dataset(dataset ds, size_t kmer_length, int n_threads, int protein) {
  (void)n_threads;
  kmer_frequencies f;
  const int k = kmer_length > 64 ? 0 : (int)kmer_length;
  f.n_kmers = mnhn_kmer_columns(k, protein != 0);
  f.n_seq = (size_t)(ds.n_values > 0 ? ds.n_values : 0);
  if (f.n_kmers == 0) {
    printf("k-mer length %zu is not supported (1 to %d)\n", kmer_length, protein ? 7 : 15);
    fflush(stdout);
    _exit(1);
  }
  f.frequencies = (short int **)malloc(sizeof(short int *) * (f.n_seq ? f.n_seq : 1));
  short int *block = (short int *)malloc(sizeof(short int) * (f.n_seq ? f.n_seq * f.n_kmers : 1));
  if (!f.frequencies || !block) {
    printf("Could not allocate the k-mer frequency table\n");
    fflush(stdout);
    _exit(1);
  }
  for (size_t i = 0; i < f.n_seq; i++) f.frequencies[i] = block + i * f.n_kmers;
  if (f.n_seq == 0) {
    f.frequencies[0] = block;
    return f;
  }
  int rc = mnhn_kmer_frequencies((const char *const *)ds.sequences, ds.sequence_lengths, ds.n_values, k,
                                 protein != 0, (int16_t *)block);
  if (rc) {
    printf("mnhn_kmer_frequencies failed: %s\n", mnhn_sw_strerror(rc));
    fflush(stdout);
    _exit(1);
  }
  return f;
}

/* Decimal text of v (as printf("%hi") writes it) at p; returns the end. */
static char *put_short(char *p, int v) {
  if (v < 0) {
    *p++ = '-';
    v = -v;
  }
  char tmp[8];
  int n = 0;
  do {
    tmp[n++] = (char)('0' + v % 10);
    v /= 10;
  } while (v);
  while (n) *p++ = tmp[--n];
  return p;
}

void write_kmer_base(FILE *f, kmer_frequencies freq) {
  if (freq.n_kmers == 0) return;
  /* "sequence_" + int + "\t" and at most 7 bytes per count ("-32768\t") */
  char *line = (char *)malloc(32 + 7 * freq.n_kmers);
  for (size_t i = 0; i < freq.n_seq; i++) {
    char *p = line;
    p += sprintf(p, "sequence_%i\t", (int)i);
    const short int *row = freq.frequencies[i];
    for (size_t j = 0; j + 1 < freq.n_kmers; j++) {
      p = put_short(p, row[j]);
      *p++ = '\t';
    }
    p = put_short(p, row[freq.n_kmers - 1]);
    *p++ = '\n';
    fwrite(line, 1, (size_t)(p - line), f);
  }
  free(line);
}

void free_kmer_frequencies(kmer_frequencies freq) {
  if (freq.frequencies) free(freq.frequencies[0]);
  free(freq.frequencies);
}
