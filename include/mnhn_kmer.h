/*
 * mnhn_kmer.h -- C ABI of the gfx950 k-mer counter behind fasta2kmer: the
 * frequency table of frequencies_from_dataset (kmers.c:288-376 of the
 * reference), computed for all sequences in one pass on the GPU instead of
 * one scan of the sequence per k-mer (frequence_of_kmer_in_sequence,
 * kmers.c:159-243).
 *
 * Counting rules (the reference's, including its quirks):
 *   - a k-mer is counted at the start positions i = 0 .. len - k - 1: the
 *     window ending at the last residue is never counted, and sequences with
 *     len <= k count nothing;
 *   - bytes are compared as they are (the FASTA reader upper-cases); a
 *     window with any byte outside the alphabet matches no k-mer;
 *   - counts are kept in a short int and wrap modulo 2^16.
 * Column order (the order of the reference's generated k-mer tables):
 *   - DNA (alphabet A C G T = 0..3): column sum_p code(s[p]) * 4^p, the first
 *     base in the low bits (gen_dna_string_from_binary, kmers.c:26-65);
 *   - protein (alphabet ACDEFGHIKLMNPQRSTVWY = 0..19): column
 *     sum_p code(s[p]) * 20^(k-1-p), the first residue varying slowest
 *     (recursive_protein_kmer_generator, kmers.c:105-157).
 */
#ifndef MNHN_KMER_H
#define MNHN_KMER_H

#include <stddef.h>
#include <stdint.h>

#include "mnhn_sw.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Columns of the table: 4^k (DNA, 1 <= k <= 15) or 20^k (protein,
 * 1 <= k <= 7); 0 outside those ranges.  number_of_kmers_from_length /
 * number_of_protein_kmers_from_length (kmers.c:67-84). */
size_t mnhn_kmer_columns(int kmer_length, int protein);

/* out[s * columns + c]: count of k-mer c in sequence s (n rows).  Returns
 * MNHN_ERR_ARG for a k outside the ranges above, MNHN_ERR_NOMEM when one
 * sequence's row does not fit in device memory. */
int mnhn_kmer_frequencies(const char *const *seqs, const size_t *lens, int n, int kmer_length,
                          int protein, int16_t *out);

#ifdef __cplusplus
}
#endif

#endif /* MNHN_KMER_H */
