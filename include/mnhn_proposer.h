/*
 * mnhn_proposer.h -- learned candidate proposer for epsilon-neighbour search.
 *
 * No counterpart in the reference: its region queries compute every distance
 * (dbscan.c:471-561).  Here a small encoder maps each sequence's 5-mer profile
 * (1024 counts over the 2-bit residues, L1-normalised) to a 128-d embedding:
 *   MLP 1024 -> 512 -> 256 -> 128, GELU, bf16 weights, fp32 accumulation,
 * trained on Smith-Waterman distances computed by the exact engine so that
 * ||e_i - e_j|| approximates d(i, j) / (match max(l_i, l_j)) (match = 4 with
 * the default scores).  It only PROPOSES
 * pairs: every candidate is re-scored by the exact integer kernel, and only
 * exact distances enter the epsilon lists.  All kernels are hand-written
 * gfx950 HIP (bf16 MFMA GEMMs with fused bias/GELU, pairwise distance with
 * fused candidate emission).  Data sets must use <= 4 residue codes.
 */
#ifndef MNHN_PROPOSER_H
#define MNHN_PROPOSER_H

#include <stdint.h>

#include "mnhn_sw.h"

#ifdef __cplusplus
extern "C" {
#endif

#define MNHN_KMER_DIM 1024
#define MNHN_EMBED_DIM 128

typedef struct mnhn_encoder mnhn_encoder;

typedef struct {
  int32_t epochs;        /* passes over the labelled pairs */
  int32_t batch;         /* pairs per step (multiple of 32) */
  float lr;              /* Adam learning rate */
  int32_t eps_train;     /* distance below which a pair counts as a neighbour */
  float neighbor_weight; /* weight of the hinge keeping neighbours inside their radius */
  uint64_t seed;         /* pair shuffling */
  float weight_decay;    /* decoupled (AdamW) decay of the weights, 0 = plain Adam */
  int32_t lr_decay;      /* nonzero: rate decays linearly to 10% over the run */
  float far_weight;      /* weight of the hinge keeping non-neighbours beyond radius + far_gap */
  float far_gap;
} mnhn_train_params;

typedef struct {
  double first_loss, final_loss; /* mean loss of the first and last epoch */
  double seconds;                /* wall time of the training loop */
  int64_t steps;
} mnhn_train_stats;

typedef struct {
  int64_t candidates;     /* unordered pairs proposed (i < j) */
  int64_t neighbours;     /* of those, exact d <= eps_max */
  double profile_ms, embed_ms, propose_ms, verify_ms, csr_ms; /* device time per phase */
  int64_t tiles;          /* 128 x 128 embedding tiles the length band leaves to compare */
  double cells;           /* sum of la * lb over the candidates (verification work) */
  double verify_kernel_ms; /* SW kernel time of the verification (HIP events) */
  int64_t stage2_tiles;    /* band tiles that ran the cosine stage (counted on the device) */
  double propose_kernel_ms; /* pairwise kernel time alone (HIP events around its launch) */
  int32_t counts_int8;     /* 1: the cosine stage ran on int8 counts (int8 MFMA), 0: bf16 */
  int32_t reserved;
} mnhn_prefilter_stats;

/* 5-mer profiles of every sequence (n x 1024 fp32, host buffer). */
int mnhn_kmer_profiles(mnhn_sw_ctx *ctx, float *out);

/* Encoder with weights initialised from seed (He-uniform). */
int mnhn_encoder_create(uint64_t seed, mnhn_encoder **out);
void mnhn_encoder_destroy(mnhn_encoder *enc);

/* Train on m labelled pairs (ij[2k], ij[2k+1]) of ctx's sequences with exact
 * distances dist[k] (e.g. from mnhn_sw_pairs). */
int mnhn_encoder_train(mnhn_encoder *enc, mnhn_sw_ctx *ctx, const int32_t *ij,
                       const int32_t *dist, int64_t m, const mnhn_train_params *p,
                       mnhn_train_stats *stats);

/* Embeddings of every sequence (n x 128 fp32, host buffer). */
int mnhn_encoder_embed(mnhn_encoder *enc, mnhn_sw_ctx *ctx, float *out);

/* Radius slack: a pair is proposed when
 *   ||e_i - e_j|| <= eps / (match max(l_i, l_j)) + margin
 * and match |l_i - l_j| <= eps (an exact lower bound of d). */
int mnhn_encoder_set_margin(mnhn_encoder *enc, float margin);

/* Margin per length bucket: a pair whose longer sequence has length l uses
 * margin[b] of the first bucket b with l <= max_len[b] (margin[n-1] above the
 * last bound).  1 <= n <= 16, max_len strictly ascending, margins >= 0.
 * mnhn_encoder_set_margin(m) is the one-bucket table. */
int mnhn_encoder_set_margin_table(mnhn_encoder *enc, const int32_t *max_len, const float *margin,
                                  int n);

/* Second proposal test, per length bucket of the pair's longer sequence (same
 * bucket rule as the margin table): the cosine of the two sequences' raw 5-mer
 * count vectors must be >= cos_min[b]; cos_min[b] < 0 switches the test off
 * for that bucket (the default: off everywhere).  Computed on the same
 * 128 x 128 MFMA tiles as the embedding distance. */
int mnhn_encoder_set_cosine_table(mnhn_encoder *enc, const int32_t *max_len, const float *cos_min,
                                  int n);

/* Largest margin of the table. */
float mnhn_encoder_margin(const mnhn_encoder *enc);

/* Save / load the weights, the margin table and the cosine table
 * (little-endian fp32 blob with a header; older files without a cosine table
 * load with the test off, files with a single margin as a one-bucket table). */
int mnhn_encoder_save(const mnhn_encoder *enc, const char *path);
int mnhn_encoder_load(mnhn_encoder *enc, const char *path);

/* Learned-prefilter epsilon lists: propose pairs from embedding distances,
 * re-score them with the exact kernel and keep d <= eps_max.  Same CSR form as
 * mnhn_sw_eps_csr; a pair the encoder does not propose is missing. */
int mnhn_prefilter_eps_csr(mnhn_encoder *enc, mnhn_sw_ctx *ctx, int32_t eps_max, mnhn_csr *out,
                           mnhn_prefilter_stats *stats);

#ifdef __cplusplus
}
#endif

#endif /* MNHN_PROPOSER_H */
