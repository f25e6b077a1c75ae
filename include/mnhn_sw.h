/*
 * mnhn_sw.h -- C ABI of the MI355X (gfx950) Smith-Waterman distance engine.
 *
 * Plain C types only (pointers, sizes, int return codes); every entry point is
 * extern "C" and never throws.  Output buffers are owned by the caller unless
 * stated otherwise.  One context is used by one host thread at a time.
 *
 * Distance definition (identical to the reference's score(),
 * smith-waterman.c:16-81): with match +4, mismatch -4 and linear gap 3,
 *   d(A, B) = 4*max(la, lb) - max_{i in [1,la], j in [1,lb]} H(i, j),
 * where H is the local-alignment matrix comparing A[i] with B[j], index 0 of
 * each sequence is never compared, and A[la] / B[lb] are string terminators
 * that match each other and nothing else.
 */
#ifndef MNHN_SW_H
#define MNHN_SW_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- return codes ---- */
#define MNHN_OK 0
#define MNHN_ERR_ARG 1       /* invalid argument (NULL, index out of range, ...) */
#define MNHN_ERR_DEVICE 2    /* HIP runtime / kernel launch error */
#define MNHN_ERR_NOMEM 3     /* device or host allocation failed */
#define MNHN_ERR_NODEVICE 4  /* no usable GPU */
#define MNHN_ERR_RANGE 5     /* an epsilon list bound above 65535 where some pair's
                                  distance can exceed 16 bits, or a work list exceeds
                                  its 32-bit offsets */

/* Scoring parameters.  The reference hard-codes match 4 / mismatch -4 / gap 3
 * (smith-waterman.c:4-14); NULL selects those.  Requires match > 0,
 * mismatch <= 0, gap >= 0. */
typedef struct {
  int32_t match;
  int32_t mismatch;
  int32_t gap;
} mnhn_sw_params;

/* Opaque engine context: sequences uploaded once to device memory (2-bit
 * packed when the residues at index >= 1 use at most four distinct bytes,
 * 16-bit codes otherwise), sorted by length into groups of 128 targets. */
typedef struct mnhn_sw_ctx mnhn_sw_ctx;

/* epsilon-neighbour lists in CSR form (library-owned; free with
 * mnhn_csr_free only: the arrays of large lists are pinned host memory, which
 * free() must not release).  Row i lists every j != i with d(i, j) <= eps_max, in
 * ascending j, with its distance (unsigned 16-bit: listed distances are at
 * most 65535, see mnhn_sw_eps_csr). */
typedef struct {
  int32_t n_rows;
  int32_t eps_max;
  int64_t nnz;
  int64_t *row_ptr; /* [n_rows + 1] */
  int32_t *col;     /* [nnz] */
  uint16_t *dist;   /* [nnz]; eps_max <= 65535 keeps every listed d in range */
} mnhn_csr;

/* Timing of the last engine call (HIP events on the engine's stream). */
typedef struct {
  double kernel_ms;   /* Smith-Waterman kernel(s) only */
  double total_ms;    /* kernel + epsilon-list compaction/sort on device */
  double cells;       /* algorithmic cells: sum of la*lb over computed pairs */
  double pairs;       /* pairs computed */
  int32_t lane_bits;  /* 16 (packed), or 32 when some pair ran the 32-bit form */
  int32_t kernel;     /* 1 = register profile, 2 = LDS profile (int16),
                         3 = LDS profile, exact fp16 where it applies */
  int32_t strip;      /* register strip width used */
  int32_t waves;      /* persistent waves launched */
  int32_t items;      /* work items in the queue */
  int32_t mixed_launch; /* 1: forms [0] and [1] ran as one launch, whose time
                           is form_ms[0] (form_ms[1] is 0) */
  /* per lane format: [0] exact fp16 form, [1] packed int16 forms, [2] 32-bit
   * form -- the cells each computed (sum of la*lb over its pairs) and its
   * kernel time (HIP events around its launch) */
  double form_cells[3];
  double form_ms[3];
} mnhn_sw_timing;

/* Upload n sequences (bytes as read from FASTA; not necessarily
 * NUL-terminated) and build the context.  Replaces the OpenCL set-up of the
 * reference's GPU build: opencl_initialization(dataset), dbscan.c:97-284. */
int mnhn_sw_create(const char *const *seqs, const size_t *lens, int n,
                   const mnhn_sw_params *params, mnhn_sw_ctx **out);

/* Release device memory and the context.  Replaces opencl_destroy,
 * dbscan.c:286-298. */
void mnhn_sw_destroy(mnhn_sw_ctx *ctx);

/* dist_out[j] = d(seq[j], seq[query]) for all j in [0, n).  One launch; the
 * drop-in for the per-query kernel launch + readback of the reference's GPU
 * region_query (dbscan.c:471-513, smith-waterman-gpu.cl:13-98). */
int mnhn_sw_one_vs_all(mnhn_sw_ctx *ctx, int query, int32_t *dist_out);

/* mnhn_sw_one_vs_all without any allocation: dist_out is DEVICE memory of n
 * int32 (the caller's), and dist_out[j] = d(seq[j], seq[query]) as
 * mnhn_sw_one_vs_all writes them to host memory, from the same kernels.  No
 * counterpart in the reference: for callers that cannot allocate (code under
 * stream capture, tight serving loops, harnesses that forbid hipMalloc).  The
 * call makes no hipMalloc / hipFree and no host heap allocation; its
 * work-queue counters and strip-boundary workspaces live in a static device
 * array of 96 MiB (plus 256 B of counters) that each GPU running the library
 * holds until the process ends.  Limit: the context's longest sequence must be
 * at most 65,534 residues (one workgroup's workspaces, 4 x (maxlen + 2 +
 * maxlen / 2 + 1) x 256 B, fit the array), else MNHN_ERR_RANGE.  All calls on one GPU share that array: calls
 * that overlap on one GPU (several host threads, several contexts) race and
 * must be serialised by the caller.  Returns once dist_out holds the
 * distances; mnhn_sw_last_timing is not updated. */
int mnhn_sw_one_vs_all_noalloc(mnhn_sw_ctx *ctx, int query, int32_t *dist_out);

/* The reference's scalar seam, score() (smith-waterman.h:1,
 * smith-waterman.c:16-81), with the same signature and result: one pair
 * computed on the GPU (a two-sequence context per call; work is accepted and
 * unused).  Reference objects that call score() bind to it unchanged with
 * -Wl,--defsym,score=mnhn_score.  There is no error path in the reference's
 * signature: a failure prints the reason and _exit(1)s.  One call costs a
 * context set-up, so batch callers should use mnhn_sw_pairs. */
unsigned int mnhn_score(const char *A, const char *B, size_t len_a, size_t len_b, int *work);

/* dist_out[k] = d(seq[ij[2k]], seq[ij[2k+1]]) for k in [0, m).  Replaces the
 * loops of score() calls in the reference's CPU region_query
 * (dbscan.c:546-555) and other batch callers of score() (smith-waterman.h:1). */
int mnhn_sw_pairs(mnhn_sw_ctx *ctx, const int32_t *ij, int64_t m,
                  int32_t *dist_out);

/* Rectangular block: dist_out[r * nc + k] = d(seq[rows[r]], seq[cols[k]]) for
 * r < nr, k < nc (host buffer of nr * nc int32).  Replaces the nested score()
 * loops of smith_waterman_distances_matrix and the comparison statistics
 * (comparison.c:20-313) and the one-against-many scans of
 * find_closest_sequence_SW.c:26-63 and filter.c:52-71; rows and columns may
 * repeat or overlap.  MNHN_ERR_RANGE when nr * ceil(nc / 128) * 128 exceeds
 * 2^31 (split the rows). */
int mnhn_sw_block(mnhn_sw_ctx *ctx, const int32_t *rows, int nr, const int32_t *cols, int nc,
                  int32_t *dist_out);

/* Full n x n distance matrix (row-major, host buffer of n*n int32).  Every
 * unordered pair is computed once and mirrored; the diagonal is 0.  For
 * moderate n (tests, small data sets): MNHN_ERR_RANGE when the unordered
 * pairs exceed 2^31 / 128 work slots. */
int mnhn_sw_matrix(mnhn_sw_ctx *ctx, int32_t *dist_out);

/* All-pairs epsilon-neighbour lists with d <= eps_max, computed in one
 * persistent launch over unique pairs and compacted on the device; no n x n
 * matrix is stored.  eps_max above 65535 is accepted when match * (longest
 * length) <= 65535 (no distance can exceed it) and returns MNHN_ERR_RANGE
 * otherwise; distances of any size come from mnhn_sw_one_vs_all / _pairs.  Feeds every region_query of a DBSCAN pass (and of an
 * adaptive epsilon sweep up to eps_max) -- the reference recomputes n
 * distances per region_query (dbscan.c:471-561). */
int mnhn_sw_eps_csr(mnhn_sw_ctx *ctx, int32_t eps_max, mnhn_csr *out);

/* Same computation as mnhn_sw_eps_csr but keeps the lists on the device and
 * returns only their count (for timing the device path alone). */
int mnhn_sw_eps_count(mnhn_sw_ctx *ctx, int32_t eps_max, int64_t *nnz_out);

void mnhn_csr_free(mnhn_csr *csr);

/* Number of sequences, alphabet mode (4 = 2-bit residues, 0 = 16-bit codes). */
int mnhn_sw_size(const mnhn_sw_ctx *ctx);
int mnhn_sw_alphabet(const mnhn_sw_ctx *ctx);

/* 64-bit FNV-1a of the scores, the sequence count and every sequence's length
 * and bytes: identifies the data a saved epsilon-list file belongs to. */
uint64_t mnhn_sw_data_hash(const mnhn_sw_ctx *ctx);

/* Timing of the last one_vs_all / pairs / matrix / eps call. */
int mnhn_sw_last_timing(const mnhn_sw_ctx *ctx, mnhn_sw_timing *out);

/* Force the kernel variant (1 = register profile, any alphabet; 2 = LDS
 * profile in int16, <= 4 residue codes; 3 = as 2 but pairs with
 * match * min(la, lb) <= 2048 run in exact fp16), register-strip width (16,
 * 32 or 64; 64 only for variants 2 and 3) and waves per CU used by later
 * calls; 0 restores the tuned defaults. */
int mnhn_sw_set_tuning(mnhn_sw_ctx *ctx, int kernel, int strip, int waves_per_cu);

/* Settings chosen by mnhn_sw_autotune. */
typedef struct {
  int32_t kernel;
  int32_t strip;
  int32_t waves_per_cu;
  double gcups;        /* measured on the timing sample */
  int32_t from_cache;  /* 1 if read from the cache without timing */
} mnhn_sw_tuning;

/* Autotuner: times each candidate kernel variant / strip width / waves-per-CU
 * setting on a strided sample of this context's all-pairs work and keeps the
 * fastest for later calls.  Choices are cached as JSON lines in
 * $MNHN_SW_TUNE_CACHE (default $HOME/.cache/mnhn_sw_tune.jsonl), keyed by GPU
 * architecture, CU count, alphabet mode, length profile and scores; with
 * use_cache != 0 a cached choice is applied without timing.  No counterpart
 * in the reference (whose OpenCL launch lets the runtime pick the local size,
 * dbscan.c:497-501). */
int mnhn_sw_autotune(mnhn_sw_ctx *ctx, int use_cache, mnhn_sw_tuning *out);

const char *mnhn_sw_strerror(int code);

#ifdef __cplusplus
}
#endif

#endif /* MNHN_SW_H */
