/*
 * mnhn_host.h -- host-side C API of the MI355X DBSCAN path.
 *
 * The types and entry points keep the reference's names, layouts and
 * argument meaning so that a caller of the reference's GPU build can switch
 * by replacing `opencl_stuff` with `mnhn_gpu_stuff`:
 *   dataset / dataset_from_fasta          dataset.h:27-35, :74; dataset.c:264-344
 *   cluster / split_set / connections     cluster.h:10-31
 *   store_split_set / read_split_set      cluster.h:140,146; cluster_io.c:595-669
 *   create_cluster_files                  cluster.h:67; cluster_io.c:672-720
 *   data_not_in_clusters                  cluster_io.c:564-593
 *   generate_split_set_relation           cluster.h:131; cluster_io.c:407-445
 *   intersection_of_clusters              cluster_io.c:447-501
 *   opencl_initialization/destroy       -> mnhn_gpu_initialization/destroy
 *                                          (dbscan.h:32-37; dbscan.c:97-298)
 *   dbscan_SW_GPU / dbscan_SW             dbscan.h:30, 36-37; dbscan.c:950-1047
 *   dbscan_L1 / dbscan_L2                 dbscan.h:30-31; dbscan.c:563-780, 950-1047
 *   adaptive_dbscan                       dbscan.h:38-48; dbscan.c:1117-1380
 *   kmer_frequencies / frequencies_from_dataset / write_kmer_base
 *                                         kmers.h:19-23, 72-113; kmers.c:67-84, 288-398
 * The Smith-Waterman distances all come from the gfx950 engine
 * (mnhn_sw.h) and the L1 / L2 vector distances from the gfx950 vector
 * engine (mnhn_vec.h); there is no CPU distance path in this library.
 */
#ifndef MNHN_HOST_H
#define MNHN_HOST_H

#include <stddef.h>
#include <stdint.h>
#include <stdio.h>

#include "mnhn_sw.h"
#include "mnhn_vec.h"
#include "mnhn_kmer.h"

#ifdef __cplusplus
extern "C" {
#endif

/* dataset.h:27-35 (same field order) */
typedef struct {
  char **sequences;
  size_t *sequence_lengths;
  void **binary_sequences;
  int n_values;
  float **values;
  int n_dimensions;
  size_t max_sequence_length;
} dataset;

/* cluster.h:10-23 */
typedef struct {
  int *members;
  int n_members;
  int id;
} cluster;

typedef struct {
  int n_clusters;
  cluster *clusters;
} split_set;

/* cluster.h:27-31 */
typedef struct {
  int *connections;
  int n_connections;
} cluster_connections;

/* dataset.h:37-40 */
typedef struct {
  size_t n_features;
  size_t n_samples;
} data_shape;

/* Device context of a dataset: the engine plus the cached epsilon-neighbour
 * lists of the dataset.  Passed by value like the reference's opencl_stuff;
 * copies share the same engine and cache.  Sequence datasets use sw; vector
 * datasets (mnhn_gpu_initialization_values) use vec with norm L1 or L2. */
typedef struct mnhn_eps_cache mnhn_eps_cache;
typedef struct {
  mnhn_sw_ctx *sw;
  mnhn_eps_cache *cache;
  int num_devs;
  mnhn_vec_ctx *vec;
  int norm; /* MNHN_NORM_L1 / MNHN_NORM_L2 for vector datasets, 0 otherwise */
} mnhn_gpu_stuff;

/* ---- dataset ---- */
dataset dataset_from_fasta(FILE *in);
void free_sequences_from_dataset(dataset ds);

/* ---- split sets / clusters ---- */
void store_split_set(char *filename, split_set s);
split_set read_split_set(char *filename);
void create_single_cluster_file(char *filename, cluster cl, dataset ds);
void create_cluster_files(char *prefix, split_set s, dataset ds);
cluster data_not_in_clusters(split_set s, dataset ds);
cluster intersection_of_clusters(cluster a, cluster b);
cluster_connections *generate_split_set_relation(split_set ancient, split_set new_set);
void free_split_set_and_associated_clusters(split_set s);

/* ---- device context (replaces the OpenCL set-up) ----
 * On failure (no GPU, HIP error) these print the reason and _exit(1), the
 * reference's error convention (dbscan.c:149-152). */
mnhn_gpu_stuff mnhn_gpu_initialization(dataset ds);
/* Context over ds.values (n_dimensions x n_values) for the L1 or L2 scans. */
mnhn_gpu_stuff mnhn_gpu_initialization_values(dataset ds, int norm);
void mnhn_gpu_destroy(mnhn_gpu_stuff g);

/* epsilon-neighbour lists are computed once up to eps + lookahead and reused
 * by every later pass with a smaller or equal epsilon. */
void mnhn_gpu_set_lookahead(mnhn_gpu_stuff g, float lookahead);

/* ---- clustering ---- */
split_set dbscan_SW_GPU(dataset ds, float epsilon, int minpts, mnhn_gpu_stuff g);
/* dbscan.h:30, the CPU build's scanner (cluster_dbscan_SW.c:62): the same pass
 * with a device context of its own, so that callers built against the
 * reference's dbscan.h link against this library unchanged. */
split_set dbscan_SW(dataset ds, float epsilon, int minpts);

/* One DBSCAN pass (same semantics as dbscan_SW_GPU, without the -1 default
 * warnings) over precomputed epsilon lists, e.g. from mnhn_sw_eps_csr or the
 * learned prefilter; requires eps <= csr->eps_max.  No counterpart in the
 * reference, whose passes always recompute distances. */
split_set mnhn_dbscan_csr(const mnhn_csr *csr, float epsilon, int minpts);

/* L1 / L2 DBSCAN over ds.values: region_query of dbscan.c:563-780 (every j,
 * the point itself included, whose accumulator is <= eps (L1) or <= eps * eps
 * (L2)), then the same seeding and expansion as dbscan_SW_GPU.  The _GPU forms
 * take a context from mnhn_gpu_initialization_values with the matching norm
 * and are the dbscanner arguments of adaptive_dbscan for vector datasets. */
split_set dbscan_L1_GPU(dataset ds, float epsilon, int minpts, mnhn_gpu_stuff g);
split_set dbscan_L2_GPU(dataset ds, float epsilon, int minpts, mnhn_gpu_stuff g);
/* dbscan.h:30-31: one pass with a context of its own */
split_set dbscan_L1(dataset ds, float epsilon, int minpts);
split_set dbscan_L2(dataset ds, float epsilon, int minpts);
/* One L1 / L2 DBSCAN pass over lists from mnhn_vec_eps_csr; requires
 * eps <= csr->thr (L1) or eps * eps <= csr->thr (L2).  No counterpart in the
 * reference. */
split_set mnhn_dbscan_vcsr(const mnhn_vcsr *csr, float epsilon, int minpts);

/* n_eps DBSCAN passes over the same lists, out[k] the pass at epsilon[k]
 * (each as mnhn_dbscan_csr / mnhn_dbscan_vcsr would return it).  The stored
 * rows are ordered by distance once, so each pass costs what its expansion
 * reads rather than a filter of every stored entry; when that copy of the
 * lists would exceed a quarter of host memory the passes filter as the
 * single-pass calls do.  symmetric != 0 declares that every stored (p, q, d)
 * has its (q, p, d) with the same d -- true of all lists the engines return
 * (mnhn_sw_eps_csr, mnhn_prefilter_eps_csr, mnhn_vec_eps_csr, .ecsr files of
 * them) -- and lets large clusters find the points still outside them from
 * those points' own short lists; lists declared symmetric that are not give
 * unspecified clusters.  Returns MNHN_ERR_ARG for a NULL csr, or for NULL
 * epsilon / out with n_eps > 0.  No counterpart in the reference, whose
 * adaptive sweep (dbscan.c:1061-1115) recomputes distances every pass. */
int mnhn_dbscan_csr_sweep(const mnhn_csr *csr, const float *epsilon, int n_eps, int minpts, int symmetric,
                          split_set *out);
int mnhn_dbscan_vcsr_sweep(const mnhn_vcsr *csr, const float *epsilon, int n_eps, int minpts, int symmetric,
                           split_set *out);

/* ---- vector inputs and outputs ---- */
/* dataset.c:346-375: features = tabs on the first line, samples = line
 * breaks from the end of the first line on (the first one included) */
data_shape shape_from_kmer_file(int infile);
/* dataset.c:378-425: per sample a label token, then n_features floats
 * (values[feature][sample]); sequences are left empty */
dataset load_kmer_from_file_into_dataset(FILE *in_file, data_shape shape);
/* cluster_io.c:343-379: "%f\t" per value and "%f\n" after the last, files
 * prefix-%03d */
void create_single_cluster_file_with_values(char *filename, cluster cl, dataset ds);
void create_cluster_files_with_values(char *prefix, split_set s, dataset ds);
/* dataset.c:646-649 */
void free_dataset(dataset ds);

/* ---- k-mer frequencies (fasta2kmer) ----
 * kmers.h:19-23: frequencies[sequence][k-mer] in the column order of
 * mnhn_kmer.h.  Here the rows share one block: release the table with
 * free_kmer_frequencies, not row by row. */
typedef struct {
  short int **frequencies;
  size_t n_kmers;
  size_t n_seq;
} kmer_frequencies;
/* kmers.c:67-84: 4^k and 20^k; 0 for k = 0 (the reference loops there) */
size_t number_of_kmers_from_length(size_t kmer_length);
size_t number_of_protein_kmers_from_length(size_t protein_kmer_length);
/* kmers.c:288-376, counted on the GPU (mnhn_kmer_frequencies); n_threads is
 * accepted and unused.  A k outside 1..15 (DNA) / 1..7 (protein) or a device
 * failure prints the reason and exits. */
kmer_frequencies frequencies_from_dataset(dataset ds, size_t kmer_length, int n_threads, int protein);
/* kmers.c:378-389: "sequence_%i\t", then the counts "%hi" tab-separated and
 * "\n" after the last */
void write_kmer_base(FILE *f, kmer_frequencies freq);
/* kmers.c:391-398 */
void free_kmer_frequencies(kmer_frequencies freq);

/* ---- the other Smith-Waterman consumers (distances from the gfx950 engine;
 * n_threads arguments are accepted and unused: one device pass) ---- */
/* dataset.c:85-110: ">sequence_<j>" records, lines wrapped at 50 residues */
void dataset_to_fasta(FILE *f, dataset ds);
/* filter.c:18-48: lengths in [size - minus, size + plus], size_t arithmetic */
dataset filter_ds_by_size(dataset in_set, size_t size, size_t plus, size_t minus);
/* filter.c:73-140: sequences with d(seq, target) <= max_distance, in order */
dataset filter_by_SW_distance(dataset in_set, char *target, size_t target_length,
                              unsigned int max_distance, int n_threads);
/* comparison.c:77-103: n1 x n2 distances, row-major */
unsigned long *smith_waterman_distances_matrix(dataset ds_one, dataset ds_two);
/* comparison.c:152-188, 240-277, 279-313 (accumulators as in the reference's
 * one-thread run) */
double mean_from_smith_waterman_datasets(dataset ds_one, dataset ds_two, size_t n_threads);
double sigma_from_smith_waterman_datasets(double mean, dataset ds_one, dataset ds_two,
                                          size_t n_threads);
double silhouette_from_smith_waterman_datasets(dataset ds_one, dataset ds_two, size_t n_threads);
/* comparison.c:315-341, 389-403 (host only) */
double mean_from_smith_waterman_distance_matrix(unsigned long *matrix, dataset ds_one,
                                                dataset ds_two);
double sigma_from_smith_waterman_distance_matrix(unsigned long *matrix, double mean,
                                                 dataset ds_one, dataset ds_two);
void shortest_longest_distance_in_matrix(unsigned long *matrix, unsigned long *min,
                                         unsigned long *max, dataset ds_one, dataset ds_two);
/* dataset.c:427-469, 629-635: PCA projections (values[dim][row]) */
void load_projections_from_file_into_dataset(FILE *projections, size_t dimensions, dataset *ds);
void free_values_from_dataset(dataset ds);
/* comparison.c:343-387: "i j SW L1 L2" header, then every pair j < i */
void print_SW_PCA_L_comparison(FILE *f, dataset ds);

/* ---- persisted epsilon lists (.ecsr) ----
 * No counterpart in the reference, which recomputes every distance per run.
 * File: "MNHNECSR", uint32 version 1, int32 n_rows, int32 eps_max, int32 0,
 * int64 nnz, uint64 data_hash (mnhn_sw_data_hash), int64 row_ptr[n_rows+1],
 * int32 col[nnz], uint16 dist[nnz]; little-endian.  Loaded lists are released
 * with mnhn_csr_free.  Both return MNHN_OK or MNHN_ERR_ARG (I/O error, bad or
 * truncated file).  With MNHN_ECSR=<path> in the environment, the DBSCAN
 * entry points load the lists from that file when it matches the data and
 * covers the epsilon, and otherwise compute them and write the file. */
int mnhn_csr_save(const mnhn_csr *csr, uint64_t data_hash, const char *path);
int mnhn_csr_load(const char *path, uint64_t *data_hash, mnhn_csr *out);

void adaptive_dbscan(split_set (*dbscanner)(dataset, float, int, mnhn_gpu_stuff),
                     dataset ds, float epsilon_start, float epsilon_inc, int minpts,
                     char *split_files_prefix, int n_threads);

#ifdef __cplusplus
}
#endif

#endif /* MNHN_HOST_H */
