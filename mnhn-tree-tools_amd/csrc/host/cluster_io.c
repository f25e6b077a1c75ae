/*
 * cluster_io.c -- split-set files, cluster FASTA files and layer relations,
 * byte-compatible with the reference (cluster_io.c of the reference):
 *   store_split_set / read_split_set (:595-669): "SPLIT_SET", int32 count, then
 *     per cluster int32 n_members, int32 id, int32 members[]; host byte order;
 *     created with mode 0644 and opened without O_TRUNC, so an existing longer
 *     file keeps its tail bytes, as the reference's does.
 *   create_single_cluster_file (:672-699): ">sequence_<index>" and a line break
 *     before every 50th residue after the first.
 *   intersection_of_clusters (:447-501): the larger cluster (the first one on a
 *     tie) is sorted in place when it has more than 20 members -- a side effect
 *     that changes the member order later written by store_split_set.
 *   generate_split_set_relation (:407-445): new cluster i is connected to old
 *     cluster j when |old_j & new_i| > 0.8 |old_j|, with the sort side effect
 *     of every pair's intersection (computed in one labelled pass when each
 *     split set holds every point at most once).
 */
#include <fcntl.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <sys/stat.h>
#include <sys/types.h>
#include <unistd.h>

#include "mnhn_host.h"

static const char kSplitMagic[9] = {'S', 'P', 'L', 'I', 'T', '_', 'S', 'E', 'T'};

void store_split_set(char *filename, split_set s) {
  int fd = open(filename, O_WRONLY | O_CREAT, S_IRUSR | S_IWUSR | S_IRGRP | S_IROTH);
  size_t expect = 9 + sizeof(int), wrote = 0;
  ssize_t w;
  if ((w = write(fd, kSplitMagic, 9)) > 0) wrote += (size_t)w;
  if ((w = write(fd, &s.n_clusters, sizeof(int))) > 0) wrote += (size_t)w;
  for (int i = 0; i < s.n_clusters; i++) {
    int hdr[2] = {s.clusters[i].n_members, s.clusters[i].id};
    if ((w = write(fd, hdr, sizeof(hdr))) > 0) wrote += (size_t)w;
    expect += sizeof(hdr);
    size_t bytes = sizeof(int) * (size_t)s.clusters[i].n_members;
    if (bytes) {
      if ((w = write(fd, s.clusters[i].members, bytes)) > 0) wrote += (size_t)w;
      expect += bytes;
    }
  }
  if (wrote != expect) printf("There was an error in creating the binary clusterfile %s", filename);
  close(fd);
}

static int read_all(int fd, void *buf, size_t n) {
  size_t got = 0;
  while (got < n) {
    ssize_t r = read(fd, (char *)buf + got, n - got);
    if (r <= 0) return -1;
    got += (size_t)r;
  }
  return 0;
}

split_set read_split_set(char *filename) {
  split_set s = {0, NULL};
  char magic[9];
  int fd = open(filename, O_RDONLY);
  if (fd < 0 || read_all(fd, magic, 9) || memcmp(magic, kSplitMagic, 9) != 0) {
    printf("Error %s is not a split_set\n", filename);
    _exit(1);
  }
  int ok = read_all(fd, &s.n_clusters, sizeof(int)) == 0 && s.n_clusters >= 0;
  if (ok) s.clusters = (cluster *)malloc(sizeof(cluster) * (size_t)(s.n_clusters ? s.n_clusters : 1));
  for (int i = 0; ok && i < s.n_clusters; i++) {
    int hdr[2];
    ok = read_all(fd, hdr, sizeof(hdr)) == 0 && hdr[0] >= 0;
    if (!ok) {
      s.n_clusters = i;
      break;
    }
    s.clusters[i].n_members = hdr[0];
    s.clusters[i].id = hdr[1];
    s.clusters[i].members = (int *)malloc(sizeof(int) * (size_t)(hdr[0] ? hdr[0] : 1));
    ok = read_all(fd, s.clusters[i].members, sizeof(int) * (size_t)hdr[0]) == 0;
    if (!ok) s.n_clusters = i + 1;
  }
  if (!ok) printf("There was an error reading the binary clusterfile %s", filename);
  close(fd);
  return s;
}

void create_single_cluster_file(char *filename, cluster cl, dataset ds) {
  FILE *f = fopen(filename, "w");
  if (!f) return;
  for (int j = 0; j < cl.n_members; j++) {
    const char *seq = ds.sequences[cl.members[j]];
    size_t len = strlen(seq);
    fprintf(f, ">sequence_%i\n", cl.members[j]);
    for (size_t k = 0; k < len; k++) {
      if (k != 0 && k % 50 == 0) fputc('\n', f);
      fputc(seq[k], f);
    }
    fputc('\n', f);
  }
  fclose(f);
}

void create_cluster_files(char *prefix, split_set s, dataset ds) {
  for (int i = 0; i < s.n_clusters; i++) {
    size_t n = strlen(prefix) + 16;
    char *name = (char *)malloc(n);
    snprintf(name, n, "%s-%03d", prefix, i);
    create_single_cluster_file(name, s.clusters[i], ds);
    free(name);
  }
}

cluster data_not_in_clusters(split_set s, dataset ds) {
  cluster out;
  unsigned char *covered = (unsigned char *)calloc((size_t)ds.n_values + 1, 1);
  out.members = (int *)malloc(sizeof(int) * (size_t)(ds.n_values ? ds.n_values : 1));
  for (int i = 0; i < s.n_clusters; i++)
    for (int j = 0; j < s.clusters[i].n_members; j++) covered[s.clusters[i].members[j]] = 1;
  int count = 0;
  for (int i = 0; i < ds.n_values; i++)
    if (!covered[i]) out.members[count++] = i;
  out.n_members = count;
  out.id = -1;
  free(covered);
  return out;
}

static int cmp_int(const void *a, const void *b) {
  int x = *(const int *)a, y = *(const int *)b;
  return (x > y) - (x < y);
}

static int in_sorted(const int *v, int n, int x) {
  int lo = 0, hi = n - 1;
  while (lo <= hi) {
    int mid = lo + (hi - lo) / 2;
    if (v[mid] < x) lo = mid + 1;
    else if (v[mid] > x) hi = mid - 1;
    else return 1;
  }
  return 0;
}

cluster intersection_of_clusters(cluster a, cluster b) {
  cluster big = a.n_members >= b.n_members ? a : b;
  cluster small = a.n_members >= b.n_members ? b : a;
  cluster r;
  r.members = (int *)malloc(sizeof(int) * (size_t)(small.n_members ? small.n_members : 1));
  int count = 0;
  if (big.n_members > 20) {
    qsort(big.members, (size_t)big.n_members, sizeof(int), cmp_int);
    for (int i = 0; i < small.n_members; i++)
      if (in_sorted(big.members, big.n_members, small.members[i])) r.members[count++] = small.members[i];
  } else {
    for (int i = 0; i < small.n_members; i++)
      for (int j = 0; j < big.n_members; j++)
        if (small.members[i] == big.members[j]) r.members[count++] = small.members[i];
  }
  r.n_members = count;
  r.id = -1;
  return r;
}

/* Largest member + 1 of s when every member is >= 0 and appears in only one
 * cluster, once (as in every split set DBSCAN produces); -1 otherwise. */
static int members_once(split_set s) {
  int mx = -1;
  int64_t total = 0;
  for (int i = 0; i < s.n_clusters; i++)
    for (int k = 0; k < s.clusters[i].n_members; k++) {
      const int x = s.clusters[i].members[k];
      if (x < 0) return -1;
      mx = x > mx ? x : mx;
      total++;
    }
  /* members are point indices; sparse ids beyond that take the pairwise path */
  if ((int64_t)mx > 16 * total + (1 << 20)) return -1;
  unsigned char *seen = (unsigned char *)calloc((size_t)mx + 1, 1);
  if (!seen) return -1;
  for (int i = 0; i < s.n_clusters; i++)
    for (int k = 0; k < s.clusters[i].n_members; k++) {
      const int x = s.clusters[i].members[k];
      if (seen[x]) {
        free(seen);
        return -1;
      }
      seen[x] = 1;
    }
  free(seen);
  return mx + 1;
}

/* The relation when each split set holds every point at most once: one pass
 * over the old clusters against a label per point gives every |old_j & new_i|
 * at once.  The in-place sorts of intersection_of_clusters are idempotent and
 * leave the counts unchanged, so their end state is reproduced directly: a
 * cluster of more than 20 members ends sorted when it is the larger side
 * (the old one on a tie) of at least one pair. */
static void relation_by_labels(split_set ancient, split_set new_set, int n_points, cluster_connections *c) {
  int *label = (int *)malloc(sizeof(int) * (size_t)(n_points ? n_points : 1));
  int *cnt = (int *)calloc((size_t)(new_set.n_clusters ? new_set.n_clusters : 1), sizeof(int));
  int *touched = (int *)malloc(sizeof(int) * (size_t)(new_set.n_clusters ? new_set.n_clusters : 1));
  for (int x = 0; x < n_points; x++) label[x] = -1;
  for (int i = 0; i < new_set.n_clusters; i++)
    for (int k = 0; k < new_set.clusters[i].n_members; k++) label[new_set.clusters[i].members[k]] = i;
  for (int j = 0; j < ancient.n_clusters; j++) {
    const cluster o = ancient.clusters[j];
    int nt = 0;
    for (int k = 0; k < o.n_members; k++) {
      const int x = o.members[k];
      const int i = x < n_points ? label[x] : -1;
      if (i >= 0 && cnt[i]++ == 0) touched[nt++] = i;
    }
    for (int t = 0; t < nt; t++) {
      const int i = touched[t];
      if ((double)cnt[i] > (double)o.n_members * 0.8) c[i].connections[c[i].n_connections++] = j;
      cnt[i] = 0;
    }
  }
  free(label);
  free(cnt);
  free(touched);
  if (ancient.n_clusters == 0 || new_set.n_clusters == 0) return;
  int min_old = ancient.clusters[0].n_members, min_new = new_set.clusters[0].n_members;
  for (int j = 1; j < ancient.n_clusters; j++)
    min_old = ancient.clusters[j].n_members < min_old ? ancient.clusters[j].n_members : min_old;
  for (int i = 1; i < new_set.n_clusters; i++)
    min_new = new_set.clusters[i].n_members < min_new ? new_set.clusters[i].n_members : min_new;
  for (int j = 0; j < ancient.n_clusters; j++) {
    const cluster o = ancient.clusters[j];
    if (o.n_members > 20 && o.n_members >= min_new) qsort(o.members, (size_t)o.n_members, sizeof(int), cmp_int);
  }
  for (int i = 0; i < new_set.n_clusters; i++) {
    const cluster w = new_set.clusters[i];
    if (w.n_members > 20 && w.n_members > min_old) qsort(w.members, (size_t)w.n_members, sizeof(int), cmp_int);
  }
}

cluster_connections *generate_split_set_relation(split_set ancient, split_set new_set) {
  cluster_connections *c =
      (cluster_connections *)malloc(sizeof(cluster_connections) * (size_t)(new_set.n_clusters ? new_set.n_clusters : 1));
  const int na = members_once(ancient), nn = na < 0 ? -1 : members_once(new_set);
  if (na >= 0 && nn >= 0) {
    for (int i = 0; i < new_set.n_clusters; i++) {
      c[i].connections = (int *)malloc(sizeof(int) * (size_t)(ancient.n_clusters ? ancient.n_clusters : 1));
      c[i].n_connections = 0;
    }
    relation_by_labels(ancient, new_set, nn, c);
    return c;
  }
  for (int i = 0; i < new_set.n_clusters; i++) {
    c[i].connections = (int *)malloc(sizeof(int) * (size_t)(ancient.n_clusters ? ancient.n_clusters : 1));
    int count = 0;
    for (int j = 0; j < ancient.n_clusters; j++) {
      cluster x = intersection_of_clusters(ancient.clusters[j], new_set.clusters[i]);
      if ((double)x.n_members > (double)ancient.clusters[j].n_members * 0.8) c[i].connections[count++] = j;
      free(x.members);
    }
    c[i].n_connections = count;
  }
  return c;
}

void free_split_set_and_associated_clusters(split_set s) {
  for (int i = 0; i < s.n_clusters; i++) free(s.clusters[i].members);
  free(s.clusters);
}

/* cluster_io.c:343-358: one line per member, its values "%f\t"-separated
 * with "%f\n" after the last dimension */
void create_single_cluster_file_with_values(char *filename, cluster cl, dataset ds) {
  FILE *f = fopen(filename, "w");
  if (!f) return;
  for (int j = 0; j < cl.n_members; j++) {
    for (int k = 0; k < ds.n_dimensions - 1; k++) fprintf(f, "%f\t", ds.values[k][cl.members[j]]);
    if (ds.n_dimensions > 0) fprintf(f, "%f\n", ds.values[ds.n_dimensions - 1][cl.members[j]]);
  }
  fclose(f);
}

/* cluster_io.c:360-379: prefix-%03d per cluster */
void create_cluster_files_with_values(char *prefix, split_set s, dataset ds) {
  size_t plen = strlen(prefix);
  char *name = (char *)malloc(plen + 32);
  for (int i = 0; i < s.n_clusters; i++) {
    snprintf(name, plen + 32, "%s-%03d", prefix, i);
    create_single_cluster_file_with_values(name, s.clusters[i], ds);
  }
  free(name);
}
