/*
 * dbscan.c -- DBSCAN and the adaptive epsilon sweep on top of the gfx950
 * Smith-Waterman engine.
 *
 * Semantics follow the reference's dbscan.c:
 *   region_query (:471-561): ascending list of every j with (float)d <= eps,
 *     the point itself included (d(p, p) = 0);
 *   dbscan_SW_GPU (:950-1047): points in index order; an unvisited point seeds
 *     a cluster when minpts < |N(p)| (strict); cluster ids are 0, 1, ... in seed
 *     order; eps == -1 becomes 0.015 and minpts == -1 becomes 5 with the same
 *     warnings on stdout;
 *   expand_cluster (:783-948): breadth-first queue that starts with N(p) in
 *     ascending order; an unvisited queue element q is marked visited and, when
 *     |N(q)| >= minpts, the members of N(q) not yet in the union of merged
 *     neighbourhoods are appended in ascending order; every queue element not
 *     yet in a cluster joins this one, in queue order.  (The union is kept as a
 *     membership stamp instead of the reference's sorted-merge arrays; the
 *     queue order is the same.)
 *   adaptive_dbscan (:1117-1380) with one scanning thread: the stdout report,
 *     the halving retries, float32 epsilon accumulation, the strictly
 *     decreasing layer filter, the relation side effect and %04d split files.
 *
 * Vector datasets (dbscan_L1 / dbscan_L2, dbscan.c:563-780) follow the same
 * pass with the accumulator lists of the vector engine (mnhn_vec_eps_csr),
 * computed once up to the threshold of eps + lookahead.
 *
 * Where the distances come from: every region query of a pass reads the
 * epsilon-neighbour lists (CSR) that the engine computes once for all unique
 * pairs up to eps + lookahead (mnhn_sw_eps_csr); later passes with a smaller
 * or equal epsilon reuse them.  During a sweep the stored rows are also
 * ordered by distance once, and each pass reads N(p) as a prefix of that
 * order.  With MNHN_REGION_QUERY=launch each region query instead launches
 * the one-vs-all kernel, as the reference's GPU build launches its kernel
 * once per query.
 */
#include <math.h>
#include <pthread.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>
#include <unistd.h>

#include "mnhn_host.h"

struct mnhn_eps_cache {
  mnhn_csr csr;
  int have;
  float lookahead;
  int per_query_launch;
  int32_t *dist_row; /* per-query launch buffer [n] */
  mnhn_vcsr vcsr;    /* vector datasets: lists up to vcsr.thr */
  int have_v;
  int sweep;         /* many passes follow: order the stored lists by distance */
  void *index;       /* the distance-ordered index of the stored lists, or NULL */
  int index_tried;   /* built, or found too large, for the stored lists */
};

static void drop_index(mnhn_eps_cache *c);

/* MNHN_TIMING=1: phase times on stderr (stdout is unchanged) */
static double now_s(void) {
  struct timespec ts;
  clock_gettime(CLOCK_MONOTONIC, &ts);
  return (double)ts.tv_sec + 1e-9 * (double)ts.tv_nsec;
}

static int timing_on(void) {
  static int on = -1;
  if (on < 0) {
    const char *v = getenv("MNHN_TIMING");
    on = v && *v && strcmp(v, "0") != 0;
  }
  return on;
}

static void die_engine(const char *what, int rc) {
  printf("%s failed: %s\n", what, mnhn_sw_strerror(rc));
  fflush(stdout);
  _exit(1);
}

mnhn_gpu_stuff mnhn_gpu_initialization(dataset ds) {
  mnhn_gpu_stuff g;
  memset(&g, 0, sizeof(g));
  int rc = mnhn_sw_create((const char *const *)ds.sequences, ds.sequence_lengths, ds.n_values,
                          NULL, &g.sw);
  if (rc) die_engine("mnhn_sw_create", rc);
  g.cache = (mnhn_eps_cache *)calloc(1, sizeof(mnhn_eps_cache));
  const char *mode = getenv("MNHN_REGION_QUERY");
  g.cache->per_query_launch = mode && strcmp(mode, "launch") == 0;
  g.num_devs = 1;
  return g;
}

mnhn_gpu_stuff mnhn_gpu_initialization_values(dataset ds, int norm) {
  mnhn_gpu_stuff g;
  memset(&g, 0, sizeof(g));
  if (norm != MNHN_NORM_L1 && norm != MNHN_NORM_L2) die_engine("mnhn_gpu_initialization_values", MNHN_ERR_ARG);
  int rc = mnhn_vec_create((const float *const *)ds.values, ds.n_dimensions, ds.n_values, &g.vec);
  if (rc) die_engine("mnhn_vec_create", rc);
  g.cache = (mnhn_eps_cache *)calloc(1, sizeof(mnhn_eps_cache));
  g.norm = norm;
  g.num_devs = 1;
  return g;
}

void mnhn_gpu_destroy(mnhn_gpu_stuff g) {
  if (g.cache) {
    drop_index(g.cache);
    if (g.cache->have) mnhn_csr_free(&g.cache->csr);
    if (g.cache->have_v) mnhn_vcsr_free(&g.cache->vcsr);
    free(g.cache->dist_row);
    free(g.cache);
  }
  mnhn_sw_destroy(g.sw);
  mnhn_vec_destroy(g.vec);
}

void mnhn_gpu_set_lookahead(mnhn_gpu_stuff g, float lookahead) {
  if (g.cache) g.cache->lookahead = lookahead > 0 ? lookahead : 0;
}

/* Largest integer distance d with (float)d <= eps (distances are >= 0 and far
 * below 2^24, so the float compare is exact); -1 when no distance qualifies. */
static int32_t eps_floor(float eps) {
  if (!(eps >= 0)) return -1;
  if (eps >= 2.0e9f) return INT32_MAX;
  return (int32_t)floorf(eps);
}

/* Switch the context to one one-vs-all launch per region query (the
 * reference's GPU scheme, dbscan.c:471-513): O(n) memory whatever epsilon
 * and n are. */
static void use_per_query_launch(mnhn_eps_cache *c, int n) {
  c->per_query_launch = 1;
  if (!c->dist_row) c->dist_row = (int32_t *)malloc(sizeof(int32_t) * (size_t)(n ? n : 1));
  if (!c->dist_row) {
    printf("out of memory for a region query row\n");
    fflush(stdout);
    _exit(1);
  }
}

/* The stored lists cover eps after this call, or -- when they do not fit in
 * device or host memory (MNHN_ERR_NOMEM), or eps lies beyond the 16-bit list
 * distances for these lengths (MNHN_ERR_RANGE) -- the context has switched to
 * per-query launches (returns 0). */
static int ensure_lists(mnhn_gpu_stuff g, float eps) {
  mnhn_eps_cache *c = g.cache;
  const int32_t need = eps_floor(eps);
  if (c->have && c->csr.eps_max >= need) return 1;
  drop_index(c);
  if (c->have) mnhn_csr_free(&c->csr);
  c->have = 0;
  double want = (double)need + (double)c->lookahead;
  if (need < 0) want = -1;
  int32_t eps_max = want > 2.0e9 ? INT32_MAX : (int32_t)floor(want);
  const char *path = getenv("MNHN_ECSR");
  const uint64_t hash = mnhn_sw_data_hash(g.sw);
  if (path && *path) {
    uint64_t fh = 0;
    if (mnhn_csr_load(path, &fh, &c->csr) == MNHN_OK) {
      if (fh == hash && c->csr.n_rows == mnhn_sw_size(g.sw) && c->csr.eps_max >= need) {
        fprintf(stderr, "epsilon lists (eps_max %d) read from %s\n", (int)c->csr.eps_max, path);
        c->have = 1;
        return 1;
      }
      mnhn_csr_free(&c->csr);
    }
  }
  const double t0 = now_s();
  int rc = mnhn_sw_eps_csr(g.sw, eps_max, &c->csr);
  if ((rc == MNHN_ERR_NOMEM || rc == MNHN_ERR_RANGE) && eps_max > need && need >= 0) {
    /* the lookahead may be what does not fit: lists up to eps alone */
    eps_max = need;
    rc = mnhn_sw_eps_csr(g.sw, eps_max, &c->csr);
  }
  if (rc == MNHN_ERR_NOMEM || rc == MNHN_ERR_RANGE) {
    fprintf(stderr, "epsilon lists up to %d: %s; region queries recompute their distances from here on\n",
            (int)eps_max, mnhn_sw_strerror(rc));
    use_per_query_launch(c, mnhn_sw_size(g.sw));
    return 0;
  }
  if (rc) die_engine("mnhn_sw_eps_csr", rc);
  c->have = 1;
  if (timing_on()) {
    mnhn_sw_timing tm;
    memset(&tm, 0, sizeof(tm));
    (void)mnhn_sw_last_timing(g.sw, &tm);
    fprintf(stderr, "[mnhn] epsilon lists up to %d: %lld entries, %.3f s (SW kernels %.3f s, device lists %.3f s)\n",
            (int)eps_max, (long long)c->csr.nnz, now_s() - t0, tm.kernel_ms / 1e3, (tm.total_ms - tm.kernel_ms) / 1e3);
  }
  if (path && *path) {
    if (mnhn_csr_save(&c->csr, hash, path) == MNHN_OK)
      fprintf(stderr, "epsilon lists (eps_max %d) written to %s\n", (int)eps_max, path);
    else
      fprintf(stderr, "could not write epsilon lists to %s\n", path);
  }
  return 1;
}

/* N(p) for threshold eps into out[] (ascending, p included) with one
 * one-vs-all launch (MNHN_REGION_QUERY=launch); returns |N(p)|.  Passes over
 * stored lists read their rows through pass_lists instead. */
static int region_query(mnhn_gpu_stuff g, int n, int p, float eps, int *out) {
  int k = 0;
  int rc = mnhn_sw_one_vs_all(g.sw, p, g.cache->dist_row);
  if (rc) die_engine("mnhn_sw_one_vs_all", rc);
  for (int j = 0; j < n; j++)
    if ((float)g.cache->dist_row[j] <= eps) out[k++] = j;
  return k;
}

/* A region query fills out[] (or points *res at a stored row) with N(p),
 * ascending, and returns |N(p)|. */
typedef int (*region_fn)(void *src, int n, int p, float eps, int *out, const int **res);

static int rq_gpu(void *src, int n, int p, float eps, int *out, const int **res) {
  *res = out;
  return region_query(*(mnhn_gpu_stuff *)src, n, p, eps, out);
}

/* ---- the lists of one pass, materialised in parallel ----
 * N(p) of every point for one epsilon: the stored lists filtered by the
 * pass's threshold ((float)dist <= eps for Smith-Waterman lists, acc <= thr
 * for vector lists), with p itself inserted in order where the stored lists
 * leave it out.  Rows are independent, so host threads build them (count,
 * scan, fill) and the sequential expansion then reads each row in place. */
typedef struct {
  int64_t *rp;
  int32_t *col;
} pass_lists;

typedef struct {
  const int64_t *rp;
  const int32_t *col;
  const uint16_t *dist; /* Smith-Waterman lists, or */
  const float *val;    /* vector lists */
  float thr;
  int self; /* insert p itself */
  int64_t *cnt;
  pass_lists *out;
  int r0, r1, phase;
} pass_job;

static inline int keep_entry(const pass_job *jb, int64_t t) {
  return jb->dist ? (float)jb->dist[t] <= jb->thr : jb->val[t] <= jb->thr;
}

static void *pass_worker(void *arg) {
  pass_job *jb = (pass_job *)arg;
  for (int p = jb->r0; p < jb->r1; p++) {
    const int64_t b = jb->rp[p], e = jb->rp[p + 1];
    if (jb->phase == 0) {
      int64_t k = jb->self;
      for (int64_t t = b; t < e; t++) k += keep_entry(jb, t);
      jb->cnt[p] = k;
    } else {
      int32_t *o = jb->out->col + jb->out->rp[p];
      int placed = !jb->self;
      for (int64_t t = b; t < e; t++) {
        const int j = jb->col[t];
        if (!placed && j > p) {
          *o++ = p;
          placed = 1;
        }
        if (keep_entry(jb, t)) *o++ = j;
      }
      if (!placed) *o++ = p;
    }
  }
  return NULL;
}

static int host_threads(int64_t work) {
  if (work < (1 << 20)) return 1;
  const char *v = getenv("MNHN_HOST_THREADS");
  long t = v && *v ? strtol(v, NULL, 10) : sysconf(_SC_NPROCESSORS_ONLN);
  if (t < 1) t = 1;
  if (t > 16) t = 16;
  return (int)t;
}

static pass_lists build_pass_lists(int n, const int64_t *rp, const int32_t *col, const uint16_t *dist,
                                   const float *val, float thr, int self) {
  pass_lists pl;
  pl.rp = (int64_t *)malloc(sizeof(int64_t) * ((size_t)n + 1));
  int64_t *cnt = (int64_t *)malloc(sizeof(int64_t) * (size_t)(n ? n : 1));
  const int T = host_threads(rp[n]);
  pthread_t th[16];
  pass_job jobs[16];
  /* rows split into T ranges of about equal stored entries */
  int bounds[17];
  bounds[0] = 0;
  for (int k = 1; k < T; k++) {
    const int64_t want = rp[n] / T * k;
    int lo = bounds[k - 1], hi = n;
    while (lo < hi) {
      const int mid = lo + (hi - lo) / 2;
      if (rp[mid] < want) lo = mid + 1;
      else hi = mid;
    }
    bounds[k] = lo;
  }
  bounds[T] = n;
  for (int phase = 0; phase < 2; phase++) {
    if (phase == 1) {
      pl.rp[0] = 0;
      for (int p = 0; p < n; p++) pl.rp[p + 1] = pl.rp[p] + cnt[p];
      pl.col = (int32_t *)malloc(sizeof(int32_t) * (size_t)(pl.rp[n] ? pl.rp[n] : 1));
      if (!pl.col) {
        printf("out of memory building the neighbour lists of a pass\n");
        fflush(stdout);
        _exit(1);
      }
    }
    int started[16];
    for (int k = 0; k < T; k++) {
      jobs[k] = (pass_job){rp, col, dist, val, thr, self, cnt, &pl, bounds[k], bounds[k + 1], phase};
      /* a thread that cannot be started does its rows here */
      started[k] = T > 1 && pthread_create(&th[k], NULL, pass_worker, &jobs[k]) == 0;
      if (!started[k]) pass_worker(&jobs[k]);
    }
    for (int k = 0; k < T; k++)
      if (started[k]) pthread_join(th[k], NULL);
  }
  free(cnt);
  return pl;
}

static void free_pass_lists(pass_lists *pl) {
  free(pl->rp);
  free(pl->col);
}

static int rq_pass(void *src, int n, int p, float eps, int *out, const int **res) {
  (void)n;
  (void)eps;
  (void)out;
  const pass_lists *pl = (const pass_lists *)src;
  *res = pl->col + pl->rp[p];
  return (int)(pl->rp[p + 1] - pl->rp[p]);
}

/* x in the ascending row a[0..m)? */
static inline int in_row(const int *a, int m, int x) {
  int lo = 0, hi = m;
  while (lo < hi) {
    const int mid = lo + (hi - lo) / 2;
    if (a[mid] < x) lo = mid + 1;
    else hi = mid;
  }
  return lo < m && a[lo] == x;
}

/* ---- distance-ordered index: many passes over the same stored lists ----
 * For an epsilon sweep every stored row is kept a second time with its
 * entries ordered by distance (ties in ascending column).  For any threshold
 * the stored lists cover, |N(p)| is then one binary search and the stored
 * entries of N(p) are a prefix of the row, so a pass reads what its expansion
 * uses instead of filtering every stored entry. */
typedef struct {
  int n;
  const int64_t *rp;
  const int32_t *col;  /* stored rows, ascending columns */
  const uint16_t *dist; /* Smith-Waterman lists, or */
  const float *val;    /* vector lists */
  int32_t *dcol;       /* the same rows ordered by distance */
  uint16_t *ddist;
  float *dval;
  int symmetric;       /* every stored (p, q, d) has its (q, p, d) */
} dist_index;

/* one pass over the index: keep entries <= thr, insert p itself when self.
 * all: the threshold keeps every stored entry; the index is then a view of the
 * stored rows themselves (column order is a valid order of a row kept whole),
 * so a pass reads the lists in place with no copy or filter. */
typedef struct {
  const dist_index *ix;
  float thr;
  int self;
  int all;
} ix_pass;

static dist_index rows_view(int n, const int64_t *rp, const int32_t *col, const uint16_t *dist, const float *val,
                            int symmetric) {
  dist_index v;
  memset(&v, 0, sizeof(v));
  v.n = n;
  v.rp = rp;
  v.col = col;
  v.dist = dist;
  v.val = val;
  v.dcol = (int32_t *)col; /* read only in passes */
  v.symmetric = symmetric;
  return v;
}

/* sort key whose unsigned order is the order of the distances: the uint16
 * distance itself, or float with every NaN last */
static inline uint32_t dist_key(const dist_index *ix, int64_t t) {
  if (ix->dist) return (uint32_t)ix->dist[t];
  uint32_t u;
  memcpy(&u, &ix->val[t], sizeof(u));
  if ((u & 0x7fffffffu) > 0x7f800000u) return 0xffffffffu;
  return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}

typedef struct {
  dist_index *ix;
  int r0, r1;
  int64_t max_len;
} index_job;

static void *index_worker(void *arg) {
  index_job *jb = (index_job *)arg;
  dist_index *ix = jb->ix;
  const int bytes = ix->dist ? 2 : 4;
  const size_t m = (size_t)(jb->max_len ? jb->max_len : 1);
  uint32_t *k0 = (uint32_t *)malloc(sizeof(uint32_t) * m), *k1 = (uint32_t *)malloc(sizeof(uint32_t) * m);
  int32_t *i0 = (int32_t *)malloc(sizeof(int32_t) * m), *i1 = (int32_t *)malloc(sizeof(int32_t) * m);
  int *hist = ix->dist ? (int *)malloc(sizeof(int) * 65536) : NULL;
  if (!k0 || !k1 || !i0 || !i1 || (ix->dist && !hist)) {
    printf("out of memory ordering the epsilon lists by distance\n");
    fflush(stdout);
    _exit(1);
  }
  for (int p = jb->r0; p < jb->r1; p++) {
    const int64_t b = ix->rp[p];
    const int len = (int)(ix->rp[p + 1] - b);
    if (len == 0) continue;
    if (ix->dist && len > 32) {
      /* integer distances over a short range: one counting sort, written in
       * place (stable, so ties keep ascending columns) */
      int lo = ix->dist[b], hi = lo;
      for (int r = 1; r < len; r++) {
        const int d = ix->dist[b + r];
        lo = d < lo ? d : lo;
        hi = d > hi ? d : hi;
      }
      const int range = hi - lo + 1;
      if (range <= 4 * len + 256) {
        memset(hist, 0, sizeof(int) * (size_t)range);
        for (int r = 0; r < len; r++) hist[ix->dist[b + r] - lo]++;
        int sum = 0;
        for (int d = 0; d < range; d++) {
          const int c = hist[d];
          hist[d] = sum;
          sum += c;
        }
        for (int r = 0; r < len; r++) {
          const int d = ix->dist[b + r];
          const int64_t o = b + hist[d - lo]++;
          ix->dcol[o] = ix->col[b + r];
          ix->ddist[o] = (uint16_t)d;
        }
        continue;
      }
    }
    uint32_t *ka = k0, *kb = k1;
    int32_t *ia = i0, *ib = i1;
    for (int r = 0; r < len; r++) {
      ka[r] = dist_key(ix, b + r);
      ia[r] = r;
    }
    if (len <= 32) {
      /* insertion sort: stable, ties keep ascending columns */
      for (int r = 1; r < len; r++) {
        const uint32_t k = ka[r];
        const int32_t v = ia[r];
        int s = r - 1;
        while (s >= 0 && ka[s] > k) {
          ka[s + 1] = ka[s];
          ia[s + 1] = ia[s];
          s--;
        }
        ka[s + 1] = k;
        ia[s + 1] = v;
      }
    } else {
      /* least significant digit first, 8 bits a pass: stable */
      for (int sh = 0; sh < 8 * bytes; sh += 8) {
        int cnt[256];
        memset(cnt, 0, sizeof(cnt));
        for (int r = 0; r < len; r++) cnt[(ka[r] >> sh) & 255u]++;
        if (cnt[(ka[0] >> sh) & 255u] == len) continue; /* one digit value: order unchanged */
        int sum = 0;
        for (int d = 0; d < 256; d++) {
          const int c = cnt[d];
          cnt[d] = sum;
          sum += c;
        }
        for (int r = 0; r < len; r++) {
          const int o = cnt[(ka[r] >> sh) & 255u]++;
          kb[o] = ka[r];
          ib[o] = ia[r];
        }
        uint32_t *tk = ka;
        ka = kb;
        kb = tk;
        int32_t *ti = ia;
        ia = ib;
        ib = ti;
      }
    }
    for (int r = 0; r < len; r++) {
      const int64_t t = b + ia[r];
      ix->dcol[b + r] = ix->col[t];
      if (ix->dist) ix->ddist[b + r] = ix->dist[t];
      else ix->dval[b + r] = ix->val[t];
    }
  }
  free(k0);
  free(k1);
  free(i0);
  free(i1);
  free(hist);
  return NULL;
}

/* Builds the index of stored lists; 0 (and no index) when its copy of the
 * lists would take more than a quarter of the host's memory.  symmetric:
 * the caller vouches that every stored (p, q, d) has its (q, p, d). */
static int build_dist_index(dist_index *ix, int n, const int64_t *rp, const int32_t *col, const uint16_t *dist,
                            const float *val, int symmetric) {
  memset(ix, 0, sizeof(*ix));
  const int64_t nnz = rp[n];
  const size_t per = sizeof(int32_t) + (dist ? sizeof(uint16_t) : sizeof(float));
  const double phys = (double)sysconf(_SC_PHYS_PAGES) * (double)sysconf(_SC_PAGE_SIZE);
  if (phys > 0 && (double)nnz * (double)per > phys / 4) return 0;
  ix->n = n;
  ix->rp = rp;
  ix->col = col;
  ix->dist = dist;
  ix->val = val;
  const size_t cap = (size_t)(nnz ? nnz : 1);
  ix->dcol = (int32_t *)malloc(sizeof(int32_t) * cap);
  if (dist) ix->ddist = (uint16_t *)malloc(sizeof(uint16_t) * cap);
  else ix->dval = (float *)malloc(sizeof(float) * cap);
  if (!ix->dcol || (dist && !ix->ddist) || (!dist && !ix->dval)) {
    free(ix->dcol);
    free(ix->ddist);
    free(ix->dval);
    memset(ix, 0, sizeof(*ix));
    return 0;
  }
  const int T = host_threads(nnz);
  pthread_t th[16];
  index_job jobs[16];
  int started[16];
  int lo = 0;
  for (int k = 0; k < T; k++) {
    /* rows split into T ranges of about equal stored entries */
    int hi = n;
    if (k + 1 < T) {
      const int64_t want = nnz / T * (k + 1);
      int a = lo, z = n;
      while (a < z) {
        const int mid = a + (z - a) / 2;
        if (rp[mid] < want) a = mid + 1;
        else z = mid;
      }
      hi = a;
    }
    int64_t ml = 0;
    for (int p = lo; p < hi; p++)
      if (rp[p + 1] - rp[p] > ml) ml = rp[p + 1] - rp[p];
    jobs[k] = (index_job){ix, lo, hi, ml};
    lo = hi;
  }
  for (int k = 0; k < T; k++) {
    started[k] = T > 1 && pthread_create(&th[k], NULL, index_worker, &jobs[k]) == 0;
    if (!started[k]) index_worker(&jobs[k]);
  }
  for (int k = 0; k < T; k++)
    if (started[k]) pthread_join(th[k], NULL);
  ix->symmetric = symmetric != 0;
  return 1;
}

static void free_dist_index(dist_index *ix) {
  free(ix->dcol);
  free(ix->ddist);
  free(ix->dval);
  memset(ix, 0, sizeof(*ix));
}

/* MNHN_SWEEP_INDEX=0: every pass filters the stored lists (no index) */
static int index_allowed(void) {
  const char *v = getenv("MNHN_SWEEP_INDEX");
  return !(v && strcmp(v, "0") == 0);
}

static void drop_index(mnhn_eps_cache *c) {
  if (c->index) {
    free_dist_index((dist_index *)c->index);
    free(c->index);
  }
  c->index = NULL;
  c->index_tried = 0;
}

/* The index of the cache's stored lists for a sweep, built on first use;
 * NULL outside a sweep or when it does not fit. */
static const dist_index *cache_index(mnhn_eps_cache *c, int n, const int64_t *rp, const int32_t *col,
                                     const uint16_t *dist, const float *val) {
  if (!c->sweep || !index_allowed()) return NULL;
  if (!c->index_tried) {
    c->index_tried = 1;
    const double t0 = now_s();
    dist_index *ix = (dist_index *)malloc(sizeof(dist_index));
    /* the engines' lists hold both directions of every pair */
    if (ix && build_dist_index(ix, n, rp, col, dist, val, 1)) {
      c->index = ix;
      if (timing_on())
        fprintf(stderr, "[mnhn] epsilon lists ordered by distance: %.3f s\n", now_s() - t0);
    } else {
      free(ix);
    }
  }
  return (const dist_index *)c->index;
}

static inline int ix_keep_sorted(const dist_index *ix, int64_t t, float thr) {
  return ix->ddist ? (float)ix->ddist[t] <= thr : ix->dval[t] <= thr;
}

/* |N(p)|: the kept prefix of the distance-ordered row, plus p itself */
static int ix_count(const ix_pass *ps, int p) {
  const dist_index *ix = ps->ix;
  if (ps->all) return (int)(ix->rp[p + 1] - ix->rp[p]) + ps->self;
  int64_t lo = ix->rp[p], hi = ix->rp[p + 1];
  const int64_t b = lo;
  while (lo < hi) {
    const int64_t mid = lo + (hi - lo) / 2;
    if (ix_keep_sorted(ix, mid, ps->thr)) lo = mid + 1;
    else hi = mid;
  }
  return (int)(lo - b) + ps->self;
}

/* x in N(q)?  A binary search of the row in column order. */
static int ix_member(const ix_pass *ps, int q, int x) {
  const dist_index *ix = ps->ix;
  if (x == q && ps->self) return 1;
  int64_t lo = ix->rp[q], hi = ix->rp[q + 1];
  const int64_t e = hi;
  while (lo < hi) {
    const int64_t mid = lo + (hi - lo) / 2;
    if (ix->col[mid] < x) lo = mid + 1;
    else hi = mid;
  }
  if (lo == e || ix->col[lo] != x) return 0;
  if (ps->all) return 1;
  return ix->dist ? (float)ix->dist[lo] <= ps->thr : ix->val[lo] <= ps->thr;
}

/* ---- triggers: the points outside a large cluster, seen from their side ----
 * With symmetric lists, x is in N(q) exactly when q is in N(x).  Points still
 * outside the cluster whose own N(x) is short are registered once under each
 * of their neighbours; a core point q then finds them in its trigger row
 * instead of a binary search of N(q) per outside point.  Outside points with
 * long N(x) keep the binary search. */
typedef struct {
  int *px;     /* [n] stored prefix length of each outside point */
  int64_t *rp; /* [n + 1] trigger rows */
  int *col;    /* outside points, ascending within each row */
  int64_t cap;
} triggers;

/* Registers the outside points rest[0..n_rest) with short lists and returns
 * how many (those with long lists, still ascending) remain in rest[]. */
static int build_triggers(const ix_pass *ps, int n, int *rest, int n_rest, triggers *tg) {
  const dist_index *ix = ps->ix;
  for (int t = 0; t < n_rest; t++) tg->px[t] = ix_count(ps, rest[t]) - ps->self;
  /* the longest list registered: as long as the triggers fit in cap */
  int K = 4096;
  for (;;) {
    int64_t sum = 0;
    for (int t = 0; t < n_rest; t++)
      if (tg->px[t] <= K) sum += tg->px[t];
    if (sum <= tg->cap || K == 0) break;
    K >>= 1;
  }
  memset(tg->rp, 0, sizeof(int64_t) * ((size_t)n + 1));
  for (int t = 0; t < n_rest; t++) {
    if (tg->px[t] > K) continue;
    const int x = rest[t];
    const int32_t *d = ix->dcol + ix->rp[x];
    for (int u = 0; u < tg->px[t]; u++)
      if (d[u] != x) tg->rp[d[u] + 1]++;
  }
  for (int y = 0; y < n; y++) tg->rp[y + 1] += tg->rp[y];
  for (int t = 0; t < n_rest; t++) {
    if (tg->px[t] > K) continue;
    const int x = rest[t];
    const int32_t *d = ix->dcol + ix->rp[x];
    for (int u = 0; u < tg->px[t]; u++)
      if (d[u] != x) tg->col[tg->rp[d[u]]++] = x;
  }
  for (int y = n; y > 0; y--) tg->rp[y] = tg->rp[y - 1];
  tg->rp[0] = 0;
  int k = 0;
  for (int t = 0; t < n_rest; t++)
    if (tg->px[t] > K) rest[k++] = rest[t];
  return k;
}

/* a[0..k) ascending (values >= 0); tmp holds k ints */
static void sort_ascending(int *a, int k, int *tmp) {
  if (k <= 48) {
    for (int r = 1; r < k; r++) {
      const int v = a[r];
      int s = r - 1;
      while (s >= 0 && a[s] > v) {
        a[s + 1] = a[s];
        s--;
      }
      a[s + 1] = v;
    }
    return;
  }
  int mx = 0;
  for (int r = 0; r < k; r++)
    if (a[r] > mx) mx = a[r];
  int *src = a, *dst = tmp;
  for (int sh = 0; sh < 31 && (mx >> sh) > 0; sh += 11) {
    int cnt[2048];
    memset(cnt, 0, sizeof(cnt));
    for (int r = 0; r < k; r++) cnt[(src[r] >> sh) & 2047]++;
    int sum = 0;
    for (int d = 0; d < 2048; d++) {
      const int c = cnt[d];
      cnt[d] = sum;
      sum += c;
    }
    for (int r = 0; r < k; r++) dst[cnt[(src[r] >> sh) & 2047]++] = src[r];
    int *t = src;
    src = dst;
    dst = t;
  }
  if (src != a) memcpy(a, src, sizeof(int) * (size_t)k);
}

/* ---- parallel row scans for a batch of queue elements ----
 * The expansion appends, for each core point q of the queue in order, the
 * members of N(q) still outside the union, ascending.  For a batch of
 * consecutive queue elements, host threads scan their rows in parallel
 * against the union as it stands at the batch's start; at each element's turn
 * the expansion keeps the scanned points still outside (the union only
 * grows), which appends the same points in the same order as scanning the row
 * then.  Rows come from the materialised pass lists or from the index. */
typedef struct scanner scanner;
typedef struct {
  scanner *sc;
  int w;
} scan_arg;

struct scanner {
  int nth;      /* shares of a batch */
  int started[16];
  scan_arg args[16];
  pthread_t th[16];
  pthread_mutex_t mu;
  pthread_cond_t cv;
  int armed; /* barriers initialised for the threads that started */
  pthread_barrier_t start, done;
  int quit;
  const pass_lists *pl; /* materialised rows (ascending, p included), or */
  const ix_pass *ix;    /* index rows (kept prefix) */
  int64_t cap;          /* scanned entries per batch */
  int max_elems;
  /* the batch */
  const int *qs;
  int nq;
  const int *stamp;
  int mark;
  int *cq;      /* [max_elems] |N(q)| */
  int64_t *off; /* [max_elems + 1] */
  int *len;     /* [max_elems] scanned points, -1: not scanned */
  int *out;     /* [cap] */
  int *tmp[16]; /* per-share sort scratch, index rows */
};

/* Share w of the batch: the elements whose output starts in the w-th part of
 * the batch's scanned entries. */
static void scan_share(scanner *sc, int w) {
  const int64_t tot = sc->off[sc->nq];
  const int64_t lo = tot / sc->nth * w, hi = w + 1 == sc->nth ? tot : tot / sc->nth * (w + 1);
  for (int k = 0; k < sc->nq; k++) {
    if (sc->len[k] < 0 || sc->off[k] < lo || sc->off[k] >= hi) continue;
    const int q = sc->qs[k];
    const int32_t *row;
    int m;
    if (sc->pl) {
      row = sc->pl->col + sc->pl->rp[q];
      m = (int)(sc->pl->rp[q + 1] - sc->pl->rp[q]);
    } else {
      row = sc->ix->ix->dcol + sc->ix->ix->rp[q];
      m = sc->cq[k] - sc->ix->self;
    }
    int *o = sc->out + sc->off[k];
    int c = 0;
    for (int t = 0; t < m; t++) {
      const int x = row[t];
      if (sc->stamp[x] != sc->mark) o[c++] = x;
    }
    if (sc->ix && !sc->ix->all) sort_ascending(o, c, sc->tmp[w]);
    sc->len[k] = c;
  }
}

static void *scan_worker(void *arg) {
  scan_arg *a = (scan_arg *)arg;
  scanner *sc = a->sc;
  pthread_mutex_lock(&sc->mu);
  while (!sc->armed) pthread_cond_wait(&sc->cv, &sc->mu);
  pthread_mutex_unlock(&sc->mu);
  for (;;) {
    pthread_barrier_wait(&sc->start);
    if (sc->quit) break;
    scan_share(sc, a->w);
    pthread_barrier_wait(&sc->done);
  }
  return NULL;
}

static void scanner_free(scanner *sc) {
  for (int w = 0; w < 16; w++) free(sc->tmp[w]);
  free(sc->cq);
  free(sc->off);
  free(sc->len);
  free(sc->out);
  free(sc);
}

/* A scanner for rows of total stored entries; NULL when one thread would do
 * (small lists) or no helper thread starts.  Shares whose thread did not start
 * are run by the caller. */
static scanner *scanner_create(int n, int64_t total, const pass_lists *pl, const ix_pass *ix) {
  const int T = host_threads(total >> 2);
  if (T < 2 || n <= 0) return NULL;
  scanner *sc = (scanner *)calloc(1, sizeof(scanner));
  if (!sc) return NULL;
  sc->nth = T;
  sc->pl = pl;
  sc->ix = ix;
  sc->cap = (int64_t)1 << 22;
  if (sc->cap < (int64_t)n + 1) sc->cap = (int64_t)n + 1;
  sc->max_elems = n;
  sc->cq = (int *)malloc(sizeof(int) * (size_t)n);
  sc->off = (int64_t *)malloc(sizeof(int64_t) * ((size_t)n + 1));
  sc->len = (int *)malloc(sizeof(int) * (size_t)n);
  sc->out = (int *)malloc(sizeof(int) * (size_t)sc->cap);
  int ok = sc->cq && sc->off && sc->len && sc->out;
  for (int w = 0; w < T && ok && ix && !ix->all; w++) ok = (sc->tmp[w] = (int *)malloc(sizeof(int) * (size_t)n)) != NULL;
  if (!ok || pthread_mutex_init(&sc->mu, NULL) != 0) {
    scanner_free(sc);
    return NULL;
  }
  if (pthread_cond_init(&sc->cv, NULL) != 0) {
    pthread_mutex_destroy(&sc->mu);
    scanner_free(sc);
    return NULL;
  }
  int nstarted = 0;
  for (int w = 1; w < T; w++) {
    sc->args[w] = (scan_arg){sc, w};
    sc->started[w] = pthread_create(&sc->th[w], NULL, scan_worker, &sc->args[w]) == 0;
    nstarted += sc->started[w];
  }
  if (nstarted == 0) {
    pthread_cond_destroy(&sc->cv);
    pthread_mutex_destroy(&sc->mu);
    scanner_free(sc);
    return NULL;
  }
  if (pthread_barrier_init(&sc->start, NULL, (unsigned)(nstarted + 1)) != 0 ||
      pthread_barrier_init(&sc->done, NULL, (unsigned)(nstarted + 1)) != 0) {
    printf("could not set up host threads for the DBSCAN expansion\n");
    fflush(stdout);
    _exit(1);
  }
  pthread_mutex_lock(&sc->mu);
  sc->armed = 1;
  pthread_cond_broadcast(&sc->cv);
  pthread_mutex_unlock(&sc->mu);
  return sc;
}

static void scanner_destroy(scanner *sc) {
  if (!sc) return;
  sc->quit = 1;
  pthread_barrier_wait(&sc->start);
  for (int w = 1; w < sc->nth; w++)
    if (sc->started[w]) pthread_join(sc->th[w], NULL);
  pthread_barrier_destroy(&sc->start);
  pthread_barrier_destroy(&sc->done);
  pthread_cond_destroy(&sc->cv);
  pthread_mutex_destroy(&sc->mu);
  scanner_free(sc);
}

/* Plans the batch starting at queue element qs[0] (avail elements queued)
 * and scans its rows in parallel; returns how many elements it covers.  An
 * element is scanned when it is unvisited, a core point, and its N(q) is not
 * long enough for the walk over the points outside the union (that test only
 * becomes more likely as the union grows, so a row not scanned here is not
 * needed later in the batch either). */
static int scan_batch(scanner *sc, const int *qs, int avail, const unsigned char *visited, const int *stamp,
                      int mark, int in_union, int n, int minpts) {
  int64_t tot = 0;
  int k = 0;
  for (; k < avail && k < sc->max_elems; k++) {
    const int q = qs[k];
    int c = -1;
    if (!visited[q]) {
      const int cq = sc->ix ? ix_count(sc->ix, q) : (int)(sc->pl->rp[q + 1] - sc->pl->rp[q]);
      sc->cq[k] = cq;
      if (cq >= minpts && cq <= 8 * (n - in_union)) c = sc->ix ? cq - sc->ix->self : cq;
    }
    if (c > 0 && k > 0 && tot + c > sc->cap) break;
    sc->off[k] = tot;
    sc->len[k] = c < 0 ? -1 : 0;
    if (c > 0) tot += c;
  }
  sc->off[k] = tot;
  sc->nq = k;
  if (tot < (1 << 16)) {
    for (int t = 0; t < k; t++) sc->len[t] = -1; /* not worth the threads: scan at each turn */
    return k;
  }
  sc->qs = qs;
  sc->stamp = stamp;
  sc->mark = mark;
  pthread_barrier_wait(&sc->start);
  scan_share(sc, 0);
  for (int w = 1; w < sc->nth; w++)
    if (!sc->started[w]) scan_share(sc, w);
  pthread_barrier_wait(&sc->done);
  return k;
}

/* One DBSCAN pass.  N(p) comes from rq (materialised pass lists, or one
 * one-vs-all launch per query) or, with ix, from stored rows read through an
 * index pass (distance-ordered rows, or the rows themselves when the pass
 * keeps every entry).  With long lists, host threads scan batches of queue
 * rows ahead of the expansion (sc). */
static split_set dbscan_core(int n, float epsilon, int minpts, region_fn rq, void *src, const ix_pass *ix) {
  split_set ret;
  ret.n_clusters = 0;
  ret.clusters = (cluster *)malloc(sizeof(cluster) * (size_t)(n ? n : 1));
  if (n == 0) return ret;

  unsigned char *visited = (unsigned char *)calloc((size_t)n, 1);
  unsigned char *member = (unsigned char *)calloc((size_t)n, 1);
  int *stamp = (int *)calloc((size_t)n, sizeof(int)); /* union membership, per cluster */
  int *nb = (int *)malloc(sizeof(int) * (size_t)n);
  int *queue = (int *)malloc(sizeof(int) * (size_t)n);
  int *rest = (int *)malloc(sizeof(int) * (size_t)n); /* points outside the union, ascending */
  /* the members of the cluster being expanded; copied out at its exact size
   * when it is done (one n-int buffer per cluster would cost a page mapping
   * and a shrinking realloc each) */
  int *grow = (int *)malloc(sizeof(int) * (size_t)n);
  int *tmp = ix ? (int *)malloc(sizeof(int) * (size_t)n) : NULL;
  triggers tg = {NULL, NULL, NULL, 0};
  const int use_triggers = ix && ix->ix->symmetric;
  scanner *sc = NULL;
  if (ix) sc = scanner_create(n, ix->ix->rp[n], NULL, ix);
  else if (rq == rq_pass && src) sc = scanner_create(n, ((const pass_lists *)src)->rp[n], (const pass_lists *)src, NULL);
  if (use_triggers) {
    tg.cap = (int64_t)n * 32 + 4096;
    tg.px = (int *)malloc(sizeof(int) * (size_t)n);
    tg.rp = (int64_t *)malloc(sizeof(int64_t) * ((size_t)n + 1));
    tg.col = (int *)malloc(sizeof(int) * (size_t)tg.cap);
  }

  for (int i = 0; i < n; i++) {
    if (visited[i]) continue;
    visited[i] = 1;
    const int *nbi = nb;
    int cnt = ix ? ix_count(ix, i) : rq(src, n, i, epsilon, nb, &nbi);
    if (!(minpts < cnt)) continue;
    if (ix) {
      /* N(i) in ascending order: the kept prefix, i itself, sorted */
      const int m = cnt - ix->self;
      memcpy(nb, ix->ix->dcol + ix->ix->rp[i], sizeof(int) * (size_t)m);
      if (ix->self) nb[m] = i;
      sort_ascending(nb, cnt, tmp);
    }
    /* expand_cluster */
    const int id = ret.n_clusters;
    const int mark = id + 1;
    cluster *cl = &ret.clusters[id];
    cl->id = id;
    cl->members = grow;
    cl->n_members = 0;
    cl->members[cl->n_members++] = i;
    member[i] = 1;
    int qn = 0;
    for (int t = 0; t < cnt; t++) {
      queue[qn++] = nbi[t];
      stamp[nbi[t]] = mark;
    }
    int in_union = cnt; /* the union holds queue[0 .. qn) */
    int n_rest = -1;    /* rest[] not built yet for this cluster */
    int bstart = 0, bend = 0; /* queue[bstart .. bend): the batch sc scanned */
    for (int qi = 0; qi < qn; qi++) {
      if (sc && qi == bend) {
        bstart = qi;
        bend = qi + scan_batch(sc, queue + qi, qn - qi, visited, stamp, mark, in_union, n, minpts);
      }
      const int q = queue[qi];
      if (!visited[q]) {
        visited[q] = 1;
        const int *nbq = nb;
        const int bk = qi - bstart;
        const int pre = sc && sc->len[bk] >= 0; /* row scanned with the batch */
        int cq = pre ? sc->cq[bk] : ix ? ix_count(ix, q) : rq(src, n, q, epsilon, nb, &nbq);
        if (pre && !ix) nbq = ((const pass_lists *)src)->col + ((const pass_lists *)src)->rp[q];
        if (cq >= minpts) {
          /* the members of N(q) outside the union, in ascending order: by
           * scanning N(q), or -- when N(q) is much longer than the set of
           * points still outside -- by scanning that set (kept ascending)
           * and looking each point up in N(q); both append the same points
           * in the same order */
          if (cq > 8 * (n - in_union)) {
            if (n_rest < 0) {
              n_rest = 0;
              for (int x = 0; x < n; x++)
                if (stamp[x] != mark) rest[n_rest++] = x;
              if (use_triggers) n_rest = build_triggers(ix, n, rest, n_rest, &tg);
            } else {
              int k = 0;
              for (int t = 0; t < n_rest; t++)
                if (stamp[rest[t]] != mark) rest[k++] = rest[t];
              n_rest = k;
            }
            if (use_triggers) {
              /* outside points with long lists by binary search, those with
               * short lists from q's trigger row; both ascending, disjoint */
              int kl = 0, ks = 0;
              for (int t = 0; t < n_rest; t++)
                if (ix_member(ix, q, rest[t])) nb[kl++] = rest[t];
              for (int64_t u = tg.rp[q]; u < tg.rp[q + 1]; u++)
                if (stamp[tg.col[u]] != mark) tmp[ks++] = tg.col[u];
              int a = 0, c = 0;
              while (a < kl || c < ks) {
                const int x = c == ks || (a < kl && nb[a] < tmp[c]) ? nb[a++] : tmp[c++];
                stamp[x] = mark;
                queue[qn++] = x;
              }
              in_union += kl + ks;
            } else {
              for (int t = 0; t < n_rest; t++) {
                const int x = rest[t];
                if (ix ? ix_member(ix, q, x) : in_row(nbq, cq, x)) {
                  stamp[x] = mark;
                  queue[qn++] = x;
                  in_union++;
                }
              }
            }
          } else if (pre) {
            /* the batch's scan, less what joined the union since */
            const int *o = sc->out + sc->off[bk];
            for (int t = 0; t < sc->len[bk]; t++) {
              const int x = o[t];
              if (stamp[x] != mark) {
                stamp[x] = mark;
                queue[qn++] = x;
                in_union++;
              }
            }
          } else if (ix) {
            /* the kept prefix outside the union, sorted; q itself is in it */
            const int32_t *d = ix->ix->dcol + ix->ix->rp[q];
            const int m = cq - ix->self;
            int k = 0;
            for (int t = 0; t < m; t++) {
              const int x = d[t];
              if (stamp[x] != mark) {
                stamp[x] = mark;
                nb[k++] = x;
              }
            }
            if (!ix->all) sort_ascending(nb, k, tmp); /* a whole row is already ascending */
            memcpy(queue + qn, nb, sizeof(int) * (size_t)k);
            qn += k;
            in_union += k;
          } else {
            for (int t = 0; t < cq; t++) {
              const int x = nbq[t];
              if (stamp[x] != mark) {
                stamp[x] = mark;
                queue[qn++] = x;
                in_union++;
              }
            }
          }
        }
      }
      if (!member[q]) {
        cl->members[cl->n_members++] = q;
        member[q] = 1;
      }
    }
    cl->members = (int *)malloc(sizeof(int) * (size_t)(cl->n_members ? cl->n_members : 1));
    if (!cl->members) {
      printf("out of memory storing a cluster\n");
      fflush(stdout);
      _exit(1);
    }
    memcpy(cl->members, grow, sizeof(int) * (size_t)cl->n_members);
    ret.n_clusters++;
  }
  free(visited);
  free(member);
  free(stamp);
  free(nb);
  free(queue);
  free(rest);
  free(grow);
  free(tmp);
  free(tg.px);
  free(tg.rp);
  free(tg.col);
  scanner_destroy(sc);
  return ret;
}

/* dbscan.c:980-989: the -1 defaults and their warnings */
static void apply_defaults(float *epsilon, int *minpts) {
  if (*epsilon == -1) {
    printf("Warning @dbscan: epsilon value is: %f\n", *epsilon);
    *epsilon = 0.015f;
    printf("Warning @dbscan: epsilon was changed to: %f\n", *epsilon);
  }
  if (*minpts == -1) {
    printf("Warning @dbscan: minpts value is: %i\n", *minpts);
    *minpts = 5;
    printf("Warning @dbscan: minpts was changed to: %i\n", *minpts);
  }
}

/* One pass over stored lists.  all_kept: the threshold keeps every stored
 * entry, so the rows are read in place (no filtered copy); otherwise each
 * pass materialises its lists. */
static split_set stored_pass(int n, float epsilon, int minpts, const int64_t *rp, const int32_t *col,
                             const uint16_t *dist, const float *val, float thr, int self, int all_kept,
                             int symmetric) {
  if (all_kept) {
    const dist_index v = rows_view(n, rp, col, dist, val, symmetric);
    const ix_pass ps = {&v, thr, self, 1};
    return dbscan_core(n, epsilon, minpts, NULL, NULL, &ps);
  }
  pass_lists pl = build_pass_lists(n, rp, col, dist, val, thr, self);
  split_set s = dbscan_core(n, epsilon, minpts, rq_pass, &pl, NULL);
  free_pass_lists(&pl);
  return s;
}

split_set dbscan_SW_GPU(dataset ds, float epsilon, int minpts, mnhn_gpu_stuff g) {
  const int n = ds.n_values;
  apply_defaults(&epsilon, &minpts);
  if (n > 0) {
    if (g.cache->per_query_launch) use_per_query_launch(g.cache, n);
    else (void)ensure_lists(g, epsilon);
  }
  const double t0 = now_s();
  split_set s;
  if (n > 0 && !g.cache->per_query_launch) {
    const mnhn_csr *c = &g.cache->csr;
    const dist_index *ix = cache_index(g.cache, n, c->row_ptr, c->col, c->dist, NULL);
    if (ix) {
      const ix_pass ps = {ix, epsilon, (float)0 <= epsilon, 0};
      s = dbscan_core(n, epsilon, minpts, NULL, NULL, &ps);
    } else {
      s = stored_pass(n, epsilon, minpts, c->row_ptr, c->col, c->dist, NULL, epsilon, (float)0 <= epsilon,
                      eps_floor(epsilon) >= c->eps_max, 1);
    }
  } else {
    s = dbscan_core(n, epsilon, minpts, rq_gpu, &g, NULL);
  }
  if (timing_on())
    fprintf(stderr, "[mnhn] dbscan pass eps %g: %d clusters, %.3f s\n", epsilon, s.n_clusters, now_s() - t0);
  return s;
}

/* The reference's CPU-build entry point (dbscan.h:30): one pass with a device
 * context of its own. */
split_set dbscan_SW(dataset ds, float epsilon, int minpts) {
  mnhn_gpu_stuff g = mnhn_gpu_initialization(ds);
  split_set s = dbscan_SW_GPU(ds, epsilon, minpts, g);
  mnhn_gpu_destroy(g);
  return s;
}

split_set mnhn_dbscan_csr(const mnhn_csr *csr, float epsilon, int minpts) {
  const int n = csr ? csr->n_rows : 0;
  if (n == 0) return dbscan_core(0, epsilon, minpts, rq_pass, NULL, NULL);
  return stored_pass(n, epsilon, minpts, csr->row_ptr, csr->col, csr->dist, NULL, epsilon, (float)0 <= epsilon,
                     eps_floor(epsilon) >= csr->eps_max, 0);
}

/* ---- vector datasets (dbscan.c:563-780) ---- */

/* The threshold the accumulator is compared with: eps * eps in float for L2
 * (dbscan.c:577), eps for L1. */
static float vec_thr(int norm, float eps) { return norm == MNHN_NORM_L2 ? eps * eps : eps; }

/* Lists up to the threshold of eps + lookahead (never below eps's own);
 * passes whose threshold is <= the lists' reuse them. */
static void ensure_vec_lists(mnhn_gpu_stuff g, float eps) {
  mnhn_eps_cache *c = g.cache;
  const float thr = vec_thr(g.norm, eps);
  if (c->have_v && thr <= c->vcsr.thr) return;
  drop_index(c);
  if (c->have_v) mnhn_vcsr_free(&c->vcsr);
  c->have_v = 0;
  float t_max = vec_thr(g.norm, eps + c->lookahead);
  if (!(t_max >= thr)) t_max = thr;
  int rc = mnhn_vec_eps_csr(g.vec, g.norm, t_max, &c->vcsr);
  if (rc) die_engine("mnhn_vec_eps_csr", rc);
  c->have_v = 1;
}

static split_set dbscan_vec(dataset ds, float epsilon, int minpts, mnhn_gpu_stuff g, int norm) {
  if (!g.vec || g.norm != norm) {
    printf("dbscan_L%d_GPU needs a context from mnhn_gpu_initialization_values with the same norm\n", norm);
    fflush(stdout);
    _exit(1);
  }
  apply_defaults(&epsilon, &minpts);
  const int n = ds.n_values;
  if (n == 0) return dbscan_core(0, epsilon, minpts, rq_pass, NULL, NULL);
  ensure_vec_lists(g, epsilon);
  /* the stored rows list p itself when its own accumulator qualifies */
  const mnhn_vcsr *c = &g.cache->vcsr;
  const dist_index *ix = cache_index(g.cache, n, c->row_ptr, c->col, NULL, c->val);
  if (ix) {
    const ix_pass ps = {ix, vec_thr(norm, epsilon), 0, 0};
    return dbscan_core(n, epsilon, minpts, NULL, NULL, &ps);
  }
  const float thr = vec_thr(norm, epsilon);
  return stored_pass(n, epsilon, minpts, c->row_ptr, c->col, NULL, c->val, thr, 0, thr >= c->thr, 1);
}

split_set dbscan_L1_GPU(dataset ds, float epsilon, int minpts, mnhn_gpu_stuff g) {
  return dbscan_vec(ds, epsilon, minpts, g, MNHN_NORM_L1);
}

split_set dbscan_L2_GPU(dataset ds, float epsilon, int minpts, mnhn_gpu_stuff g) {
  return dbscan_vec(ds, epsilon, minpts, g, MNHN_NORM_L2);
}

split_set dbscan_L1(dataset ds, float epsilon, int minpts) {
  mnhn_gpu_stuff g = mnhn_gpu_initialization_values(ds, MNHN_NORM_L1);
  split_set s = dbscan_L1_GPU(ds, epsilon, minpts, g);
  mnhn_gpu_destroy(g);
  return s;
}

split_set dbscan_L2(dataset ds, float epsilon, int minpts) {
  mnhn_gpu_stuff g = mnhn_gpu_initialization_values(ds, MNHN_NORM_L2);
  split_set s = dbscan_L2_GPU(ds, epsilon, minpts, g);
  mnhn_gpu_destroy(g);
  return s;
}

split_set mnhn_dbscan_vcsr(const mnhn_vcsr *csr, float epsilon, int minpts) {
  const int n = csr ? csr->n_rows : 0;
  if (n == 0) return dbscan_core(0, epsilon, minpts, rq_pass, NULL, NULL);
  const float thr = vec_thr(csr->norm, epsilon);
  return stored_pass(n, epsilon, minpts, csr->row_ptr, csr->col, NULL, csr->val, thr, 0, thr >= csr->thr, 0);
}

/* Several passes over the same lists: through the distance-ordered index when
 * it fits, else one materialised pass each. */
static int dbscan_sweep(int n, const int64_t *rp, const int32_t *col, const uint16_t *dist, const float *val,
                        int norm, const float *eps, int n_eps, int minpts, int symmetric, split_set *out) {
  if (n_eps < 0 || (n_eps > 0 && (!eps || !out))) return MNHN_ERR_ARG;
  dist_index ix;
  double t0 = now_s();
  const int have = n > 0 && n_eps > 0 && index_allowed() && build_dist_index(&ix, n, rp, col, dist, val, symmetric);
  if (timing_on() && have) fprintf(stderr, "[mnhn] epsilon lists ordered by distance: %.3f s\n", now_s() - t0);
  for (int k = 0; k < n_eps; k++) {
    t0 = now_s();
    const float thr = dist ? eps[k] : vec_thr(norm, eps[k]);
    const int self = dist ? (float)0 <= eps[k] : 0;
    if (n == 0) {
      out[k] = dbscan_core(0, eps[k], minpts, rq_pass, NULL, NULL);
    } else if (have) {
      const ix_pass ps = {&ix, thr, self, 0};
      out[k] = dbscan_core(n, eps[k], minpts, NULL, NULL, &ps);
    } else {
      pass_lists pl = build_pass_lists(n, rp, col, dist, val, thr, self);
      out[k] = dbscan_core(n, eps[k], minpts, rq_pass, &pl, NULL);
      free_pass_lists(&pl);
    }
    if (timing_on())
      fprintf(stderr, "[mnhn] dbscan pass eps %g: %d clusters, %.3f s\n", eps[k], out[k].n_clusters, now_s() - t0);
  }
  if (have) free_dist_index(&ix);
  return MNHN_OK;
}

int mnhn_dbscan_csr_sweep(const mnhn_csr *csr, const float *epsilon, int n_eps, int minpts, int symmetric,
                          split_set *out) {
  if (!csr) return MNHN_ERR_ARG;
  return dbscan_sweep(csr->n_rows, csr->row_ptr, csr->col, csr->dist, NULL, 0, epsilon, n_eps, minpts, symmetric,
                      out);
}

int mnhn_dbscan_vcsr_sweep(const mnhn_vcsr *csr, const float *epsilon, int n_eps, int minpts, int symmetric,
                           split_set *out) {
  if (!csr) return MNHN_ERR_ARG;
  return dbscan_sweep(csr->n_rows, csr->row_ptr, csr->col, NULL, csr->val, csr->norm, epsilon, n_eps, minpts,
                      symmetric, out);
}

void adaptive_dbscan(split_set (*dbscanner)(dataset, float, int, mnhn_gpu_stuff),
                     dataset ds, float epsilon_start, float epsilon_inc, int minpts,
                     char *split_files_prefix, int n_threads) {
  (void)n_threads; /* epsilon values are scanned in order by one thread */
  const double t_init = now_s();
  mnhn_gpu_stuff g = dbscanner == dbscan_L2_GPU   ? mnhn_gpu_initialization_values(ds, MNHN_NORM_L2)
                     : dbscanner == dbscan_L1_GPU ? mnhn_gpu_initialization_values(ds, MNHN_NORM_L1)
                                                  : mnhn_gpu_initialization(ds);
  /* neighbour lists are computed up to 8 increments beyond the epsilon asked
   * for, so the sweep's next passes reuse them */
  mnhn_gpu_set_lookahead(g, 8.0f * (epsilon_inc > 0 ? epsilon_inc : 0));
  g.cache->sweep = 1;
  if (timing_on()) fprintf(stderr, "[mnhn] device context: %.3f s\n", now_s() - t_init);

  printf("Performing adaptive clustering with parameters: \n"
         "minpts: %i\n"
         "epsilon_start: %f\n"
         "epsilon_inc: %f\n",
         minpts, epsilon_start, epsilon_inc);

  int cap = 16;
  split_set *sets = (split_set *)malloc(sizeof(split_set) * (size_t)cap);
  int nsets = 0;
  split_set s0 = dbscanner(ds, epsilon_start, minpts, g);
  printf("Initial set obtained with %i clusters\n", s0.n_clusters);
  int tries = 0;
  while (s0.n_clusters == 1) {
    if (tries == 20 || epsilon_start == 0) {
      printf("Error did not find a sufficient starting position in 20 tries \n");
      fflush(stdout);
      _exit(1);
    }
    epsilon_start = epsilon_start / 2;
    printf("Trying new starting point \n");
    free_split_set_and_associated_clusters(s0);
    s0 = dbscanner(ds, epsilon_start, minpts, g);
    tries++;
  }
  printf("Sarting with %i clusters\n", s0.n_clusters);
  cluster nc = data_not_in_clusters(s0, ds);
  printf("Coverage at initial point: %f\n", (float)1.f - (float)nc.n_members / (float)ds.n_values);
  free(nc.members);
  if (s0.n_clusters == 1) {
    printf("All clusters fusioned in one step, decrease epsilon increment\n");
    fflush(stdout);
    _exit(1);
  }
  sets[nsets++] = s0;

  /* the scanning thread of dbscan.c:1061-1115, run inline */
  float eps_now = epsilon_start + epsilon_inc;
  for (;;) {
    float eps = eps_now;
    eps_now = eps_now + epsilon_inc;
    split_set s = dbscanner(ds, eps, minpts, g);
    if (nsets == cap) {
      cap *= 2;
      sets = (split_set *)realloc(sets, sizeof(split_set) * (size_t)cap);
    }
    sets[nsets++] = s;
    if (s.n_clusters == 1) break;
  }
  const int stop = nsets - 1;

  const double t_layers = now_s();
  split_set *kept = (split_set *)malloc(sizeof(split_set) * (size_t)nsets);
  cluster_connections **conn = (cluster_connections **)malloc(sizeof(*conn) * (size_t)nsets);
  unsigned char *is_kept = (unsigned char *)calloc((size_t)nsets, 1);
  int count = 0;
  kept[0] = sets[0];
  is_kept[0] = 1;
  for (int i = 1; i <= stop; i++) {
    split_set ns = sets[i];
    if (ns.n_clusters < kept[count].n_clusters) {
      cluster nc2 = data_not_in_clusters(ns, ds);
      printf("Layer %i has %i clusters\n", count + 1, ns.n_clusters);
      printf("  Coverage at layer %i: %f\n", count + 1,
             1.f - (float)nc2.n_members / (float)ds.n_values);
      free(nc2.members);
      printf("  Epsilon at layer %i: %f\n", count + 1, epsilon_start + (i + 1) * (float)epsilon_inc);
      conn[count] = generate_split_set_relation(kept[count], ns);
      count++;
      kept[count] = ns;
      is_kept[i] = 1;
    }
  }
  if (timing_on()) fprintf(stderr, "[mnhn] layers and relations: %.3f s\n", now_s() - t_layers);
  const double t_free = now_s();
  mnhn_gpu_destroy(g);
  if (timing_on()) fprintf(stderr, "[mnhn] device context and lists released: %.3f s\n", now_s() - t_free);
  const double t_out = now_s();

  printf("Connections found: \n");
  size_t plen = strlen(split_files_prefix);
  char *fname = (char *)malloc(plen + 16);
  for (int i = 0; i < count + 1; i++) {
    snprintf(fname, plen + 16, "%s%04d", split_files_prefix, i);
    store_split_set(fname, kept[i]);
  }
  free(fname);
  for (int i = count; i > 0; i--) {
    printf("Layer %i:\n", i);
    for (int j = 0; j < kept[i].n_clusters; j++) {
      printf("  Cluster %i connected to: \n  ", j);
      for (int k = 0; k < conn[i - 1][j].n_connections; k++) printf("%i ", conn[i - 1][j].connections[k]);
      printf("\n");
    }
  }
  for (int i = count; i > 0; i--) {
    for (int j = 0; j < kept[i].n_clusters; j++) free(conn[i - 1][j].connections);
    free(conn[i - 1]);
  }
  for (int i = 0; i < nsets; i++) free_split_set_and_associated_clusters(sets[i]);
  if (timing_on()) fprintf(stderr, "[mnhn] split sets and relations written: %.3f s\n", now_s() - t_out);
  free(is_kept);
  free(conn);
  free(kept);
  free(sets);
}
