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
 * or equal epsilon reuse them.  With MNHN_REGION_QUERY=launch each region
 * query instead launches the one-vs-all kernel, as the reference's GPU build
 * launches its kernel once per query.
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
};

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

static void ensure_lists(mnhn_gpu_stuff g, float eps) {
  mnhn_eps_cache *c = g.cache;
  const int32_t need = eps_floor(eps);
  if (c->have && c->csr.eps_max >= need) return;
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
        return;
      }
      mnhn_csr_free(&c->csr);
    }
  }
  const double t0 = now_s();
  int rc = mnhn_sw_eps_csr(g.sw, eps_max, &c->csr);
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
  const int16_t *dist; /* Smith-Waterman lists, or */
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

static pass_lists build_pass_lists(int n, const int64_t *rp, const int32_t *col, const int16_t *dist,
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

static split_set dbscan_core(int n, float epsilon, int minpts, region_fn rq, void *src) {
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

  for (int i = 0; i < n; i++) {
    if (visited[i]) continue;
    visited[i] = 1;
    const int *nbi;
    int cnt = rq(src, n, i, epsilon, nb, &nbi);
    if (!(minpts < cnt)) continue;
    /* expand_cluster */
    const int id = ret.n_clusters;
    const int mark = id + 1;
    cluster *cl = &ret.clusters[id];
    cl->id = id;
    cl->members = (int *)malloc(sizeof(int) * (size_t)n);
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
    for (int qi = 0; qi < qn; qi++) {
      const int q = queue[qi];
      if (!visited[q]) {
        visited[q] = 1;
        const int *nbq;
        int cq = rq(src, n, q, epsilon, nb, &nbq);
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
            } else {
              int k = 0;
              for (int t = 0; t < n_rest; t++)
                if (stamp[rest[t]] != mark) rest[k++] = rest[t];
              n_rest = k;
            }
            for (int t = 0; t < n_rest; t++) {
              const int x = rest[t];
              if (in_row(nbq, cq, x)) {
                stamp[x] = mark;
                queue[qn++] = x;
                in_union++;
              }
            }
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
    cl->members = (int *)realloc(cl->members, sizeof(int) * (size_t)(cl->n_members ? cl->n_members : 1));
    ret.n_clusters++;
  }
  free(visited);
  free(member);
  free(stamp);
  free(nb);
  free(queue);
  free(rest);
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

split_set dbscan_SW_GPU(dataset ds, float epsilon, int minpts, mnhn_gpu_stuff g) {
  const int n = ds.n_values;
  apply_defaults(&epsilon, &minpts);
  if (n > 0) {
    if (g.cache->per_query_launch) {
      if (!g.cache->dist_row) g.cache->dist_row = (int32_t *)malloc(sizeof(int32_t) * (size_t)n);
    } else {
      ensure_lists(g, epsilon);
    }
  }
  const double t0 = now_s();
  split_set s;
  if (n > 0 && !g.cache->per_query_launch) {
    const mnhn_csr *c = &g.cache->csr;
    pass_lists pl = build_pass_lists(n, c->row_ptr, c->col, c->dist, NULL, epsilon, (float)0 <= epsilon);
    s = dbscan_core(n, epsilon, minpts, rq_pass, &pl);
    free_pass_lists(&pl);
  } else {
    s = dbscan_core(n, epsilon, minpts, rq_gpu, &g);
  }
  if (timing_on())
    fprintf(stderr, "[mnhn] dbscan pass eps %g: %d clusters, %.3f s\n", epsilon, s.n_clusters, now_s() - t0);
  return s;
}

split_set mnhn_dbscan_csr(const mnhn_csr *csr, float epsilon, int minpts) {
  const int n = csr ? csr->n_rows : 0;
  if (n == 0) return dbscan_core(0, epsilon, minpts, rq_pass, NULL);
  pass_lists pl = build_pass_lists(n, csr->row_ptr, csr->col, csr->dist, NULL, epsilon, (float)0 <= epsilon);
  split_set s = dbscan_core(n, epsilon, minpts, rq_pass, &pl);
  free_pass_lists(&pl);
  return s;
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
  if (n == 0) return dbscan_core(0, epsilon, minpts, rq_pass, NULL);
  ensure_vec_lists(g, epsilon);
  /* the stored rows list p itself when its own accumulator qualifies */
  const mnhn_vcsr *c = &g.cache->vcsr;
  pass_lists pl = build_pass_lists(n, c->row_ptr, c->col, NULL, c->val, vec_thr(norm, epsilon), 0);
  split_set s = dbscan_core(n, epsilon, minpts, rq_pass, &pl);
  free_pass_lists(&pl);
  return s;
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
  if (n == 0) return dbscan_core(0, epsilon, minpts, rq_pass, NULL);
  pass_lists pl = build_pass_lists(n, csr->row_ptr, csr->col, NULL, csr->val, vec_thr(csr->norm, epsilon), 0);
  split_set s = dbscan_core(n, epsilon, minpts, rq_pass, &pl);
  free_pass_lists(&pl);
  return s;
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
  mnhn_gpu_destroy(g);
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
