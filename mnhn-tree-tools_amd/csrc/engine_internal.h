// engine_internal.h -- shared internals of the gfx950 engine library (not
// part of the C ABI): the context layout, device-memory helpers and the entry
// points the learned-proposer code reuses.
#pragma once
#include <hip/hip_runtime.h>

#include <chrono>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <algorithm>
#include <vector>

#include "mnhn_sw.h"

// MNHN_TIMING=2: wall time of each step of a list build on stderr.  Every
// mark waits for the stream, which perturbs the timing it reports, so timed
// runs leave the level at 0 or 1.
struct mnhn_phase_clock {
  bool on = false;
  hipStream_t st = nullptr;
  const char *who = "";
  std::chrono::steady_clock::time_point t;
  mnhn_phase_clock(hipStream_t s, const char *w) : st(s), who(w) {
    const char *v = getenv("MNHN_TIMING");
    on = v && atoi(v) >= 2;
    t = std::chrono::steady_clock::now();
  }
  void mark(const char *what) {
    if (!on) return;
    (void)hipStreamSynchronize(st);
    const auto now = std::chrono::steady_clock::now();
    fprintf(stderr, "[mnhn] %s %-30s %9.3f ms\n", who, what,
            std::chrono::duration<double, std::milli>(now - t).count());
    t = now;
  }
};

// Pinned staging buffers for large device-to-host copies (allocated on first
// use, 16 MB each; released by stage_free).
struct mnhn_detail_stage {
  void *h[2] = {nullptr, nullptr};
  hipEvent_t ev[2] = {nullptr, nullptr};
};

struct mnhn_sw_ctx {
  int n = 0;
  bool acgt = true;
  int match = 4, mismatch = -4, gap = 3;
  std::vector<int32_t> len, order, rank;
  std::vector<uint32_t> woff;
  int groups = 0;
  std::vector<int32_t> gtgt, gmax, gmin;
  // per target group, suffix sums of its slots' lengths: gsuf[g * 129 + t] =
  // sum of len over slots t..127 (empty slots count 0)
  std::vector<double> gsuf;
  int32_t maxlen = 0;
  int device = 0, num_cu = 0;
  int kernel = 0, strip = 0, waves_per_cu = 0;  // 0 = tuned default
  uint32_t *d_words = nullptr, *d_woff = nullptr;
  int32_t *d_len = nullptr, *d_rank = nullptr, *d_gtgt = nullptr, *d_gmax = nullptr,
          *d_gmin = nullptr;
  int *d_counter = nullptr;
  unsigned long long *d_nkeys = nullptr;
  uint32_t *d_bound = nullptr;
  size_t bound_bytes = 0;
  hipStream_t stream = nullptr;
  hipEvent_t ev[4] = {nullptr, nullptr, nullptr, nullptr};
  hipEvent_t pev[12] = {};  // around each kernel pass of launch_tiles (created on first use)
  mnhn_sw_timing last{};
  // device-resident epsilon lists of the last mnhn_sw_eps_count call
  unsigned long long *d_keys = nullptr;
  size_t keys_cap = 0;
  uint64_t data_hash = 0;  // scores, count, lengths and bytes of the sequences
  void *scratch = nullptr;  // this device's static scratch (mnhn_sw_one_vs_all_noalloc)
  mnhn_detail_stage stage;
  // pinned staging of host work lists for their copy to the device (grown on
  // demand, released with the context)
  void *h_items = nullptr;
  size_t h_items_bytes = 0;
};

namespace mnhn_detail {

#define HIP_TRY(expr)                                  \
  do {                                                 \
    hipError_t e_ = (expr);                            \
    if (e_ != hipSuccess) return MNHN_ERR_DEVICE;      \
  } while (0)

// Wave-uniform copy of a 64-bit value held by the first active lane.  The
// readfirstlane builtin works on 32-bit ints, so each half is reinterpreted
// as unsigned before it is widened (a signed low half would sign-extend).
__device__ __forceinline__ unsigned long long uniform_u64(unsigned long long v) {
  const unsigned lo = (unsigned)__builtin_amdgcn_readfirstlane((int)(unsigned)v);
  const unsigned hi = (unsigned)__builtin_amdgcn_readfirstlane((int)(unsigned)(v >> 32));
  return ((unsigned long long)hi << 32) | lo;
}

// Workgroups for a grid-stride loop over n elements: element launches stay far
// below 2^31 work-items whatever n is.
inline unsigned stride_blocks(long long n, int tb) {
  const long long b = (n + tb - 1) / tb;
  return (unsigned)std::min<long long>(std::max<long long>(b, 1), 1 << 16);
}

template <typename T>
int dev_alloc(T **p, size_t count) {
  *p = nullptr;
  if (count == 0) count = 1;
  if (hipMalloc((void **)p, count * sizeof(T)) != hipSuccess) {
    *p = nullptr;
    return MNHN_ERR_NOMEM;
  }
  return MNHN_OK;
}

template <typename T>
void dev_free(T *&p) {
  if (p) (void)hipFree(p);
  p = nullptr;
}

struct Work {
  std::vector<int4> items;
  std::vector<int32_t> qids;
  // qids is the context's length order (qids[k] = order[k], so rank[qids[k]]
  // = k and lengths fall along it): launch_tiles then takes an item's longest
  // query and its cells from prefix sums instead of walking its queries
  bool qids_order = false;
  std::vector<int32_t> gtgt, gmax, gmin;  // empty = use the context's length groups
  double cells = 0, pairs = 0;
  int max_q_len = 0;
  // device-resident work (used instead of the vectors when dev is set)
  bool dev = false;
  const int4 *d_items = nullptr;
  int n_items = 0;
  const int32_t *d_qids = nullptr, *d_gtgt = nullptr, *d_gmax = nullptr, *d_gmin = nullptr;
  int dev_kernel = 0;  // kernel variant for device work (KERNEL_LDS, KERNEL_LDS_F16 or KERNEL_REG)
  // KERNEL_LDS_F16 device work: every pair of every item satisfies
  // match * min(la, lb) <= 2048 (otherwise the int16 LDS kernel runs it)
  bool dev_f16_ok = false;
  int dev_strip = 0;  // strip width for device work (0: the context's)
  // target groups shared by several rows (epsilon mode, KERNEL_LDS or
  // KERNEL_LDS_F16; the item's qids range holds its rows): 3 = two rows, one
  // per half of the lanes' packed pairs (slots 0-63 and 64-127), 4 = four
  // rows, (half-wave, half) (slots 0-31, 64-95, 32-63, 96-127 in qids order)
  int dev_split = 1;
};


// Launch the all-pairs / pairs / one-vs-all kernels over a work list (host
// item list, or device-resident items and target groups when w.dev is set).
int launch_tiles(mnhn_sw_ctx *c, const Work &w, int mode, int triangle, int eps_max,
                 int32_t *d_dist, unsigned long long *d_keys, unsigned long long key_cap);

// Sort each row segment [d_rp[r], d_rp[r+1]) of distinct column values in
// [0, n_cols) ascending into out (same positions), moving the optional 32-bit
// payload along (pay_in / pay_out may be null), on stream st.
int sort_row_segments(hipStream_t st, const long long *d_rp, const int32_t *seg, int32_t *out,
                      const uint32_t *pay_in, uint32_t *pay_out, int n_rows, int n_cols);

// Host allocation for large result arrays, released with host_free: pinned
// (copy_d2h then copies into it directly) within a budget of pinned bytes
// outstanding, else 2 MB-aligned pages marked for transparent huge pages.
void *host_alloc(size_t bytes);
void host_free(void *p);

// Register-strip width the context's launches use (tuning, environment or default).
int context_strip(const mnhn_sw_ctx *c);

// Device-to-host copy on stream st: straight into a pinned host_alloc
// buffer, otherwise through two pinned staging buffers (chunk k+1 in flight
// while chunk k is copied to dst).
int copy_d2h(hipStream_t st, mnhn_detail_stage &s, void *dst, const void *src, size_t bytes);
inline int copy_d2h(mnhn_sw_ctx *c, void *dst, const void *src, size_t bytes) {
  return copy_d2h(c->stream, c->stage, dst, src, bytes);
}
void stage_free(mnhn_detail_stage &s);

// Group the ctx's epsilon keys (c->d_keys, nk of them) into CSR rows.
int csr_from_keys(mnhn_sw_ctx *c, unsigned long long nk, mnhn_csr *out);

// The epsilon bound an all-pairs or prefilter list pass runs with: eps_max,
// or 65535 when eps_max is larger and no distance can exceed 65535 (keys and
// lists carry d in 16 bits); MNHN_ERR_RANGE when eps_max is larger and some
// distance can exceed it.
int list_bound(const mnhn_sw_ctx *c, int32_t eps_max, int32_t *eff);

// Make sure c->d_keys holds at least cap keys.
int ensure_keys(mnhn_sw_ctx *c, unsigned long long cap);

// Epsilon-pass work over the unique pairs of the sequences at length-order
// positions p0 .. n-1 (the shortest ones; p0 = 0: all pairs), with the
// length-bound pruning of the all-pairs pass; items finalised for launch.
Work suffix_work(const mnhn_sw_ctx *c, int32_t eps, int p0);

// Keys an epsilon pass of work w will emit (both directions of every listed
// pair), estimated from a systematic 0.5 % sample of its items run into
// c->d_keys (overwritten), times 1.3; w's pair bound (2 keys per pair) when
// it has fewer than 4,096 items.
int eps_keys_estimate(mnhn_sw_ctx *c, const Work &w, int32_t eps, unsigned long long *est);

// KERNEL_LDS_MIXED: one launch for the exact fp16 form's and the int16 form's
// items (epsilon mode, host work lists), not a tuning choice of its own
enum { KERNEL_AUTO = 0, KERNEL_REG = 1, KERNEL_LDS = 2, KERNEL_LDS_F16 = 3, KERNEL_I32 = 4, KERNEL_LDS_MIXED = 5 };
enum { MODE_DIST = 0, MODE_EPS = 1 };
constexpr int kGroupSize = 128;

}  // namespace mnhn_detail
