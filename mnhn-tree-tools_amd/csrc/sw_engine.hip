// sw_engine.hip -- gfx950 Smith-Waterman distance engine behind include/mnhn_sw.h.
//
// Computes the reference's distance (smith-waterman.c:16-81)
//     d(A, B) = 4*max(la, lb) - max H(i, j),  i in [1, la], j in [1, lb]
// with match +4 / mismatch -4 / linear gap 3, index 0 never compared and the
// string terminators A[la], B[lb] matching only each other.
//
// Kernel design (inter-sequence parallelism, one pair per 16-bit lane half):
//  * A work item is (target group of 128 sequences, range of queries).  Each
//    wave64 lane holds two targets, one per 16-bit half of a 32-bit VGPR, and
//    the query is wave-uniform, so its residues come through scalar loads.
//  * Scores are kept biased by -32768 in signed 16-bit halves.  Saturating
//    v_pk_add_i16 then applies the local-alignment floor max(0, .) for free:
//        D = sat(Hdiag + s), E = sat(max(Hleft, Hup) - gap), H = max(D, E)
//    plus best = max(best, H): five packed VALU ops per two cells.
//  * The target dimension is cut into register strips of S columns.  For
//    each strip the lane builds a query profile P[c][jj] = s(c, target[jj])
//    for the four residue codes c; the wave-uniform query code selects which
//    P is used (scalar branch), so no per-cell lookup is needed.  The strip's
//    right-most H column is handed to the next strip through a per-wave
//    global workspace laid out [row][lane] (coalesced 256 B per row).
//  * The terminators need no DP row or column: every cell in the terminator
//    row/column is bounded by its neighbours minus 3 except the corner
//    (la, lb), whose value is H(la-1, lb-1) + 4.  So the DP runs over the
//    residues only (rows 1..la-1, columns 1..lb-1) and adds the corner term
//    in closed form.  Columns past a target's end are padding that mismatches
//    every query residue; such cells never exceed their neighbours, so they
//    cannot change the maximum.
//  * Persistent grid: waves pull work items from an atomic queue (items are
//    sorted by cost, longest first).
//  * epsilon mode emits (row, col, d) keys for d <= eps_max for both
//    directions of each unordered pair.  A counting pass, an exclusive scan
//    and a scatter group them by row; each row is then put in ascending
//    column order by ranking its columns (comparison ranks for short rows,
//    a column bitmap + popcount prefix in LDS for long ones).
#include <hip/hip_runtime.h>
#include <rocprim/device/device_radix_sort.hpp>
#include <sys/mman.h>
#include <unistd.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <string>
#include <chrono>
#include <cstdint>
#include <type_traits>
#include <cstdlib>
#include <cstring>
#include <map>
#include <mutex>
#include <thread>
#include <new>
#include <numeric>
#include <vector>

#include <sys/stat.h>

#include "mnhn_sw.h"
#include "engine_internal.h"

using namespace mnhn_detail;

typedef short s16x2 __attribute__((ext_vector_type(2)));
typedef unsigned short u16x2 __attribute__((ext_vector_type(2)));

namespace {

constexpr int kWave = 64;
constexpr int kGroup = 128;          // targets per work item (2 per lane)
constexpr int kBlock = 256;          // threads per workgroup (4 waves)
constexpr short kFloor = -32768;     // biased representation of H = 0
constexpr int kBias = 32768;
constexpr long long kI16ScoreMax = 65535;  // largest score the biased 16-bit halves hold
constexpr int kI32Strip = 16;              // register strip of the 32-bit kernel
constexpr size_t kWorkspaceBytes = 8ull << 30;  // strip-boundary workspaces of one launch


struct KArgs {
  const uint32_t *words;   // packed residues of all sequences
  const uint32_t *woff;    // first word of each sequence
  const int32_t *len;      // sequence lengths
  const int32_t *rank;     // position of each sequence in length order
  const int32_t *gtgt;     // [groups][128] target ids, -1 = empty slot
  const int32_t *gmax;     // [groups] longest target of the group
  const int32_t *gmin;     // [groups] shortest non-empty target of the group
  const int4 *items;       // (group, q_begin, q_end, out_base)
  const int32_t *qids;     // query ids referenced by item ranges
  int *counter;            // work-queue head
  uint32_t *bound;         // per-wave strip-boundary workspace
  int32_t *dist_out;       // MODE_DIST output
  unsigned long long *keys;      // MODE_EPS output
  unsigned long long *nkeys;     // MODE_EPS key counter
  unsigned long long key_cap;
  int n_items;
  int bound_stride;        // uint32 words per wave
  int list_row;            // LDS kernels: first workspace row of the row-address list
  int item_base;           // group-item kernels: item k is target group item_base + k
  int query;               // against this query
  int mode;
  int triangle;            // only pairs with rank[q] < rank[t]
  int eps_max;
  int match, mismatch, gap;
};

// Work item k: from the item list, or (kG: group items, the allocation-free
// one-vs-all) target group item_base + k against the single query.  kG is a
// template parameter so that the item-list kernels carry none of its code.
template <bool kG>
__device__ __forceinline__ int4 load_item(const KArgs &a, int k) {
  if constexpr (kG) return make_int4(a.item_base + k, 0, 1, 0);
  else return a.items[k];
}
template <bool kG>
__device__ __forceinline__ int load_qid(const KArgs &a, int qi) {
  if constexpr (kG) return a.query;
  else return a.qids[qi];
}

// A wave's workspace as rows of 64 words, [row][lane], behind a buffer
// resource: the row goes in the wave-uniform scalar offset and the lane's
// byte offset is fixed, so a row access costs no vector address arithmetic.
// (Word 3 of the resource is the raw-buffer format of the gfx9 family.)
// The LDS kernels make every workspace access through it, so the compiler
// keeps them in program order; a buffer store also keeps the next row's LDS
// profile reads behind it, so the rows of a strip run one after another
// (the registers of one row's profile suffice).
constexpr int kRawBufferWord3 = 0x00020000;
struct WaveRows {
  __amdgpu_buffer_rsrc_t r;
  int lane4;
  __device__ __forceinline__ WaveRows(uint32_t *base, int words, int lane)
      : r(__builtin_amdgcn_make_buffer_rsrc(base, (short)0, words * 4, kRawBufferWord3)), lane4(lane * 4) {}
  __device__ __forceinline__ uint32_t load(int row) const {
    return __builtin_amdgcn_raw_buffer_load_b32(r, lane4, row * (kWave * 4), 0);
  }
  __device__ __forceinline__ void store(int row, uint32_t v) const {
    __builtin_amdgcn_raw_buffer_store_b32(v, r, lane4, row * (kWave * 4), 0);
  }
  // a row that differs between lanes
  __device__ __forceinline__ uint32_t load_lane_row(int row) const {
    return __builtin_amdgcn_raw_buffer_load_b32(r, lane4 + row * (kWave * 4), 0, 0);
  }
};

typedef __attribute__((address_space(3))) uint32_t lds_u32;
typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));
typedef __attribute__((address_space(3))) const u32x4 lds_u4;

// Profile-table row of a lane's target code pair (lo, hi) (codes 0-3 = A, C,
// G, T; 4 = past the target's end).  The 16 pairs of residues come first
// (lo * 4 + hi): with the table's row stride of S + 4 words (an odd number of
// 16-byte bank groups mod 64 banks) the rows the lanes of one DP row read lie
// in distinct bank groups, so profile reads of different rows do not
// conflict; the 9 pairs with a past-the-end code follow: (4, hi) -> 16 + hi,
// (lo < 4, 4) -> 21 + lo.
__device__ __forceinline__ uint32_t tab_row(uint32_t lo, uint32_t hi) {
  return ((lo | hi) & 4u) ? (lo == 4u ? 16u + hi : 21u + lo) : lo * 4u + hi;
}
__device__ __forceinline__ void tab_codes(int p, uint32_t &lo, uint32_t &hi) {
  if (p < 16) {
    lo = (uint32_t)p >> 2;
    hi = (uint32_t)p & 3u;
  } else if (p <= 20) {
    lo = 4u;
    hi = (uint32_t)(p - 16);
  } else {
    lo = (uint32_t)(p - 21);
    hi = 4u;
  }
}

// LDS kernels: the DP rows of a work item's target group, decoded once per
// item instead of once per strip and query.  Word m of the row-address list
// (workspace row list_row + m) holds, for DP rows j = 2m + 1 (low half) and
// 2m + 2 (high half), the LDS byte address of the profile row that the
// lane's two target residues select, tab_row(c_lo, c_hi) * row_bytes past the
// wave's table, with code 4 for rows past a target's end.  Rows are the
// target residues 1..nrows; the words read are those of residues
// 0..16 * ((npairs - 1) / 8) + 15, inside the words the row loop of the
// per-strip decode read.
__device__ __forceinline__ void build_row_list(const KArgs &a, const WaveRows &ws, int nrows, uint32_t wo_lo,
                                               uint32_t wo_hi, int rows_lo, int rows_hi, uint32_t tab_lds,
                                               uint32_t row_bytes) {
  const int npairs = (nrows + 1) >> 1;
  for (int mb = 0; mb * 8 < npairs; ++mb) {
    const uint32_t w_lo = a.words[wo_lo + (uint32_t)mb], w_hi = a.words[wo_hi + (uint32_t)mb];
#pragma unroll
    for (int k = 0; k < 8; ++k) {
      const int m = mb * 8 + k;
      if (m >= npairs) break;
      const int j = 2 * m + 1;  // residue j - 1 = 16 * mb + 2k of both words
      uint32_t c0l = (w_lo >> (4 * k)) & 3u, c1l = (w_lo >> (4 * k + 2)) & 3u;
      uint32_t c0h = (w_hi >> (4 * k)) & 3u, c1h = (w_hi >> (4 * k + 2)) & 3u;
      c0l = j > rows_lo ? 4u : c0l;
      c1l = j + 1 > rows_lo ? 4u : c1l;
      c0h = j > rows_hi ? 4u : c0h;
      c1h = j + 1 > rows_hi ? 4u : c1h;
      const uint32_t a0 = tab_lds + tab_row(c0l, c0h) * row_bytes;
      const uint32_t a1 = tab_lds + tab_row(c1l, c1h) * row_bytes;
      ws.store(a.list_row + m, a0 | (a1 << 16));
    }
  }
}

__device__ __forceinline__ s16x2 as_s16x2(uint32_t v) {
  return __builtin_bit_cast(s16x2, v);
}
__device__ __forceinline__ uint32_t as_u32(s16x2 v) {
  return __builtin_bit_cast(uint32_t, v);
}

// One DP row over a register strip, query residue profile P.  The diagonal
// term of column jj+1 is formed from the old H[jj] before H[jj] is replaced,
// so every new H value can take its predecessor's register (no copies).
template <int S>
__device__ __forceinline__ void dp_row(s16x2 (&H)[S], const s16x2 (&P)[S],
                                       s16x2 left, s16x2 diag, s16x2 gapv,
                                       s16x2 &best) {
  s16x2 dnext = __builtin_elementwise_add_sat(diag, P[0]);
#pragma unroll
  for (int jj = 0; jj < S; ++jj) {
    const s16x2 D = dnext;
    if (jj + 1 < S) dnext = __builtin_elementwise_add_sat(H[jj], P[jj + 1]);
    const s16x2 M = __builtin_elementwise_max(left, H[jj]);
    const s16x2 E = __builtin_elementwise_add_sat(M, gapv);
    const s16x2 h = __builtin_elementwise_max(D, E);
    best = __builtin_elementwise_max(best, h);
    H[jj] = h;
    left = h;
  }
}

// dp_row with the running maximum split over two accumulators by column
// parity: the compiler sinks the maxima to the row's end, and one chain of
// dependent packed ops there needs a wait state between each pair on gfx950
template <int S>
__device__ __forceinline__ void dp_row2(s16x2 (&H)[S], const s16x2 (&P)[S],
                                        s16x2 left, s16x2 diag, s16x2 gapv,
                                        s16x2 &best, s16x2 &best2) {
  s16x2 dnext = __builtin_elementwise_add_sat(diag, P[0]);
#pragma unroll
  for (int jj = 0; jj < S; ++jj) {
    const s16x2 D = dnext;
    if (jj + 1 < S) dnext = __builtin_elementwise_add_sat(H[jj], P[jj + 1]);
    const s16x2 M = __builtin_elementwise_max(left, H[jj]);
    const s16x2 E = __builtin_elementwise_add_sat(M, gapv);
    const s16x2 h = __builtin_elementwise_max(D, E);
    if (jj & 1) best2 = __builtin_elementwise_max(best2, h);
    else best = __builtin_elementwise_max(best, h);
    H[jj] = h;
    left = h;
  }
}

// Generic alphabet: substitution score computed from 16-bit codes per cell.
template <int S>
__device__ __forceinline__ s16x2 code_score(uint32_t tc, uint32_t qq, s16x2 pm, s16x2 dmm) {
  const u16x2 x = __builtin_bit_cast(u16x2, tc ^ qq);
  const u16x2 m01 = __builtin_elementwise_min(x, (u16x2){1, 1});
  return pm + __builtin_bit_cast(s16x2, m01) * dmm;
}

template <int S>
__device__ __forceinline__ void dp_row_codes(s16x2 (&H)[S], const uint32_t (&TC)[S],
                                             uint32_t qq, s16x2 pm, s16x2 dmm,
                                             s16x2 left, s16x2 diag, s16x2 gapv,
                                             s16x2 &best) {
  s16x2 dnext = __builtin_elementwise_add_sat(diag, code_score<S>(TC[0], qq, pm, dmm));
#pragma unroll
  for (int jj = 0; jj < S; ++jj) {
    const s16x2 D = dnext;
    if (jj + 1 < S)
      dnext = __builtin_elementwise_add_sat(H[jj], code_score<S>(TC[jj + 1], qq, pm, dmm));
    const s16x2 M = __builtin_elementwise_max(left, H[jj]);
    const s16x2 E = __builtin_elementwise_add_sat(M, gapv);
    const s16x2 h = __builtin_elementwise_max(D, E);
    best = __builtin_elementwise_max(best, h);
    H[jj] = h;
    left = h;
  }
}

// Distances of the lane's two targets against query q and their output:
// MODE_DIST writes them, MODE_EPS emits keys for d <= eps_max.
// The high halves may run another query (q_hi, of length la_hi).
template <bool kG>
__device__ __forceinline__ void finish_query_int(const KArgs &a, int lane, int q, int qrel, int ob,
                                                 int la, int t_lo, int t_hi, int lb_lo, int lb_hi,
                                                 int rk_lo, int rk_hi, int best_lo, int best_hi,
                                                 int cor_lo, int cor_hi, int q_hi, int la_hi) {
  int tot_lo = 0, tot_hi = 0;
  if (la >= 1 && lb_lo >= 1) tot_lo = max(best_lo, cor_lo + a.match);
  if (la_hi >= 1 && lb_hi >= 1) tot_hi = max(best_hi, cor_hi + a.match);
  const int d_lo = a.match * max(la, lb_lo) - tot_lo;
  const int d_hi = a.match * max(la_hi, lb_hi) - tot_hi;

  if constexpr (kG) {  // group items: distances by target id
    if (t_lo >= 0) a.dist_out[t_lo] = d_lo;
    if (t_hi >= 0) a.dist_out[t_hi] = d_hi;
    return;
  }
  const int rq = a.triangle ? a.rank[q] : 0;
  const int rq_hi = a.triangle ? a.rank[q_hi] : 0;
  const bool ok_lo = t_lo >= 0 && (!a.triangle || rq < rk_lo);
  const bool ok_hi = t_hi >= 0 && (!a.triangle || rq_hi < rk_hi);
  if (a.mode == MODE_DIST) {
    int32_t *o = a.dist_out + (size_t)ob + (size_t)qrel * kGroup;
    if (ok_lo) o[lane] = d_lo;
    if (ok_hi) o[kWave + lane] = d_hi;
    return;
  }
  const bool h_lo = ok_lo && d_lo <= a.eps_max;
  const bool h_hi = ok_hi && d_hi <= a.eps_max;
  const unsigned long long b_lo = __ballot(h_lo);
  const unsigned long long b_hi = __ballot(h_hi);
  const int total = 2 * (__popcll(b_lo) + __popcll(b_hi));
  if (total == 0) return;
  unsigned long long base = 0;
  if (lane == 0) base = atomicAdd(a.nkeys, (unsigned long long)total);
  base = uniform_u64(base);
  const unsigned mlo = __builtin_amdgcn_mbcnt_hi((unsigned)(b_lo >> 32),
                        __builtin_amdgcn_mbcnt_lo((unsigned)b_lo, 0u));
  const unsigned mhi = __builtin_amdgcn_mbcnt_hi((unsigned)(b_hi >> 32),
                        __builtin_amdgcn_mbcnt_lo((unsigned)b_hi, 0u));
  unsigned long long pos = base + 2ull * (mlo + mhi);
  const unsigned long long uq = (unsigned long long)q;
  if (h_lo) {
    const unsigned long long ut = (unsigned long long)t_lo, ud = (unsigned)d_lo;
    if (pos < a.key_cap) a.keys[pos] = (uq << 40) | (ut << 16) | ud;
    if (pos + 1 < a.key_cap) a.keys[pos + 1] = (ut << 40) | (uq << 16) | ud;
    pos += 2;
  }
  if (h_hi) {
    const unsigned long long ut = (unsigned long long)t_hi, ud = (unsigned)d_hi;
    const unsigned long long uq = (unsigned long long)q_hi;
    if (pos < a.key_cap) a.keys[pos] = (uq << 40) | (ut << 16) | ud;
    if (pos + 1 < a.key_cap) a.keys[pos + 1] = (ut << 40) | (uq << 16) | ud;
  }
}

// int16 kernels: best / corner are biased by -32768
template <bool kG>
__device__ __forceinline__ void finish_query(const KArgs &a, int lane, int q, int qrel, int ob,
                                             int la, int t_lo, int t_hi, int lb_lo, int lb_hi,
                                             int rk_lo, int rk_hi, s16x2 best, s16x2 corner,
                                             int q_hi, int la_hi) {
  finish_query_int<kG>(a, lane, q, qrel, ob, la, t_lo, t_hi, lb_lo, lb_hi, rk_lo, rk_hi,
                   (int)best.x + kBias, (int)best.y + kBias, (int)corner.x + kBias,
                   (int)corner.y + kBias, q_hi, la_hi);
}

template <int S, bool kAcgt, bool kG = false>
__global__ __launch_bounds__(kBlock) void sw_tiles_kernel(KArgs a) {
  const int lane = threadIdx.x & (kWave - 1);
  const int wave = __builtin_amdgcn_readfirstlane(blockIdx.x * (kBlock / kWave) + (threadIdx.x >> 6));
  // the packed residues are read-only for the whole launch: reading them
  // through the constant address space lets uniform loads use s_load
  const __attribute__((address_space(4))) uint32_t *cwords =
      (const __attribute__((address_space(4))) uint32_t *)a.words;
  uint32_t *bnd = a.bound + (size_t)wave * a.bound_stride;

  const s16x2 floor2 = {kFloor, kFloor};
  const s16x2 gapv = {(short)-a.gap, (short)-a.gap};
  const short smat = (short)a.match, smis = (short)a.mismatch;
  const s16x2 pm = {smat, smat};
  const s16x2 dmm = {(short)(smis - smat), (short)(smis - smat)};

  for (;;) {
    int item_idx = 0;
    if (lane == 0) item_idx = atomicAdd(a.counter, 1);
    item_idx = __builtin_amdgcn_readfirstlane(item_idx);
    if (item_idx >= a.n_items) break;
    const int4 it = load_item<kG>(a, item_idx);
    const int g = __builtin_amdgcn_readfirstlane(it.x);
    const int qb = __builtin_amdgcn_readfirstlane(it.y);
    const int qe = __builtin_amdgcn_readfirstlane(it.z);
    const int ob = __builtin_amdgcn_readfirstlane(it.w);

    const int t_lo = a.gtgt[(size_t)g * kGroup + lane];
    const int t_hi = a.gtgt[(size_t)g * kGroup + kWave + lane];
    const int lb_lo = t_lo >= 0 ? a.len[t_lo] : 0;
    const int lb_hi = t_hi >= 0 ? a.len[t_hi] : 0;
    const uint32_t wo_lo = t_lo >= 0 ? a.woff[t_lo] : 0u;
    const uint32_t wo_hi = t_hi >= 0 ? a.woff[t_hi] : 0u;
    const int rk_lo = t_lo >= 0 ? a.rank[t_lo] : 0;
    const int rk_hi = t_hi >= 0 ? a.rank[t_hi] : 0;
    const int gm = __builtin_amdgcn_readfirstlane(a.gmax[g]);
    // inner columns are 1..lb-1
    const int nstrips = gm > 1 ? (gm - 1 + S - 1) / S : 0;

    for (int qi = qb; qi < qe; ++qi) {
      const int q = __builtin_amdgcn_readfirstlane(load_qid<kG>(a, qi));
      const int la = __builtin_amdgcn_readfirstlane(a.len[q]);
      const uint32_t qwo = __builtin_amdgcn_readfirstlane(a.woff[q]);
      const int nrows = la > 1 ? la - 1 : 0;
      s16x2 best = floor2;
      s16x2 corner = floor2;

      for (int s = 0; s < nstrips && nrows > 0; ++s) {
        const int j0 = 1 + s * S;  // DP column of strip slot 0
        s16x2 H[S];
#pragma unroll
        for (int jj = 0; jj < S; ++jj) H[jj] = floor2;

        // ---- target residues of this strip (residue index j0-1 .. j0+S-2)
        constexpr int kPerWord = kAcgt ? 16 : 2;
        constexpr int kWords = S / kPerWord;
        uint32_t wl[kWords], wh[kWords];
        const uint32_t wbase = (uint32_t)(s * S) / kPerWord;
#pragma unroll
        for (int w = 0; w < kWords; ++w) {
          wl[w] = a.words[wo_lo + wbase + w];
          wh[w] = a.words[wo_hi + wbase + w];
        }
        // strip-local index of the last inner column of each target
        const int cl_lo = lb_lo - 1 - j0;  // corner column slot, may be out of range
        const int cl_hi = lb_hi - 1 - j0;

        if constexpr (kAcgt) {
          s16x2 P0[S], P1[S], P2[S], P3[S];
#pragma unroll
          for (int jj = 0; jj < S; ++jj) {
            const int sh = 2 * (jj % 16);
            // residue valid while column j0+jj <= lb-1, i.e. jj <= cl
            uint32_t c_lo = jj <= cl_lo ? (wl[jj / 16] >> sh) & 3u : 4u;
            uint32_t c_hi = jj <= cl_hi ? (wh[jj / 16] >> sh) & 3u : 4u;
            P0[jj] = (s16x2){c_lo == 0u ? smat : smis, c_hi == 0u ? smat : smis};
            P1[jj] = (s16x2){c_lo == 1u ? smat : smis, c_hi == 1u ? smat : smis};
            P2[jj] = (s16x2){c_lo == 2u ? smat : smis, c_hi == 2u ? smat : smis};
            P3[jj] = (s16x2){c_lo == 3u ? smat : smis, c_hi == 3u ? smat : smis};
          }
          s16x2 diag_b = floor2;
          uint32_t next_b = as_u32(floor2);
          if (s > 0) next_b = bnd[kWave + (uint32_t)lane];
          // query residues: 16 per word, fetched one word ahead
          uint32_t qword = cwords[qwo];
          uint32_t qnext = cwords[qwo + 1];
          for (int i = 1; i <= nrows; ++i) {
            const s16x2 left_b = as_s16x2(next_b);
            if (s > 0 && i < nrows) next_b = bnd[(uint32_t)(i + 1) * kWave + (uint32_t)lane];
            const int r = i - 1;  // residue index of the query row
            if ((r & 15) == 0 && r > 0) {
              qword = qnext;
              qnext = cwords[qwo + (uint32_t)(r >> 4) + 1];
            }
            const uint32_t code = (qword >> (2 * (r & 15))) & 3u;
            // each case ends in a distinct asm marker so the four row bodies
            // are not merged into one body fed by register copies
            switch (code) {
              case 0: dp_row<S>(H, P0, left_b, diag_b, gapv, best); asm volatile("; row c0"); break;
              case 1: dp_row<S>(H, P1, left_b, diag_b, gapv, best); asm volatile("; row c1"); break;
              case 2: dp_row<S>(H, P2, left_b, diag_b, gapv, best); asm volatile("; row c2"); break;
              default: dp_row<S>(H, P3, left_b, diag_b, gapv, best); asm volatile("; row c3"); break;
            }
            if (s + 1 < nstrips) bnd[(uint32_t)i * kWave + (uint32_t)lane] = as_u32(H[S - 1]);
            diag_b = left_b;
          }
        } else {
          uint32_t TC[S];
#pragma unroll
          for (int jj = 0; jj < S; ++jj) {
            const int sh = 16 * (jj & 1);
            uint32_t c_lo = jj <= cl_lo ? (wl[jj / 2] >> sh) & 0xffffu : 0xffffu;
            uint32_t c_hi = jj <= cl_hi ? (wh[jj / 2] >> sh) & 0xffffu : 0xffffu;
            TC[jj] = c_lo | (c_hi << 16);
          }
          s16x2 diag_b = floor2;
          uint32_t next_b = as_u32(floor2);
          if (s > 0) next_b = bnd[kWave + (uint32_t)lane];
          uint32_t qword = cwords[qwo];
          uint32_t qnext = cwords[qwo + 1];
          for (int i = 1; i <= nrows; ++i) {
            const s16x2 left_b = as_s16x2(next_b);
            if (s > 0 && i < nrows) next_b = bnd[(uint32_t)(i + 1) * kWave + (uint32_t)lane];
            const int r = i - 1;
            if ((r & 1) == 0 && r > 0) {
              qword = qnext;
              qnext = cwords[qwo + (uint32_t)(r >> 1) + 1];
            }
            const uint32_t code = (qword >> (16 * (r & 1))) & 0xffffu;
            const uint32_t qq = code | (code << 16);
            dp_row_codes<S>(H, TC, qq, pm, dmm, left_b, diag_b, gapv, best);
            if (s + 1 < nstrips) bnd[(uint32_t)i * kWave + (uint32_t)lane] = as_u32(H[S - 1]);
            diag_b = left_b;
          }
        }
        // corner H(la-1, lb-1) of each half, if its column is in this strip
        uint32_t cu = as_u32(corner);
#pragma unroll
        for (int jj = 0; jj < S; ++jj) {
          const uint32_t hv = as_u32(H[jj]);
          if (jj == cl_lo) cu = (cu & 0xffff0000u) | (hv & 0x0000ffffu);
          if (jj == cl_hi) cu = (cu & 0x0000ffffu) | (hv & 0xffff0000u);
        }
        corner = as_s16x2(cu);
      }

      finish_query<kG>(a, lane, q, qi - qb, ob, la, t_lo, t_hi, lb_lo, lb_hi, rk_lo, rk_hi,
                   best, corner, q, la);
    }
  }
}

// 32-bit form, for pairs whose local score can leave the 16-bit lane range
// (match * min(la, lb) > 65535: both sequences longer than 16,383 residues at
// the default scores).  The reference keeps its whole matrix in int32
// (smith-waterman.c:16-81); here each lane takes its low-half target and then
// its high-half target, one 32-bit DP at a time.  The query is the
// wave-uniform DP row, a register strip holds S target columns, and the
// strip's right-most H column is handed to the next strip through the
// per-wave workspace [row][lane].  With at most four residue codes the lane
// keeps the four substitution rows of its strip and the query code selects
// one per DP row (uniform branch); otherwise each cell compares 16-bit codes.
template <int S, bool kAcgt>
__device__ __forceinline__ void dp_row_i32(int (&H)[S], const int (&P)[S], int left, int diag, int gap,
                                           int &best) {
  int dnext = diag + P[0];
#pragma unroll
  for (int jj = 0; jj < S; ++jj) {
    const int D = dnext;
    if (jj + 1 < S) dnext = H[jj] + P[jj + 1];
    const int E = max(left, H[jj]) - gap;
    const int h = max(max(D, E), 0);
    best = max(best, h);
    H[jj] = h;
    left = h;
  }
}

template <int S, bool kAcgt, bool kG = false>
__global__ __launch_bounds__(kBlock) void sw_i32_kernel(KArgs a) {
  const int lane = threadIdx.x & (kWave - 1);
  const int wave = __builtin_amdgcn_readfirstlane(blockIdx.x * (kBlock / kWave) + (threadIdx.x >> 6));
  const __attribute__((address_space(4))) uint32_t *cwords =
      (const __attribute__((address_space(4))) uint32_t *)a.words;
  uint32_t *bnd = a.bound + (size_t)wave * a.bound_stride;
  constexpr int kPerWord = kAcgt ? 16 : 2;
  constexpr int kWords = S / kPerWord;

  for (;;) {
    int item_idx = 0;
    if (lane == 0) item_idx = atomicAdd(a.counter, 1);
    item_idx = __builtin_amdgcn_readfirstlane(item_idx);
    if (item_idx >= a.n_items) break;
    const int4 it = load_item<kG>(a, item_idx);
    const int g = __builtin_amdgcn_readfirstlane(it.x);
    const int qb = __builtin_amdgcn_readfirstlane(it.y);
    const int qe = __builtin_amdgcn_readfirstlane(it.z);
    const int ob = __builtin_amdgcn_readfirstlane(it.w);

    const int t_lo = a.gtgt[(size_t)g * kGroup + lane];
    const int t_hi = a.gtgt[(size_t)g * kGroup + kWave + lane];
    const int lb_lo = t_lo >= 0 ? a.len[t_lo] : 0;
    const int lb_hi = t_hi >= 0 ? a.len[t_hi] : 0;
    const uint32_t wo_lo = t_lo >= 0 ? a.woff[t_lo] : 0u;
    const uint32_t wo_hi = t_hi >= 0 ? a.woff[t_hi] : 0u;
    const int rk_lo = t_lo >= 0 ? a.rank[t_lo] : 0;
    const int rk_hi = t_hi >= 0 ? a.rank[t_hi] : 0;
    // strips of each half: its longest target's inner columns 1..lb-1 (a half
    // without targets runs none)
    int mlo = lb_lo, mhi = lb_hi;
    for (int off = kWave / 2; off > 0; off >>= 1) {
      mlo = max(mlo, __shfl_xor(mlo, off));
      mhi = max(mhi, __shfl_xor(mhi, off));
    }
    const int ns_lo = __builtin_amdgcn_readfirstlane(mlo > 1 ? (mlo - 1 + S - 1) / S : 0);
    const int ns_hi = __builtin_amdgcn_readfirstlane(mhi > 1 ? (mhi - 1 + S - 1) / S : 0);

    for (int qi = qb; qi < qe; ++qi) {
      const int q = __builtin_amdgcn_readfirstlane(load_qid<kG>(a, qi));
      const int la = __builtin_amdgcn_readfirstlane(a.len[q]);
      const uint32_t qwo = __builtin_amdgcn_readfirstlane(a.woff[q]);
      const int nrows = la > 1 ? la - 1 : 0;
      int best2[2] = {0, 0}, corner2[2] = {0, 0};
#pragma unroll 1
      for (int half = 0; half < 2; ++half) {
        const int lb = half ? lb_hi : lb_lo;
        const uint32_t wo = half ? wo_hi : wo_lo;
        const int nstrips = half ? ns_hi : ns_lo;
        int best = 0, corner = 0;
        for (int s = 0; s < nstrips && nrows > 0; ++s) {
          const int j0 = 1 + s * S;     // DP column of strip slot 0
          const int cl = lb - 1 - j0;   // strip slot of the target's last inner column
          uint32_t wd[kWords];
          const uint32_t wbase = (uint32_t)(s * S) / kPerWord;
#pragma unroll
          for (int w = 0; w < kWords; ++w) wd[w] = a.words[wo + wbase + w];
          int H[S];
#pragma unroll
          for (int jj = 0; jj < S; ++jj) H[jj] = 0;
          int diag_b = 0;
          int next_b = 0;
          if (s > 0) next_b = (int)bnd[kWave + (uint32_t)lane];
          if constexpr (kAcgt) {
            // the lane's four substitution rows; columns past the target's
            // end mismatch every query residue
            int P0[S], P1[S], P2[S], P3[S];
#pragma unroll
            for (int jj = 0; jj < S; ++jj) {
              const uint32_t c = jj <= cl ? (wd[jj / 16] >> (2 * (jj % 16))) & 3u : 4u;
              P0[jj] = c == 0u ? a.match : a.mismatch;
              P1[jj] = c == 1u ? a.match : a.mismatch;
              P2[jj] = c == 2u ? a.match : a.mismatch;
              P3[jj] = c == 3u ? a.match : a.mismatch;
            }
            uint32_t qword = cwords[qwo];
            uint32_t qnext = cwords[qwo + 1];
            for (int i = 1; i <= nrows; ++i) {
              const int left_b = next_b;
              if (s > 0 && i < nrows) next_b = (int)bnd[(uint32_t)(i + 1) * kWave + (uint32_t)lane];
              const int r = i - 1;
              if ((r & 15) == 0 && r > 0) {
                qword = qnext;
                qnext = cwords[qwo + (uint32_t)(r >> 4) + 1];
              }
              const uint32_t code = (qword >> (2 * (r & 15))) & 3u;
              switch (code) {
                case 0: dp_row_i32<S, true>(H, P0, left_b, diag_b, a.gap, best); asm volatile("; i32 row c0"); break;
                case 1: dp_row_i32<S, true>(H, P1, left_b, diag_b, a.gap, best); asm volatile("; i32 row c1"); break;
                case 2: dp_row_i32<S, true>(H, P2, left_b, diag_b, a.gap, best); asm volatile("; i32 row c2"); break;
                default: dp_row_i32<S, true>(H, P3, left_b, diag_b, a.gap, best); asm volatile("; i32 row c3"); break;
              }
              if (s + 1 < nstrips) bnd[(uint32_t)i * kWave + (uint32_t)lane] = (uint32_t)H[S - 1];
              diag_b = left_b;
            }
          } else {
            uint32_t TC[S];
#pragma unroll
            for (int jj = 0; jj < S; ++jj)
              TC[jj] = jj <= cl ? (wd[jj / 2] >> (16 * (jj & 1))) & 0xffffu : 0xffffffffu;
            uint32_t qword = cwords[qwo];
            uint32_t qnext = cwords[qwo + 1];
            for (int i = 1; i <= nrows; ++i) {
              const int left_b = next_b;
              if (s > 0 && i < nrows) next_b = (int)bnd[(uint32_t)(i + 1) * kWave + (uint32_t)lane];
              const int r = i - 1;
              if ((r & 1) == 0 && r > 0) {
                qword = qnext;
                qnext = cwords[qwo + (uint32_t)(r >> 1) + 1];
              }
              const uint32_t code = (qword >> (16 * (r & 1))) & 0xffffu;
              int P[S];
#pragma unroll
              for (int jj = 0; jj < S; ++jj) P[jj] = TC[jj] == code ? a.match : a.mismatch;
              dp_row_i32<S, false>(H, P, left_b, diag_b, a.gap, best);
              if (s + 1 < nstrips) bnd[(uint32_t)i * kWave + (uint32_t)lane] = (uint32_t)H[S - 1];
              diag_b = left_b;
            }
          }
          // corner H(la-1, lb-1): the last DP row at the target's last column
#pragma unroll
          for (int jj = 0; jj < S; ++jj)
            if (jj == cl) corner = H[jj];
        }
        best2[half] = best;
        corner2[half] = corner;
      }
      finish_query_int<kG>(a, lane, q, qi - qb, ob, la, t_lo, t_hi, lb_lo, lb_hi, rk_lo, rk_hi, best2[0], best2[1],
                       corner2[0], corner2[1], q, la);
    }
  }
}

// Query-profile-in-LDS kernel (residues in at most four codes).  The
// wave-uniform query is the register strip (S columns, right-aligned so its
// last column is always strip slot S-1), and each lane streams its two targets
// as DP rows.  Per strip the wave writes a table of packed substitution scores
// into its own LDS slice: one row per (code_lo, code_hi) pair of the lane's two
// target residues (codes A, C, G, T and 4 = past the target's end; row
// tab_row(code_lo, code_hi)), S entries per row.  A DP row then reads its S scores with S/4 ds_read_b128 -- no
// per-row selection, no register copies.  Columns left of the query's first
// residue ("pre-padding" of the first strip) hold -32768, which keeps their
// cells at the floor.
// Register budgets: 3 waves per SIMD at strip 64 (168 VGPRs), 5 at strip 32
// (96).  Measured on MI355X, the extra resident waves hide the wait states
// between one row's dependent packed ops better than interleaving two rows
// in one wave (which needs registers for both rows' profiles).
//
// kQ / kLH: several queries per wave, as in sw_lds_f16_kernel (below).
// One work item of sw_lds_kernel: target group it.x against the queries
// qids[it.y .. it.z) (it.w: output base), with the wave's profile table(s)
// at tab and its workspace ws.
template <int S, bool kG, int kQ, bool kLH>
__device__ __forceinline__ void sw_lds_item(const KArgs &a, uint32_t *tab, const WaveRows &ws, int lane,
                                        const int4 it) {
  static_assert((kQ == 1 && !kLH) || !kG, "several queries per wave: item lists only");
  static_assert(kQ == 1 || (kQ == 2 && kLH), "two tables per wave: four queries");
  constexpr int kPer = kLH ? 2 : 1;  // queries per table
  constexpr int NQ = kQ * kPer;      // queries per wave
  // queries of the second half-wave's table (kQ = 2): low halves, high halves
  constexpr int kSegLo = kQ == 2 ? kPer : 0, kSegHi = kQ == 2 ? NQ - 1 : kPer - 1;
  constexpr int kPairs = 25;
  constexpr int RS = S + 4;  // table row stride in words; the +4 shifts banks per row
  constexpr int kTab = kPairs * RS;  // words of one query's table
  const int seg = kQ == 2 ? lane >> 5 : 0;  // the lane's table
  const uint32_t tab_lds = (uint32_t)(uintptr_t)(lds_u32 *)tab + (uint32_t)seg * (kTab * 4u);

  const s16x2 floor2 = {kFloor, kFloor};
  const s16x2 gapv = {(short)-a.gap, (short)-a.gap};
  const uint32_t smat = (uint32_t)(uint16_t)(short)a.match;
  const uint32_t smis = (uint32_t)(uint16_t)(short)a.mismatch;

  const int g = __builtin_amdgcn_readfirstlane(it.x);
  const int qb = __builtin_amdgcn_readfirstlane(it.y);
  const int qe = __builtin_amdgcn_readfirstlane(it.z);
  const int ob = __builtin_amdgcn_readfirstlane(it.w);

  const int gm = __builtin_amdgcn_readfirstlane(a.gmax[g]);
  const int t_lo = a.gtgt[(size_t)g * kGroup + lane];
  const int t_hi = a.gtgt[(size_t)g * kGroup + kWave + lane];
  const int lb_lo = t_lo >= 0 ? a.len[t_lo] : 0;
  const int lb_hi = t_hi >= 0 ? a.len[t_hi] : 0;
  // empty slots are treated as full length (their results are discarded)
  const int rows_lo = t_lo >= 0 ? lb_lo - 1 : gm;
  const int rows_hi = t_hi >= 0 ? lb_hi - 1 : gm;
  const uint32_t wo_lo = t_lo >= 0 ? a.woff[t_lo] : 0u;
  const uint32_t wo_hi = t_hi >= 0 ? a.woff[t_hi] : 0u;
  const int rk_lo = t_lo >= 0 ? a.rank[t_lo] : 0;
  const int rk_hi = t_hi >= 0 ? a.rank[t_hi] : 0;
  const int nrows = gm > 1 ? gm - 1 : 0;     // DP rows = inner residues of the longest target
  build_row_list(a, ws, nrows, wo_lo, wo_hi, rows_lo, rows_hi, tab_lds, RS * 4u);

  for (int qi = qb; qi < qe; qi += NQ) {
    // the wave's queries: qv[u], its length, first word and left pre-padding
    int qv[NQ], lav[NQ], shv[NQ];
    uint32_t qwov[NQ];
    int ncols = 0;
#pragma unroll
    for (int u = 0; u < NQ; ++u) {
      qv[u] = __builtin_amdgcn_readfirstlane(u == 0 ? load_qid<kG>(a, qi) : a.qids[qi + u]);
      lav[u] = __builtin_amdgcn_readfirstlane(a.len[qv[u]]);
      qwov[u] = __builtin_amdgcn_readfirstlane(a.woff[qv[u]]);
      ncols = max(ncols, lav[u] > 1 ? lav[u] - 1 : 0);
    }
#pragma unroll
    for (int u = 0; u < NQ; ++u) shv[u] = ncols - (lav[u] > 1 ? lav[u] - 1 : 0);
    // the lane's queries: table seg's, low halves q, high halves q_hi
    const int q = (kQ == 2 && seg) ? qv[kSegLo] : qv[0];
    const int q_hi = (kQ == 2 && seg) ? qv[kSegHi] : qv[kPer - 1];
    const int la = (kQ == 2 && seg) ? lav[kSegLo] : lav[0];
    const int la_hi = (kQ == 2 && seg) ? lav[kSegHi] : lav[kPer - 1];
    const int nstrips = (ncols + S - 1) / S;
    s16x2 best = floor2, best2 = floor2;
    s16x2 corner = floor2;

    // strip s of the query; first (compile-time): s == 0, no strip to its left
    auto strip = [&](int s, auto firstc) {
      constexpr bool first = decltype(firstc)::value;
      const int j0 = ncols - (nstrips - s) * S + 1;  // query column of slot 0
      // The first list word and left boundaries are loaded before the
      // table build, whose own loads are waited for, so nothing is pending
      // when the row loop starts: the loop's waits then count only the
      // loads and stores of one iteration.
      uint32_t ra = ws.load(a.list_row);
      uint32_t bA = as_u32(floor2), bB = as_u32(floor2);
      if (!first) {
        bA = ws.load(1);
        if (nrows >= 2) bB = ws.load(2);
      }
      // ---- score table of this strip
      __builtin_amdgcn_wave_barrier();
      for (int e = lane; e < kQ * kPairs * S; e += kWave) {
        const int sg = kQ == 2 && e >= kPairs * S ? 1 : 0;
        const int e1 = e - sg * (kPairs * S);
        const int p = e1 / S, jj = e1 - p * S;
        uint32_t clo, chi;
        tab_codes(p, clo, chi);
        uint32_t v = 0x80008000u;
        if constexpr (!kLH) {  // one query: one code for both halves
          const int k = j0 + jj - shv[0];
          if (k >= 1) {
            const uint32_t r = (uint32_t)(k - 1);
            const uint32_t qc = (a.words[qwov[0] + (r >> 4)] >> (2u * (r & 15u))) & 3u;
            v = (clo == qc ? smat : smis) | ((chi == qc ? smat : smis) << 16);
          }
        } else {
          // table sg: low halves score query sg * kPer, high halves sg * kPer + 1
          const int kl = j0 + jj - (kQ == 2 && sg ? shv[kSegLo] : shv[0]);
          const int kh = j0 + jj - (kQ == 2 && sg ? shv[kSegHi] : shv[kPer - 1]);
          if (kl >= 1) {
            const uint32_t r = (uint32_t)(kl - 1);
            const uint32_t wo = kQ == 2 && sg ? qwov[kSegLo] : qwov[0];
            const uint32_t qc = (a.words[wo + (r >> 4)] >> (2u * (r & 15u))) & 3u;
            v = (v & 0xffff0000u) | (clo == qc ? smat : smis);
          }
          if (kh >= 1) {
            const uint32_t r = (uint32_t)(kh - 1);
            const uint32_t wo = kQ == 2 && sg ? qwov[kSegHi] : qwov[kPer - 1];
            const uint32_t qc = (a.words[wo + (r >> 4)] >> (2u * (r & 15u))) & 3u;
            v = (v & 0xffffu) | ((chi == qc ? smat : smis) << 16);
          }
        }
        tab[sg * kTab + p * RS + jj] = v;
      }
      __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
      __builtin_amdgcn_wave_barrier();

      s16x2 H[S];
#pragma unroll
      for (int jj = 0; jj < S; ++jj) H[jj] = floor2;
      s16x2 diag_b = floor2;
      // one DP row; its left boundary becomes the next row's diagonal
      auto do_row = [&](int j, uint32_t left_u, uint32_t paddr) {
        const lds_u4 *prow = (const lds_u4 *)(uintptr_t)paddr;
        s16x2 P[S];
#pragma unroll
        for (int w = 0; w < S / 4; ++w) {
          const u32x4 v = prow[w];
          P[4 * w + 0] = as_s16x2(v.x);
          P[4 * w + 1] = as_s16x2(v.y);
          P[4 * w + 2] = as_s16x2(v.z);
          P[4 * w + 3] = as_s16x2(v.w);
        }
        const s16x2 left_b = as_s16x2(left_u);
        dp_row2<S>(H, P, left_b, diag_b, gapv, best, best2);
        ws.store(j, as_u32(H[S - 1]));
        diag_b = left_b;
      };
      // Two rows per iteration; their profile-row addresses and left
      // boundaries are loaded one iteration ahead, each into the register
      // whose previous value has just been used: a load that had to be
      // copied into a loop-carried register would have to complete first.
      // The loads are unconditional, keeping the iteration one basic block:
      // boundary rows j + 2, j + 3 <= nrows + 2 <= maxlen + 1 stay inside the
      // wave's maxlen + 2 boundary rows, and list word (j + 1) / 2 <=
      // nrows / 2 inside its maxlen / 2 + 1 list rows; what is loaded past
      // nrows is never used.
      int j = 1;
      for (; j + 1 <= nrows; j += 2) {
        const uint32_t p0 = ra & 0xffffu, p1 = ra >> 16;
        ra = ws.load(a.list_row + ((j + 1) >> 1));
        __builtin_amdgcn_sched_barrier(0);  // issue the load here, a full iteration ahead
        do_row(j, bA, p0);
        if (!first) bA = ws.load(j + 2);
        do_row(j + 1, bB, p1);
        if (!first) bB = ws.load(j + 3);
      }
      if (j <= nrows) do_row(j, bA, ra & 0xffffu);  // odd row count: the strip's last row
    };
    if (nstrips > 0 && nrows > 0) {
      strip(0, std::true_type{});
      for (int s = 1; s < nstrips; ++s) strip(s, std::false_type{});
    }
    best = __builtin_elementwise_max(best, best2);
    if (nstrips > 0 && nrows > 0) {
      // corner H(la-1, lb-1): the last strip's right-most column at row lb-1
      uint32_t cu = as_u32(floor2);
      if (lb_lo >= 2) cu = (cu & 0xffff0000u) | (ws.load_lane_row(lb_lo - 1) & 0xffffu);
      if (lb_hi >= 2) cu = (cu & 0x0000ffffu) | (ws.load_lane_row(lb_hi - 1) & 0xffff0000u);
      corner = as_s16x2(cu);
    }
    finish_query<kG>(a, lane, q, qi - qb, ob, la, t_lo, t_hi, lb_lo, lb_hi, rk_lo, rk_hi,
                 best, corner, q_hi, la_hi);
  }
}

// The persistent kernel: each wave takes work items from the queue.
template <int S, bool kG = false, int kQ = 1, bool kLH = false>
__global__ __launch_bounds__(kBlock, S == 64 ? 3 : S == 32 ? 5 : 1) void sw_lds_kernel(KArgs a) {
  static_assert((kQ == 1 && !kLH) || !kG, "several queries per wave: item lists only");
  static_assert(kQ == 1 || (kQ == 2 && kLH), "two tables per wave: four queries");
  constexpr int kPairs = 25;
  constexpr int RS = S + 4;  // table row stride in words; the +4 shifts banks per row
  constexpr int kTab = kPairs * RS;  // words of one query's table
  constexpr int kWaves = kBlock / kWave;
  __shared__ __attribute__((aligned(16))) uint32_t tab_all[kWaves * kQ * kTab];
  const int lane = threadIdx.x & (kWave - 1);
  const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int wave = blockIdx.x * kWaves + wv;
  uint32_t *tab = tab_all + wv * kQ * kTab;
  uint32_t *bnd = a.bound + (size_t)wave * a.bound_stride;
  const WaveRows ws(bnd, a.bound_stride, lane);
  static_assert(sizeof(tab_all) <= 65536, "profile-row addresses are 16-bit");
  for (;;) {
    int item_idx = 0;
    if (lane == 0) item_idx = atomicAdd(a.counter, 1);
    item_idx = __builtin_amdgcn_readfirstlane(item_idx);
    if (item_idx >= a.n_items) break;
    sw_lds_item<S, kG, kQ, kLH>(a, tab, ws, lane, load_item<kG>(a, item_idx));
  }
}

// ---------------------------------------------------------------------------
// Exact fp16 form of the LDS kernel, for pairs with match * min(la, lb) <= 2048.
// Values are held as k / 2048 in packed fp16, which is exact for integers k of
// magnitude <= 2048; every inner cell has H <= match * (min(la, lb) - 1), so no
// value leaves that range.  Each column keeps G = H - gap instead of H, which
// turns the cell into
//     D = clamp01(G_diag + (s + gap))          v_pk_add_f16 ... clamp
//     H = max3(D, G_left, G_up)                v_pk_maximum3_f16
//     G = H - gap                              v_pk_add_f16
// and best = max3(best, H_even, H_odd) every two columns: 3.5 packed VALU ops
// per two cells instead of 5.  (clamp01 is the local-alignment floor; its upper
// bound 1.0 = 2048 is never exceeded.)
// ---------------------------------------------------------------------------
// compile-time strip width for the strip lambda of the fp16 LDS kernel
template <int W>
struct Width {
  static constexpr int value = W;
};

typedef _Float16 f16x2 __attribute__((ext_vector_type(2)));
constexpr float kF16Scale = 1.0f / 2048.0f;

__device__ __forceinline__ f16x2 as_f16x2(uint32_t v) { return __builtin_bit_cast(f16x2, v); }
__device__ __forceinline__ uint32_t f16_u32(f16x2 v) { return __builtin_bit_cast(uint32_t, v); }
__device__ __forceinline__ f16x2 clamp01(f16x2 x) {
  return __builtin_elementwise_min(__builtin_elementwise_max(x, (f16x2){0, 0}), (f16x2){1, 1});
}

template <int S>
__device__ __forceinline__ void dp_row_f16(f16x2 (&G)[S], const f16x2 (&P)[S], f16x2 gleft,
                                           f16x2 gdiag, f16x2 gapv, f16x2 &best, f16x2 &best2) {
  f16x2 dnext = clamp01(gdiag + P[0]);
  f16x2 hprev = best;
  // two running maxima, alternating every two columns: the compiler sinks
  // them to the row's end, where one chain of dependent v_pk_maximum3 would
  // need a wait state between each pair on gfx950 and two chains need none
#pragma unroll
  for (int jj = 0; jj < S; ++jj) {
    const f16x2 D = dnext;
    if (jj + 1 < S) dnext = clamp01(G[jj] + P[jj + 1]);
    const f16x2 H = __builtin_elementwise_maximum(__builtin_elementwise_maximum(D, gleft), G[jj]);
    const f16x2 g = H + gapv;
    G[jj] = g;
    gleft = g;
    if ((jj & 3) == 1) best = __builtin_elementwise_maximum(__builtin_elementwise_maximum(best, hprev), H);
    else if ((jj & 3) == 3) best2 = __builtin_elementwise_maximum(__builtin_elementwise_maximum(best2, hprev), H);
    else hprev = H;
  }
}

// Register budgets as for sw_lds_kernel.
//
// Several queries per wave (kLH; epsilon mode, item lists), for target
// groups shared by several rows: the low and high halves of the lanes run
// different queries.  A table word holds the low query's score in its low
// half and the high query's in its high half, so one table serves both.
// - kQ = 1: two queries per wave at one query's LDS, qids[qi] against group
//   slots 0-63 (low halves), qids[qi + 1] against slots 64-127.
// - kQ = 2: four, one table per half-wave s (lanes 32 s .. 32 s + 31):
//   qids[qi + 2s] against slots 32 s .., qids[qi + 2s + 1] against slots
//   64 + 32 s ...
// All queries of a wave share the strips, the shorter ones pre-padded on the
// left (padding columns score below every cell they could feed, so a
// query's maximum and its right-most column are unchanged); the rows are
// the group's, as for one query.
// One work item of sw_lds_f16_kernel: target group it.x against the queries
// qids[it.y .. it.z) (it.w: output base), with the wave's profile table(s)
// at tab and its workspace ws.
template <int S, bool kG, int kQ, bool kLH>
__device__ __forceinline__ void sw_lds_f16_item(const KArgs &a, uint32_t *tab, const WaveRows &ws, int lane,
                                        const int4 it) {
  static_assert((kQ == 1 && !kLH) || !kG, "several queries per wave: item lists only");
  static_assert(kQ == 1 || (kQ == 2 && kLH), "two tables per wave: four queries");
  constexpr int kPer = kLH ? 2 : 1;  // queries per table
  constexpr int NQ = kQ * kPer;      // queries per wave
  // queries of the second half-wave's table (kQ = 2): low halves, high halves
  constexpr int kSegLo = kQ == 2 ? kPer : 0, kSegHi = kQ == 2 ? NQ - 1 : kPer - 1;
  constexpr int kPairs = 25;
  constexpr int RS = S + 4;
  constexpr int kTab = kPairs * RS;  // words of one table
  static_assert(S % 2 == 0, "pairs of columns");
  const int seg = kQ == 2 ? lane >> 5 : 0;  // the lane's table
  const uint32_t tab_lds = (uint32_t)(uintptr_t)(lds_u32 *)tab + (uint32_t)seg * (kTab * 4u);

  const _Float16 gap_h = (_Float16)((float)a.gap * kF16Scale);
  const f16x2 gapv = {(_Float16)-gap_h, (_Float16)-gap_h};
  const f16x2 g0 = gapv;  // G of an H = 0 cell
  const f16x2 zero2 = {(_Float16)0.0f, (_Float16)0.0f};
  // substitution scores plus the gap (G stores H - gap), scaled
  const uint32_t smat = (uint32_t)__builtin_bit_cast(uint16_t, (_Float16)((float)(a.match + a.gap) * kF16Scale));
  const uint32_t smis = (uint32_t)__builtin_bit_cast(uint16_t, (_Float16)((float)(a.mismatch + a.gap) * kF16Scale));
  const uint32_t spre = (uint32_t)__builtin_bit_cast(uint16_t, (_Float16)-60000.0f);

  const int g = __builtin_amdgcn_readfirstlane(it.x);
  const int qb = __builtin_amdgcn_readfirstlane(it.y);
  const int qe = __builtin_amdgcn_readfirstlane(it.z);
  const int ob = __builtin_amdgcn_readfirstlane(it.w);

  const int gm = __builtin_amdgcn_readfirstlane(a.gmax[g]);
  const int t_lo = a.gtgt[(size_t)g * kGroup + lane];
  const int t_hi = a.gtgt[(size_t)g * kGroup + kWave + lane];
  const int lb_lo = t_lo >= 0 ? a.len[t_lo] : 0;
  const int lb_hi = t_hi >= 0 ? a.len[t_hi] : 0;
  const int rows_lo = t_lo >= 0 ? lb_lo - 1 : gm;
  const int rows_hi = t_hi >= 0 ? lb_hi - 1 : gm;
  const uint32_t wo_lo = t_lo >= 0 ? a.woff[t_lo] : 0u;
  const uint32_t wo_hi = t_hi >= 0 ? a.woff[t_hi] : 0u;
  const int rk_lo = t_lo >= 0 ? a.rank[t_lo] : 0;
  const int rk_hi = t_hi >= 0 ? a.rank[t_hi] : 0;
  const int nrows = gm > 1 ? gm - 1 : 0;
  build_row_list(a, ws, nrows, wo_lo, wo_hi, rows_lo, rows_hi, tab_lds, RS * 4u);

  for (int qi = qb; qi < qe; qi += NQ) {
    // the wave's queries: qv[u], its length, first word and left pre-padding
    int qv[NQ], lav[NQ], shv[NQ];
    uint32_t qwov[NQ];
    int ncols = 0;
#pragma unroll
    for (int u = 0; u < NQ; ++u) {
      qv[u] = __builtin_amdgcn_readfirstlane(u == 0 ? load_qid<kG>(a, qi) : a.qids[qi + u]);
      lav[u] = __builtin_amdgcn_readfirstlane(a.len[qv[u]]);
      qwov[u] = __builtin_amdgcn_readfirstlane(a.woff[qv[u]]);
      ncols = max(ncols, lav[u] > 1 ? lav[u] - 1 : 0);
    }
#pragma unroll
    for (int u = 0; u < NQ; ++u) shv[u] = ncols - (lav[u] > 1 ? lav[u] - 1 : 0);
    // the lane's queries: table seg's, low halves q, high halves q_hi
    const int q = (kQ == 2 && seg) ? qv[kSegLo] : qv[0];
    const int q_hi = (kQ == 2 && seg) ? qv[kSegHi] : qv[kPer - 1];
    const int la = (kQ == 2 && seg) ? lav[kSegLo] : lav[0];
    const int la_hi = (kQ == 2 && seg) ? lav[kSegHi] : lav[kPer - 1];
    const int nstrips = (ncols + S - 1) / S;
    f16x2 best = zero2, best2 = zero2;

    // One strip of W query columns whose slot 0 is column j0 (columns < 1
    // are pre-padding); first (compile-time): no strip to its left.
    auto strip = [&](auto wc, int j0, auto firstc) {
      constexpr int W = decltype(wc)::value;
      constexpr bool first = decltype(firstc)::value;
      // The first list word and left boundaries are loaded before the
      // table build, whose own loads are waited for, so nothing is pending
      // when the row loop starts: the loop's waits then count only the
      // loads and stores of one iteration.
      uint32_t ra = ws.load(a.list_row);
      uint32_t bA = f16_u32(g0), bB = f16_u32(g0);
      if (!first) {
        bA = ws.load(1);
        if (nrows >= 2) bB = ws.load(2);
      }
      __builtin_amdgcn_wave_barrier();
      for (int e = lane; e < kQ * kPairs * W; e += kWave) {
        const int sg = kQ == 2 && e >= kPairs * W ? 1 : 0;
        const int e1 = e - sg * (kPairs * W);
        const int p = e1 / W, jj = e1 - p * W;
        uint32_t clo, chi;
        tab_codes(p, clo, chi);
        uint32_t v = spre | (spre << 16);
        if constexpr (!kLH) {  // one query: one code for both halves
          const int k = j0 + jj - shv[0];
          if (k >= 1) {
            const uint32_t r = (uint32_t)(k - 1);
            const uint32_t qc = (a.words[qwov[0] + (r >> 4)] >> (2u * (r & 15u))) & 3u;
            v = (clo == qc ? smat : smis) | ((chi == qc ? smat : smis) << 16);
          }
        } else {
          // table sg: low halves score query sg * kPer, high halves sg * kPer + 1
          const int kl = j0 + jj - (kQ == 2 && sg ? shv[kSegLo] : shv[0]);
          const int kh = j0 + jj - (kQ == 2 && sg ? shv[kSegHi] : shv[kPer - 1]);
          if (kl >= 1) {
            const uint32_t r = (uint32_t)(kl - 1);
            const uint32_t wo = kQ == 2 && sg ? qwov[kSegLo] : qwov[0];
            const uint32_t qc = (a.words[wo + (r >> 4)] >> (2u * (r & 15u))) & 3u;
            v = (v & 0xffff0000u) | (clo == qc ? smat : smis);
          }
          if (kh >= 1) {
            const uint32_t r = (uint32_t)(kh - 1);
            const uint32_t wo = kQ == 2 && sg ? qwov[kSegHi] : qwov[kPer - 1];
            const uint32_t qc = (a.words[wo + (r >> 4)] >> (2u * (r & 15u))) & 3u;
            v = (v & 0xffffu) | ((chi == qc ? smat : smis) << 16);
          }
        }
        tab[sg * kTab + p * RS + jj] = v;
      }
      __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
      __builtin_amdgcn_wave_barrier();

      f16x2 G[W];
#pragma unroll
      for (int jj = 0; jj < W; ++jj) G[jj] = g0;
      f16x2 gdiag = g0;
      auto do_row = [&](int j, uint32_t left_u, uint32_t paddr) {
        const lds_u4 *prow = (const lds_u4 *)(uintptr_t)paddr;
        f16x2 P[W];
#pragma unroll
        for (int w = 0; w < W / 4; ++w) {
          const u32x4 v = prow[w];
          P[4 * w + 0] = as_f16x2(v.x);
          P[4 * w + 1] = as_f16x2(v.y);
          P[4 * w + 2] = as_f16x2(v.z);
          P[4 * w + 3] = as_f16x2(v.w);
        }
        const f16x2 gleft = as_f16x2(left_u);
        dp_row_f16<W>(G, P, gleft, gdiag, gapv, best, best2);
        ws.store(j, f16_u32(G[W - 1]));
        gdiag = gleft;
      };
      // Two rows per iteration; their profile-row addresses and left
      // boundaries are loaded one iteration ahead, each into the register
      // whose previous value has just been used: a load that had to be
      // copied into a loop-carried register would have to complete first.
      // The loads are unconditional, keeping the iteration one basic block:
      // boundary rows j + 2, j + 3 <= nrows + 2 <= maxlen + 1 stay inside the
      // wave's maxlen + 2 boundary rows, and list word (j + 1) / 2 <=
      // nrows / 2 inside its maxlen / 2 + 1 list rows; what is loaded past
      // nrows is never used.
      int j = 1;
      for (; j + 1 <= nrows; j += 2) {
        const uint32_t p0 = ra & 0xffffu, p1 = ra >> 16;
        ra = ws.load(a.list_row + ((j + 1) >> 1));
        __builtin_amdgcn_sched_barrier(0);  // issue the load here, a full iteration ahead
        // the second row's profile reads after the first row's first
        // column adds, not in one burst with the first row's
        do_row(j, bA, p0);
        if (!first) bA = ws.load(j + 2);
        do_row(j + 1, bB, p1);
        if (!first) bB = ws.load(j + 3);
      }
      if (j <= nrows) do_row(j, bA, ra & 0xffffu);
    };
    if (nstrips > 0 && nrows > 0) {
      // strips are right-aligned: the first holds the rem = 1..S leftmost
      // columns and runs in the narrowest multiple of 8 columns that holds
      // them (S = 64: 8, 16, ..., 64; S = 32: 8, 16, 24, 32), so it pads
      // fewer than 8
      const int rem = ncols - (nstrips - 1) * S;
      if constexpr (S == 64) {
        switch ((rem + 7) >> 3) {
          case 1: strip(Width<8>{}, rem - 8 + 1, std::true_type{}); break;
          case 2: strip(Width<16>{}, rem - 16 + 1, std::true_type{}); break;
          case 3: strip(Width<24>{}, rem - 24 + 1, std::true_type{}); break;
          case 4: strip(Width<32>{}, rem - 32 + 1, std::true_type{}); break;
          case 5: strip(Width<40>{}, rem - 40 + 1, std::true_type{}); break;
          case 6: strip(Width<48>{}, rem - 48 + 1, std::true_type{}); break;
          case 7: strip(Width<56>{}, rem - 56 + 1, std::true_type{}); break;
          default: strip(Width<64>{}, rem - 64 + 1, std::true_type{}); break;
        }
      } else if constexpr (S == 32) {
        if (rem <= S / 4) strip(Width<S / 4>{}, rem - S / 4 + 1, std::true_type{});
        else if (rem <= S / 2) strip(Width<S / 2>{}, rem - S / 2 + 1, std::true_type{});
        else if (rem <= 3 * S / 4) strip(Width<3 * S / 4>{}, rem - 3 * S / 4 + 1, std::true_type{});
        else strip(Width<S>{}, rem - S + 1, std::true_type{});
      } else {
        strip(Width<S>{}, rem - S + 1, std::true_type{});
      }
      for (int s = 1; s < nstrips; ++s) strip(Width<S>{}, rem + (s - 1) * S + 1, std::false_type{});
    }
    best = __builtin_elementwise_maximum(best, best2);
    // back to integers: H = 2048 * value; corner H = G + gap at row lb-1
    const float sc = 2048.0f;
    int best_lo = (int)__builtin_rintf((float)best.x * sc);
    int best_hi = (int)__builtin_rintf((float)best.y * sc);
    int cor_lo = 0, cor_hi = 0;
    if (nstrips > 0 && nrows > 0) {
      if (lb_lo >= 2) {
        const f16x2 gv = as_f16x2(ws.load_lane_row(lb_lo - 1));
        cor_lo = (int)__builtin_rintf((float)gv.x * sc) + a.gap;
      }
      if (lb_hi >= 2) {
        const f16x2 gv = as_f16x2(ws.load_lane_row(lb_hi - 1));
        cor_hi = (int)__builtin_rintf((float)gv.y * sc) + a.gap;
      }
    }
    finish_query_int<kG>(a, lane, q, qi - qb, ob, la, t_lo, t_hi, lb_lo, lb_hi, rk_lo, rk_hi,
                     best_lo, best_hi, cor_lo, cor_hi, q_hi, la_hi);
  }
}

// The persistent kernel: each wave takes work items from the queue.
template <int S, bool kG = false, int kQ = 1, bool kLH = false>
__global__ __launch_bounds__(kBlock, S == 64 ? 3 : S == 32 ? 5 : 1) void sw_lds_f16_kernel(KArgs a) {
  static_assert((kQ == 1 && !kLH) || !kG, "several queries per wave: item lists only");
  static_assert(kQ == 1 || (kQ == 2 && kLH), "two tables per wave: four queries");
  constexpr int kPairs = 25;
  constexpr int RS = S + 4;
  constexpr int kTab = kPairs * RS;  // words of one table
  constexpr int kWaves = kBlock / kWave;
  static_assert(S % 2 == 0, "pairs of columns");
  __shared__ __attribute__((aligned(16))) uint32_t tab_all[kWaves * kQ * kTab];
  const int lane = threadIdx.x & (kWave - 1);
  const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int wave = blockIdx.x * kWaves + wv;
  uint32_t *tab = tab_all + wv * kQ * kTab;
  uint32_t *bnd = a.bound + (size_t)wave * a.bound_stride;
  const WaveRows ws(bnd, a.bound_stride, lane);
  static_assert(sizeof(tab_all) <= 65536, "profile-row addresses are 16-bit");
  for (;;) {
    int item_idx = 0;
    if (lane == 0) item_idx = atomicAdd(a.counter, 1);
    item_idx = __builtin_amdgcn_readfirstlane(item_idx);
    if (item_idx >= a.n_items) break;
    sw_lds_f16_item<S, kG, kQ, kLH>(a, tab, ws, lane, load_item<kG>(a, item_idx));
  }
}

// One launch for the items of both 16-bit lane formats (epsilon mode, item
// lists, one query per wave): it.w holds the item's form, so the items that
// need the int16 lanes and those the exact fp16 form covers share one
// persistent queue in one cost order, with no tail between two launches.
// Both item bodies use the same table layout (25 rows of S + 4 words).
constexpr int kFormF16 = 0, kFormI16 = 1;
template <int S>
__global__ __launch_bounds__(kBlock, S == 64 ? 3 : S == 32 ? 5 : 1) void sw_lds_mixed_kernel(KArgs a) {
  constexpr int kTab = 25 * (S + 4);
  constexpr int kWaves = kBlock / kWave;
  __shared__ __attribute__((aligned(16))) uint32_t tab_all[kWaves * kTab];
  static_assert(sizeof(tab_all) <= 65536, "profile-row addresses are 16-bit");
  const int lane = threadIdx.x & (kWave - 1);
  const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int wave = blockIdx.x * kWaves + wv;
  uint32_t *tab = tab_all + wv * kTab;
  uint32_t *bnd = a.bound + (size_t)wave * a.bound_stride;
  const WaveRows ws(bnd, a.bound_stride, lane);
  for (;;) {
    int item_idx = 0;
    if (lane == 0) item_idx = atomicAdd(a.counter, 1);
    item_idx = __builtin_amdgcn_readfirstlane(item_idx);
    if (item_idx >= a.n_items) break;
    const int4 it = a.items[item_idx];
    const int4 it0 = make_int4(it.x, it.y, it.z, 0);
    if (__builtin_amdgcn_readfirstlane(it.w) == kFormI16) sw_lds_item<S, false, 1, false>(a, tab, ws, lane, it0);
    else sw_lds_f16_item<S, false, 1, false>(a, tab, ws, lane, it0);
  }
}

// ---- epsilon-list compaction: keys -> CSR rows in ascending column order ----

__global__ void count_rows_kernel(const unsigned long long *keys, long long n, int32_t *cnt) {
  for (long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x; k < n;
       k += (long long)gridDim.x * blockDim.x)
    atomicAdd(&cnt[(int)(keys[k] >> 40)], 1);
}

__global__ void scatter_rows_kernel(const unsigned long long *keys, long long n,
                                    const long long *row_ptr, int32_t *cursor,
                                    int32_t *seg_col, uint16_t *seg_d) {
  for (long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x; k < n;
       k += (long long)gridDim.x * blockDim.x) {
    const unsigned long long key = keys[k];
    const int row = (int)(key >> 40);
    const long long pos = row_ptr[row] + atomicAdd(&cursor[row], 1);
    seg_col[pos] = (int32_t)((key >> 16) & 0xffffffull);
    seg_d[pos] = (uint16_t)(key & 0xffffull);
  }
}

constexpr int kSortThreads = 256;
constexpr int kShortRow = 256;

// One workgroup per row: columns are distinct, so each entry's position in
// the sorted row is the number of smaller columns in the row.  seg_d / out_d
// may be null (columns only).
template <typename P>
__global__ __launch_bounds__(kSortThreads) void sort_rows_kernel(
    const long long *row_ptr, const int32_t *seg_col, const P *seg_d,
    int32_t *out_col, P *out_d, int n_cols) {
  extern __shared__ uint32_t lds[];
  const int row = blockIdx.x;
  const long long b = row_ptr[row];
  const int len = (int)(row_ptr[row + 1] - b);
  const int tid = threadIdx.x;
  if (len == 0) return;
  if (len <= kShortRow) {
    for (int t = tid; t < len; t += kSortThreads) lds[t] = (uint32_t)seg_col[b + t];
    __syncthreads();
    for (int t = tid; t < len; t += kSortThreads) {
      const uint32_t c = lds[t];
      int r = 0;
      for (int m = 0; m < len; ++m) r += lds[m] < c ? 1 : 0;
      out_col[b + r] = (int32_t)c;
      if (seg_d) out_d[b + r] = seg_d[b + t];
    }
    return;
  }
  const int W = (n_cols + 31) / 32;
  uint32_t *bits = lds;
  uint32_t *prefix = lds + W;
  uint32_t *tot = lds + 2 * W;
  for (int w = tid; w < W; w += kSortThreads) bits[w] = 0u;
  __syncthreads();
  for (int t = tid; t < len; t += kSortThreads) {
    const uint32_t c = (uint32_t)seg_col[b + t];
    atomicOr(&bits[c >> 5], 1u << (c & 31));
  }
  __syncthreads();
  const int chunk = (W + kSortThreads - 1) / kSortThreads;
  const int w0 = tid * chunk, w1 = min(W, w0 + chunk);
  uint32_t sum = 0;
  for (int w = w0; w < w1; ++w) sum += __popc(bits[w]);
  tot[tid] = sum;
  __syncthreads();
  if (tid == 0) {
    uint32_t run = 0;
    for (int t = 0; t < kSortThreads; ++t) {
      const uint32_t v = tot[t];
      tot[t] = run;
      run += v;
    }
  }
  __syncthreads();
  uint32_t run = tot[tid];
  for (int w = w0; w < w1; ++w) {
    prefix[w] = run;
    run += __popc(bits[w]);
  }
  __syncthreads();
  for (int t = tid; t < len; t += kSortThreads) {
    const uint32_t c = (uint32_t)seg_col[b + t];
    const uint32_t w = c >> 5;
    const uint32_t r = prefix[w] + __popc(bits[w] & ((1u << (c & 31)) - 1u));
    out_col[b + r] = (int32_t)c;
    if (seg_d) out_d[b + r] = seg_d[b + t];
  }
}

}  // namespace

// ----------------------------------------------------------------------------
// host side
// ----------------------------------------------------------------------------

namespace {

// Kernel variants: the LDS query-profile kernel (residues in <= 4 codes) and
// the register-profile kernel (any alphabet).

int pick_kernel(const mnhn_sw_ctx *c) {
  int k = c->kernel;
  if (k == KERNEL_AUTO) {
    const char *e = getenv("MNHN_SW_KERNEL");
    if (e && strcmp(e, "reg") == 0) k = KERNEL_REG;
    if (e && strcmp(e, "lds") == 0) k = KERNEL_LDS;
    if (e && strcmp(e, "f16") == 0) k = KERNEL_LDS_F16;
  }
  if (k == KERNEL_AUTO) k = KERNEL_LDS_F16;
  if (!c->acgt) k = KERNEL_REG;  // the LDS table needs <= 4 residue codes
  return k;
}

int pick_strip(const mnhn_sw_ctx *c) {
  const int k = pick_kernel(c);
  int v = c->strip;
  if (v == 0) {
    const char *e = getenv("MNHN_SW_STRIP");
    if (e) v = atoi(e);
  }
  // untuned default 64: measured on MI355X, the strip-64 LDS kernels run both the 20,000-sequence
  // (lengths 100-600) and the 105,000-sequence (150-300) all-pairs sweeps fastest; items whose
  // queries are all <= 128 residues still run at 32
  if (k == KERNEL_LDS || k == KERNEL_LDS_F16) return (v == 16 || v == 32 || v == 64) ? v : 64;
  return (v == 16 || v == 32) ? v : 32;
}

template <bool kG>
void *kernel_ptr_g(int k, int S, bool acgt) {
  if (k == KERNEL_I32)
    return acgt ? reinterpret_cast<void *>(&sw_i32_kernel<16, true, kG>)
                : reinterpret_cast<void *>(&sw_i32_kernel<16, false, kG>);
  if (k == KERNEL_LDS_F16) {
    if (S == 16) return reinterpret_cast<void *>(&sw_lds_f16_kernel<16, kG>);
    if (S == 64) return reinterpret_cast<void *>(&sw_lds_f16_kernel<64, kG>);
    return reinterpret_cast<void *>(&sw_lds_f16_kernel<32, kG>);
  }
  if (k == KERNEL_LDS) {
    if (S == 16) return reinterpret_cast<void *>(&sw_lds_kernel<16, kG>);
    if (S == 64) return reinterpret_cast<void *>(&sw_lds_kernel<64, kG>);
    return reinterpret_cast<void *>(&sw_lds_kernel<32, kG>);
  }
  if (S == 16) return acgt ? reinterpret_cast<void *>(&sw_tiles_kernel<16, true, kG>)
                           : reinterpret_cast<void *>(&sw_tiles_kernel<16, false, kG>);
  return acgt ? reinterpret_cast<void *>(&sw_tiles_kernel<32, true, kG>)
              : reinterpret_cast<void *>(&sw_tiles_kernel<32, false, kG>);
}

// Workspace rows per wave for kernel variant k and sequences of at most len
// residues: len + 2 strip-boundary rows, and for the kernels that decode
// their target rows once per item, len / 2 + 1 rows of the row-address list
// after them (first list row: len + 2).
int wave_rows(int len, int k) {
  return len + 2 + (k == KERNEL_LDS_F16 || k == KERNEL_LDS || k == KERNEL_LDS_MIXED ? len / 2 + 1 : 0);
}

// Kernel of variant k at strip S; group_items: the allocation-free
// one-vs-all instantiation; split = 3 / 4: the fp16 or int16 LDS kernel with
// two / four queries per wave (the lanes' low and high halves on different
// queries, four with one table per half-wave; strips 32 and 64).
void *kernel_ptr(int k, int S, bool acgt, bool group_items = false, int split = 1) {
  if (k == KERNEL_LDS_MIXED) {
    if (group_items || split != 1) return nullptr;
    if (S == 16) return reinterpret_cast<void *>(&sw_lds_mixed_kernel<16>);
    return S == 64 ? reinterpret_cast<void *>(&sw_lds_mixed_kernel<64>)
                   : reinterpret_cast<void *>(&sw_lds_mixed_kernel<32>);
  }
  if (split == 3 || split == 4) {
    if (group_items || (k != KERNEL_LDS_F16 && k != KERNEL_LDS)) return nullptr;
    if (k == KERNEL_LDS_F16) {
      if (split == 3)
        return S == 64 ? reinterpret_cast<void *>(&sw_lds_f16_kernel<64, false, 1, true>)
                       : reinterpret_cast<void *>(&sw_lds_f16_kernel<32, false, 1, true>);
      return S == 64 ? reinterpret_cast<void *>(&sw_lds_f16_kernel<64, false, 2, true>)
                     : reinterpret_cast<void *>(&sw_lds_f16_kernel<32, false, 2, true>);
    }
    if (split == 3)
      return S == 64 ? reinterpret_cast<void *>(&sw_lds_kernel<64, false, 1, true>)
                     : reinterpret_cast<void *>(&sw_lds_kernel<32, false, 1, true>);
    return S == 64 ? reinterpret_cast<void *>(&sw_lds_kernel<64, false, 2, true>)
                   : reinterpret_cast<void *>(&sw_lds_kernel<32, false, 2, true>);
  }
  if (split != 1) return nullptr;
  return group_items ? kernel_ptr_g<true>(k, S, acgt) : kernel_ptr_g<false>(k, S, acgt);
}

// Resident waves per CU for a launch of kernel variant k at strip S: the
// tuned or requested value, else the occupancy limit.
int launch_waves_per_cu(const mnhn_sw_ctx *c, int k, int S, const void *kp) {
  (void)k;
  (void)S;
  int wpc = c->waves_per_cu;
  if (wpc <= 0) {
    const char *ev = getenv("MNHN_SW_WAVES_PER_CU");
    wpc = ev ? atoi(ev) : 0;
  }
  if (wpc <= 0) {
    int blocks_per_cu = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&blocks_per_cu, kp, kBlock, 0) != hipSuccess ||
        blocks_per_cu <= 0)
      blocks_per_cu = 1;
    wpc = blocks_per_cu * (kBlock / kWave);
  }
  return wpc;
}

}  // namespace

// Device scratch of the allocation-free entry point (mnhn_sw_one_vs_all_noalloc):
// its work-queue counters (first 256 bytes) and 96 MiB of strip-boundary
// workspaces for its launches (four waves' worth at 65,534 residues, 384 B per
// residue and wave).  Each GPU that the library runs on holds its own copy,
// from the code object's load until the process ends.  In host code `scratch`
// only names the symbol for hipGetSymbolAddress: kernels and copies receive
// the device address that call returns (scratch_address), never the name.
static constexpr size_t kScratchHead = 256;
static constexpr size_t kScratchBytes = (96ull << 20) + kScratchHead;
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void *g_scratch_ptr[kMaxDevices];
bool g_scratch_done[kMaxDevices];

// The address of `scratch` on device dev, which must be current; resolved
// once per device.  nullptr if the runtime cannot resolve it.
void *scratch_address(int dev) {
  if (dev < 0 || dev >= kMaxDevices) return nullptr;
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_done[dev]) {
    void *ptr = nullptr;
    const hipError_t e = hipGetSymbolAddress((void **)&ptr, HIP_SYMBOL(scratch));
    g_scratch_ptr[dev] = e == hipSuccess ? ptr : nullptr;
    g_scratch_done[dev] = true;
  }
  return g_scratch_ptr[dev];
}

__global__ void zero_words_kernel(int *p, int n) {
  if ((int)threadIdx.x < n) p[threadIdx.x] = 0;
}
}  // namespace

int mnhn_detail::launch_tiles(mnhn_sw_ctx *c, const Work &w, int mode, int triangle, int eps_max,
                 int32_t *d_dist, unsigned long long *d_keys,
                 unsigned long long key_cap) {
  const size_t n_items_total = w.dev ? (size_t)w.n_items : w.items.size();
  if (n_items_total == 0) return MNHN_OK;
  mnhn_phase_clock clk(c->stream, "launch_tiles");
  int rc;
  int4 *d_items = nullptr;
  int32_t *d_qids = nullptr, *d_gt = nullptr, *d_gm = nullptr, *d_gn = nullptr;
  const bool own_groups = w.dev || !w.gtgt.empty();
  if (!w.dev) {
    if ((rc = dev_alloc(&d_items, w.items.size()))) return rc;
    if ((rc = dev_alloc(&d_qids, w.qids.size()))) { dev_free(d_items); return rc; }
    if (own_groups) {
      if ((rc = dev_alloc(&d_gt, w.gtgt.size())) || (rc = dev_alloc(&d_gm, w.gmax.size())) ||
          (rc = dev_alloc(&d_gn, w.gmin.size()))) {
        dev_free(d_items); dev_free(d_qids); dev_free(d_gt); dev_free(d_gm); dev_free(d_gn);
        return rc;
      }
    }
  }
  auto cleanup = [&]() {
    dev_free(d_items); dev_free(d_qids); dev_free(d_gt); dev_free(d_gm); dev_free(d_gn);
  };
  hipError_t e = hipSuccess;
  if (!w.dev) {
    e = hipMemcpyAsync(d_qids, w.qids.data(), w.qids.size() * sizeof(int32_t),
                       hipMemcpyHostToDevice, c->stream);
    if (e == hipSuccess && own_groups) {
      e = hipMemcpyAsync(d_gt, w.gtgt.data(), w.gtgt.size() * sizeof(int32_t),
                         hipMemcpyHostToDevice, c->stream);
      if (e == hipSuccess)
        e = hipMemcpyAsync(d_gm, w.gmax.data(), w.gmax.size() * sizeof(int32_t),
                           hipMemcpyHostToDevice, c->stream);
      if (e == hipSuccess)
        e = hipMemcpyAsync(d_gn, w.gmin.data(), w.gmin.size() * sizeof(int32_t),
                           hipMemcpyHostToDevice, c->stream);
    }
  }
  if (e != hipSuccess) { cleanup(); return MNHN_ERR_DEVICE; }
  const int32_t *q_ptr = w.dev ? w.d_qids : d_qids;
  const int32_t *gt_ptr = w.dev ? w.d_gtgt : (own_groups ? d_gt : c->d_gtgt);
  const int32_t *gm_ptr = w.dev ? w.d_gmax : (own_groups ? d_gm : c->d_gmax);
  const int32_t *gn_ptr = w.dev ? w.d_gmin : (own_groups ? d_gn : c->d_gmin);

  const int K = w.dev ? (w.dev_kernel ? w.dev_kernel : KERNEL_LDS) : pick_kernel(c);
  const int S = pick_strip(c);
  // passes: (kernel, items).  The fp16 variant takes the items whose pairs
  // all satisfy match * min(la, lb) <= 2048; the rest run the int16 LDS kernel.
  struct Pass {
    int kernel, strip;
    std::vector<int4> items;
    double cells = 0;
    int split = 1;
    double form_cells[2] = {0, 0};  // KERNEL_LDS_MIXED: cells of its fp16 / int16 items
  };
  std::vector<Pass> passes;
  // an item's longest query; with the length order, its first
  auto item_max_query = [&](const int4 &it) {
    if (w.qids_order) return it.z > it.y ? c->len[w.qids[it.y]] : 0;
    int m = 0;
    for (int qi = it.y; qi < it.z; ++qi) m = std::max(m, c->len[w.qids[qi]]);
    return m;
  };
  // prefix sums of the queries' lengths along qids (length order only)
  std::vector<double> qlen_prefix;
  if (!w.dev && w.qids_order) {
    qlen_prefix.assign(w.qids.size() + 1, 0.0);
    for (size_t k = 0; k < w.qids.size(); ++k) qlen_prefix[k + 1] = qlen_prefix[k] + c->len[w.qids[k]];
  }
  if (w.dev) {
    const int DK = (K == KERNEL_LDS_F16 && !w.dev_f16_ok) ? KERNEL_LDS : K;
    const int DS = DK == KERNEL_I32 ? kI32Strip : (w.dev_strip ? w.dev_strip : S);
    if (w.dev_split != 1 && (w.dev_split < 3 || w.dev_split > 4 || (DK != KERNEL_LDS_F16 && DK != KERNEL_LDS) ||
                             mode != MODE_EPS || (DS != 32 && DS != 64))) {
      cleanup();
      return MNHN_ERR_ARG;
    }
    passes.push_back({DK, DS, {}, 0.0, w.dev_split});
  } else {
    // items by the lane format their pairs need: a pair's inner scores are at
    // most match * min(la, lb), so items whose bound exceeds the 16-bit range
    // run the 32-bit kernel; with the fp16 form, items whose bound is at most
    // 2048 run exactly in fp16 and the rest in packed int16
    // With the fp16 form in epsilon mode, its items and the int16 ones run
    // in one launch (sw_lds_mixed_kernel, the form in it.w) in the work
    // list's cost order; MNHN_SW_MIXED=0 gives each form a launch of its own.
    const char *mx = getenv("MNHN_SW_MIXED");
    const bool mixed = K == KERNEL_LDS_F16 && mode == MODE_EPS && !(mx && atoi(mx) == 0);
    std::vector<int4> f16, i16, i32;
    f16.reserve(w.items.size());
    for (const int4 &it : w.items) {
      const int gm = own_groups ? w.gmax[it.x] : c->gmax[it.x];
      const int maxla = item_max_query(it);
      const long long bound = (long long)c->match * std::min(gm, maxla);
      if (bound > kI16ScoreMax) i32.push_back(it);
      else if (mixed) f16.push_back(make_int4(it.x, it.y, it.z, bound <= 2048 ? kFormF16 : kFormI16));
      else if (K == KERNEL_LDS_F16 && bound <= 2048) f16.push_back(it);
      else i16.push_back(it);
    }
    if (!f16.empty()) passes.push_back({mixed ? KERNEL_LDS_MIXED : KERNEL_LDS_F16, S, std::move(f16)});
    if (!i16.empty()) passes.push_back({K == KERNEL_LDS_F16 ? KERNEL_LDS : K, S, std::move(i16)});
    if (!i32.empty()) passes.push_back({KERNEL_I32, kI32Strip, std::move(i32)});
  }
  // algorithmic cells of each pass (sum of la * lb over the pairs its items
  // compute): with the context's groups from the targets' length suffix sums
  // (the triangle keeps targets after the query in length order), with own
  // groups from their slots' lengths
  std::vector<double> own_gsum;
  if (!w.dev && own_groups) {
    own_gsum.assign(w.gmax.size(), 0.0);
    for (size_t g = 0; g < w.gmax.size(); ++g)
      for (int t = 0; t < kGroup; ++t) {
        const int id = w.gtgt[g * kGroup + t];
        if (id >= 0) own_gsum[g] += c->len[id];
      }
  }
  auto item_cells = [&](const int4 &it) {
    double s = 0;
    if (w.qids_order && !own_groups) {
      // queries before the group's first member pair with all its slots
      // (suffix from 0); with the triangle, the group's own members pair
      // with the slots after them
      const double *gs = &c->gsuf[(size_t)it.x * (kGroup + 1)];
      const int g0 = it.x * kGroup;
      const int mid = triangle ? std::min(std::max(g0, it.y), it.z) : it.z;
      s = gs[0] * (qlen_prefix[mid] - qlen_prefix[it.y]);
      for (int k = mid; k < it.z; ++k)
        s += (double)c->len[c->order[k]] * gs[std::max(0, std::min(kGroup, k - g0 + 1))];
      return s;
    }
    for (int qi = it.y; qi < it.z; ++qi) {
      const int q = w.qids[qi];
      double tg;
      if (own_groups) {
        tg = own_gsum[it.x];
      } else {
        int from = 0;
        if (triangle) from = std::max(0, std::min(kGroup, c->rank[q] - it.x * kGroup + 1));
        tg = c->gsuf[(size_t)it.x * (kGroup + 1) + from];
      }
      s += (double)c->len[q] * tg;
    }
    return s;
  };
  // LDS kernels at S = 64: items whose queries are all at most kShortQuery
  // long run at S = 32 (right-aligned 64-column strips would pad a 90-residue
  // query to 128 columns)
  if (!w.dev && S == 64) {
    constexpr int kShortQuery = 128;
    std::vector<Pass> split;
    for (Pass &ps : passes) {
      if (ps.kernel != KERNEL_LDS && ps.kernel != KERNEL_LDS_F16) {
        split.push_back(std::move(ps));
        continue;
      }
      std::vector<int4> shorter, rest;
      rest.reserve(ps.items.size());
      for (const int4 &it : ps.items) (item_max_query(it) <= kShortQuery ? shorter : rest).push_back(it);
      if (!rest.empty()) split.push_back({ps.kernel, S, std::move(rest)});
      if (!shorter.empty()) split.push_back({ps.kernel, 32, std::move(shorter)});
    }
    passes.swap(split);
  }
  for (Pass &ps : passes) {
    if (w.dev) {
      ps.cells = w.cells;
    } else {
      for (const int4 &it : ps.items) {
        const double ci = item_cells(it);
        ps.cells += ci;
        if (ps.kernel == KERNEL_LDS_MIXED) ps.form_cells[it.w == kFormI16 ? 1 : 0] += ci;
      }
    }
  }
  if ((int)passes.size() * 2 > (int)(sizeof(c->pev) / sizeof(c->pev[0]))) {
    cleanup();
    return MNHN_ERR_ARG;
  }
  for (size_t k = 0; k < 2 * passes.size(); ++k)
    if (!c->pev[k] && hipEventCreate(&c->pev[k]) != hipSuccess) {
      c->pev[k] = nullptr;
      cleanup();
      return MNHN_ERR_DEVICE;
    }

  // strip-boundary workspace, [row][lane] per wave.  Rows are query residues
  // in the register-profile kernel and target residues in the LDS kernels, so
  // it is sized for the longest sequence of the context.
  const int ws_len = std::max(w.max_q_len, c->maxlen);
  int waves_used = 0;
  bool any_i32 = false;
  clk.mark("passes and cells");
  (void)hipEventRecord(c->ev[0], c->stream);
  for (size_t pi = 0; pi < passes.size(); ++pi) {
    Pass &pass = passes[pi];
    const int PK = pass.kernel, PS = pass.strip;
    const std::vector<int4> &items = pass.items;
    const size_t n_items = w.dev ? (size_t)w.n_items : items.size();
    void *kp = kernel_ptr(PK, PS, c->acgt, false, pass.split);
    const int stride = wave_rows(ws_len, PK) * kWave;
    // resident waves are capped so that their workspaces stay within
    // kWorkspaceBytes (long sequences: 16,384 residues take 6 MB per wave)
    const size_t ws_waves =
        std::max<size_t>(kWorkspaceBytes / ((size_t)stride * sizeof(uint32_t)), kBlock / kWave);
    const int wpc = launch_waves_per_cu(c, PK, PS, kp);
    int blocks = (c->num_cu * wpc + (kBlock / kWave) - 1) / (kBlock / kWave);
    const int max_useful = (int)((n_items + 3) / 4);
    if (blocks > max_useful) blocks = std::max(1, max_useful);
    blocks = (int)std::min<size_t>((size_t)blocks, ws_waves / (kBlock / kWave));
    any_i32 |= PK == KERNEL_I32;
    const int waves = blocks * (kBlock / kWave);
    waves_used = std::max(waves_used, waves);
    const size_t need = (size_t)waves * stride * sizeof(uint32_t);
    if (need > c->bound_bytes) {
      e = hipStreamSynchronize(c->stream);  // an earlier pass may still use it
      dev_free(c->d_bound);
      c->bound_bytes = 0;
      if (e != hipSuccess || hipMalloc((void **)&c->d_bound, need) != hipSuccess) {
        c->d_bound = nullptr;
        cleanup();
        return MNHN_ERR_NOMEM;
      }
      c->bound_bytes = need;
    }
    if (!w.dev) {
      // lists of 1 MB and more go through the context's pinned staging (an
      // 11 MB list: 3 ms from pageable memory); the previous pass has
      // finished with it (the stream is synchronized after every pass)
      const size_t ib = items.size() * sizeof(int4);
      const void *src = items.data();
      if (ib >= (1u << 20)) {
        if (c->h_items_bytes < ib) {
          if (c->h_items) (void)hipHostFree(c->h_items);
          c->h_items = nullptr;
          c->h_items_bytes = 0;
          if (hipHostMalloc(&c->h_items, ib, hipHostMallocDefault) == hipSuccess) c->h_items_bytes = ib;
          else c->h_items = nullptr;
        }
        if (c->h_items) {
          memcpy(c->h_items, items.data(), ib);
          src = c->h_items;
        }
      }
      e = hipMemcpyAsync(d_items, src, ib, hipMemcpyHostToDevice, c->stream);
    }
    if (e == hipSuccess) e = hipMemsetAsync(c->d_counter, 0, sizeof(int), c->stream);
    KArgs a{};
    a.words = c->d_words; a.woff = c->d_woff; a.len = c->d_len; a.rank = c->d_rank;
    a.gtgt = gt_ptr;
    a.gmax = gm_ptr;
    a.gmin = gn_ptr;
    a.items = w.dev ? w.d_items : d_items; a.qids = q_ptr; a.counter = c->d_counter;
    a.bound = c->d_bound; a.dist_out = d_dist; a.keys = d_keys; a.nkeys = c->d_nkeys;
    a.key_cap = key_cap; a.n_items = (int)n_items; a.bound_stride = stride; a.list_row = ws_len + 2;
    a.mode = mode; a.triangle = triangle; a.eps_max = eps_max;
    a.match = c->match; a.mismatch = c->mismatch; a.gap = c->gap;
    if (e == hipSuccess) {
      void *args[] = {&a};
      (void)hipEventRecord(c->pev[2 * pi], c->stream);
      e = hipLaunchKernel(kp, dim3(blocks), dim3(kBlock), args, 0, c->stream);
      (void)hipEventRecord(c->pev[2 * pi + 1], c->stream);
    }
    if (e == hipSuccess) e = hipGetLastError();
    if (e != hipSuccess) break;
    // the next pass reuses d_items: wait until this one has read them
    if (passes.size() > 1) e = hipStreamSynchronize(c->stream);
    if (e != hipSuccess) break;
    clk.mark(PK == KERNEL_LDS_F16    ? "pass (fp16 form)"
             : PK == KERNEL_LDS       ? "pass (int16 form)"
             : PK == KERNEL_LDS_MIXED ? "pass (fp16 and int16 forms)"
                                      : "pass");
  }
  (void)hipEventRecord(c->ev[1], c->stream);
  if (e == hipSuccess) e = hipEventSynchronize(c->ev[1]);
  float ms = 0.f;
  if (e == hipSuccess) (void)hipEventElapsedTime(&ms, c->ev[0], c->ev[1]);
  cleanup();
  if (e != hipSuccess) return MNHN_ERR_DEVICE;
  for (int f = 0; f < 3; ++f) c->last.form_cells[f] = c->last.form_ms[f] = 0;
  c->last.mixed_launch = 0;
  for (size_t pi = 0; pi < passes.size(); ++pi) {
    float pms = 0.f;
    (void)hipEventElapsedTime(&pms, c->pev[2 * pi], c->pev[2 * pi + 1]);
    if (passes[pi].kernel == KERNEL_LDS_MIXED) {  // one launch: its time goes to form 0
      c->last.form_cells[0] += passes[pi].form_cells[0];
      c->last.form_cells[1] += passes[pi].form_cells[1];
      c->last.form_ms[0] += pms;
      c->last.mixed_launch = 1;
      continue;
    }
    const int f = passes[pi].kernel == KERNEL_LDS_F16 ? 0 : passes[pi].kernel == KERNEL_I32 ? 2 : 1;
    c->last.form_cells[f] += passes[pi].cells;
    c->last.form_ms[f] += pms;
  }
  c->last.kernel_ms = ms;
  c->last.total_ms = ms;
  c->last.cells = w.cells;
  c->last.pairs = w.pairs;
  c->last.lane_bits = any_i32 ? 32 : 16;
  c->last.strip = S;
  c->last.kernel = K;
  c->last.waves = waves_used;
  c->last.items = (int)n_items_total;
  return MNHN_OK;
}

namespace {

// Work items per CU in the persistent queue (MNHN_SW_ITEMS_PER_CU, default
// 4096).  On 20,000 sequences of 100-600 residues, 256 items per CU left the
// longest-first queue's tail at ~5 % of the run (17,081 GCUPS at strip 32)
// and 4096 runs at 17,876; items stay one or more whole queries.
double items_per_cu() {
  const char *e = getenv("MNHN_SW_ITEMS_PER_CU");
  const double v = e ? atof(e) : 0.0;
  return v > 0 ? v : 4096.0;
}

// Query costs of a work list as prefix sums over its qids: P[k] = sum over
// u < k of max(len[qids[u]] - 1, 0), so that the cost of the contiguous query
// range [qb, qe) of a target group with `strips` strips of S columns is
// strips * S * (P[qe] - P[qb]) + 64 * (qe - qb) (DP columns, plus a fixed
// share per query) in O(1), instead of a walk over the range.
std::vector<double> query_prefix(const mnhn_sw_ctx *c, const Work &w) {
  std::vector<double> P(w.qids.size() + 1, 0.0);
  for (size_t u = 0; u < w.qids.size(); ++u) P[u + 1] = P[u] + (double)std::max(c->len[w.qids[u]] - 1, 0);
  return P;
}
inline double range_cost(const std::vector<double> &P, int qb, int qe, int strips, int S) {
  return (double)strips * S * (P[qe] - P[qb]) + 64.0 * (qe - qb);
}

// Split the query range [q_begin, q_end) of group g into items of similar
// cost: each item ends at the first query where its cost reaches
// target_cost (binary search on the prefix sums), the last at q_end.
void add_group_items(const mnhn_sw_ctx *c, Work &w, const std::vector<double> &P, int g, int q_begin, int q_end,
                     double target_cost, int S) {
  const int gm = c->gmax[g];
  const int strips = gm > 1 ? (gm - 1 + S - 1) / S : 0;
  int qb = q_begin;
  while (qb < q_end) {
    // first qi with cost(qb .. qi) >= target, else q_end - 1: a galloping
    // search from qb (items hold a few queries), then a binary one
    int lo = qb, hi = qb, step = 1;
    while (hi < q_end - 1 && range_cost(P, qb, hi + 1, strips, S) < target_cost) {
      lo = hi + 1;
      hi = std::min(q_end - 1, hi + step);
      step *= 2;
    }
    while (lo < hi) {
      const int mid = lo + (hi - lo) / 2;
      if (range_cost(P, qb, mid + 1, strips, S) >= target_cost) hi = mid;
      else lo = mid + 1;
    }
    w.items.push_back(make_int4(g, qb, lo + 1, 0));
    qb = lo + 1;
  }
}

double item_cost(const mnhn_sw_ctx *c, const Work &w, const std::vector<double> &P, const int4 &it, int S) {
  const int gm = w.gmax.empty() ? c->gmax[it.x] : w.gmax[it.x];
  const int strips = gm > 1 ? (gm - 1 + S - 1) / S : 0;
  return range_cost(P, it.y, it.z, strips, S);
}

// Items longest first, so the queue's last items are its smallest: a
// counting sort on 1/16-octave cost classes (stable within a class), linear
// in the item count (a comparison sort of the ~10^6 items of a large run took
// ~0.1 s of host time per call).
void finalize_items(const mnhn_sw_ctx *c, Work &w, int S, bool dense_out) {
  constexpr int kClasses = 16 * 64;
  const size_t m = w.items.size();
  std::vector<uint16_t> cls(m);
  std::vector<size_t> start(kClasses + 1, 0);
  const std::vector<double> P = query_prefix(c, w);
  // class floor(16 log2(cost)) from the exponent and a table of 2^(j / 16)
  static const double kSteps[15] = {1.0442737824274138, 1.0905077326652577, 1.1387886347566916,
                                    1.1892071150027210, 1.2418578120734840, 1.2968395546510096,
                                    1.3542555469368927, 1.4142135623730951, 1.4768261459394993,
                                    1.5422108254079407, 1.6104903319492543, 1.6817928305074290,
                                    1.7562521603732995, 1.8340080864093424, 1.9152065613971474};
  for (size_t k = 0; k < m; ++k) {
    const double cost = item_cost(c, w, P, w.items[k], S);
    int q = 0;
    if (cost > 1.0) {
      int e = 0;
      const double f = 2.0 * std::frexp(cost, &e);  // cost = f 2^(e - 1), f in [1, 2)
      int j = 0;
      while (j < 15 && f >= kSteps[j]) ++j;
      q = 16 * (e - 1) + j;
    }
    q = std::min(std::max(q, 0), kClasses - 1);
    cls[k] = (uint16_t)(kClasses - 1 - q);  // larger cost, smaller class index
    ++start[cls[k] + 1];
  }
  for (int q = 0; q < kClasses; ++q) start[q + 1] += start[q];
  std::vector<int4> sorted(m);
  for (size_t k = 0; k < m; ++k) sorted[start[cls[k]]++] = w.items[k];
  w.items.swap(sorted);
  if (dense_out) {
    int off = 0;
    for (auto &it : w.items) {
      it.w = off;
      off += (it.z - it.y) * kGroup;
    }
  }
  int mq = 0;
  for (int32_t q : w.qids) mq = std::max(mq, c->len[q]);
  w.max_q_len = mq;
}

// All unique pairs: group g holds sequences at length-order positions
// [128g, 128g+128); its queries are the positions before its last member.
// Upper-triangle work over the length-sorted order.  With eps >= 0 the pairs
// that the length bound already places beyond eps are left out: a pair's
// local score is at most match * min(la, lb), so d >= match * |la - lb|.
// Lengths fall along the order, so for target group g the queries that the
// bound removes (match * (len_p - gmax[g]) > eps) are a prefix of positions.
// With p0 > 0 only the pairs among positions p0 .. n-1 are covered: queries
// start at p0, and a target group's slots before p0 are never paired with
// them (the triangle keeps targets after the query).
Work triangle_work(const mnhn_sw_ctx *c, int S, int32_t eps, int p0 = 0) {
  Work w;
  w.qids = c->order;
  w.qids_order = true;
  std::vector<int> first(c->groups, p0);
  if (eps >= 0) {
    for (int g = 0; g < c->groups; ++g) {
      const long long gm = c->gmax[g];
      int lo = 0, hi = std::min((g + 1) * kGroup, c->n);
      while (lo < hi) {
        const int mid = (lo + hi) / 2;
        if ((long long)c->match * ((long long)c->len[c->order[mid]] - gm) > (long long)eps) lo = mid + 1;
        else hi = mid;
      }
      first[g] = std::max(lo, p0);
    }
  }
  const std::vector<double> P = query_prefix(c, w);  // qids = positions in length order
  double total = 0;
  for (int g = 0; g < c->groups; ++g) {
    const int last = std::min((g + 1) * kGroup, c->n) - 1;
    const int gm = c->gmax[g];
    const int strips = gm > 1 ? (gm - 1 + S - 1) / S : 0;
    if (last > first[g]) total += (double)strips * S * (P[last] - P[first[g]]);
  }
  const double target = std::max(total / ((double)std::max(c->num_cu, 1) * items_per_cu()), 2.0e5);
  w.items.reserve((size_t)(total / target * 1.5) + 2 * (size_t)c->groups + 16);
  for (int g = 0; g < c->groups; ++g) {
    const int last = std::min((g + 1) * kGroup, c->n) - 1;
    if (last <= first[g]) continue;
    add_group_items(c, w, P, g, first[g], last, target, S);
  }
  // algorithmic cells and pairs of the unique pairs the items cover
  std::vector<double> prefix((size_t)c->n + 1, 0.0);
  for (int p = 0; p < c->n; ++p) prefix[p + 1] = prefix[p] + c->len[c->order[p]];
  double cells = 0, pairs = 0;
  for (int t = p0; t < c->n; ++t) {
    const int f = std::min(first[t / kGroup], t);
    cells += (double)c->len[c->order[t]] * (prefix[t] - prefix[f]);
    pairs += (double)(t - f);
  }
  w.cells = cells;
  w.pairs = pairs;
  return w;
}


}  // namespace

// Epsilon-list keys carry d in 16 bits, so the lists hold d <= 65535.  Every
// distance is at most match * max(la, lb): when that bound is within 16 bits
// a larger eps_max lists the same pairs as 65535 and is lowered to it;
// otherwise it is refused (a per-query pass, mnhn_sw_one_vs_all, returns any
// distance in int32).
int mnhn_detail::list_bound(const mnhn_sw_ctx *c, int32_t eps_max, int32_t *eff) {
  *eff = eps_max;
  if (eps_max <= kI16ScoreMax) return MNHN_OK;
  if ((long long)c->match * c->maxlen > kI16ScoreMax) return MNHN_ERR_RANGE;
  *eff = (int32_t)kI16ScoreMax;
  return MNHN_OK;
}

extern "C" {

const char *mnhn_sw_strerror(int code) {
  switch (code) {
    case MNHN_OK: return "ok";
    case MNHN_ERR_ARG: return "invalid argument";
    case MNHN_ERR_DEVICE: return "HIP runtime error";
    case MNHN_ERR_NOMEM: return "out of memory";
    case MNHN_ERR_NODEVICE: return "no usable GPU";
    case MNHN_ERR_RANGE: return "epsilon above the 16-bit list range for these lengths, or work exceeds 32-bit offsets";
    default: return "unknown error";
  }
}

int mnhn_sw_create(const char *const *seqs, const size_t *lens, int n,
                   const mnhn_sw_params *params, mnhn_sw_ctx **out) {
  if (!out || n < 0 || (n > 0 && (!seqs || !lens))) return MNHN_ERR_ARG;
  *out = nullptr;
  if (n >= (1 << 24)) return MNHN_ERR_ARG;  // key format: 24-bit ids
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) return MNHN_ERR_NODEVICE;
  mnhn_sw_ctx *c = new (std::nothrow) mnhn_sw_ctx();
  if (!c) return MNHN_ERR_NOMEM;
  if (params) {
    if (params->match <= 0 || params->mismatch > 0 || params->gap < 0 ||
        params->match > 1000 || params->mismatch < -1000 || params->gap > 1000) {
      delete c;
      return MNHN_ERR_ARG;
    }
    c->match = params->match;
    c->mismatch = params->mismatch;
    c->gap = params->gap;
  }
  c->n = n;
  c->len.resize(n);
  for (int i = 0; i < n; ++i) {
    if (lens[i] > (size_t)INT32_MAX / 2) { delete c; return MNHN_ERR_ARG; }
    c->len[i] = (int32_t)lens[i];
    c->maxlen = std::max(c->maxlen, c->len[i]);
  }
  // FNV-1a over the scores, the count and every sequence's length and bytes
  {
    uint64_t h = 1469598103934665603ull;
    auto mix = [&h](const void *p, size_t k) {
      const unsigned char *b = (const unsigned char *)p;
      for (size_t t = 0; t < k; ++t) { h ^= b[t]; h *= 1099511628211ull; }
    };
    const int32_t hdr[4] = {c->match, c->mismatch, c->gap, n};
    mix(hdr, sizeof(hdr));
    for (int i = 0; i < n; ++i) {
      const uint64_t L = (uint64_t)lens[i];
      mix(&L, sizeof(L));
      if (L) mix(seqs[i], (size_t)L);
    }
    c->data_hash = h;
  }
  // alphabet of the compared residues (index >= 1)
  bool seen[256] = {false};
  int distinct = 0;
  for (int i = 0; i < n; ++i)
    for (int32_t k = 1; k < c->len[i]; ++k) {
      unsigned char b = (unsigned char)seqs[i][k];
      if (!seen[b]) { seen[b] = true; ++distinct; }
    }
  c->acgt = distinct <= 4;
  uint8_t code_of[256];
  memset(code_of, 0, sizeof(code_of));
  {
    int next = 0;
    for (int b = 0; b < 256; ++b)
      if (seen[b]) code_of[b] = (uint8_t)(next++);
  }
  // pack residues 1..L-1; each sequence padded with one spare word so that
  // strip loads past the end stay inside the buffer
  const int per_word = c->acgt ? 16 : 2;
  c->woff.resize(n);
  size_t total_words = 0;
  for (int i = 0; i < n; ++i) {
    c->woff[i] = (uint32_t)total_words;
    const size_t res = c->len[i] > 1 ? (size_t)c->len[i] - 1 : 0;
    total_words += (res + per_word - 1) / per_word;
  }
  // strip loads of a short target run up to the longest length of its group
  total_words += (size_t)(c->maxlen + per_word) / per_word + 64;
  if (total_words >= 0xffffffffull) { delete c; return MNHN_ERR_ARG; }
  std::vector<uint32_t> words(total_words, 0u);
  for (int i = 0; i < n; ++i) {
    uint32_t *w = words.data() + c->woff[i];
    for (int32_t k = 1; k < c->len[i]; ++k) {
      const unsigned char b = (unsigned char)seqs[i][k];
      const size_t r = (size_t)k - 1;
      if (c->acgt) w[r / 16] |= (uint32_t)code_of[b] << (2 * (r % 16));
      else w[r / 2] |= (uint32_t)b << (16 * (r % 2));
    }
  }
  // length order (longest first) and target groups of 128
  c->order.resize(n);
  std::iota(c->order.begin(), c->order.end(), 0);
  std::stable_sort(c->order.begin(), c->order.end(),
                   [&](int x, int y) { return c->len[x] > c->len[y]; });
  c->rank.resize(n);
  for (int p = 0; p < n; ++p) c->rank[c->order[p]] = p;
  c->groups = (n + kGroup - 1) / kGroup;
  c->gtgt.assign((size_t)c->groups * kGroup, -1);
  c->gmax.assign(c->groups, 0);
  c->gmin.assign(c->groups, INT32_MAX);
  for (int p = 0; p < n; ++p) {
    const int g = p / kGroup, t = p % kGroup;
    // slot t: lane t for t < 64 (low half), lane t-64 (high half) otherwise
    c->gtgt[(size_t)g * kGroup + t] = c->order[p];
    c->gmax[g] = std::max(c->gmax[g], c->len[c->order[p]]);
    c->gmin[g] = std::min(c->gmin[g], c->len[c->order[p]]);
  }
  c->gsuf.assign((size_t)c->groups * (kGroup + 1), 0.0);
  for (int g = 0; g < c->groups; ++g)
    for (int t = kGroup - 1; t >= 0; --t) {
      const int id = c->gtgt[(size_t)g * kGroup + t];
      c->gsuf[(size_t)g * (kGroup + 1) + t] =
          c->gsuf[(size_t)g * (kGroup + 1) + t + 1] + (id >= 0 ? (double)c->len[id] : 0.0);
    }

  int rc = MNHN_OK;
  hipDeviceProp_t prop;
  if (hipGetDevice(&c->device) != hipSuccess ||
      hipGetDeviceProperties(&prop, c->device) != hipSuccess) {
    delete c;
    return MNHN_ERR_DEVICE;
  }
  c->num_cu = prop.multiProcessorCount;
  // scratch of the allocation-free entry point on this (current) device
  c->scratch = scratch_address(c->device);
  if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) {
    delete c;
    return MNHN_ERR_DEVICE;
  }
  for (auto &e : c->ev)
    if (hipEventCreate(&e) != hipSuccess) { mnhn_sw_destroy(c); return MNHN_ERR_DEVICE; }
  if ((rc = dev_alloc(&c->d_words, words.size())) || (rc = dev_alloc(&c->d_woff, (size_t)n)) ||
      (rc = dev_alloc(&c->d_len, (size_t)n)) || (rc = dev_alloc(&c->d_rank, (size_t)n)) ||
      (rc = dev_alloc(&c->d_gtgt, c->gtgt.size())) || (rc = dev_alloc(&c->d_gmax, c->gmax.size())) ||
      (rc = dev_alloc(&c->d_gmin, c->gmin.size())) ||
      (rc = dev_alloc(&c->d_counter, 1)) || (rc = dev_alloc(&c->d_nkeys, 1))) {
    mnhn_sw_destroy(c);
    return rc;
  }
  hipError_t e = hipMemcpy(c->d_words, words.data(), words.size() * 4, hipMemcpyHostToDevice);
  if (e == hipSuccess && n) e = hipMemcpy(c->d_woff, c->woff.data(), (size_t)n * 4, hipMemcpyHostToDevice);
  if (e == hipSuccess && n) e = hipMemcpy(c->d_len, c->len.data(), (size_t)n * 4, hipMemcpyHostToDevice);
  if (e == hipSuccess && n) e = hipMemcpy(c->d_rank, c->rank.data(), (size_t)n * 4, hipMemcpyHostToDevice);
  if (e == hipSuccess && c->groups)
    e = hipMemcpy(c->d_gtgt, c->gtgt.data(), c->gtgt.size() * 4, hipMemcpyHostToDevice);
  if (e == hipSuccess && c->groups)
    e = hipMemcpy(c->d_gmax, c->gmax.data(), c->gmax.size() * 4, hipMemcpyHostToDevice);
  if (e == hipSuccess && c->groups)
    e = hipMemcpy(c->d_gmin, c->gmin.data(), c->gmin.size() * 4, hipMemcpyHostToDevice);
  if (e != hipSuccess) { mnhn_sw_destroy(c); return MNHN_ERR_DEVICE; }
  *out = c;
  return MNHN_OK;
}

void mnhn_sw_destroy(mnhn_sw_ctx *c) {
  if (!c) return;
  if (c->stream) (void)hipStreamSynchronize(c->stream);
  dev_free(c->d_words); dev_free(c->d_woff); dev_free(c->d_len); dev_free(c->d_rank);
  dev_free(c->d_gtgt); dev_free(c->d_gmax); dev_free(c->d_gmin); dev_free(c->d_counter);
  dev_free(c->d_nkeys);
  dev_free(c->d_bound); dev_free(c->d_keys);
  for (auto &e : c->ev)
    if (e) (void)hipEventDestroy(e);
  for (auto &e : c->pev)
    if (e) (void)hipEventDestroy(e);
  stage_free(c->stage);
  if (c->h_items) (void)hipHostFree(c->h_items);
  if (c->stream) (void)hipStreamDestroy(c->stream);
  delete c;
}

int mnhn_sw_size(const mnhn_sw_ctx *c) { return c ? c->n : -1; }
uint64_t mnhn_sw_data_hash(const mnhn_sw_ctx *c) { return c ? c->data_hash : 0; }
int mnhn_sw_alphabet(const mnhn_sw_ctx *c) { return c ? (c->acgt ? 4 : 0) : -1; }

int mnhn_sw_last_timing(const mnhn_sw_ctx *c, mnhn_sw_timing *out) {
  if (!c || !out) return MNHN_ERR_ARG;
  *out = c->last;
  return MNHN_OK;
}

int mnhn_sw_set_tuning(mnhn_sw_ctx *c, int kernel, int strip, int waves_per_cu) {
  if (!c || kernel < 0 || kernel > 3 || (strip != 0 && strip != 16 && strip != 32 && strip != 64) ||
      waves_per_cu < 0)
    return MNHN_ERR_ARG;
  c->kernel = kernel;
  c->strip = strip;
  c->waves_per_cu = waves_per_cu;
  return MNHN_OK;
}

int mnhn_sw_one_vs_all(mnhn_sw_ctx *c, int query, int32_t *dist_out) {
  if (!c || !dist_out || query < 0 || query >= c->n) return MNHN_ERR_ARG;
  int rc = MNHN_OK;
  const int S = pick_strip(c);
  Work w;
  w.qids.push_back(query);
  for (int g = 0; g < c->groups; ++g) w.items.push_back(make_int4(g, 0, 1, 0));
  for (int j = 0; j < c->n; ++j) w.cells += (double)c->len[query] * c->len[j];
  w.pairs = c->n;
  finalize_items(c, w, S, true);
  const size_t out_n = w.items.size() * kGroup;
  int32_t *d_out = nullptr;
  if ((rc = dev_alloc(&d_out, out_n))) return rc;
  rc = launch_tiles(c, w, MODE_DIST, 0, 0, d_out, nullptr, 0);
  if (rc == MNHN_OK) {
    std::vector<int32_t> h(out_n);
    if (hipMemcpy(h.data(), d_out, out_n * 4, hipMemcpyDeviceToHost) != hipSuccess) rc = MNHN_ERR_DEVICE;
    else
      for (const auto &it : w.items)
        for (int t = 0; t < kGroup; ++t) {
          const int id = c->gtgt[(size_t)it.x * kGroup + t];
          if (id >= 0) dist_out[id] = h[(size_t)it.w + t];
        }
  }
  dev_free(d_out);
  return rc;
}

int mnhn_sw_one_vs_all_noalloc(mnhn_sw_ctx *c, int query, int32_t *dist_out) {
  if (!c || !dist_out || query < 0 || query >= c->n) return MNHN_ERR_ARG;
  if (!c->scratch) return MNHN_ERR_DEVICE;
  // workspace rows: target residues (LDS kernels) or query residues (register
  // and 32-bit kernels), at most the longest sequence, plus the fp16 kernel's
  // row-address list; one launch needs at least one workgroup's four
  // workspaces in the scratch, checked for the largest of them
  const size_t ws_bytes = kScratchBytes - kScratchHead;
  const size_t wave_bytes_max = (size_t)wave_rows(c->maxlen, KERNEL_LDS_F16) * kWave * sizeof(uint32_t);
  if (ws_bytes / wave_bytes_max < (size_t)(kBlock / kWave)) return MNHN_ERR_RANGE;
  const int la = c->len[query];
  const int K = pick_kernel(c);
  int S = pick_strip(c);
  if ((K == KERNEL_LDS || K == KERNEL_LDS_F16) && S == 64 && la <= 128) S = 32;
  // items are the target groups in length order (longest first), the query
  // against each; a group's score bound match * min(gmax, la) falls along
  // the order, so the groups needing 32-bit lanes are a prefix and (fp16
  // form) those within 2048 a suffix
  const int G = c->groups;
  auto first_within = [&](int from, long long limit) {
    int lo = from, hi = G;
    while (lo < hi) {
      const int mid = lo + (hi - lo) / 2;
      if ((long long)c->match * std::min(c->gmax[mid], la) > limit) lo = mid + 1;
      else hi = mid;
    }
    return lo;
  };
  const int gA = first_within(0, kI16ScoreMax);
  const int gB = K == KERNEL_LDS_F16 ? first_within(gA, 2048) : G;
  struct {
    int kernel, strip, g0, g1;
  } const passes[3] = {{KERNEL_I32, kI32Strip, 0, gA},
                       {K == KERNEL_LDS_F16 ? KERNEL_LDS : K, S, gA, gB},
                       {KERNEL_LDS_F16, S, gB, G}};
  int *counters = reinterpret_cast<int *>(c->scratch);
  uint32_t *ws = reinterpret_cast<uint32_t *>(static_cast<unsigned char *>(c->scratch) + kScratchHead);
  hipLaunchKernelGGL(zero_words_kernel, dim3(1), dim3(kWave), 0, c->stream, counters, 3);
  hipError_t e = hipGetLastError();
  for (int k = 0; k < 3 && e == hipSuccess; ++k) {
    const int n_items = passes[k].g1 - passes[k].g0;
    if (n_items <= 0) continue;
    void *kp = kernel_ptr(passes[k].kernel, passes[k].strip, c->acgt, true);
    const int stride = wave_rows(c->maxlen, passes[k].kernel) * kWave;
    const size_t ws_waves = ws_bytes / ((size_t)stride * sizeof(uint32_t));
    const int wpc = launch_waves_per_cu(c, passes[k].kernel, passes[k].strip, kp);
    int blocks = (c->num_cu * wpc + (kBlock / kWave) - 1) / (kBlock / kWave);
    blocks = std::min(blocks, (n_items + 3) / 4);
    blocks = std::max(1, (int)std::min<size_t>((size_t)blocks, ws_waves / (kBlock / kWave)));
    KArgs a{};
    a.words = c->d_words; a.woff = c->d_woff; a.len = c->d_len; a.rank = c->d_rank;
    a.gtgt = c->d_gtgt; a.gmax = c->d_gmax; a.gmin = c->d_gmin;
    a.counter = counters + k;
    a.bound = ws; a.bound_stride = stride; a.list_row = c->maxlen + 2;
    a.dist_out = dist_out;
    a.n_items = n_items;
    a.item_base = passes[k].g0; a.query = query;
    a.mode = MODE_DIST; a.triangle = 0; a.eps_max = 0;
    a.match = c->match; a.mismatch = c->mismatch; a.gap = c->gap;
    void *args[] = {&a};
    e = hipLaunchKernel(kp, dim3(blocks), dim3(kBlock), args, 0, c->stream);
    if (e == hipSuccess) e = hipGetLastError();
  }
  if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
  return e == hipSuccess ? MNHN_OK : MNHN_ERR_DEVICE;
}

int mnhn_sw_block(mnhn_sw_ctx *c, const int32_t *rows, int nr, const int32_t *cols, int nc,
                  int32_t *dist_out) {
  if (!c || nr < 0 || nc < 0 || (nr > 0 && !rows) || (nc > 0 && !cols) || (nr > 0 && nc > 0 && !dist_out))
    return MNHN_ERR_ARG;
  for (int r = 0; r < nr; ++r)
    if (rows[r] < 0 || rows[r] >= c->n) return MNHN_ERR_ARG;
  for (int k = 0; k < nc; ++k)
    if (cols[k] < 0 || cols[k] >= c->n) return MNHN_ERR_ARG;
  if (nr == 0 || nc == 0) return MNHN_OK;
  int rc = MNHN_OK;
  const int S = pick_strip(c);
  // targets: the columns by length (longest first) in groups of 128; queries:
  // the rows by length (longest first)
  std::vector<int> cperm(nc), rperm(nr);
  std::iota(cperm.begin(), cperm.end(), 0);
  std::iota(rperm.begin(), rperm.end(), 0);
  std::stable_sort(cperm.begin(), cperm.end(), [&](int x, int y) { return c->len[cols[x]] > c->len[cols[y]]; });
  std::stable_sort(rperm.begin(), rperm.end(), [&](int x, int y) { return c->len[rows[x]] > c->len[rows[y]]; });
  const int G = (nc + kGroup - 1) / kGroup;
  if ((size_t)G * kGroup * (size_t)nr > (size_t)INT32_MAX) return MNHN_ERR_RANGE;  // 32-bit output offsets
  Work w;
  w.gtgt.assign((size_t)G * kGroup, -1);
  w.gmax.assign(G, 0);
  w.gmin.assign(G, INT32_MAX);
  for (int k = 0; k < nc; ++k) {
    const int g = k / kGroup, id = cols[cperm[k]];
    w.gtgt[(size_t)g * kGroup + (k % kGroup)] = id;
    w.gmax[g] = std::max(w.gmax[g], c->len[id]);
    w.gmin[g] = std::min(w.gmin[g], c->len[id]);
  }
  for (int r = 0; r < nr; ++r) w.qids.push_back(rows[rperm[r]]);
  double total = 0, sum_lb = 0, sum_la = 0;
  for (int g = 0; g < G; ++g) {
    const int strips = w.gmax[g] > 1 ? (w.gmax[g] - 1 + S - 1) / S : 0;
    for (int r = 0; r < nr; ++r) total += (double)std::max(c->len[w.qids[r]] - 1, 0) * strips * S;
  }
  for (int k = 0; k < nc; ++k) sum_lb += c->len[cols[k]];
  for (int r = 0; r < nr; ++r) sum_la += c->len[rows[r]];
  const double target = std::max(total / ((double)std::max(c->num_cu, 1) * items_per_cu()), 2.0e5);
  for (int g = 0; g < G; ++g) {
    const int strips = w.gmax[g] > 1 ? (w.gmax[g] - 1 + S - 1) / S : 0;
    int qb = 0;
    double acc = 0;
    for (int qi = 0; qi < nr; ++qi) {
      acc += (double)std::max(c->len[w.qids[qi]] - 1, 0) * strips * S + 64.0;
      if (acc >= target || qi + 1 == nr) {
        w.items.push_back(make_int4(g, qb, qi + 1, 0));
        qb = qi + 1;
        acc = 0;
      }
    }
  }
  w.cells = sum_la * sum_lb;
  w.pairs = (double)nr * nc;
  finalize_items(c, w, S, true);
  const size_t out_n = (size_t)G * kGroup * nr;
  int32_t *d_out = nullptr;
  if ((rc = dev_alloc(&d_out, out_n))) return rc;
  rc = launch_tiles(c, w, MODE_DIST, 0, 0, d_out, nullptr, 0);
  if (rc == MNHN_OK) {
    std::vector<int32_t> h(out_n);
    if (copy_d2h(c, h.data(), d_out, out_n * sizeof(int32_t)) != MNHN_OK) {
      rc = MNHN_ERR_DEVICE;
    } else {
      for (const auto &it : w.items)
        for (int qi = it.y; qi < it.z; ++qi) {
          int32_t *orow = dist_out + (size_t)rperm[qi] * nc;
          const int32_t *hrow = h.data() + (size_t)it.w + (size_t)(qi - it.y) * kGroup;
          for (int t = 0; t < kGroup; ++t) {
            const int k = it.x * kGroup + t;
            if (k < nc) orow[cperm[k]] = hrow[t];
          }
        }
    }
  }
  dev_free(d_out);
  return rc;
}

unsigned int mnhn_score(const char *A, const char *B, size_t len_a, size_t len_b, int *work) {
  (void)work;
  const char *seqs[2] = {A, B};
  const size_t lens[2] = {len_a, len_b};
  const int32_t ij[2] = {0, 1};
  int32_t d = 0;
  mnhn_sw_ctx *c = nullptr;
  int rc = mnhn_sw_create(seqs, lens, 2, nullptr, &c);
  if (rc == MNHN_OK) rc = mnhn_sw_pairs(c, ij, 1, &d);
  mnhn_sw_destroy(c);
  if (rc != MNHN_OK) {
    printf("mnhn_score failed: %s\n", mnhn_sw_strerror(rc));
    fflush(stdout);
    _exit(1);
  }
  return (unsigned int)d;
}

int mnhn_sw_pairs(mnhn_sw_ctx *c, const int32_t *ij, int64_t m, int32_t *dist_out) {
  if (!c || m < 0 || (m > 0 && (!ij || !dist_out))) return MNHN_ERR_ARG;
  for (int64_t k = 0; k < 2 * m; ++k)
    if (ij[k] < 0 || ij[k] >= c->n) return MNHN_ERR_ARG;
  if (m == 0) return MNHN_OK;
  int rc = MNHN_OK;
  const int S = pick_strip(c);
  // group pairs by first index; the first index is the wave-uniform query
  std::vector<int64_t> idx(m);
  std::iota(idx.begin(), idx.end(), 0);
  std::stable_sort(idx.begin(), idx.end(), [&](int64_t x, int64_t y) {
    if (ij[2 * x] != ij[2 * y]) return ij[2 * x] < ij[2 * y];
    return c->len[ij[2 * x + 1]] > c->len[ij[2 * y + 1]];
  });
  Work w;
  std::vector<int64_t> slot_pair;  // pair index of each (group, slot)
  for (int64_t k = 0; k < m;) {
    const int q = ij[2 * idx[k]];
    int64_t e = k;
    while (e < m && ij[2 * idx[e]] == q) ++e;
    w.qids.push_back(q);
    const int qi = (int)w.qids.size() - 1;
    for (int64_t s = k; s < e; s += kGroup) {
      const int g = (int)w.gmax.size();
      int gm = 0, gn = INT32_MAX;
      for (int t = 0; t < kGroup; ++t) {
        const int64_t p = s + t;
        const int id = p < e ? ij[2 * idx[p] + 1] : -1;
        w.gtgt.push_back(id);
        slot_pair.push_back(p < e ? idx[p] : -1);
        if (id >= 0) {
          gm = std::max(gm, c->len[id]);
          gn = std::min(gn, c->len[id]);
        }
      }
      w.gmax.push_back(gm);
      w.gmin.push_back(gn);
      w.items.push_back(make_int4(g, qi, qi + 1, g * kGroup));
    }
    k = e;
  }
  if (w.gmax.size() * (size_t)kGroup > (size_t)INT32_MAX) return MNHN_ERR_RANGE;  // 32-bit output offsets
  for (int64_t k = 0; k < m; ++k) w.cells += (double)c->len[ij[2 * k]] * c->len[ij[2 * k + 1]];
  w.pairs = (double)m;
  finalize_items(c, w, S, false);
  const size_t out_n = w.gmax.size() * kGroup;
  int32_t *d_out = nullptr;
  if ((rc = dev_alloc(&d_out, out_n))) return rc;
  rc = launch_tiles(c, w, MODE_DIST, 0, 0, d_out, nullptr, 0);
  if (rc == MNHN_OK) {
    std::vector<int32_t> h(out_n);
    if (hipMemcpy(h.data(), d_out, out_n * 4, hipMemcpyDeviceToHost) != hipSuccess) rc = MNHN_ERR_DEVICE;
    else
      for (size_t s = 0; s < out_n; ++s)
        if (slot_pair[s] >= 0) dist_out[slot_pair[s]] = h[s];
  }
  dev_free(d_out);
  return rc;
}

int mnhn_sw_matrix(mnhn_sw_ctx *c, int32_t *dist_out) {
  if (!c || (c->n > 0 && !dist_out)) return MNHN_ERR_ARG;
  const int n = c->n;
  if (n == 0) return MNHN_OK;
  int rc = MNHN_OK;
  const int S = pick_strip(c);
  Work w = triangle_work(c, S, -1);
  finalize_items(c, w, S, true);
  size_t out_n = 0;
  for (const auto &it : w.items) out_n += (size_t)(it.z - it.y) * kGroup;
  if (out_n > (size_t)INT32_MAX) return MNHN_ERR_RANGE;  // item output offsets are 32-bit
  int32_t *d_out = nullptr;
  if ((rc = dev_alloc(&d_out, out_n))) return rc;
  rc = launch_tiles(c, w, MODE_DIST, 1, 0, d_out, nullptr, 0);
  if (rc == MNHN_OK) {
    std::vector<int32_t> h(out_n);
    if (hipMemcpy(h.data(), d_out, out_n * 4, hipMemcpyDeviceToHost) != hipSuccess) {
      rc = MNHN_ERR_DEVICE;
    } else {
      for (int i = 0; i < n; ++i) dist_out[(size_t)i * n + i] = 0;
      for (const auto &it : w.items)
        for (int qi = it.y; qi < it.z; ++qi) {
          const int q = w.qids[qi];
          for (int t = 0; t < kGroup; ++t) {
            const int id = c->gtgt[(size_t)it.x * kGroup + t];
            if (id < 0 || c->rank[q] >= c->rank[id]) continue;
            const int32_t d = h[(size_t)it.w + (size_t)(qi - it.y) * kGroup + t];
            dist_out[(size_t)q * n + id] = d;
            dist_out[(size_t)id * n + q] = d;
          }
        }
    }
  }
  dev_free(d_out);
  return rc;
}

}  // extern "C"

namespace {

// Run the all-pairs epsilon pass; on return the context's d_keys holds nkeys
// unsorted keys (two per unordered pair with d <= eps_max).
// MNHN_SW_LENGTH_PRUNE=0 keeps every unique pair in the epsilon passes (the
// bound is exact, so the lists are the same either way).
bool length_prune() {
  const char *e = getenv("MNHN_SW_LENGTH_PRUNE");
  return !(e && e[0] == '0');
}

// Largest number of list entries (keys) a pass may hold; more returns
// MNHN_ERR_NOMEM (the DBSCAN host then recomputes per region query).
// MNHN_SW_LIST_MAX_ENTRIES sets it; by default only device and host memory
// bound the lists.
unsigned long long list_entry_cap() {
  const char *e = getenv("MNHN_SW_LIST_MAX_ENTRIES");
  if (e && *e) return strtoull(e, nullptr, 10);
  return ~0ull;
}

int eps_pass(mnhn_sw_ctx *c, int32_t eps_req, unsigned long long *nkeys_out) {
  int32_t eps_max = eps_req;
  int rc = list_bound(c, eps_req, &eps_max);
  if (rc) return rc;
  mnhn_phase_clock clk(c->stream, "eps pass");
  const int S = pick_strip(c);
  Work w = triangle_work(c, S, length_prune() ? eps_max : -1);
  finalize_items(c, w, S, false);
  clk.mark("work list");
  const unsigned long long all = (unsigned long long)c->n * (unsigned long long)(c->n > 0 ? c->n - 1 : 0);
  unsigned long long cap = c->keys_cap;
  if (cap == 0) {
    const char *e = getenv("MNHN_SW_KEY_BUDGET_GB");
    const double gb = e ? atof(e) : 32.0;
    const unsigned long long budget = (unsigned long long)(gb * 1e9 / 8.0);
    cap = std::max<unsigned long long>(std::min(all, budget), 1024ull);
    if (all > budget && w.items.size() >= 4096) {
      // the lists may not fit the default buffer: size it from a systematic
      // sample of 0.5 % of the items (keys per cell, times 1.3; items are in
      // cost order, so the sample spans every length class) instead of
      // overflowing and computing every pair twice
      Work smp;
      smp.qids = w.qids;
      const size_t stride = 200;
      for (size_t k = 0; k < w.items.size(); k += stride) smp.items.push_back(w.items[k]);
      smp.max_q_len = w.max_q_len;
      double smp_cost = 0, all_cost = 0;
      const std::vector<double> P = query_prefix(c, w);
      for (size_t k = 0; k < w.items.size(); ++k) {
        const double ck = item_cost(c, w, P, w.items[k], S);
        all_cost += ck;
        if (k % stride == 0) smp_cost += ck;
      }
      if ((rc = ensure_keys(c, cap))) return rc;
      HIP_TRY(hipMemsetAsync(c->d_nkeys, 0, sizeof(unsigned long long), c->stream));
      rc = launch_tiles(c, smp, MODE_EPS, 1, eps_max, nullptr, c->d_keys, c->keys_cap);
      if (rc) return rc;
      unsigned long long sk = 0;
      HIP_TRY(hipMemcpy(&sk, c->d_nkeys, sizeof(sk), hipMemcpyDeviceToHost));
      const double est = (double)sk * all_cost / std::max(smp_cost, 1.0) * 1.3;
      if (est > (double)cap) cap = std::min<unsigned long long>(all, (unsigned long long)est);
    }
  }
  const unsigned long long max_entries = list_entry_cap();
  for (int attempt = 0; attempt < 2; ++attempt) {
    cap = std::max<unsigned long long>(std::min(cap, max_entries), 1);
    if ((rc = ensure_keys(c, cap))) return rc;
    HIP_TRY(hipMemsetAsync(c->d_nkeys, 0, sizeof(unsigned long long), c->stream));
    clk.mark("key buffer");
    rc = launch_tiles(c, w, MODE_EPS, 1, eps_max, nullptr, c->d_keys, c->keys_cap);
    if (rc) return rc;
    clk.mark("launch_tiles");
    unsigned long long nk = 0;
    HIP_TRY(hipMemcpy(&nk, c->d_nkeys, sizeof(nk), hipMemcpyDeviceToHost));
    if (nk <= c->keys_cap) {
      *nkeys_out = nk;
      return MNHN_OK;
    }
    if (nk > max_entries) return MNHN_ERR_NOMEM;
    cap = nk;  // overflowed: grow to the exact size and recompute once
  }
  return MNHN_ERR_NOMEM;
}

}  // namespace

extern "C" {

int mnhn_sw_eps_count(mnhn_sw_ctx *c, int32_t eps_max, int64_t *nnz_out) {
  if (!c || !nnz_out) return MNHN_ERR_ARG;
  unsigned long long nk = 0;
  int rc = eps_pass(c, eps_max, &nk);
  if (rc == MNHN_OK) *nnz_out = (int64_t)nk;
  return rc;
}

int mnhn_sw_eps_csr(mnhn_sw_ctx *c, int32_t eps_max, mnhn_csr *out) {
  if (!c || !out) return MNHN_ERR_ARG;
  memset(out, 0, sizeof(*out));
  const int n = c->n;
  out->n_rows = n;
  out->eps_max = eps_max;
  out->row_ptr = (int64_t *)calloc((size_t)n + 1, sizeof(int64_t));
  if (!out->row_ptr) return MNHN_ERR_NOMEM;
  if (n == 0) return MNHN_OK;
  unsigned long long nk = 0;
  int rc = eps_pass(c, eps_max, &nk);
  if (rc) { mnhn_csr_free(out); return rc; }
  return csr_from_keys(c, nk, out);
}

}  // extern "C"

Work mnhn_detail::suffix_work(const mnhn_sw_ctx *c, int32_t eps, int p0) {
  const int S = pick_strip(c);
  Work w = triangle_work(c, S, length_prune() ? eps : -1, std::max(0, std::min(p0, c->n)));
  finalize_items(c, w, S, false);
  return w;
}

int mnhn_detail::eps_keys_estimate(mnhn_sw_ctx *c, const Work &w, int32_t eps, unsigned long long *est) {
  *est = (unsigned long long)(2.0 * w.pairs);
  if (w.items.size() < 4096) return MNHN_OK;
  // items are cut to a common cost (add_group_items), so the sample stands
  // for the work by its share of the items
  Work smp;
  smp.qids = w.qids;
  smp.max_q_len = w.max_q_len;
  const size_t stride = 200;
  for (size_t k = 0; k < w.items.size(); k += stride) smp.items.push_back(w.items[k]);
  // the sample's keys are at most twice its pairs
  double smp_pairs = 0;
  for (const int4 &it : smp.items) smp_pairs += (double)(it.z - it.y) * kGroup;
  int rc = ensure_keys(c, std::max<unsigned long long>((unsigned long long)(2.0 * smp_pairs), 1024ull));
  if (rc) return rc;
  HIP_TRY(hipMemsetAsync(c->d_nkeys, 0, sizeof(unsigned long long), c->stream));
  if ((rc = launch_tiles(c, smp, MODE_EPS, 1, eps, nullptr, c->d_keys, c->keys_cap))) return rc;
  unsigned long long sk = 0;
  HIP_TRY(hipMemcpy(&sk, c->d_nkeys, sizeof(sk), hipMemcpyDeviceToHost));
  const double e = (double)sk * (double)w.items.size() / (double)smp.items.size() * 1.3;
  *est = std::min(*est, (unsigned long long)e + 1024ull);
  return MNHN_OK;
}

int mnhn_detail::ensure_keys(mnhn_sw_ctx *c, unsigned long long cap) {
  if (cap <= c->keys_cap && c->d_keys) return MNHN_OK;
  dev_free(c->d_keys);
  c->keys_cap = 0;
  if (hipMalloc((void **)&c->d_keys, std::max<unsigned long long>(cap, 1) * sizeof(unsigned long long)) !=
      hipSuccess) {
    c->d_keys = nullptr;
    return MNHN_ERR_NOMEM;
  }
  c->keys_cap = cap;
  return MNHN_OK;
}

// out->row_ptr must be allocated ([n+1], zeroed) and out->n_rows set.
namespace {

// Radix-sort form of the key -> CSR step, for any n (the row-ranking form
// needs an LDS bitmap of n bits): keys row<<40 | col<<16 | d become
// row<<24 | col with d as payload, are sorted over their 24 + b significant
// bits (n < 2^b), checked on the device, and cut into rows by binary search.
__global__ void split_keys_kernel(const unsigned long long *__restrict__ keys, long long nk,
                                  unsigned long long *__restrict__ k_out, uint16_t *__restrict__ d_out) {
  for (long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x; k < nk;
       k += (long long)gridDim.x * blockDim.x) {
    const unsigned long long key = keys[k];
    k_out[k] = key >> 16;
    d_out[k] = (uint16_t)(key & 0xffffull);
  }
}

__global__ void sorted_check_kernel(const unsigned long long *__restrict__ keys, long long nk,
                                    unsigned long long *__restrict__ bad) {
  for (long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x; k + 1 < nk;
       k += (long long)gridDim.x * blockDim.x)
    if (keys[k] >= keys[k + 1]) atomicOr(bad, 1ull);
}

__global__ void cut_rows_kernel(const unsigned long long *__restrict__ keys, long long nk, int n,
                                long long *__restrict__ row_ptr) {
  for (long long r = (long long)blockIdx.x * blockDim.x + threadIdx.x; r <= n;
       r += (long long)gridDim.x * blockDim.x) {
    const unsigned long long want = (unsigned long long)r << 24;
    long long lo = 0, hi = nk;
    while (lo < hi) {
      const long long mid = (lo + hi) >> 1;
      if (keys[mid] < want) lo = mid + 1;
      else hi = mid;
    }
    row_ptr[r] = lo;
  }
}

__global__ void low_col_kernel(const unsigned long long *__restrict__ keys, long long nk,
                               int32_t *__restrict__ col) {
  for (long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x; k < nk;
       k += (long long)gridDim.x * blockDim.x)
    col[k] = (int32_t)(keys[k] & 0xffffffull);
}

int csr_from_keys_radix(mnhn_sw_ctx *c, unsigned long long nk, mnhn_csr *out) {
  const int n = c->n;
  const double kernel_ms = c->last.kernel_ms;
  int b = 1;
  while ((1ll << b) <= (long long)n) ++b;
  unsigned long long *d_k0 = nullptr, *d_k1 = nullptr, *d_flag = nullptr;
  uint16_t *d_d0 = nullptr, *d_d1 = nullptr;
  long long *d_rp = nullptr;
  void *d_tmp = nullptr;
  auto cleanup = [&]() {
    dev_free(d_k0); dev_free(d_k1); dev_free(d_d0); dev_free(d_d1); dev_free(d_rp); dev_free(d_flag);
    if (d_tmp) (void)hipFree(d_tmp);
    d_tmp = nullptr;
  };
  int rc;
  if ((rc = dev_alloc(&d_k0, nk)) || (rc = dev_alloc(&d_k1, nk)) || (rc = dev_alloc(&d_d0, nk)) ||
      (rc = dev_alloc(&d_d1, nk)) || (rc = dev_alloc(&d_rp, (size_t)n + 1)) || (rc = dev_alloc(&d_flag, 1))) {
    cleanup(); mnhn_csr_free(out); return rc;
  }
  const int tb = 256;
  (void)hipEventRecord(c->ev[2], c->stream);
  hipError_t e = hipSuccess;
  if (nk > 0) {
    hipLaunchKernelGGL(split_keys_kernel, dim3(stride_blocks((long long)nk, tb)), dim3(tb), 0, c->stream, c->d_keys,
                       (long long)nk, d_k0, d_d0);
    e = hipGetLastError();
  }
  rocprim::double_buffer<unsigned long long> kbuf(d_k0, d_k1);
  rocprim::double_buffer<uint16_t> dbuf(d_d0, d_d1);
  size_t tmp_bytes = 0;
  if (e == hipSuccess && nk > 0)
    e = rocprim::radix_sort_pairs(nullptr, tmp_bytes, kbuf, dbuf, (size_t)nk, 0u, (unsigned)(24 + b), c->stream);
  if (e == hipSuccess && nk > 0) e = hipMalloc(&d_tmp, std::max<size_t>(tmp_bytes, 1));
  if (e != hipSuccess) d_tmp = nullptr;
  if (e == hipSuccess && nk > 0)
    e = rocprim::radix_sort_pairs(d_tmp, tmp_bytes, kbuf, dbuf, (size_t)nk, 0u, (unsigned)(24 + b), c->stream);
  if (e == hipSuccess) e = hipMemsetAsync(d_flag, 0, sizeof(unsigned long long), c->stream);
  if (e == hipSuccess && nk > 1) {
    hipLaunchKernelGGL(sorted_check_kernel, dim3(stride_blocks((long long)nk, tb)), dim3(tb), 0, c->stream,
                       kbuf.current(), (long long)nk, d_flag);
    e = hipGetLastError();
  }
  if (e == hipSuccess) {
    hipLaunchKernelGGL(cut_rows_kernel, dim3(stride_blocks((long long)n + 1, tb)), dim3(tb), 0, c->stream,
                       kbuf.current(), (long long)nk, n, d_rp);
    e = hipGetLastError();
  }
  int32_t *d_col = reinterpret_cast<int32_t *>(kbuf.alternate());
  if (e == hipSuccess && nk > 0) {
    hipLaunchKernelGGL(low_col_kernel, dim3(stride_blocks((long long)nk, tb)), dim3(tb), 0, c->stream,
                       kbuf.current(), (long long)nk, d_col);
    e = hipGetLastError();
  }
  unsigned long long unsorted = 0;
  if (e == hipSuccess) e = hipMemcpyAsync(&unsorted, d_flag, sizeof(unsorted), hipMemcpyDeviceToHost, c->stream);
  (void)hipEventRecord(c->ev[3], c->stream);
  if (e == hipSuccess) e = hipEventSynchronize(c->ev[3]);
  if (e == hipSuccess && unsorted) e = hipErrorUnknown;
  float post_ms = 0.f;
  if (e == hipSuccess) (void)hipEventElapsedTime(&post_ms, c->ev[2], c->ev[3]);
  if (e == hipSuccess) {
    out->col = (int32_t *)host_alloc(std::max<size_t>(nk, 1) * sizeof(int32_t));
    out->dist = (uint16_t *)host_alloc(std::max<size_t>(nk, 1) * sizeof(uint16_t));
    if (!out->col || !out->dist) { cleanup(); mnhn_csr_free(out); return MNHN_ERR_NOMEM; }
    rc = copy_d2h(c, out->row_ptr, d_rp, ((size_t)n + 1) * sizeof(long long));
    if (rc == MNHN_OK && nk) rc = copy_d2h(c, out->col, d_col, nk * sizeof(int32_t));
    if (rc == MNHN_OK && nk) rc = copy_d2h(c, out->dist, dbuf.current(), nk * sizeof(uint16_t));
    if (rc != MNHN_OK) e = hipErrorUnknown;
  }
  cleanup();
  if (e != hipSuccess) { mnhn_csr_free(out); return MNHN_ERR_DEVICE; }
  out->nnz = (int64_t)nk;
  c->last.kernel_ms = kernel_ms;
  c->last.total_ms = kernel_ms + post_ms;
  return MNHN_OK;
}

}  // namespace

int mnhn_detail::csr_from_keys(mnhn_sw_ctx *c, unsigned long long nk, mnhn_csr *out) {
  const int n = c->n;
  int rc = MNHN_OK;
  // host lists take 6 bytes per entry: refuse lists beyond half of the host's
  // physical memory instead of running it out
  const double phys = (double)sysconf(_SC_PHYS_PAGES) * (double)sysconf(_SC_PAGE_SIZE);
  if (phys > 0 && (double)nk * (sizeof(int32_t) + sizeof(uint16_t)) > 0.5 * phys) {
    mnhn_csr_free(out);
    return MNHN_ERR_NOMEM;
  }
  const double kernel_ms = c->last.kernel_ms;
  // LDS of the long-row path: column bitmap + per-word prefix + thread totals
  const size_t lds_bytes = std::max<size_t>(
      (size_t)(2 * ((n + 31) / 32) + kSortThreads) * sizeof(uint32_t), kShortRow * sizeof(uint32_t));
  const char *form = getenv("MNHN_CSR_SORT");
  if (lds_bytes > 160 * 1024 || (form && strcmp(form, "radix") == 0)) return csr_from_keys_radix(c, nk, out);
  int32_t *d_cnt = nullptr, *d_seg_col = nullptr, *d_col = nullptr;
  uint16_t *d_seg_d = nullptr, *d_dist = nullptr;
  long long *d_rp = nullptr;
  auto cleanup = [&]() {
    dev_free(d_cnt); dev_free(d_seg_col); dev_free(d_col); dev_free(d_seg_d); dev_free(d_dist);
    dev_free(d_rp);
  };
  if ((rc = dev_alloc(&d_cnt, (size_t)n)) || (rc = dev_alloc(&d_rp, (size_t)n + 1)) ||
      (rc = dev_alloc(&d_seg_col, nk)) || (rc = dev_alloc(&d_seg_d, nk)) ||
      (rc = dev_alloc(&d_col, nk)) || (rc = dev_alloc(&d_dist, nk))) {
    cleanup(); mnhn_csr_free(out); return rc;
  }
  mnhn_phase_clock clk(c->stream, "csr");
  const int tb = 256;
  const unsigned kblocks = stride_blocks((long long)nk, tb);
  std::vector<int32_t> cnt(n);
  (void)hipEventRecord(c->ev[2], c->stream);
  hipError_t e = hipMemsetAsync(d_cnt, 0, (size_t)n * sizeof(int32_t), c->stream);
  if (e == hipSuccess && nk > 0) {
    hipLaunchKernelGGL(count_rows_kernel, dim3(kblocks), dim3(tb), 0, c->stream,
                       c->d_keys, (long long)nk, d_cnt);
    e = hipGetLastError();
  }
  if (e == hipSuccess)
    e = hipMemcpyAsync(cnt.data(), d_cnt, (size_t)n * sizeof(int32_t), hipMemcpyDeviceToHost, c->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
  if (e == hipSuccess) {
    for (int i = 0; i < n; ++i) out->row_ptr[i + 1] = out->row_ptr[i] + cnt[i];
    e = hipMemcpyAsync(d_rp, out->row_ptr, ((size_t)n + 1) * sizeof(long long),
                       hipMemcpyHostToDevice, c->stream);
  }
  if (e == hipSuccess) e = hipMemsetAsync(d_cnt, 0, (size_t)n * sizeof(int32_t), c->stream);
  if (e == hipSuccess && nk > 0) {
    hipLaunchKernelGGL(scatter_rows_kernel, dim3(kblocks), dim3(tb), 0, c->stream,
                       c->d_keys, (long long)nk, d_rp, d_cnt, d_seg_col, d_seg_d);
    e = hipGetLastError();
  }
  if (e == hipSuccess && lds_bytes > 64 * 1024)
    e = hipFuncSetAttribute(reinterpret_cast<const void *>(&sort_rows_kernel<uint16_t>),
                            hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds_bytes);
  if (e == hipSuccess && nk > 0) {
    hipLaunchKernelGGL(sort_rows_kernel<uint16_t>, dim3((unsigned)n), dim3(kSortThreads), lds_bytes,
                       c->stream, d_rp, d_seg_col, d_seg_d, d_col, d_dist, n);
    e = hipGetLastError();
  }
  (void)hipEventRecord(c->ev[3], c->stream);
  if (e == hipSuccess) e = hipEventSynchronize(c->ev[3]);
  float post_ms = 0.f;
  if (e == hipSuccess) (void)hipEventElapsedTime(&post_ms, c->ev[2], c->ev[3]);
  clk.mark("count, scatter, row sort");
  if (e == hipSuccess) {
    out->col = (int32_t *)host_alloc(std::max<size_t>(nk, 1) * sizeof(int32_t));
    out->dist = (uint16_t *)host_alloc(std::max<size_t>(nk, 1) * sizeof(uint16_t));
    if (!out->col || !out->dist) { cleanup(); mnhn_csr_free(out); return MNHN_ERR_NOMEM; }
    rc = copy_d2h(c, out->col, d_col, nk * sizeof(int32_t));
    clk.mark("columns to host");
    if (rc == MNHN_OK) rc = copy_d2h(c, out->dist, d_dist, nk * sizeof(uint16_t));
    clk.mark("distances to host");
    if (rc != MNHN_OK) e = hipErrorUnknown;
  }
  cleanup();
  clk.mark("release");
  if (e != hipSuccess) { mnhn_csr_free(out); return MNHN_ERR_DEVICE; }
  out->nnz = (int64_t)nk;
  c->last.kernel_ms = kernel_ms;
  c->last.total_ms = kernel_ms + post_ms;
  return MNHN_OK;
}

int mnhn_detail::context_strip(const mnhn_sw_ctx *c) { return pick_strip(c); }

int mnhn_detail::sort_row_segments(hipStream_t st, const long long *d_rp, const int32_t *seg, int32_t *out,
                                   const uint32_t *pay_in, uint32_t *pay_out, int n_rows, int n_cols) {
  const size_t lds_bytes = std::max<size_t>(
      (size_t)(2 * ((n_cols + 31) / 32) + kSortThreads) * sizeof(uint32_t), kShortRow * sizeof(uint32_t));
  if (lds_bytes > 160 * 1024) return MNHN_ERR_RANGE;
  hipError_t e = hipSuccess;
  if (lds_bytes > 64 * 1024)
    e = hipFuncSetAttribute(reinterpret_cast<const void *>(&sort_rows_kernel<uint32_t>),
                            hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds_bytes);
  if (e == hipSuccess && n_rows > 0) {
    hipLaunchKernelGGL(sort_rows_kernel<uint32_t>, dim3((unsigned)n_rows), dim3(kSortThreads), lds_bytes, st,
                       d_rp, seg, pay_in, out, pay_out, n_cols);
    e = hipGetLastError();
  }
  return e == hipSuccess ? MNHN_OK : MNHN_ERR_DEVICE;
}

// Host buffers of the CSR results (released with host_free).  From 16 MB on
// they are pinned while the pinned buffers outstanding stay within
// kPinnedBudget: the device then copies into them at the link's rate
// (68 MB in 1.3 ms on MI355X, against 4 ms through the staging buffers into
// fresh pageable memory).  Beyond the budget they are 2 MB-aligned and marked
// for transparent huge pages, so filling them takes a page fault per 2 MB
// instead of per 4 KB.  The pinned ones are listed (start -> size) so that
// host_free and copy_d2h can tell them apart.
namespace {
constexpr size_t kPinnedBudget = 16ull << 30;
std::mutex g_pinned_mu;
std::map<const char *, size_t> g_pinned;
size_t g_pinned_bytes = 0;

// the pinned buffer holding [p, p + bytes), if any
bool pinned_range(const void *p, size_t bytes) {
  std::lock_guard<std::mutex> lk(g_pinned_mu);
  auto it = g_pinned.upper_bound((const char *)p);
  if (it == g_pinned.begin()) return false;
  --it;
  return (const char *)p + bytes <= it->first + it->second;
}
}  // namespace

void *mnhn_detail::host_alloc(size_t bytes) {
  constexpr size_t kHuge = 2u << 20;
  if (bytes < 8 * kHuge) return malloc(bytes);
  {
    std::lock_guard<std::mutex> lk(g_pinned_mu);
    if (g_pinned_bytes + bytes <= kPinnedBudget) {
      void *p = nullptr;
      if (hipHostMalloc(&p, bytes, hipHostMallocDefault) == hipSuccess && p) {
        g_pinned[(const char *)p] = bytes;
        g_pinned_bytes += bytes;
        return p;
      }
      (void)hipGetLastError();  // clear the failure; fall back to pageable memory
    }
  }
  const size_t rounded = (bytes + kHuge - 1) / kHuge * kHuge;
  void *p = aligned_alloc(kHuge, rounded);
  if (p) (void)madvise(p, rounded, MADV_HUGEPAGE);
  return p;
}

void mnhn_detail::host_free(void *p) {
  if (!p) return;
  {
    std::lock_guard<std::mutex> lk(g_pinned_mu);
    auto it = g_pinned.find((const char *)p);
    if (it != g_pinned.end()) {
      g_pinned_bytes -= it->second;
      g_pinned.erase(it);
      (void)hipHostFree(p);
      return;
    }
  }
  free(p);
}

int mnhn_detail::copy_d2h(hipStream_t st, mnhn_detail_stage &s, void *dst, const void *src, size_t bytes) {
  // copies of 8 MB and more go through two pinned 16 MB stages (DMA of one
  // chunk while host threads copy the previous one out); measured on MI355X,
  // a plain copy into pageable memory moved 34 MB at 6.7 GB/s
  constexpr size_t kStageBytes = 16u << 20;
  if (bytes < (8u << 20)) {
    return hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost) == hipSuccess ? MNHN_OK : MNHN_ERR_DEVICE;
  }
  if (pinned_range(dst, bytes))
    return hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, st) == hipSuccess &&
                   hipStreamSynchronize(st) == hipSuccess
               ? MNHN_OK
               : MNHN_ERR_DEVICE;
  for (int k = 0; k < 2; ++k) {
    if (!s.h[k] && hipHostMalloc(&s.h[k], kStageBytes, hipHostMallocDefault) != hipSuccess) {
      s.h[k] = nullptr;
      return hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost) == hipSuccess ? MNHN_OK : MNHN_ERR_DEVICE;
    }
    if (!s.ev[k] && hipEventCreateWithFlags(&s.ev[k], hipEventDisableTiming) != hipSuccess) {
      s.ev[k] = nullptr;
      return MNHN_ERR_DEVICE;
    }
  }
  const size_t chunks = (bytes + kStageBytes - 1) / kStageBytes;
  auto issue = [&](size_t k) -> bool {
    const size_t off = k * kStageBytes, len = std::min(kStageBytes, bytes - off);
    return hipMemcpyAsync(s.h[k & 1], (const char *)src + off, len, hipMemcpyDeviceToHost, st) == hipSuccess &&
           hipEventRecord(s.ev[k & 1], st) == hipSuccess;
  };
  // the pinned -> pageable copy of a chunk is split over host threads (the
  // destination pages are touched for the first time here, which one thread
  // does at about 10 GB/s) while the next chunk's DMA is in flight
  const unsigned hw = std::max(1u, std::thread::hardware_concurrency());
  const int nt = (int)std::min<unsigned>(8u, hw);
  if (!issue(0)) return MNHN_ERR_DEVICE;
  for (size_t k = 0; k < chunks; ++k) {
    if (k + 1 < chunks && !issue(k + 1)) return MNHN_ERR_DEVICE;
    if (hipEventSynchronize(s.ev[k & 1]) != hipSuccess) return MNHN_ERR_DEVICE;
    const size_t off = k * kStageBytes, len = std::min(kStageBytes, bytes - off);
    char *d = (char *)dst + off;
    const char *src_h = (const char *)s.h[k & 1];
    if (nt > 1 && len >= (4u << 20)) {
      const size_t part = (len / nt + 4095) / 4096 * 4096;
      std::vector<std::thread> th;
      size_t done_from = len;  // parts from here on are copied by threads
      for (int t = nt - 1; t >= 1; --t) {
        const size_t b = (size_t)t * part;
        if (b >= len) continue;
        try {
          th.emplace_back([=]() { memcpy(d + b, src_h + b, std::min(part, len - b)); });
          done_from = b;
        } catch (...) {
          break;  // no more threads: the rest is copied below
        }
      }
      memcpy(d, src_h, std::min(done_from, len));
      for (auto &x : th) x.join();
    } else {
      memcpy(d, src_h, len);
    }
  }
  return MNHN_OK;
}

void mnhn_detail::stage_free(mnhn_detail_stage &s) {
  for (int k = 0; k < 2; ++k) {
    if (s.h[k]) (void)hipHostFree(s.h[k]);
    if (s.ev[k]) (void)hipEventDestroy(s.ev[k]);
    s.h[k] = nullptr;
    s.ev[k] = nullptr;
  }
}

extern "C" {

void mnhn_csr_free(mnhn_csr *csr) {
  if (!csr) return;
  host_free(csr->row_ptr);
  host_free(csr->col);
  host_free(csr->dist);
  csr->row_ptr = nullptr;
  csr->col = nullptr;
  csr->dist = nullptr;
  csr->nnz = 0;
}

}  // extern "C"

// ---------------------------------------------------------------------------
// autotuner: times every candidate (kernel variant, strip width, waves per
// CU) on a strided sample of this context's all-pairs work items and keeps
// the fastest; choices are cached as JSON lines keyed by device and length
// profile.
// ---------------------------------------------------------------------------
namespace {

struct Cand {
  int kernel, strip, wpc;
};

std::string tune_cache_path() {
  const char *e = getenv("MNHN_SW_TUNE_CACHE");
  if (e && *e) return e;
  const char *home = getenv("HOME");
  return std::string(home && *home ? home : "/tmp") + "/.cache/mnhn_sw_tune.jsonl";
}

std::string tune_key(const mnhn_sw_ctx *c) {
  hipDeviceProp_t prop;
  std::string arch = "unknown";
  if (hipGetDeviceProperties(&prop, c->device) == hipSuccess) arch = prop.gcnArchName;
  std::vector<int32_t> L(c->len);
  std::sort(L.begin(), L.end());
  const int med = L.empty() ? 0 : L[L.size() / 2];
  const int p90 = L.empty() ? 0 : L[(L.size() * 9) / 10];
  // lengths bucketed on a quarter-octave scale
  auto bucket = [](int x) { return x <= 1 ? 0 : (int)std::lround(4.0 * std::log2((double)x)); };
  char buf[256];
  snprintf(buf, sizeof(buf), "%s|cu%d|%s|m%d|p%d|g%d|mm%d", arch.c_str(), c->num_cu,
           c->acgt ? "acgt" : "codes", bucket(med), bucket(p90), c->gap, c->match - c->mismatch);
  return buf;
}

bool read_tune_cache(const std::string &key, Cand *out, double *gcups) {
  FILE *f = fopen(tune_cache_path().c_str(), "r");
  if (!f) return false;
  char line[1024];
  bool found = false;
  const std::string needle = "\"key\": \"" + key + "\"";
  while (fgets(line, sizeof(line), f)) {
    if (!strstr(line, needle.c_str())) continue;
    Cand t{0, 0, 0};
    double g = 0;
    const char *k = strstr(line, "\"kernel\": ");
    const char *st = strstr(line, "\"strip\": ");
    const char *w = strstr(line, "\"waves_per_cu\": ");
    const char *gc = strstr(line, "\"gcups\": ");
    if (k && st && w && gc && sscanf(k + 10, "%d", &t.kernel) == 1 && sscanf(st + 9, "%d", &t.strip) == 1 &&
        sscanf(w + 16, "%d", &t.wpc) == 1 && sscanf(gc + 9, "%lf", &g) == 1) {
      *out = t;  // the last matching line wins
      *gcups = g;
      found = true;
    }
  }
  fclose(f);
  return found;
}

void write_tune_cache(const std::string &key, const Cand &t, double gcups) {
  const std::string path = tune_cache_path();
  const size_t slash = path.rfind('/');
  if (slash != std::string::npos && slash > 0) {
    const std::string dir = path.substr(0, slash);
    (void)mkdir(dir.c_str(), 0755);
  }
  FILE *f = fopen(path.c_str(), "a");
  if (!f) return;
  fprintf(f, "{\"key\": \"%s\", \"kernel\": %d, \"strip\": %d, \"waves_per_cu\": %d, \"gcups\": %.1f}\n",
          key.c_str(), t.kernel, t.strip, t.wpc, gcups);
  fclose(f);
}

}  // namespace

extern "C" int mnhn_sw_autotune(mnhn_sw_ctx *c, int use_cache, mnhn_sw_tuning *out) {
  if (!c) return MNHN_ERR_ARG;
  int rc = MNHN_OK;
  const std::string key = tune_key(c);
  Cand best{0, 0, 0};
  double best_gcups = 0;
  if (use_cache && read_tune_cache(key, &best, &best_gcups)) {
    c->kernel = best.kernel;
    c->strip = best.strip;
    c->waves_per_cu = best.wpc;
    if (out) *out = mnhn_sw_tuning{best.kernel, best.strip, best.wpc, best_gcups, 1};
    return MNHN_OK;
  }
  std::vector<Cand> cands;
  // resident-wave limits of the LDS kernels: 3 per SIMD at strip 64, 5 at
  // 32, 6 at 16 (register budgets)
  if (c->acgt) {
    for (int w : {12, 16, 20}) cands.push_back({KERNEL_LDS_F16, 32, w});
    for (int w : {16, 24}) cands.push_back({KERNEL_LDS_F16, 16, w});
    for (int w : {8, 12}) cands.push_back({KERNEL_LDS_F16, 64, w});
    for (int w : {16, 24}) cands.push_back({KERNEL_LDS, 16, w});
    for (int w : {12, 16, 20}) cands.push_back({KERNEL_LDS, 32, w});
    for (int w : {8, 12}) cands.push_back({KERNEL_LDS, 64, w});
  }
  for (int w : {4, 8}) cands.push_back({KERNEL_REG, 32, w});
  for (int w : {8, 16}) cands.push_back({KERNEL_REG, 16, w});
  // sample of the all-pairs work: single-query items (one query against one
  // target group), about 32 per resident wave at the largest candidate
  // occupancy (so that the queue's tail weighs little) -- a few hundred
  // milliseconds of work per candidate whatever the data-set size.  They are
  // strided over the cells of all (group, query) pairs, so the sample has the
  // length mix of the work it stands for.
  std::vector<double> gsum(c->groups, 0.0);
  double total_cells = 0;
  for (int g = 0; g < c->groups; ++g) {
    for (int t = 0; t < kGroup; ++t) {
      const int id = c->gtgt[(size_t)g * kGroup + t];
      if (id >= 0) gsum[g] += c->len[id];
    }
    const int last = std::min((g + 1) * kGroup, c->n) - 1;
    for (int p = 0; p < last; ++p) total_cells += (double)c->len[c->order[p]] * gsum[g];
  }
  const double want = (double)std::max(c->num_cu, 1) * 24 * 32;
  const double stride = total_cells / want;
  Work sample;
  sample.qids = c->order;
  double next = 0, u = 0;
  for (int g = 0; g < c->groups; ++g) {
    const int last = std::min((g + 1) * kGroup, c->n) - 1;
    for (int p = 0; p < last; ++p) {
      const double cells = (double)c->len[c->order[p]] * gsum[g];
      u += cells;
      if (u <= next || cells <= 0) continue;
      while (next < u) next += stride;  // once per pair, however many strides it spans
      sample.items.push_back(make_int4(g, p, p + 1, 0));
      sample.cells += cells;
    }
  }
  finalize_items(c, sample, 32, false);
  const int saved_k = c->kernel, saved_s = c->strip, saved_w = c->waves_per_cu;
  for (const Cand &cd : cands) {
    c->kernel = cd.kernel;
    c->strip = cd.strip;
    c->waves_per_cu = cd.wpc;
    double ms = 1e30;
    for (int rep = 0; rep < 3; ++rep) {  // first run warms caches, code and clocks
      rc = launch_tiles(c, sample, MODE_EPS, 1, -1, nullptr, nullptr, 0);
      if (rc) break;
      ms = std::min(ms, c->last.kernel_ms);
    }
    if (rc) break;
    const double g = sample.cells / std::max(ms, 1e-6) / 1e6;
    const char *tv = getenv("MNHN_TIMING");
    if (tv && *tv && strcmp(tv, "0") != 0)
      fprintf(stderr, "[mnhn] autotune kernel %d strip %d waves/CU %d: %.0f GCUPS on %.3g sample cells\n",
              cd.kernel, cd.strip, cd.wpc, g, sample.cells);
    // a narrower strip must win by 1 % per halving: within the sample's
    // noise the wider strip moves half the workspace bytes per cell and
    // carries half the per-row bookkeeping, and the choice stays stable
    // from run to run
    double need = best_gcups;
    for (int sw = cd.strip; sw < best.strip; sw *= 2) need *= 1.01;
    for (int sw = best.strip; best.kernel && sw < cd.strip; sw *= 2) need /= 1.01;
    if (g > need) {
      best_gcups = g;
      best = cd;
    }
  }
  if (rc || best.kernel == 0) {
    c->kernel = saved_k;
    c->strip = saved_s;
    c->waves_per_cu = saved_w;
    return rc ? rc : MNHN_ERR_DEVICE;
  }
  c->kernel = best.kernel;
  c->strip = best.strip;
  c->waves_per_cu = best.wpc;
  write_tune_cache(key, best, best_gcups);
  if (out) *out = mnhn_sw_tuning{best.kernel, best.strip, best.wpc, best_gcups, 0};
  return MNHN_OK;
}
