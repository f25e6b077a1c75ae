// vec_engine.hip -- gfx950 epsilon-neighbour lists of float vectors (PCA
// projections, k-mer count vectors) under the reference's L1 / L2 scans
// (dbscan.c:563-780), exported through include/mnhn_vec.h.
//
// Exactness.  For a query p and a point x the reference accumulates over the
// dimensions in order, in float32, acc = fma(p_k - x_k, p_k - x_k, acc) (L2,
// its AVX2 + FMA build, dbscan.c:648-659) or acc = |p_k - x_k| + acc (L1,
// dbscan.c:660-664), and keeps x when acc <= eps * eps (L2, float product) or
// acc <= eps (L1) (dbscan.c:676-680).  p_k - x_k = -(x_k - p_k) exactly, so the
// accumulator of (p, x) equals that of (x, p) bit for bit: the kernel
// evaluates every unordered pair once, with the same operations in the same
// order, and lists it under both rows.  Zero-padded dimensions add +0, which
// leaves every accumulator (>= +0 or NaN) unchanged.
//
// Kernel.  The upper triangle (diagonal included) of 128 x 128 point tiles is
// walked by a grid-stride loop; a 256-thread workgroup stages up to 16
// dimensions of both tiles in LDS per step and every thread keeps an 8 x 8 block of
// accumulators in registers (two VALU ops per pair and dimension: the
// subtraction and the fma / abs-add); the next chunk -- across the tile
// boundary too -- is loaded into registers while the current one computes.
// A lane whose smallest accumulator passes the threshold hands its 8 x 8 block
// to the whole wave through LDS, one accumulator per lane; pairs under the
// threshold are collected per wave in an LDS buffer and appended as
// (i << 32 | j, acc) with one atomic reservation per flush of up to 512
// pairs, so sparse lists do not serialise on the output counter.
//
// Lists.  Pairs are expanded to both directions as keys row * 2^b + col, radix
// sorted (rocPRIM) with the accumulator as payload, and cut into CSR rows by a
// binary search per row.
#include <hip/hip_runtime.h>
#include <rocprim/device/device_radix_sort.hpp>

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "engine_internal.h"
#include "mnhn_vec.h"

using namespace mnhn_detail;

struct mnhn_vec_ctx {
  int n = 0, dims = 0, dims_pad = 0;
  long long n_pad = 0;
  int device = 0, num_cu = 0;
  float *d_x = nullptr;  // [dims_pad][n_pad], dimension-major, zero-padded
  hipStream_t stream = nullptr;
  hipEvent_t ev[2] = {nullptr, nullptr};
  mnhn_detail_stage stage;
  unsigned long long *d_pairs = nullptr;  // (i << 32 | j), i <= j
  float *d_pval = nullptr;
  size_t pair_cap = 0;
  unsigned long long *d_count = nullptr;  // [0] pairs found, [1] diagonal pairs
  double last_ms = 0, last_pairs = 0;
  unsigned long long last_hits = 0;
};

namespace {

constexpr int kTile = 128;
constexpr int kThreads = 256;
constexpr int kChunk = 16;
constexpr int kWaveBuf = 512;  // listed pairs a wave collects in LDS before one reservation

__device__ __forceinline__ long long tri_start(long long t) { return t * (t + 1) / 2; }

// Upper-triangle tile t (column-major: tiles (0,0), (0,1), (1,1), (0,2), ...).
// A float estimate of the column (off by at most a few for any tile count
// that fits the output keys), then exact integer correction.
__device__ __forceinline__ void tile_of(long long t, int &ti, int &tj) {
  long long c = (long long)((sqrtf(8.0f * (float)t + 1.0f) - 1.0f) * 0.5f);
  if (c < 0) c = 0;
  while (tri_start(c) > t) --c;
  while (tri_start(c + 1) <= t) ++c;
  tj = (int)c;
  ti = (int)(t - tri_start(c));
}

template <int NORM>
__device__ __forceinline__ float step(float acc, float a, float b) {
  const float d = a - b;
  if (NORM == MNHN_NORM_L2) return fmaf(d, d, acc);
  return acc + fabsf(d);
}

// Staging of one chunk (up to kChunk dimensions of both tiles): thread tid
// holds rows lk0 and lk0 + 8 of float4 column lc of each tile.
struct Stage {
  float4 a0, a1, b0, b1;
};

__device__ __forceinline__ Stage load_chunk(const float *__restrict__ x, long long n_pad, int dims_pad, int i0, int j0,
                                            int k0, int lk0, int lc) {
  const float4 zero4 = make_float4(0.f, 0.f, 0.f, 0.f);
  Stage g{zero4, zero4, zero4, zero4};
  const float4 *xa = reinterpret_cast<const float4 *>(x + i0);
  const float4 *xb = reinterpret_cast<const float4 *>(x + j0);
  const long long row4 = n_pad / 4;
  if (k0 + lk0 < dims_pad) {
    g.a0 = xa[(long long)(k0 + lk0) * row4 + lc];
    g.b0 = xb[(long long)(k0 + lk0) * row4 + lc];
  }
  if (k0 + lk0 + 8 < dims_pad) {
    g.a1 = xa[(long long)(k0 + lk0 + 8) * row4 + lc];
    g.b1 = xb[(long long)(k0 + lk0 + 8) * row4 + lc];
  }
  return g;
}

template <int NORM>
__global__ __launch_bounds__(kThreads) __attribute__((amdgpu_waves_per_eu(3))) void vec_tiles_kernel(
    const float *__restrict__ x, long long n_pad, int n, int dims_pad, long long n_tiles, float thr,
    unsigned long long *__restrict__ pairs, float *__restrict__ pval, unsigned long long cap,
    unsigned long long *__restrict__ count) {
  __shared__ float4 sa[kChunk][kTile / 4];
  __shared__ float4 sb[kChunk][kTile / 4];
  __shared__ unsigned long long bkey[kThreads / 64][kWaveBuf];
  __shared__ float bval[kThreads / 64][kWaveBuf];
  __shared__ float4 spill[kThreads / 64][16];  // one lane's 8 x 8 accumulators, row-major
  const int tid = threadIdx.x;
  const int tx = tid & 15, ty = tid >> 4;
  const int lane = tid & 63, wave = tid >> 6;
  const int lk0 = tid >> 5, lc = tid & 31;
  // pairs buffered by this wave (the same value in every lane)
  int bcount = 0;
  auto flush = [&]() {
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    __builtin_amdgcn_wave_barrier();
    unsigned long long base = 0;
    if (lane == 0) base = atomicAdd(count, (unsigned long long)bcount);
    base = __shfl(base, 0, 64);
    for (int e = lane; e < bcount; e += 64) {
      if (base + (unsigned long long)e < cap) {
        pairs[base + e] = bkey[wave][e];
        pval[base + e] = bval[wave][e];
      }
    }
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    __builtin_amdgcn_wave_barrier();
    bcount = 0;
  };

  long long t = blockIdx.x;
  if (t >= n_tiles) return;
  int ti, tj;
  tile_of(t, ti, tj);
  // the first chunk of the next tile is loaded while this tile computes
  Stage g = load_chunk(x, n_pad, dims_pad, ti * kTile, tj * kTile, 0, lk0, lc);
  for (;;) {
    const int i0 = ti * kTile, j0 = tj * kTile;
    const long long t_next = t + gridDim.x;
    int ti_n = 0, tj_n = 0;
    if (t_next < n_tiles) tile_of(t_next, ti_n, tj_n);
    float acc[8][8];
#pragma unroll
    for (int r = 0; r < 8; ++r)
#pragma unroll
      for (int c = 0; c < 8; ++c) acc[r][c] = 0.f;
    for (int k0 = 0; k0 < dims_pad; k0 += kChunk) {
      __syncthreads();
      sa[lk0][lc] = g.a0;
      sa[lk0 + 8][lc] = g.a1;
      sb[lk0][lc] = g.b0;
      sb[lk0 + 8][lc] = g.b1;
      __syncthreads();
      if (k0 + kChunk < dims_pad)
        g = load_chunk(x, n_pad, dims_pad, i0, j0, k0 + kChunk, lk0, lc);
      else if (t_next < n_tiles)
        g = load_chunk(x, n_pad, dims_pad, ti_n * kTile, tj_n * kTile, 0, lk0, lc);
      const int kc = min(kChunk, dims_pad - k0);
#pragma unroll 2
      for (int k = 0; k < kc; ++k) {
        const float4 a0 = sa[k][ty], a1 = sa[k][16 + ty];
        const float4 b0 = sb[k][tx], b1 = sb[k][16 + tx];
        const float av[8] = {a0.x, a0.y, a0.z, a0.w, a1.x, a1.y, a1.z, a1.w};
        const float bv[8] = {b0.x, b0.y, b0.z, b0.w, b1.x, b1.y, b1.z, b1.w};
#pragma unroll
        for (int r = 0; r < 8; ++r)
#pragma unroll
          for (int c = 0; c < 8; ++c) acc[r][c] = step<NORM>(acc[r][c], av[r], bv[c]);
      }
    }

    // emission: pair (i, j) with i <= j < n and acc <= thr, into the wave's
    // buffer.  The smallest accumulator first (NaN never passes a threshold
    // and fminf skips it): most waves of a sparse list stop here.  Otherwise
    // each lane that may hold a pair spills its 8 x 8 block to LDS in turn and
    // the wave tests it one accumulator per lane, so the cost follows the
    // lanes with pairs rather than 64 tests in every lane.
    float mn = acc[0][0];
#pragma unroll
    for (int r = 0; r < 8; ++r)
#pragma unroll
      for (int c = 0; c < 8; ++c) mn = fminf(mn, acc[r][c]);
    unsigned long long cand = __ballot(mn <= thr);
    while (cand) {
      const int h = __builtin_ctzll(cand);
      cand &= cand - 1ull;
      if (lane == h) {
#pragma unroll
        for (int r = 0; r < 8; ++r) {
          spill[wave][r * 2 + 0] = make_float4(acc[r][0], acc[r][1], acc[r][2], acc[r][3]);
          spill[wave][r * 2 + 1] = make_float4(acc[r][4], acc[r][5], acc[r][6], acc[r][7]);
        }
      }
      __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
      __builtin_amdgcn_wave_barrier();
      // lane e tests accumulator (r, c) = (e / 8, e % 8) of lane h
      const int r = lane >> 3, c = lane & 7;
      const float v = reinterpret_cast<const float *>(&spill[wave][0])[r * 8 + c];
      const int hy = wave * 4 + (h >> 4), hx = h & 15;  // ty, tx of lane h
      const int i = i0 + (r < 4 ? hy * 4 + r : 64 + hy * 4 + r - 4);
      const int j = j0 + (c < 4 ? hx * 4 + c : 64 + hx * 4 + c - 4);
      const bool hit = i <= j && j < n && v <= thr;
      const unsigned long long hm = __ballot(hit);
      const int cnt = __popcll(hm);
      if (cnt > 0) {
        if (bcount + cnt > kWaveBuf) flush();
        if (hit) {
          const int p = bcount + __popcll(hm & ((1ull << lane) - 1ull));
          bkey[wave][p] = ((unsigned long long)(unsigned)i << 32) | (unsigned long long)(unsigned)j;
          bval[wave][p] = v;
        }
        bcount += cnt;
      }
      __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
      __builtin_amdgcn_wave_barrier();
    }
    if (t_next >= n_tiles) break;
    t = t_next;
    ti = ti_n;
    tj = tj_n;
  }
  if (bcount > 0) flush();
}

// Both directions of every pair as row * 2^b + col; the second copy of a
// diagonal pair becomes a sentinel that sorts after every real key.
__global__ void expand_kernel(const unsigned long long *__restrict__ pairs, const float *__restrict__ pval,
                              long long m, int b, unsigned long long *__restrict__ keys,
                              float *__restrict__ vals, unsigned long long *__restrict__ diag) {
  for (long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x; k < m;
       k += (long long)gridDim.x * blockDim.x) {
    const unsigned long long p = pairs[k];
    const unsigned long long i = p >> 32, j = p & 0xffffffffull;
    const float v = pval[k];
    keys[2 * k] = (i << b) | j;
    vals[2 * k] = v;
    if (i == j) {
      keys[2 * k + 1] = ~0ull;
      atomicAdd(diag, 1ull);
    } else {
      keys[2 * k + 1] = (j << b) | i;
    }
    vals[2 * k + 1] = v;
  }
}

// row_ptr[r] = first sorted position whose row is >= r, r = 0 .. n.
__global__ void row_ptr_kernel(const unsigned long long *__restrict__ keys, long long nnz, int n, int b,
                               long long *__restrict__ row_ptr) {
  for (long long r = (long long)blockIdx.x * blockDim.x + threadIdx.x; r <= n;
       r += (long long)gridDim.x * blockDim.x) {
    const unsigned long long want = (unsigned long long)r << b;
    long long lo = 0, hi = nnz;
    while (lo < hi) {
      const long long mid = (lo + hi) >> 1;
      if (keys[mid] < want) lo = mid + 1;
      else hi = mid;
    }
    row_ptr[r] = lo;
  }
}

// Sets *bad when the sorted keys are not strictly increasing (a checked
// postcondition of the radix sort: keys are distinct).
__global__ void check_sorted_kernel(const unsigned long long *__restrict__ keys, long long nnz,
                                    unsigned long long *__restrict__ bad) {
  for (long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x; k + 1 < nnz;
       k += (long long)gridDim.x * blockDim.x)
    if (keys[k] >= keys[k + 1]) atomicOr(bad, 1ull);
}

__global__ void col_kernel(const unsigned long long *__restrict__ keys, long long nnz, int b,
                           int32_t *__restrict__ col) {
  const unsigned long long mask = (1ull << b) - 1ull;
  for (long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x; k < nnz;
       k += (long long)gridDim.x * blockDim.x)
    col[k] = (int32_t)(keys[k] & mask);
}

int bits_for(int n) {
  int b = 1;
  while ((1ll << b) <= (long long)n) ++b;
  return b;  // n < 2^b
}

int ensure_pairs(mnhn_vec_ctx *c, size_t cap) {
  if (cap <= c->pair_cap && c->d_pairs) return MNHN_OK;
  dev_free(c->d_pairs);
  dev_free(c->d_pval);
  c->pair_cap = 0;
  int rc;
  if ((rc = dev_alloc(&c->d_pairs, cap)) || (rc = dev_alloc(&c->d_pval, cap))) {
    dev_free(c->d_pairs);
    dev_free(c->d_pval);
    return rc;
  }
  c->pair_cap = cap;
  return MNHN_OK;
}

int launch_tiles(mnhn_vec_ctx *c, int norm, float thr) {
  const long long side = c->n_pad / kTile;
  const long long n_tiles = side * (side + 1) / 2;
  const unsigned grid = (unsigned)std::min<long long>(n_tiles, (long long)c->num_cu * 16);
  if (hipMemsetAsync(c->d_count, 0, 2 * sizeof(unsigned long long), c->stream) != hipSuccess)
    return MNHN_ERR_DEVICE;
  if (n_tiles == 0) return MNHN_OK;
  if (norm == MNHN_NORM_L2)
    hipLaunchKernelGGL(vec_tiles_kernel<MNHN_NORM_L2>, dim3(grid), dim3(kThreads), 0, c->stream, c->d_x,
                       c->n_pad, c->n, c->dims_pad, n_tiles, thr, c->d_pairs, c->d_pval,
                       (unsigned long long)c->pair_cap, c->d_count);
  else
    hipLaunchKernelGGL(vec_tiles_kernel<MNHN_NORM_L1>, dim3(grid), dim3(kThreads), 0, c->stream, c->d_x,
                       c->n_pad, c->n, c->dims_pad, n_tiles, thr, c->d_pairs, c->d_pval,
                       (unsigned long long)c->pair_cap, c->d_count);
  return hipGetLastError() == hipSuccess ? MNHN_OK : MNHN_ERR_DEVICE;
}

}  // namespace

extern "C" {

int mnhn_vec_create(const float *const *values, int dims, int n, mnhn_vec_ctx **out) {
  if (!out) return MNHN_ERR_ARG;
  *out = nullptr;
  if (n < 0 || dims < 0 || (n > 0 && dims > 0 && !values)) return MNHN_ERR_ARG;
  if (n > 0x7fffff00) return MNHN_ERR_RANGE;
  for (int k = 0; k < dims && n > 0; ++k)
    if (!values[k]) return MNHN_ERR_ARG;
  mnhn_vec_ctx *c = new mnhn_vec_ctx();
  c->n = n;
  c->dims = dims;
  c->dims_pad = std::max(1, dims);  // no padding: a zero dimension would cost a step per pair
  c->n_pad = ((long long)n + kTile - 1) / kTile * kTile;
  int rc = MNHN_OK;
  auto fail = [&](int code) {
    mnhn_vec_destroy(c);
    return code;
  };
  if (hipGetDevice(&c->device) != hipSuccess) return fail(MNHN_ERR_DEVICE);
  if (hipDeviceGetAttribute(&c->num_cu, hipDeviceAttributeMultiprocessorCount, c->device) != hipSuccess)
    return fail(MNHN_ERR_DEVICE);
  if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) {
    c->stream = nullptr;
    return fail(MNHN_ERR_DEVICE);
  }
  for (auto &e : c->ev)
    if (hipEventCreate(&e) != hipSuccess) {
      e = nullptr;
      return fail(MNHN_ERR_DEVICE);
    }
  if ((rc = dev_alloc(&c->d_count, 2))) return fail(rc);
  const size_t elems = (size_t)c->dims_pad * (size_t)c->n_pad;
  if ((rc = dev_alloc(&c->d_x, elems))) return fail(rc);
  std::vector<float> h(elems, 0.f);
  for (int k = 0; k < dims; ++k) memcpy(&h[(size_t)k * (size_t)c->n_pad], values[k], (size_t)n * sizeof(float));
  if (hipMemcpy(c->d_x, h.data(), elems * sizeof(float), hipMemcpyHostToDevice) != hipSuccess)
    return fail(MNHN_ERR_DEVICE);
  *out = c;
  return MNHN_OK;
}

void mnhn_vec_destroy(mnhn_vec_ctx *c) {
  if (!c) return;
  if (c->stream) (void)hipStreamSynchronize(c->stream);
  dev_free(c->d_x);
  dev_free(c->d_pairs);
  dev_free(c->d_pval);
  dev_free(c->d_count);
  for (auto &e : c->ev)
    if (e) (void)hipEventDestroy(e);
  stage_free(c->stage);
  if (c->stream) (void)hipStreamDestroy(c->stream);
  delete c;
}

int mnhn_vec_eps_csr(mnhn_vec_ctx *c, int norm, float thr, mnhn_vcsr *out) {
  if (!c || !out) return MNHN_ERR_ARG;
  memset(out, 0, sizeof(*out));
  if (norm != MNHN_NORM_L1 && norm != MNHN_NORM_L2) return MNHN_ERR_ARG;
  const int n = c->n;
  out->n_rows = n;
  out->norm = norm;
  out->thr = thr;
  out->row_ptr = (int64_t *)calloc((size_t)n + 1, sizeof(int64_t));
  if (!out->row_ptr) return MNHN_ERR_NOMEM;
  if (n == 0) return MNHN_OK;

  // pairs: first attempt with the capacity of the last call (or 32 pairs per
  // point), then once more with the exact count if that overflowed
  size_t cap = std::max<size_t>({c->pair_cap, (size_t)c->last_hits + (size_t)c->last_hits / 4,
                                 (size_t)n * 32, (size_t)1 << 20});
  int rc = ensure_pairs(c, cap);
  if (rc) { mnhn_vcsr_free(out); return rc; }
  unsigned long long hits = 0;
  float ms = 0.f;
  for (int attempt = 0; attempt < 2; ++attempt) {
    (void)hipEventRecord(c->ev[0], c->stream);
    if ((rc = launch_tiles(c, norm, thr))) { mnhn_vcsr_free(out); return rc; }
    (void)hipEventRecord(c->ev[1], c->stream);
    if (hipMemcpyAsync(&hits, c->d_count, sizeof(hits), hipMemcpyDeviceToHost, c->stream) != hipSuccess ||
        hipStreamSynchronize(c->stream) != hipSuccess) {
      mnhn_vcsr_free(out);
      return MNHN_ERR_DEVICE;
    }
    (void)hipEventElapsedTime(&ms, c->ev[0], c->ev[1]);
    if (hits <= c->pair_cap) break;
    if ((rc = ensure_pairs(c, (size_t)hits))) { mnhn_vcsr_free(out); return rc; }
  }
  if (hits > c->pair_cap) { mnhn_vcsr_free(out); return MNHN_ERR_DEVICE; }
  c->last_hits = hits;
  c->last_ms = ms;
  c->last_pairs = (double)n * ((double)n + 1) / 2;

  const int b = bits_for(n);
  const long long m = (long long)hits, m2 = 2 * m;
  unsigned long long *d_k0 = nullptr, *d_k1 = nullptr;
  float *d_v0 = nullptr, *d_v1 = nullptr;
  long long *d_rp = nullptr;
  int32_t *d_col = nullptr;
  void *d_tmp = nullptr;
  auto cleanup = [&]() {
    dev_free(d_k0); dev_free(d_k1); dev_free(d_v0); dev_free(d_v1); dev_free(d_rp); dev_free(d_col);
    if (d_tmp) (void)hipFree(d_tmp);
    d_tmp = nullptr;
  };
  if ((rc = dev_alloc(&d_k0, (size_t)m2)) || (rc = dev_alloc(&d_k1, (size_t)m2)) ||
      (rc = dev_alloc(&d_v0, (size_t)m2)) || (rc = dev_alloc(&d_v1, (size_t)m2)) ||
      (rc = dev_alloc(&d_rp, (size_t)n + 1))) {
    cleanup(); mnhn_vcsr_free(out); return rc;
  }
  const int tb = 256;
  hipError_t e = hipSuccess;
  if (m > 0) {
    hipLaunchKernelGGL(expand_kernel, dim3(stride_blocks(m, tb)), dim3(tb), 0, c->stream, c->d_pairs, c->d_pval,
                       m, b, d_k0, d_v0, c->d_count + 1);
    e = hipGetLastError();
  }
  rocprim::double_buffer<unsigned long long> kbuf(d_k0, d_k1);
  rocprim::double_buffer<float> vbuf(d_v0, d_v1);
  size_t tmp_bytes = 0;
  if (e == hipSuccess && m2 > 0)
    e = rocprim::radix_sort_pairs(nullptr, tmp_bytes, kbuf, vbuf, (size_t)m2, 0u, (unsigned)(2 * b), c->stream);
  if (e == hipSuccess && m2 > 0) e = hipMalloc(&d_tmp, std::max<size_t>(tmp_bytes, 1));
  if (e != hipSuccess) d_tmp = nullptr;
  if (e == hipSuccess && m2 > 0)
    e = rocprim::radix_sort_pairs(d_tmp, tmp_bytes, kbuf, vbuf, (size_t)m2, 0u, (unsigned)(2 * b), c->stream);
  unsigned long long diag = 0;
  if (e == hipSuccess)
    e = hipMemcpyAsync(&diag, c->d_count + 1, sizeof(diag), hipMemcpyDeviceToHost, c->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
  const long long nnz = m2 - (long long)diag;
  // d_count[0] is free again: reuse it as the sortedness flag
  if (e == hipSuccess) e = hipMemsetAsync(c->d_count, 0, sizeof(unsigned long long), c->stream);
  if (e == hipSuccess && nnz > 1) {
    hipLaunchKernelGGL(check_sorted_kernel, dim3(stride_blocks(nnz, tb)), dim3(tb), 0, c->stream, kbuf.current(),
                       nnz, c->d_count);
    e = hipGetLastError();
  }
  unsigned long long unsorted = 0;
  if (e == hipSuccess)
    e = hipMemcpyAsync(&unsorted, c->d_count, sizeof(unsorted), hipMemcpyDeviceToHost, c->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
  if (e == hipSuccess && unsorted) e = hipErrorUnknown;
  if (e == hipSuccess) {
    hipLaunchKernelGGL(row_ptr_kernel, dim3(stride_blocks((long long)n + 1, tb)), dim3(tb), 0, c->stream,
                       kbuf.current(), nnz, n, b, d_rp);
    e = hipGetLastError();
  }
  // columns into the idle key buffer (an int32 view of it)
  int32_t *col_dev = reinterpret_cast<int32_t *>(kbuf.alternate());
  if (e == hipSuccess && nnz > 0) {
    hipLaunchKernelGGL(col_kernel, dim3(stride_blocks(nnz, tb)), dim3(tb), 0, c->stream, kbuf.current(), nnz, b,
                       col_dev);
    e = hipGetLastError();
  }
  if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
  if (e != hipSuccess) { cleanup(); mnhn_vcsr_free(out); return MNHN_ERR_DEVICE; }
  out->nnz = nnz;
  out->col = (int32_t *)host_alloc(std::max<size_t>((size_t)nnz, 1) * sizeof(int32_t));
  out->val = (float *)host_alloc(std::max<size_t>((size_t)nnz, 1) * sizeof(float));
  if (!out->col || !out->val) { cleanup(); mnhn_vcsr_free(out); return MNHN_ERR_NOMEM; }
  rc = copy_d2h(c->stream, c->stage, out->row_ptr, d_rp, ((size_t)n + 1) * sizeof(long long));
  if (rc == MNHN_OK && nnz > 0) rc = copy_d2h(c->stream, c->stage, out->col, col_dev, (size_t)nnz * sizeof(int32_t));
  if (rc == MNHN_OK && nnz > 0)
    rc = copy_d2h(c->stream, c->stage, out->val, vbuf.current(), (size_t)nnz * sizeof(float));
  cleanup();
  if (rc != MNHN_OK) { mnhn_vcsr_free(out); return rc; }
  return MNHN_OK;
}

void mnhn_vcsr_free(mnhn_vcsr *csr) {
  if (!csr) return;
  host_free(csr->row_ptr);
  host_free(csr->col);
  host_free(csr->val);
  csr->row_ptr = nullptr;
  csr->col = nullptr;
  csr->val = nullptr;
  csr->nnz = 0;
}

int mnhn_vec_last_timing(const mnhn_vec_ctx *c, double *kernel_ms, double *pairs) {
  if (!c) return MNHN_ERR_ARG;
  if (kernel_ms) *kernel_ms = c->last_ms;
  if (pairs) *pairs = c->last_pairs;
  return MNHN_OK;
}

}  // extern "C"
