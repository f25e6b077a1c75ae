// proposer.hip -- learned candidate proposer (include/mnhn_proposer.h).
//
// Pipeline, all on the device:
//   5-mer profiles (kmer_profile_kernel)          1024-d, L1-normalised, bf16
//   encoder MLP 1024 -> 512 -> 256 -> 128           gemm_nt_kernel (bf16 MFMA
//                                                   16x16x32, fp32 accumulate)
//                                                   with fused bias + GELU
//   training: pair loss, backward GEMMs, Adam       fp32 master weights
//   pairwise embedding distance + candidates      pair_candidates_kernel
//                                                   (bf16 MFMA, fused radius
//                                                   test and emission)
//   exact verification                            the engine's int16 SW kernel
//                                                   over candidate groups,
//                                                   epsilon keys -> CSR
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cfloat>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <new>
#include <type_traits>
#include <random>
#include <vector>

#include "engine_internal.h"
#include "mnhn_proposer.h"

using namespace mnhn_detail;

namespace {

typedef __bf16 bf16;
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef int i32x4 __attribute__((ext_vector_type(4)));

constexpr int D0 = MNHN_KMER_DIM, D1 = 512, D2 = 256, D3 = MNHN_EMBED_DIM;
constexpr float kProfileScale = 64.0f;  // profile entries scaled from 1/L to O(0.1)

// ------------------------------------------------------------------ profiles
// One workgroup per sequence (grid-stride): 5-mer counts of the stored
// residues (sequence positions 1 .. L-1) in an LDS histogram, then L1-normalised.
__global__ __launch_bounds__(256) void kmer_profile_kernel(const uint32_t *__restrict__ words,
                                                           const uint32_t *__restrict__ woff,
                                                           const int32_t *__restrict__ len, int n,
                                                           bf16 *__restrict__ X, float *__restrict__ Xf) {
  __shared__ int hist[D0];
  for (int s = blockIdx.x; s < n; s += gridDim.x) {
    for (int c = threadIdx.x; c < D0; c += 256) hist[c] = 0;
    __syncthreads();
    const int R = len[s] > 1 ? len[s] - 1 : 0;
    const int nk = R >= 5 ? R - 4 : 0;
    const uint32_t wo = woff[s];
    for (int p = threadIdx.x; p < nk; p += 256) {
      uint32_t code = 0;
#pragma unroll
      for (int i = 0; i < 5; ++i) {
        const uint32_t r = (uint32_t)(p + i);
        code = (code << 2) | ((words[wo + (r >> 4)] >> (2u * (r & 15u))) & 3u);
      }
      atomicAdd(&hist[code], 1);
    }
    __syncthreads();
    const float inv = nk > 0 ? kProfileScale / (float)nk : 0.0f;
    for (int c = threadIdx.x; c < D0; c += 256) {
      const float v = (float)hist[c] * inv;
      X[(size_t)s * D0 + c] = (bf16)v;
      if (Xf) Xf[(size_t)s * D0 + c] = v / kProfileScale;
    }
    __syncthreads();
  }
}

// ---------------------------------------------------------------------- GEMM
// C[m][n] = sum_k A[m][k] * B[n][k]  (both operands k-contiguous), bf16 in,
// fp32 accumulate, 64 x 64 block tile, 4 waves of 32 x 32, MFMA 16x16x32.
// Requires K, lda, ldb multiples of 8 (16-byte rows).
enum { EPI_F32 = 0, EPI_BIAS_GELU = 1, EPI_BIAS_OUT = 2, EPI_GELU_BWD = 3 };

struct Epi {
  float *c32 = nullptr;       // fp32 output (F32: C, BIAS_GELU: pre-activation, BIAS_OUT: value)
  int ldc = 0;
  bf16 *c16 = nullptr;        // bf16 output (BIAS_GELU: gelu(pre), BIAS_OUT: value, GELU_BWD: dpre)
  int ldc16 = 0;
  const float *bias = nullptr;
  const float *pre = nullptr; // GELU_BWD: saved pre-activation
  int ldp = 0;
};

__device__ __forceinline__ float gelu_f(float x) { return 0.5f * x * (1.0f + erff(x * 0.70710678118f)); }
__device__ __forceinline__ float gelu_grad(float x) {
  const float cdf = 0.5f * (1.0f + erff(x * 0.70710678118f));
  const float pdf = 0.39894228040f * expf(-0.5f * x * x);
  return cdf + x * pdf;
}

template <int EPI>
__global__ __launch_bounds__(256) void gemm_nt_kernel(const bf16 *__restrict__ A, int lda,
                                                      const bf16 *__restrict__ B, int ldb, int M,
                                                      int N, int K, Epi ep) {
  constexpr int BM = 64, BN = 64, BK = 32, LK = BK + 8;
  __shared__ __attribute__((aligned(16))) bf16 sA[BM * LK];
  __shared__ __attribute__((aligned(16))) bf16 sB[BN * LK];
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  const int wm = w >> 1, wn = w & 1;
  const int m0 = blockIdx.y * BM, n0 = blockIdx.x * BN;
  f32x4 acc[2][2];
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j) acc[i][j] = (f32x4){0.f, 0.f, 0.f, 0.f};
  const int lr = tid >> 2, lk = (tid & 3) * 8;
  const int fr = lane & 15, fk = 8 * (lane >> 4);
  for (int k0 = 0; k0 < K; k0 += BK) {
    bf16x8 va, vb;
#pragma unroll
    for (int t = 0; t < 8; ++t) { va[t] = (bf16)0.0f; vb[t] = (bf16)0.0f; }
    if (m0 + lr < M && k0 + lk < K) va = *reinterpret_cast<const bf16x8 *>(A + (size_t)(m0 + lr) * lda + k0 + lk);
    if (n0 + lr < N && k0 + lk < K) vb = *reinterpret_cast<const bf16x8 *>(B + (size_t)(n0 + lr) * ldb + k0 + lk);
    __syncthreads();
    *reinterpret_cast<bf16x8 *>(sA + lr * LK + lk) = va;
    *reinterpret_cast<bf16x8 *>(sB + lr * LK + lk) = vb;
    __syncthreads();
    bf16x8 af[2], bfr[2];
#pragma unroll
    for (int i = 0; i < 2; ++i) af[i] = *reinterpret_cast<const bf16x8 *>(sA + (32 * wm + 16 * i + fr) * LK + fk);
#pragma unroll
    for (int j = 0; j < 2; ++j) bfr[j] = *reinterpret_cast<const bf16x8 *>(sB + (32 * wn + 16 * j + fr) * LK + fk);
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
      for (int j = 0; j < 2; ++j)
        acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(af[i], bfr[j], acc[i][j], 0, 0, 0);
  }
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int m = m0 + 32 * wm + 16 * i + 4 * (lane >> 4) + r;
        const int nn = n0 + 32 * wn + 16 * j + (lane & 15);
        if (m >= M || nn >= N) continue;
        const float v = acc[i][j][r];
        if constexpr (EPI == EPI_F32) {
          ep.c32[(size_t)m * ep.ldc + nn] = v;
        } else if constexpr (EPI == EPI_BIAS_GELU) {
          const float pre = v + ep.bias[nn];
          if (ep.c32) ep.c32[(size_t)m * ep.ldc + nn] = pre;
          ep.c16[(size_t)m * ep.ldc16 + nn] = (bf16)gelu_f(pre);
        } else if constexpr (EPI == EPI_BIAS_OUT) {
          const float o = v + ep.bias[nn];
          if (ep.c32) ep.c32[(size_t)m * ep.ldc + nn] = o;
          if (ep.c16) ep.c16[(size_t)m * ep.ldc16 + nn] = (bf16)o;
        } else {
          ep.c16[(size_t)m * ep.ldc16 + nn] = (bf16)(v * gelu_grad(ep.pre[(size_t)m * ep.ldp + nn]));
        }
      }
}

// bf16 [R][C] -> [C][R]
__global__ void transpose_bf16_kernel(const bf16 *__restrict__ in, int R, int C, bf16 *__restrict__ out) {
  __shared__ bf16 t[32][33];
  const int c0 = blockIdx.x * 32, r0 = blockIdx.y * 32;
  for (int k = threadIdx.y; k < 32; k += blockDim.y) {
    const int r = r0 + k, c = c0 + threadIdx.x;
    if (r < R && c < C) t[k][threadIdx.x] = in[(size_t)r * C + c];
  }
  __syncthreads();
  for (int k = threadIdx.y; k < 32; k += blockDim.y) {
    const int c = c0 + k, r = r0 + threadIdx.x;
    if (r < R && c < C) out[(size_t)c * R + r] = t[threadIdx.x][k];
  }
}

__global__ void gather_rows_kernel(const bf16 *__restrict__ X, const int32_t *__restrict__ idx, int rows,
                                   bf16 *__restrict__ out) {
  const int r = blockIdx.x;
  if (r >= rows) return;
  const bf16x8 *src = reinterpret_cast<const bf16x8 *>(X + (size_t)idx[r] * D0);
  bf16x8 *dst = reinterpret_cast<bf16x8 *>(out + (size_t)r * D0);
  for (int t = threadIdx.x; t < D0 / 8; t += blockDim.x) dst[t] = src[t];
}

// One wave per pair: ||e_a - e_b|| regressed onto the normalised distance,
// plus a hinge keeping neighbours inside their radius r and, with fw > 0, one
// keeping non-neighbours beyond r + gap.  radius[k] = r >= 0 marks a
// neighbour, radius[k] = -1 - r a non-neighbour.
__global__ __launch_bounds__(256) void pair_loss_kernel(const float *__restrict__ emb,
                                                        const float *__restrict__ target,
                                                        const float *__restrict__ radius, int npairs,
                                                        float nbw, float fw, float gap, float inv_batch,
                                                        bf16 *__restrict__ demb, float *__restrict__ loss) {
  const int lane = threadIdx.x & 63;
  const int k = blockIdx.x * 4 + (threadIdx.x >> 6);
  if (k >= npairs) return;
  const float *ea = emb + (size_t)(2 * k) * D3, *eb = emb + (size_t)(2 * k + 1) * D3;
  const float e0 = ea[lane] - eb[lane], e1 = ea[lane + 64] - eb[lane + 64];
  float ss = e0 * e0 + e1 * e1;
#pragma unroll
  for (int o = 32; o >= 1; o >>= 1) ss += __shfl_xor(ss, o);
  const float dist = sqrtf(ss + 1e-12f);
  const float diff = dist - target[k];
  float L = diff * diff, g = 2.0f * diff;
  const float rad = radius[k];
  if (rad >= 0.0f && dist > rad) {
    const float h = dist - rad;
    L += nbw * h * h;
    g += 2.0f * nbw * h;
  } else if (rad < 0.0f && fw > 0.0f) {
    const float h = (-1.0f - rad) + gap - dist;
    if (h > 0.0f) {
      L += fw * h * h;
      g -= 2.0f * fw * h;
    }
  }
  const float s = g * inv_batch / dist;
  bf16 *da = demb + (size_t)(2 * k) * D3, *db = demb + (size_t)(2 * k + 1) * D3;
  da[lane] = (bf16)(s * e0);
  da[lane + 64] = (bf16)(s * e1);
  db[lane] = (bf16)(-s * e0);
  db[lane + 64] = (bf16)(-s * e1);
  if (lane == 0) loss[k] = L;  // per pair; the host adds them in order
}

// part[y][c] = sum over the rows of block row y (128 rows) of dY[r][c].
// 256 threads = 64 columns x 4 row lanes per block; each sum runs in a fixed
// order, and col_reduce_kernel adds the blocks' partials in block order, so
// the column sums (the bias gradients) are the same on every run.
__global__ __launch_bounds__(256) void col_sum_kernel(const bf16 *__restrict__ dY, int R, int C,
                                                      float *__restrict__ part_out) {
  __shared__ float part[4][64];
  const int cl = threadIdx.x & 63, rl = threadIdx.x >> 6;
  const int c = blockIdx.x * 64 + cl;
  const int r0 = blockIdx.y * 128, r1 = min(R, r0 + 128);
  float s = 0.f;
  if (c < C)
    for (int r = r0 + rl; r < r1; r += 4) s += (float)dY[(size_t)r * C + c];
  part[rl][cl] = s;
  __syncthreads();
  if (rl == 0 && c < C) part_out[(size_t)blockIdx.y * C + c] = part[0][cl] + part[1][cl] + part[2][cl] + part[3][cl];
}

// out[c] = sum over y < nb of part[y][c], in order of y
__global__ void col_reduce_kernel(const float *__restrict__ part, int nb, int C, float *__restrict__ out) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= C) return;
  float s = 0.f;
  for (int y = 0; y < nb; ++y) s += part[(size_t)y * C + c];
  out[c] = s;
}

// Adam with decoupled weight decay wd (0 for biases)
__global__ void adam_kernel(float *__restrict__ p, const float *__restrict__ g, float *__restrict__ m,
                            float *__restrict__ v, int n, float lr, float b1, float b2, float c1,
                            float c2, float wd) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float gi = g[i];
  const float mi = b1 * m[i] + (1.f - b1) * gi;
  const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
  m[i] = mi;
  v[i] = vi;
  const float pi = p[i];
  p[i] = pi - lr * ((mi / c1) / (sqrtf(vi / c2) + 1e-8f) + wd * pi);
}

// fp32 master Wt [out][in] -> bf16 Wt [out][in] and bf16 W [in][out]
__global__ void refresh_weights_kernel(const float *__restrict__ wt, int out_d, int in_d,
                                       bf16 *__restrict__ wt16, bf16 *__restrict__ w16) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= out_d * in_d) return;
  const int o = i / in_d, c = i - o * in_d;
  const bf16 b = (bf16)wt[i];
  wt16[i] = b;
  w16[(size_t)c * out_d + o] = b;
}

// -------------------------------------------------------- pairwise proposals
// The pairwise stage works in length order (longest first): row p of Es is
// sequence order[p].  A pair (p, q), p < q, is proposed when
//   match |l_p - l_q| <= eps                        (exact bound: d >= match |dl|)
//   ||e_p - e_q||^2 <= rad^2,  rad = eps / (match l_long) + margin(l_long)
// where l_long is the longer length of the pair and margin(l) comes from the
// encoder's margin table (first bucket whose upper length bound holds l).
constexpr int kMaxMarginBuckets = 16;
struct MarginTable {
  int n;
  int upper[kMaxMarginBuckets];
  float margin[kMaxMarginBuckets];
};

__device__ __forceinline__ float bucket_value(const MarginTable &t, int L) {
  float m = t.margin[t.n - 1];
  for (int k = t.n - 1; k >= 0; --k)
    if (L <= t.upper[k]) m = t.margin[k];
  return m;
}

// One wave per sorted row: copy of the bf16 embedding, its squared norm (the
// same bf16 values the MFMA tiles read), length, the row's radius term and its
// minimum profile cosine (both taken from the row's length bucket).
__global__ __launch_bounds__(256) void sorted_rows_kernel(const bf16 *__restrict__ E,
                                                          const int32_t *__restrict__ order,
                                                          const int32_t *__restrict__ len, int n,
                                                          float eps_over_match, MarginTable mt, MarginTable ct,
                                                          bf16 *__restrict__ Es, float *__restrict__ norm,
                                                          float *__restrict__ srad, float *__restrict__ cmin,
                                                          int32_t *__restrict__ lens) {
  const int lane = threadIdx.x & 63;
  const int p = blockIdx.x * 4 + (threadIdx.x >> 6);
  if (p >= n) return;
  const int id = order[p];
  const bf16 x0 = E[(size_t)id * D3 + lane], x1 = E[(size_t)id * D3 + lane + 64];
  Es[(size_t)p * D3 + lane] = x0;
  Es[(size_t)p * D3 + lane + 64] = x1;
  float ss = (float)x0 * (float)x0 + (float)x1 * (float)x1;
#pragma unroll
  for (int o = 32; o >= 1; o >>= 1) ss += __shfl_xor(ss, o);
  if (lane == 0) {
    const int L = len[id];
    norm[p] = ss;
    lens[p] = L;
    srad[p] = eps_over_match / (float)max(L, 1) + bucket_value(mt, L);
    cmin[p] = bucket_value(ct, L);
  }
}

// Raw 5-mer counts of every sequence in length order (row p = sequence
// order[p]) and their norms, for the count cosine of the pairwise stage.
// T = bf16 (integers exact up to 256), or
// int8 for the int8 MFMA stage: counts above 127 set *over (the host then
// builds the bf16 rows instead).
template <typename T>
__global__ __launch_bounds__(256) void kmer_counts_sorted_kernel(const uint32_t *__restrict__ words,
                                                                 const uint32_t *__restrict__ woff,
                                                                 const int32_t *__restrict__ len,
                                                                 const int32_t *__restrict__ order, int n,
                                                                 T *__restrict__ Xs, float *__restrict__ pnorm,
                                                                 int *__restrict__ over) {
  __shared__ int hist[D0];
  __shared__ float red[256];
  for (int p = blockIdx.x; p < n; p += gridDim.x) {
    const int s = order[p];
    for (int c = threadIdx.x; c < D0; c += 256) hist[c] = 0;
    __syncthreads();
    const int R = len[s] > 1 ? len[s] - 1 : 0;
    const int nk = R >= 5 ? R - 4 : 0;
    const uint32_t wo = woff[s];
    for (int q = threadIdx.x; q < nk; q += 256) {
      uint32_t code = 0;
#pragma unroll
      for (int i = 0; i < 5; ++i) {
        const uint32_t r = (uint32_t)(q + i);
        code = (code << 2) | ((words[wo + (r >> 4)] >> (2u * (r & 15u))) & 3u);
      }
      atomicAdd(&hist[code], 1);
    }
    __syncthreads();
    float ss = 0.f;
    for (int c = threadIdx.x; c < D0; c += 256) {
      const int h = hist[c];
      float f;
      if constexpr (std::is_same<T, int8_t>::value) {
        if (h > 127) atomicOr(over, 1);
        Xs[(size_t)p * D0 + c] = (int8_t)min(h, 127);
        f = (float)h;
      } else {
        const bf16 v = (bf16)(float)h;
        Xs[(size_t)p * D0 + c] = v;
        f = (float)v;
      }
      ss += f * f;
    }
    red[threadIdx.x] = ss;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
      if ((int)threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
      __syncthreads();
    }
    if (threadIdx.x == 0) pnorm[p] = sqrtf(red[0]);
    __syncthreads();
  }
}

// Buffer resource over rows [r0, r0 + 128) of a row-major matrix with n rows
// of row_bytes each.  Its record count ends at row n, so loads of rows past n
// return zeros (the hardware's range check) and need no clamp; the 32-bit
// offsets stay below 128 rows whatever n is.  0x00020000: the third word of a
// raw buffer on gfx9-family parts.
__device__ __forceinline__ __amdgpu_buffer_rsrc_t tile_rows_rsrc(const void *base, int r0, int n, int row_bytes) {
  const char *p = static_cast<const char *>(base) + (size_t)r0 * row_bytes;
  return __builtin_amdgcn_make_buffer_rsrc(const_cast<char *>(p), (short)0, min(n - r0, 128) * row_bytes, 0x00020000);
}

// acc += A[i0 .. i0+128) . B[j0 .. j0+128)^T over k in [0, K) (row stride K),
// staged through LDS 64 k at a time; each wave owns a 64 x 64 sub-tile.
// Global loads run two chunks ahead of the MFMAs (below).
template <int K>
__device__ __forceinline__ void tile_dot(const bf16 *__restrict__ A, const bf16 *__restrict__ B, int n,
                                         int i0, int j0, bf16 *sA, bf16 *sB, f32x4 (&acc)[4][4]) {
  // sA / sB hold two stages of 128 x (64 + 8) each: chunk c computes from
  // stage c & 1 while chunk c + 1 is loaded and stored into the other, so one
  // barrier per chunk suffices
  constexpr int KH = 64, LK = KH + 8, STG = 128 * LK;
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  const int wm = w >> 1, wn = w & 1;
  const int fr = lane & 15, fk = 8 * (lane >> 4);
  // 128 rows x 64 k = 1024 chunks of 8 per operand; 4 per thread (thread
  // tid takes rows tid / 8 + 32 t, k (tid % 8) * 8).  Rows past n read zeros
  // (tile_rows_rsrc): their products are never used, as the callers' tests
  // reject every row or column at or beyond n.  One 32-bit offset per thread;
  // the row step and the chunk's k go into the scalar offset.
  // two register sets: chunk c + 2 is loaded while chunk c is multiplied
  // and chunk c + 1 (loaded one chunk earlier) goes into the other LDS
  // stage, so a load has two chunks of math and barriers to arrive.  Chunk k
  // travels in register set k & 1 and LDS stage k & 1.
  constexpr int NC = K / KH, RB = K * 2;
  static_assert(K % KH == 0 && NC % 2 == 0, "whole pairs of chunks");
  const __amdgpu_buffer_rsrc_t ra = tile_rows_rsrc(A, i0, n, RB), rb = tile_rows_rsrc(B, j0, n, RB);
  const int voff = (tid >> 3) * RB + (tid & 7) * 16;
  bf16x8 va0[4], vb0[4], va1[4], vb1[4];
  auto load = [&](bf16x8 (&va)[4], bf16x8 (&vb)[4], int kh) {
#pragma unroll
    for (int t = 0; t < 4; ++t) {
      va[t] = __builtin_bit_cast(bf16x8, __builtin_amdgcn_raw_buffer_load_b128(ra, voff, 32 * t * RB + 2 * kh, 0));
      vb[t] = __builtin_bit_cast(bf16x8, __builtin_amdgcn_raw_buffer_load_b128(rb, voff, 32 * t * RB + 2 * kh, 0));
    }
  };
  auto store = [&](const bf16x8 (&va)[4], const bf16x8 (&vb)[4], int stage) {
#pragma unroll
    for (int t = 0; t < 4; ++t) {
      const int ch = tid + 256 * t;
      const int r = ch >> 3, k = (ch & 7) * 8;
      *reinterpret_cast<bf16x8 *>(sA + stage * STG + r * LK + k) = va[t];
      *reinterpret_cast<bf16x8 *>(sB + stage * STG + r * LK + k) = vb[t];
    }
  };
  auto mfma = [&](int stage) {
    const bf16 *a = sA + stage * STG, *b = sB + stage * STG;
#pragma unroll
    for (int ks = 0; ks < KH; ks += 32) {
      bf16x8 af[4], bfr[4];
#pragma unroll
      for (int i = 0; i < 4; ++i) af[i] = *reinterpret_cast<const bf16x8 *>(a + (64 * wm + 16 * i + fr) * LK + ks + fk);
#pragma unroll
      for (int j = 0; j < 4; ++j) bfr[j] = *reinterpret_cast<const bf16x8 *>(b + (64 * wn + 16 * j + fr) * LK + ks + fk);
#pragma unroll
      for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j)
          acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(af[i], bfr[j], acc[i][j], 0, 0, 0);
    }
  };
  __syncthreads();  // the previous user of both stages is done
  load(va0, vb0, 0);
  load(va1, vb1, KH);
  store(va0, vb0, 0);
  __syncthreads();
#pragma unroll
  for (int c = 0; c < NC; c += 2) {
    if (c + 2 < NC) load(va0, vb0, (c + 2) * KH);  // set 0 held chunk c, already in stage 0
    __builtin_amdgcn_sched_barrier(0);  // issue the loads here, ahead of the math
    mfma(0);
    store(va1, vb1, 1);  // chunk c + 1
    __syncthreads();
    if (c + 3 < NC) load(va1, vb1, (c + 3) * KH);
    __builtin_amdgcn_sched_barrier(0);
    mfma(1);
    if (c + 2 < NC) store(va0, vb0, 0);  // chunk c + 2
    __syncthreads();
  }
}

// tile_dot over int8 operands (raw 5-mer counts <= 127) on the int8 MFMA
// (v_mfma_i32_16x16x64_i8): int32 sums, exact.  Chunks of 128 bytes per row
// (two MFMA k-steps of 64) in the same LDS footprint, double buffering and
// two-chunk-ahead loads as the bf16 form, so a chunk moves the same bytes for
// twice the k.  Rows past n read zeros, as in tile_dot.
template <int K>
__device__ __forceinline__ void tile_dot_i8(const int8_t *__restrict__ A, const int8_t *__restrict__ B, int n,
                                            int i0, int j0, int8_t *sA, int8_t *sB, i32x4 (&acc)[4][4]) {
  constexpr int KH = 128, LK = KH + 16, STG = 128 * LK;  // bytes
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  const int wm = w >> 1, wn = w & 1;
  const int fr = lane & 15, fk = 16 * (lane >> 4);
  // 128 rows x 128 bytes = 1024 pieces of 16 bytes per operand; 4 per thread
  // (rows tid / 8 + 32 t, bytes (tid % 8) * 16), addressed as in tile_dot
  constexpr int NC = K / KH;  // two register sets, as in tile_dot
  static_assert(K % KH == 0 && NC % 2 == 0, "whole pairs of chunks");
  const __amdgpu_buffer_rsrc_t ra = tile_rows_rsrc(A, i0, n, K), rb = tile_rows_rsrc(B, j0, n, K);
  const int voff = (tid >> 3) * K + (tid & 7) * 16;
  i32x4 va0[4], vb0[4], va1[4], vb1[4];  // native vectors: registers, not a stack object
  auto load = [&](i32x4 (&va)[4], i32x4 (&vb)[4], int kh) {
#pragma unroll
    for (int t = 0; t < 4; ++t) {
      va[t] = __builtin_bit_cast(i32x4, __builtin_amdgcn_raw_buffer_load_b128(ra, voff, 32 * t * K + kh, 0));
      vb[t] = __builtin_bit_cast(i32x4, __builtin_amdgcn_raw_buffer_load_b128(rb, voff, 32 * t * K + kh, 0));
    }
  };
  auto store = [&](const i32x4 (&va)[4], const i32x4 (&vb)[4], int stage) {
#pragma unroll
    for (int t = 0; t < 4; ++t) {
      const int ch = tid + 256 * t;
      const int r = ch >> 3, k = (ch & 7) * 16;
      *reinterpret_cast<i32x4 *>(sA + stage * STG + r * LK + k) = va[t];
      *reinterpret_cast<i32x4 *>(sB + stage * STG + r * LK + k) = vb[t];
    }
  };
  auto mfma = [&](int stage) {
    const int8_t *a = sA + stage * STG, *b = sB + stage * STG;
#pragma unroll
    for (int ks = 0; ks < KH; ks += 64) {
      i32x4 af[4], bfr[4];
#pragma unroll
      for (int i = 0; i < 4; ++i) af[i] = *reinterpret_cast<const i32x4 *>(a + (64 * wm + 16 * i + fr) * LK + ks + fk);
#pragma unroll
      for (int j = 0; j < 4; ++j) bfr[j] = *reinterpret_cast<const i32x4 *>(b + (64 * wn + 16 * j + fr) * LK + ks + fk);
#pragma unroll
      for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j)
          acc[i][j] = __builtin_amdgcn_mfma_i32_16x16x64_i8(af[i], bfr[j], acc[i][j], 0, 0, 0);
    }
  };
  __syncthreads();  // the previous user of both stages is done
  load(va0, vb0, 0);
  load(va1, vb1, KH);
  store(va0, vb0, 0);
  __syncthreads();
#pragma unroll
  for (int c = 0; c < NC; c += 2) {
    if (c + 2 < NC) load(va0, vb0, (c + 2) * KH);  // set 0 held chunk c, already in stage 0
    __builtin_amdgcn_sched_barrier(0);  // issue the loads here, ahead of the math
    mfma(0);
    store(va1, vb1, 1);  // chunk c + 1
    __syncthreads();
    if (c + 3 < NC) load(va1, vb1, (c + 3) * KH);
    __builtin_amdgcn_sched_barrier(0);
    mfma(1);
    if (c + 2 < NC) store(va0, vb0, 0);  // chunk c + 2
    __syncthreads();
  }
}

// 128 x 128 tiles over the length band: the host lists the band tiles (ti, tj)
// (row tile ti, column tiles ti .. that the length bound can reach) in
// tile_list, in blocks of kTileBlock row tiles and, within a block, by column
// tile: consecutive tiles share their column operand and cycle through the
// block's row operands.  4 waves of 64 x 64 per tile.  Stage 1: Es Es^T
// (K = 128) and the radius test.  Stage 2 (kCos, skipped for a tile with no
// stage-1 survivor): Xs Xs^T over the 1024 raw 5-mer counts and the cosine
// test dot >= cmin * |x_p| |x_q|, cmin = cmin of the longer sequence (< 0:
// no test).  Pairs whose longer sequence is at most exact_len long are left
// to the exact pass over the shortest sequences (exact_len < 0: none).  Hits
// are counted first and reserved with one atomic per wave, then written
// (sorted positions p << 32 | q) at consecutive addresses.
constexpr int kTileBlock = 8;  // row tiles per block of the band-tile order

template <bool kCos, bool kI8>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2))) void pair_candidates_kernel(
    const bf16 *__restrict__ Es, const float *__restrict__ norm, const float *__restrict__ srad,
    const void *__restrict__ Xv, const float *__restrict__ pnorm, const float *__restrict__ cmin,
    const int32_t *__restrict__ lens, int n, const int2 *__restrict__ tile_list, int max_dl, int exact_len,
    unsigned long long *__restrict__ out, unsigned long long *__restrict__ count, unsigned long long cap,
    unsigned long long *__restrict__ stage2) {
  constexpr int BT = 128, LK = 64 + 8;
  __shared__ __attribute__((aligned(16))) bf16 sA[2 * BT * LK];
  __shared__ __attribute__((aligned(16))) bf16 sB[2 * BT * LK];
  __shared__ int any_hit;
  // Workgroups are dealt to the 8 XCDs in turn (workgroup b to XCD b % 8),
  // each XCD with its own L2.  Consecutive tiles of the list share operands
  // (the workgroups resident on one XCD at a time span a few column tiles and
  // at most kTileBlock row tiles, whose operands fit its L2), so each XCD
  // takes a contiguous run of the list: workgroup b computes tile
  // (b % 8) * run + b / 8 (runs of per or per + 1 tiles).
  int blk;
  {
    constexpr int kXcds = 8;
    const int nb = (int)gridDim.x, per = nb / kXcds, rem = nb % kXcds;
    const int x = (int)blockIdx.x % kXcds, k = (int)blockIdx.x / kXcds;
    blk = x < rem ? x * (per + 1) + k : rem * (per + 1) + (x - rem) * per + k;
  }
  const int2 tt = tile_list[blk];
  const int ti = __builtin_amdgcn_readfirstlane(tt.x), tj = __builtin_amdgcn_readfirstlane(tt.y);
  const int i0 = ti * BT, j0 = tj * BT;
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  const int wm = w >> 1, wn = w & 1;
  if (tid == 0) any_hit = 0;
  // Stage 1 is folded into two 64-bit masks per lane (bit 16 i + 4 r + j):
  // the radius test, and which side is the longer sequence; its accumulators
  // and row / column terms are dead before stage 2 starts.  In an
  // off-diagonal tile every row precedes every column in length order, so the
  // row is the longer sequence of each pair (equal lengths have equal radii
  // and cosine bounds): its radius and bound decide, and the length test
  // 0 <= l_i - l_j <= max_dl is one unsigned compare.  Rows and columns past n
  // get lengths that make l_i - l_j negative.
  const bool diag = ti == tj;
  unsigned long long m1 = 0ull, mlong = 0ull;
  {
    f32x4 acc[4][4];
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
      for (int j = 0; j < 4; ++j) acc[i][j] = (f32x4){0.f, 0.f, 0.f, 0.f};
    tile_dot<D3>(Es, Es, n, i0, j0, sA, sB, acc);
    // row / column terms are loaded at clamped indices (no branches)
    float nj[4], sj2[4];
    int lj[4];
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const int gj = j0 + 64 * wn + 16 * j + (lane & 15);
      const int cj = min(gj, n - 1);
      const float sj = srad[cj];
      nj[j] = norm[cj];
      sj2[j] = sj * sj;
      lj[j] = gj < n ? lens[cj] : (1 << 29);
    }
    if (!diag) {
      // |e_i - e_j|^2 <= s_i^2  <=>  n_j - 2 e_i.e_j <= s_i^2 - n_i
#pragma unroll
      for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const int gi = i0 + 64 * wm + 16 * i + 4 * (lane >> 4) + r;
          const int ci = min(gi, n - 1);
          const float si = srad[ci];
          const int li = gi < n ? lens[ci] : -(1 << 28);
          const float ri = li > exact_len ? si * si - norm[ci] : -FLT_MAX;
#pragma unroll
          for (int j = 0; j < 4; ++j) {
            const bool h = (unsigned)(li - lj[j]) <= (unsigned)max_dl && nj[j] - 2.0f * acc[i][j][r] <= ri;
            m1 |= h ? 1ull << (16 * i + 4 * r + j) : 0ull;
          }
        }
      mlong = ~0ull;
    } else {
#pragma unroll
      for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const int gi = i0 + 64 * wm + 16 * i + 4 * (lane >> 4) + r;
          const int ci = min(gi, n - 1);
          const float ni = norm[ci];
          const float si = srad[ci];
          const float si2 = si * si;
          const int li = gi < n ? lens[ci] : -(1 << 28);
#pragma unroll
          for (int j = 0; j < 4; ++j) {
            const int gj = j0 + 64 * wn + 16 * j + (lane & 15);
            const int dl = li - lj[j];
            const float rad2 = dl >= 0 ? si2 : sj2[j];
            const float d2 = ni + nj[j] - 2.0f * acc[i][j][r];
            const bool h = gi < gj && abs(dl) <= max_dl && d2 <= rad2 && max(li, lj[j]) > exact_len;
            const int bit = 16 * i + 4 * r + j;
            m1 |= h ? 1ull << bit : 0ull;
            mlong |= dl >= 0 ? 1ull << bit : 0ull;
          }
        }
    }
  }
  unsigned long long m2 = m1;  // hits after both stages
  if constexpr (kCos) {
    if (__ballot(m1 != 0ull) != 0ull && lane == 0) any_hit = 1;
    __syncthreads();
    if (any_hit == 0) return;  // uniform over the block
    if (tid == 0) atomicAdd(stage2, 1ull);
    // cosine test dot >= c |x_i| |x_j| with c the longer sequence's bound
    // (c < 0: no test); pnorm holds the norms |x| of the count rows, and the
    // counts are nonnegative, so dot >= 0
    f32x4 accx[4][4];
    if constexpr (kI8) {
      i32x4 acci[4][4];
#pragma unroll
      for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j) acci[i][j] = (i32x4){0, 0, 0, 0};
      const int8_t *X8 = static_cast<const int8_t *>(Xv);
      tile_dot_i8<D0>(X8, X8, n, i0, j0, reinterpret_cast<int8_t *>(sA), reinterpret_cast<int8_t *>(sB), acci);
#pragma unroll
      for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j)
#pragma unroll
          for (int r = 0; r < 4; ++r) accx[i][j][r] = (float)acci[i][j][r];
    } else {
#pragma unroll
      for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j) accx[i][j] = (f32x4){0.f, 0.f, 0.f, 0.f};
      const bf16 *Xs = static_cast<const bf16 *>(Xv);
      tile_dot<D0>(Xs, Xs, n, i0, j0, sA, sB, accx);
    }
    float vj[4], uj[4];  // |x_j| and the column's bound times it (-FLT_MAX: no test)
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const int gj = j0 + 64 * wn + 16 * j + (lane & 15);
      const int cc = min(gj, n - 1);
      const float c = gj < n ? cmin[cc] : -1.f;
      vj[j] = pnorm[cc];
      uj[j] = c >= 0.0f ? c * vj[j] : -FLT_MAX;
    }
    // off-diagonal tiles (kDiag false) take the row's bound throughout; the
    // branch is uniform and taken once per tile
    auto cos_test = [&](auto kDiag) {
#pragma unroll
      for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const int gi = i0 + 64 * wm + 16 * i + 4 * (lane >> 4) + r;
          const int cc = min(gi, n - 1);
          const float vi = pnorm[cc];
          const float c = gi < n ? cmin[cc] : -1.f;
          const float ui = c >= 0.0f ? c * vi : -FLT_MAX;
#pragma unroll
          for (int j = 0; j < 4; ++j) {
            const int bit = 16 * i + 4 * r + j;
            const float thr = decltype(kDiag)::value && !((mlong >> bit) & 1ull) ? uj[j] * vi : ui * vj[j];
            m2 &= accx[i][j][r] < thr ? ~(1ull << bit) : ~0ull;
          }
        }
    };
    if (diag) cos_test(std::integral_constant<bool, true>{});
    else cos_test(std::integral_constant<bool, false>{});
  }
  // Emission: each lane's hit count, a wave-wide prefix sum of the counts,
  // one reservation per wave, then each lane writes its own hits (ascending
  // bits) at consecutive positions.  Candidates are later grouped by row and
  // sorted within rows, so their order in the buffer does not matter.
  // The prefix sum is taken digit by digit: counts are at most 64, so
  // prefix = sum over b of 2^b * (lanes below with bit b of their count set).
  const unsigned cnt = (unsigned)__popcll(m2);
  unsigned below = 0, total = 0;
#pragma unroll
  for (int b = 0; b < 7; ++b) {
    const unsigned long long bal = __ballot((cnt >> b) & 1u);
    below += __builtin_amdgcn_mbcnt_hi((unsigned)(bal >> 32), __builtin_amdgcn_mbcnt_lo((unsigned)bal, 0u)) << b;
    total += (unsigned)__popcll(bal) << b;
  }
  if (total == 0) return;
  unsigned long long base = 0;
  if (lane == 0) base = atomicAdd(count, (unsigned long long)total);
  base = uniform_u64(base);
  unsigned long long pos = base + below;
  unsigned long long m = m2;
  while (m) {
    const int bit = __builtin_ctzll(m);
    m &= m - 1ull;
    const int i = bit >> 4, r = (bit >> 2) & 3, j = bit & 3;
    const int gi = i0 + 64 * wm + 16 * i + 4 * (lane >> 4) + r;
    const int gj = j0 + 64 * wn + 16 * j + (lane & 15);
    if (pos < cap) out[pos] = ((unsigned long long)gi << 32) | (unsigned)gj;
    ++pos;
  }
}

// Wave-aggregated per-row atomics: candidates next to each other in the
// buffer mostly share a row (a wave's 64 hits of one emission step cover at
// most 4 rows), so each distinct row of the wave takes one atomic; every lane
// gets its own slot (base + its rank among the lanes of its row).
__device__ __forceinline__ int row_slot(int32_t *__restrict__ ctr, int row, bool valid) {
  const int lane = threadIdx.x & 63;
  unsigned long long todo = __ballot(valid);
  int slot = 0;
  while (todo) {
    const int leader = __ffsll((long long)todo) - 1;
    const int lrow = __shfl(row, leader);
    const unsigned long long same = __ballot(valid && row == lrow);
    int base = 0;
    if (lane == leader) base = atomicAdd(&ctr[lrow], __popcll(same));
    base = __shfl(base, leader);
    if (valid && row == lrow)
      slot = base + (int)__builtin_amdgcn_mbcnt_hi((unsigned)(same >> 32),
                                                   __builtin_amdgcn_mbcnt_lo((unsigned)same, 0u));
    todo &= ~same;
  }
  return slot;
}

// candidate keys (sorted positions p << 32 | q) -> per-row counts of the
// longer sequence's id order[p].  The loop runs whole waves to the end (the
// aggregation needs every lane).
__global__ void cand_count_kernel(const unsigned long long *__restrict__ c, long long m,
                                  const int32_t *__restrict__ order, int32_t *__restrict__ cnt) {
  const long long stride = (long long)gridDim.x * blockDim.x;
  const long long first = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const long long wave0 = first - (threadIdx.x & 63);
  for (long long w0 = wave0; w0 < m; w0 += stride) {
    const long long k = w0 + (threadIdx.x & 63);
    const bool valid = k < m;
    const int i = valid ? order[(int)(c[k] >> 32)] : 0;
    (void)row_slot(cnt, i, valid);
  }
}

// scatter candidates (sorted positions p << 32 | q) into the row segments of
// the longer sequence's id order[p]: seg[rs[i] ..] holds the q of row i
__global__ void cand_scatter_kernel(const unsigned long long *__restrict__ c, long long m,
                                    const int32_t *__restrict__ order, const long long *__restrict__ rs,
                                    int32_t *__restrict__ cursor, int32_t *__restrict__ seg) {
  const long long stride = (long long)gridDim.x * blockDim.x;
  const long long first = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const long long wave0 = first - (threadIdx.x & 63);
  for (long long w0 = wave0; w0 < m; w0 += stride) {
    const long long k = w0 + (threadIdx.x & 63);
    const bool valid = k < m;
    const unsigned long long key = valid ? c[k] : 0ull;
    const int i = valid ? order[(int)(key >> 32)] : 0;
    const int slot = row_slot(cursor, i, valid);
    if (valid) seg[rs[i] + slot] = (int32_t)(key & 0xffffffffull);
  }
}

// Per row i, the targets' length bands of its candidates (ascending q =
// descending target length, so each band is a run): kband[2i] targets longer
// than t2 (pairs beyond the 16-bit lanes), kband[2i + 1] targets longer than
// t1 (pairs beyond the exact fp16 form).  The row's sequence is the longer of
// each of its pairs, so the target's length bounds the pair's score.
__global__ void band_counts_kernel(const int32_t *__restrict__ sorted, const long long *__restrict__ rs,
                                   const int32_t *__restrict__ order, const int32_t *__restrict__ len, int n,
                                   int t1, int t2, int32_t *__restrict__ kband) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const long long b = rs[i];
  const int cnt = (int)(rs[i + 1] - b);
  auto first_at_most = [&](int t) {  // first k with len <= t
    int lo = 0, hi = cnt;
    while (lo < hi) {
      const int mid = (lo + hi) >> 1;
      if (len[order[sorted[b + mid]]] > t) lo = mid + 1;
      else hi = mid;
    }
    return lo;
  };
  kband[2 * i] = first_at_most(t2);
  kband[2 * i + 1] = first_at_most(t1);
}

// A row's k targets of band t in target groups: `full` groups of 128, then,
// when groups are shared (split >= 1) and the band is the int16 (t = 1) or
// the fp16 one (t = 2), the remainder in pieces of shared groups: a half
// piece of up to 64 targets (two rows per group, one per half of the lanes'
// packed pairs) and, with split = 2, a quarter piece of up to 32 (four rows
// per group), chosen so that fewer than 32 of the row's slots stay empty:
// remainder r in (96, 128): one more full group; (64, 96]: a full half piece
// and r - 64 in a quarter; (32, 64]: a half piece; (0, 32]: a quarter.
// split = 1: half pieces only (r <= 64).
struct Pieces {
  int full, half, quarter;  // full groups; targets in the half / quarter piece
};
__device__ __host__ inline Pieces band_pieces(int k, int t, int split) {
  const int r = k % kGroupSize, f = k / kGroupSize;
  if (!split || t == 0 || r == 0) return {(k + kGroupSize - 1) / kGroupSize, 0, 0};
  if (split == 1) return r <= kGroupSize / 2 ? Pieces{f, r, 0} : Pieces{f + 1, 0, 0};
  if (r > 3 * kGroupSize / 4) return {f + 1, 0, 0};
  if (r > kGroupSize / 2) return {f, kGroupSize / 2, r - kGroupSize / 2};
  if (r > kGroupSize / 4) return {f, r, 0};
  return {f, 0, r};
}

// One workgroup per row i: its ascending q (= descending target length) in
// target groups of 128 per length band (band 0: k < kband[2i], band 1: up to
// kband[2i + 1], band 2: the rest), so each group holds targets of similar
// length and one lane format; the group's first target is its longest, its
// last the shortest.  The shared pieces (band_pieces) follow the band's full
// groups: piece h = hgs[4i + x] (x = 0: fp16 half, 1: int16 half, 2: fp16
// quarter, 3: int16 quarter) is sub-list h & 3 of the shared group h >> 2;
// its u-th target goes to the slot the kernels read for it: half s:
// s * 64 + u; quarter s: (s & 1) * 64 + (s >> 1) * 32 + u.  The rows of a
// shared group set its length bounds by atomics (gmax / gmin of shared
// groups start at 0 / INT_MAX).  tlen[7i + band] = sum of the band's target
// lengths in full groups, tlen[7i + 3 + x] in its piece of kind x.
__global__ __launch_bounds__(256) void row_groups_kernel(const int32_t *__restrict__ sorted,
                                                         const long long *__restrict__ rs,
                                                         const int32_t *__restrict__ gbase,
                                                         const int32_t *__restrict__ kband,
                                                         const int32_t *__restrict__ hgs, int split,
                                                         const int32_t *__restrict__ order,
                                                         const int32_t *__restrict__ len, int32_t *__restrict__ gtgt,
                                                         int32_t *__restrict__ gmax, int32_t *__restrict__ gmin,
                                                         unsigned long long *__restrict__ tlen) {
  __shared__ unsigned long long red[7][256];
  const int i = blockIdx.x;
  const long long b = rs[i];
  const int cnt = (int)(rs[i + 1] - b);
  const int e0 = kband[2 * i], e1 = kband[2 * i + 1];
  const Pieces p1 = band_pieces(e1 - e0, 1, split), p2 = band_pieces(cnt - e1, 2, split);
  const int g0 = band_pieces(e0, 0, split).full, g1 = p1.full;
  // band-1 / band-2 targets in full groups
  const int full1 = g1 * kGroupSize, full2 = p2.full * kGroupSize;
  unsigned long long sum[7] = {0, 0, 0, 0, 0, 0, 0};
  for (int k = threadIdx.x; k < cnt; k += 256) {
    const int band = k < e0 ? 0 : k < e1 ? 1 : 2;
    const int from = band == 0 ? 0 : band == 1 ? e0 : e1;
    const int to = band == 0 ? e0 : band == 1 ? e1 : cnt;
    const int r = k - from;
    const int j = order[sorted[b + k]];
    const int lj = len[j];
    const int full = band == 1 ? full1 : full2;
    if (band != 0 && r >= full) {  // a shared piece
      const Pieces &pp = band == 1 ? p1 : p2;
      int u = r - full, h, slot, x;
      bool last = k == to - 1;
      if (u < pp.half) {
        x = band == 1 ? 1 : 0;
        h = hgs[4 * i + x];
        slot = (h & 3) * (kGroupSize / 2) + u;
        last = u == pp.half - 1;
      } else {
        x = band == 1 ? 3 : 2;
        u -= pp.half;
        h = hgs[4 * i + x];
        slot = (h & 1) * (kGroupSize / 2) + ((h & 3) >> 1) * (kGroupSize / 4) + u;
      }
      const int g = h >> 2;
      gtgt[(size_t)g * kGroupSize + slot] = j;
      if (u == 0) atomicMax(&gmax[g], lj);
      if (last) atomicMin(&gmin[g], lj);
      sum[3 + x] += (unsigned long long)lj;
      continue;
    }
    const int gfirst = gbase[i] + (band == 0 ? 0 : band == 1 ? g0 : g0 + g1);
    const int g = gfirst + r / kGroupSize, slot = r % kGroupSize;
    gtgt[(size_t)g * kGroupSize + slot] = j;
    if (slot == 0) gmax[g] = lj;
    if (slot == kGroupSize - 1 || k == to - 1 || (band != 0 && r == full - 1)) gmin[g] = lj;
    sum[band] += (unsigned long long)lj;
  }
  for (int t = 0; t < 7; ++t) red[t][threadIdx.x] = sum[t];
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if ((int)threadIdx.x < o)
      for (int t = 0; t < 7; ++t) red[t][threadIdx.x] += red[t][threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x < 7) tlen[7 * i + threadIdx.x] = red[threadIdx.x][0];
}

// items of row i, one per full target group: band t's groups at ibase[3i + t] ..
__global__ void cand_items_kernel(const int32_t *__restrict__ gbase, const int32_t *__restrict__ kband,
                                  const long long *__restrict__ rs, const int32_t *__restrict__ ibase, int n,
                                  int split, int4 *__restrict__ items) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int cnt = (int)(rs[i + 1] - rs[i]);
  const int e[3] = {kband[2 * i], kband[2 * i + 1] - kband[2 * i], cnt - kband[2 * i + 1]};
  int g = gbase[i];
  for (int t = 0; t < 3; ++t) {
    const int ng = band_pieces(e[t], t, split).full;
    for (int q = 0; q < ng; ++q, ++g) items[ibase[3 * i + t] + q] = make_int4(g, i, i + 1, 0);
  }
}

// Cells (la * lb) of each verification part, from the rows' target-length
// sums tlen (row_groups_kernel): full groups of band t go to part
// 3 (t = 0), 2 (t = 1), or for t = 2 part 0 when the row is at most
// short_len long, else 1; shared pieces to their group's part gpart[g - Gf].
// cells[10] must start at zero.
__global__ __launch_bounds__(256) void part_cells_kernel(const unsigned long long *__restrict__ tlen,
                                                         const int32_t *__restrict__ len,
                                                         const int32_t *__restrict__ hgs,
                                                         const int32_t *__restrict__ gpart, int Gf, int n,
                                                         int short_len, unsigned long long *__restrict__ cells) {
  constexpr int kP = 10;
  __shared__ unsigned long long red[kP];
  if (threadIdx.x < kP) red[threadIdx.x] = 0;
  __syncthreads();
  unsigned long long acc[kP] = {};
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    const unsigned long long li = (unsigned long long)len[i];
    acc[3] += li * tlen[7 * (size_t)i];
    acc[2] += li * tlen[7 * (size_t)i + 1];
    acc[len[i] <= short_len ? 0 : 1] += li * tlen[7 * (size_t)i + 2];
#pragma unroll
    for (int x = 0; x < 4; ++x) {
      const int h = hgs[4 * (size_t)i + x];
      if (h >= 0) {
        const int part = gpart[(h >> 2) - Gf];
#pragma unroll
        for (int p = 4; p < kP; ++p)
          if (p == part) acc[p] += li * tlen[7 * (size_t)i + 3 + x];
      }
    }
  }
#pragma unroll
  for (int p = 0; p < kP; ++p) {
    unsigned long long v = acc[p];
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
    if ((threadIdx.x & 63) == 0 && v) atomicAdd(&red[p], v);
  }
  __syncthreads();
  if (threadIdx.x < kP && red[threadIdx.x]) atomicAdd(&cells[threadIdx.x], red[threadIdx.x]);
}

__global__ void iota_kernel(int32_t *a, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) a[i] = i;
}

__global__ void fill_i32_kernel(int32_t *a, long long n, int32_t v) {
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (long long)gridDim.x * blockDim.x)
    a[i] = v;
}

unsigned blocks_for(long long n, int tb) { return (unsigned)std::max<long long>((n + tb - 1) / tb, 1); }

struct Layer {
  int in = 0, out = 0;
  float *wt = nullptr, *b = nullptr;        // fp32 master, Wt [out][in]
  float *gw = nullptr, *gb = nullptr;       // gradients
  float *mw = nullptr, *vw = nullptr, *mb = nullptr, *vb = nullptr;  // Adam moments
  bf16 *wt16 = nullptr, *w16 = nullptr;     // bf16 Wt [out][in] and W [in][out]
};

}  // namespace

struct mnhn_encoder {
  Layer L[3];
  MarginTable margins{1, {INT32_MAX}, {0.02f}};
  MarginTable cosmin{1, {INT32_MAX}, {-1.0f}};  // minimum profile cosine per bucket (< 0: none)
  int64_t adam_t = 0;
  int device = 0;
  hipStream_t stream = nullptr;
  hipEvent_t ev[2] = {nullptr, nullptr};
};

namespace {

int gemm(hipStream_t st, int epi, const bf16 *A, int lda, const bf16 *B, int ldb, int M, int N, int K,
         const Epi &ep) {
  if (K % 8 || lda % 8 || ldb % 8) return MNHN_ERR_ARG;
  const dim3 grid((unsigned)((N + 63) / 64), (unsigned)((M + 63) / 64));
  switch (epi) {
    case EPI_F32: hipLaunchKernelGGL(gemm_nt_kernel<EPI_F32>, grid, dim3(256), 0, st, A, lda, B, ldb, M, N, K, ep); break;
    case EPI_BIAS_GELU: hipLaunchKernelGGL(gemm_nt_kernel<EPI_BIAS_GELU>, grid, dim3(256), 0, st, A, lda, B, ldb, M, N, K, ep); break;
    case EPI_BIAS_OUT: hipLaunchKernelGGL(gemm_nt_kernel<EPI_BIAS_OUT>, grid, dim3(256), 0, st, A, lda, B, ldb, M, N, K, ep); break;
    default: hipLaunchKernelGGL(gemm_nt_kernel<EPI_GELU_BWD>, grid, dim3(256), 0, st, A, lda, B, ldb, M, N, K, ep); break;
  }
  return hipGetLastError() == hipSuccess ? MNHN_OK : MNHN_ERR_DEVICE;
}

int transpose(hipStream_t st, const bf16 *in, int R, int C, bf16 *out) {
  hipLaunchKernelGGL(transpose_bf16_kernel, dim3((unsigned)((C + 31) / 32), (unsigned)((R + 31) / 32)),
                     dim3(32, 8), 0, st, in, R, C, out);
  return hipGetLastError() == hipSuccess ? MNHN_OK : MNHN_ERR_DEVICE;
}

void free_layer(Layer &l) {
  dev_free(l.wt); dev_free(l.b); dev_free(l.gw); dev_free(l.gb); dev_free(l.mw); dev_free(l.vw);
  dev_free(l.mb); dev_free(l.vb); dev_free(l.wt16); dev_free(l.w16);
}

int refresh(mnhn_encoder *enc) {
  for (Layer &l : enc->L) {
    const int n = l.in * l.out;
    hipLaunchKernelGGL(refresh_weights_kernel, dim3(blocks_for(n, 256)), dim3(256), 0, enc->stream,
                       l.wt, l.out, l.in, l.wt16, l.w16);
  }
  return hipGetLastError() == hipSuccess ? MNHN_OK : MNHN_ERR_DEVICE;
}

// bf16 profiles of every sequence of ctx
int profiles_dev(mnhn_sw_ctx *c, bf16 **X, float *Xf_host) {
  int rc;
  if ((rc = dev_alloc(X, (size_t)std::max(c->n, 1) * D0))) return rc;
  float *d_f = nullptr;
  if (Xf_host && (rc = dev_alloc(&d_f, (size_t)std::max(c->n, 1) * D0))) { dev_free(*X); return rc; }
  if (c->n > 0)
    hipLaunchKernelGGL(kmer_profile_kernel, dim3((unsigned)std::min(c->n, 16384)), dim3(256), 0, c->stream,
                       c->d_words, c->d_woff, c->d_len, c->n, *X, d_f);
  hipError_t e = hipGetLastError();
  if (e == hipSuccess && Xf_host && c->n > 0)
    e = hipMemcpyAsync(Xf_host, d_f, (size_t)c->n * D0 * sizeof(float), hipMemcpyDeviceToHost, c->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
  dev_free(d_f);
  if (e != hipSuccess) { dev_free(*X); return MNHN_ERR_DEVICE; }
  return MNHN_OK;
}

// Forward pass of `rows` profile rows; writes fp32 / bf16 embeddings and, when
// keep != nullptr, the activations needed by the backward pass.
struct Acts {
  float *pre1 = nullptr, *pre2 = nullptr, *emb = nullptr;
  bf16 *a1 = nullptr, *a2 = nullptr, *emb16 = nullptr;
};

int forward(mnhn_encoder *enc, hipStream_t st, const bf16 *X, int rows, const Acts &a) {
  int rc;
  Epi e1; e1.c32 = a.pre1; e1.ldc = D1; e1.c16 = a.a1; e1.ldc16 = D1; e1.bias = enc->L[0].b;
  if ((rc = gemm(st, EPI_BIAS_GELU, X, D0, enc->L[0].wt16, D0, rows, D1, D0, e1))) return rc;
  Epi e2; e2.c32 = a.pre2; e2.ldc = D2; e2.c16 = a.a2; e2.ldc16 = D2; e2.bias = enc->L[1].b;
  if ((rc = gemm(st, EPI_BIAS_GELU, a.a1, D1, enc->L[1].wt16, D1, rows, D2, D1, e2))) return rc;
  Epi e3; e3.c32 = a.emb; e3.ldc = D3; e3.c16 = a.emb16; e3.ldc16 = D3; e3.bias = enc->L[2].b;
  return gemm(st, EPI_BIAS_OUT, a.a2, D2, enc->L[2].wt16, D2, rows, D3, D2, e3);
}

// bf16 embeddings of every sequence (n x 128)
int embed_dev(mnhn_encoder *enc, mnhn_sw_ctx *c, const bf16 *X, bf16 *E16, float *E32) {
  const int chunk = 32768;
  Acts a;
  int rc = MNHN_OK;
  if ((rc = dev_alloc(&a.a1, (size_t)chunk * D1)) || (rc = dev_alloc(&a.a2, (size_t)chunk * D2))) {
    dev_free(a.a1); dev_free(a.a2);
    return rc;
  }
  for (int r0 = 0; r0 < c->n && rc == MNHN_OK; r0 += chunk) {
    const int rows = std::min(chunk, c->n - r0);
    Acts b = a;
    b.emb16 = E16 + (size_t)r0 * D3;
    b.emb = E32 ? E32 + (size_t)r0 * D3 : nullptr;
    rc = forward(enc, c->stream, X + (size_t)r0 * D0, rows, b);
  }
  if (rc == MNHN_OK && hipStreamSynchronize(c->stream) != hipSuccess) rc = MNHN_ERR_DEVICE;
  dev_free(a.a1); dev_free(a.a2);
  return rc;
}

}  // namespace

extern "C" {

int mnhn_kmer_profiles(mnhn_sw_ctx *c, float *out) {
  if (!c || (!out && c->n > 0)) return MNHN_ERR_ARG;
  if (!c->acgt) return MNHN_ERR_ARG;
  bf16 *X = nullptr;
  int rc = profiles_dev(c, &X, out);
  dev_free(X);
  return rc;
}

int mnhn_encoder_create(uint64_t seed, mnhn_encoder **out) {
  if (!out) return MNHN_ERR_ARG;
  *out = nullptr;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) return MNHN_ERR_NODEVICE;
  mnhn_encoder *enc = new (std::nothrow) mnhn_encoder();
  if (!enc) return MNHN_ERR_NOMEM;
  const int dims[4] = {D0, D1, D2, D3};
  std::mt19937_64 rng(seed);
  int rc = MNHN_OK;
  if (hipGetDevice(&enc->device) != hipSuccess ||
      hipStreamCreateWithFlags(&enc->stream, hipStreamNonBlocking) != hipSuccess ||
      hipEventCreate(&enc->ev[0]) != hipSuccess || hipEventCreate(&enc->ev[1]) != hipSuccess)
    rc = MNHN_ERR_DEVICE;
  for (int l = 0; l < 3 && rc == MNHN_OK; ++l) {
    Layer &L = enc->L[l];
    L.in = dims[l];
    L.out = dims[l + 1];
    const size_t nw = (size_t)L.in * L.out;
    if ((rc = dev_alloc(&L.wt, nw)) || (rc = dev_alloc(&L.b, L.out)) || (rc = dev_alloc(&L.gw, nw)) ||
        (rc = dev_alloc(&L.gb, L.out)) || (rc = dev_alloc(&L.mw, nw)) || (rc = dev_alloc(&L.vw, nw)) ||
        (rc = dev_alloc(&L.mb, L.out)) || (rc = dev_alloc(&L.vb, L.out)) ||
        (rc = dev_alloc(&L.wt16, nw)) || (rc = dev_alloc(&L.w16, nw)))
      break;
    // He-uniform for the GELU layers, Glorot-uniform for the output layer
    const double bound = l < 2 ? std::sqrt(6.0 / L.in) : std::sqrt(6.0 / (L.in + L.out));
    std::uniform_real_distribution<double> U(-bound, bound);
    std::vector<float> w(nw);
    for (auto &x : w) x = (float)U(rng);
    std::vector<float> z(std::max(nw, (size_t)L.out), 0.f);
    if (hipMemcpy(L.wt, w.data(), nw * 4, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(L.b, z.data(), (size_t)L.out * 4, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(L.mw, z.data(), nw * 4, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(L.vw, z.data(), nw * 4, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(L.mb, z.data(), (size_t)L.out * 4, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(L.vb, z.data(), (size_t)L.out * 4, hipMemcpyHostToDevice) != hipSuccess)
      rc = MNHN_ERR_DEVICE;
  }
  if (rc == MNHN_OK) rc = refresh(enc);
  if (rc == MNHN_OK && hipStreamSynchronize(enc->stream) != hipSuccess) rc = MNHN_ERR_DEVICE;
  if (rc) {
    mnhn_encoder_destroy(enc);
    return rc;
  }
  *out = enc;
  return MNHN_OK;
}

void mnhn_encoder_destroy(mnhn_encoder *enc) {
  if (!enc) return;
  if (enc->stream) (void)hipStreamSynchronize(enc->stream);
  for (Layer &l : enc->L) free_layer(l);
  for (auto &e : enc->ev)
    if (e) (void)hipEventDestroy(e);
  if (enc->stream) (void)hipStreamDestroy(enc->stream);
  delete enc;
}

int mnhn_encoder_set_margin(mnhn_encoder *enc, float margin) {
  if (!enc || !(margin >= 0.f) || !std::isfinite(margin)) return MNHN_ERR_ARG;
  enc->margins.n = 1;
  enc->margins.upper[0] = INT32_MAX;
  enc->margins.margin[0] = margin;
  return MNHN_OK;
}

int mnhn_encoder_set_margin_table(mnhn_encoder *enc, const int32_t *max_len, const float *margin, int n) {
  if (!enc || !max_len || !margin || n < 1 || n > kMaxMarginBuckets) return MNHN_ERR_ARG;
  for (int k = 0; k < n; ++k) {
    if (!(margin[k] >= 0.f) || !std::isfinite(margin[k])) return MNHN_ERR_ARG;
    if (k > 0 && max_len[k] <= max_len[k - 1]) return MNHN_ERR_ARG;
  }
  enc->margins.n = n;
  for (int k = 0; k < n; ++k) {
    enc->margins.upper[k] = max_len[k];
    enc->margins.margin[k] = margin[k];
  }
  return MNHN_OK;
}

int mnhn_encoder_set_cosine_table(mnhn_encoder *enc, const int32_t *max_len, const float *cos_min, int n) {
  if (!enc || !max_len || !cos_min || n < 1 || n > kMaxMarginBuckets) return MNHN_ERR_ARG;
  for (int k = 0; k < n; ++k) {
    if (!(cos_min[k] <= 1.f) || !std::isfinite(cos_min[k])) return MNHN_ERR_ARG;
    if (k > 0 && max_len[k] <= max_len[k - 1]) return MNHN_ERR_ARG;
  }
  enc->cosmin.n = n;
  for (int k = 0; k < n; ++k) {
    enc->cosmin.upper[k] = max_len[k];
    enc->cosmin.margin[k] = cos_min[k] < 0.f ? -1.f : cos_min[k];
  }
  return MNHN_OK;
}

float mnhn_encoder_margin(const mnhn_encoder *enc) {
  if (!enc) return -1.f;
  float m = 0.f;
  for (int k = 0; k < enc->margins.n; ++k) m = std::max(m, enc->margins.margin[k]);
  return m;
}

int mnhn_encoder_train(mnhn_encoder *enc, mnhn_sw_ctx *c, const int32_t *ij, const int32_t *dist,
                       int64_t m, const mnhn_train_params *p, mnhn_train_stats *stats) {
  if (!enc || !c || !p || m <= 0 || !ij || !dist || p->batch < 32 || p->batch % 32 || p->epochs < 1 ||
      !(p->lr > 0.f) || !(p->weight_decay >= 0.f) || !(p->far_weight >= 0.f))
    return MNHN_ERR_ARG;
  if (!c->acgt) return MNHN_ERR_ARG;
  for (int64_t k = 0; k < 2 * m; ++k)
    if (ij[k] < 0 || ij[k] >= c->n) return MNHN_ERR_ARG;
  const auto t0 = std::chrono::steady_clock::now();
  hipStream_t st = c->stream;
  int rc;
  bf16 *X = nullptr;
  if ((rc = profiles_dev(c, &X, nullptr))) return rc;
  const int B = p->batch, R = 2 * B;
  Acts a;
  bf16 *Xb = nullptr, *XbT = nullptr, *demb = nullptr, *dembT = nullptr, *a1T = nullptr, *a2T = nullptr,
       *dp1 = nullptr, *dp1T = nullptr, *dp2 = nullptr, *dp2T = nullptr;
  int32_t *d_idx = nullptr;
  float *d_tgt = nullptr, *d_rad = nullptr, *d_loss = nullptr, *d_colpart = nullptr;
  const int col_blocks = (R + 127) / 128;  // row blocks of col_sum_kernel
  auto cleanup = [&]() {
    dev_free(X); dev_free(a.pre1); dev_free(a.pre2); dev_free(a.emb); dev_free(a.a1); dev_free(a.a2);
    dev_free(Xb); dev_free(XbT); dev_free(demb); dev_free(dembT); dev_free(a1T); dev_free(a2T);
    dev_free(dp1); dev_free(dp1T); dev_free(dp2); dev_free(dp2T); dev_free(d_idx); dev_free(d_tgt);
    dev_free(d_rad); dev_free(d_loss); dev_free(d_colpart);
  };
  if ((rc = dev_alloc(&a.pre1, (size_t)R * D1)) || (rc = dev_alloc(&a.pre2, (size_t)R * D2)) ||
      (rc = dev_alloc(&a.emb, (size_t)R * D3)) || (rc = dev_alloc(&a.a1, (size_t)R * D1)) ||
      (rc = dev_alloc(&a.a2, (size_t)R * D2)) || (rc = dev_alloc(&Xb, (size_t)R * D0)) ||
      (rc = dev_alloc(&XbT, (size_t)R * D0)) || (rc = dev_alloc(&demb, (size_t)R * D3)) ||
      (rc = dev_alloc(&dembT, (size_t)R * D3)) || (rc = dev_alloc(&a1T, (size_t)R * D1)) ||
      (rc = dev_alloc(&a2T, (size_t)R * D2)) || (rc = dev_alloc(&dp1, (size_t)R * D1)) ||
      (rc = dev_alloc(&dp1T, (size_t)R * D1)) || (rc = dev_alloc(&dp2, (size_t)R * D2)) ||
      (rc = dev_alloc(&dp2T, (size_t)R * D2)) || (rc = dev_alloc(&d_idx, (size_t)R)) ||
      (rc = dev_alloc(&d_tgt, (size_t)B)) || (rc = dev_alloc(&d_rad, (size_t)B)) ||
      (rc = dev_alloc(&d_loss, (size_t)B)) ||
      (rc = dev_alloc(&d_colpart, (size_t)col_blocks * std::max(D1, std::max(D2, D3))))) {
    cleanup();
    return rc;
  }
  // bias gradient: column sums of dY [R x C] into gb, in a fixed order
  auto col_sums = [&](const bf16 *dY, int C, float *gb) {
    hipLaunchKernelGGL(col_sum_kernel, dim3((unsigned)((C + 63) / 64), (unsigned)col_blocks), dim3(256), 0, st,
                       dY, R, C, d_colpart);
    hipLaunchKernelGGL(col_reduce_kernel, dim3((unsigned)((C + 255) / 256)), dim3(256), 0, st, d_colpart,
                       col_blocks, C, gb);
  };
  std::vector<int64_t> order(m);
  for (int64_t k = 0; k < m; ++k) order[k] = k;
  std::mt19937_64 rng(p->seed);
  std::vector<int32_t> h_idx(R);
  std::vector<float> h_tgt(B), h_rad(B), h_loss(B);
  double first_loss = 0, last_loss = 0;
  int64_t steps = 0;
  const int64_t steps_per_epoch = (m + B - 1) / B;
  const int64_t total_steps = steps_per_epoch * p->epochs;
  for (int ep = 0; ep < p->epochs && rc == MNHN_OK; ++ep) {
    std::shuffle(order.begin(), order.end(), rng);
    double ep_loss = 0;
    int64_t ep_pairs = 0;
    for (int64_t s = 0; s < steps_per_epoch && rc == MNHN_OK; ++s) {
      // a batch always holds B pairs (the last one wraps around the epoch)
      for (int k = 0; k < B; ++k) {
        const int64_t pk = order[(s * B + k) % m];
        const int ia = ij[2 * pk], ib = ij[2 * pk + 1];
        h_idx[2 * k] = ia;
        h_idx[2 * k + 1] = ib;
        const double L4 = (double)c->match * std::max(std::max(c->len[ia], c->len[ib]), 1);
        h_tgt[k] = (float)(dist[pk] / L4);
        const float r = (float)(p->eps_train / L4);
        h_rad[k] = dist[pk] <= p->eps_train ? r : -1.f - r;
      }
      hipError_t e = hipMemcpyAsync(d_idx, h_idx.data(), R * 4, hipMemcpyHostToDevice, st);
      if (e == hipSuccess) e = hipMemcpyAsync(d_tgt, h_tgt.data(), B * 4, hipMemcpyHostToDevice, st);
      if (e == hipSuccess) e = hipMemcpyAsync(d_rad, h_rad.data(), B * 4, hipMemcpyHostToDevice, st);
      if (e != hipSuccess) { rc = MNHN_ERR_DEVICE; break; }
      hipLaunchKernelGGL(gather_rows_kernel, dim3(R), dim3(128), 0, st, X, d_idx, R, Xb);
      Acts b = a;
      b.emb16 = nullptr;
      if ((rc = forward(enc, st, Xb, R, b))) break;
      hipLaunchKernelGGL(pair_loss_kernel, dim3(blocks_for(B, 4)), dim3(256), 0, st, a.emb, d_tgt, d_rad, B,
                         p->neighbor_weight, p->far_weight, p->far_gap, 1.0f / (float)B, demb, d_loss);
      // layer 3: dW3 = dEmb^T a2, db3 = colsum dEmb, dpre2 = (dEmb W3) * gelu'(pre2)
      if ((rc = transpose(st, demb, R, D3, dembT)) || (rc = transpose(st, a.a2, R, D2, a2T))) break;
      Epi g3; g3.c32 = enc->L[2].gw; g3.ldc = D2;
      if ((rc = gemm(st, EPI_F32, dembT, R, a2T, R, D3, D2, R, g3))) break;
      col_sums(demb, D3, enc->L[2].gb);
      Epi d2; d2.c16 = dp2; d2.ldc16 = D2; d2.pre = a.pre2; d2.ldp = D2;
      if ((rc = gemm(st, EPI_GELU_BWD, demb, D3, enc->L[2].w16, D3, R, D2, D3, d2))) break;
      // layer 2
      if ((rc = transpose(st, dp2, R, D2, dp2T)) || (rc = transpose(st, a.a1, R, D1, a1T))) break;
      Epi g2; g2.c32 = enc->L[1].gw; g2.ldc = D1;
      if ((rc = gemm(st, EPI_F32, dp2T, R, a1T, R, D2, D1, R, g2))) break;
      col_sums(dp2, D2, enc->L[1].gb);
      Epi d1; d1.c16 = dp1; d1.ldc16 = D1; d1.pre = a.pre1; d1.ldp = D1;
      if ((rc = gemm(st, EPI_GELU_BWD, dp2, D2, enc->L[1].w16, D2, R, D1, D2, d1))) break;
      // layer 1
      if ((rc = transpose(st, dp1, R, D1, dp1T)) || (rc = transpose(st, Xb, R, D0, XbT))) break;
      Epi g1; g1.c32 = enc->L[0].gw; g1.ldc = D0;
      if ((rc = gemm(st, EPI_F32, dp1T, R, XbT, R, D1, D0, R, g1))) break;
      col_sums(dp1, D1, enc->L[0].gb);
      // Adam
      enc->adam_t++;
      const float b1 = 0.9f, b2 = 0.999f;
      const float c1 = 1.f - (float)std::pow(b1, (double)enc->adam_t), c2 = 1.f - (float)std::pow(b2, (double)enc->adam_t);
      // optional linear decay of the rate to 10% over the run
      const double frac = (double)steps / (double)std::max<int64_t>(total_steps, 1);
      const float lr = p->lr_decay ? (float)(p->lr * (1.0 - 0.9 * frac)) : p->lr;
      for (Layer &L : enc->L) {
        const int nw = L.in * L.out;
        hipLaunchKernelGGL(adam_kernel, dim3(blocks_for(nw, 256)), dim3(256), 0, st, L.wt, L.gw, L.mw, L.vw, nw,
                           lr, b1, b2, c1, c2, p->weight_decay);
        hipLaunchKernelGGL(adam_kernel, dim3(blocks_for(L.out, 256)), dim3(256), 0, st, L.b, L.gb, L.mb, L.vb,
                           L.out, lr, b1, b2, c1, c2, 0.0f);
        hipLaunchKernelGGL(refresh_weights_kernel, dim3(blocks_for(nw, 256)), dim3(256), 0, st, L.wt, L.out,
                           L.in, L.wt16, L.w16);
      }
      e = hipGetLastError();
      if (e == hipSuccess) e = hipMemcpyAsync(h_loss.data(), d_loss, (size_t)B * 4, hipMemcpyDeviceToHost, st);
      if (e == hipSuccess) e = hipStreamSynchronize(st);
      if (e != hipSuccess) { rc = MNHN_ERR_DEVICE; break; }
      for (int k = 0; k < B; ++k) ep_loss += h_loss[k];
      ep_pairs += B;
      ++steps;
    }
    const double mean = ep_pairs ? ep_loss / ep_pairs : 0;
    if (ep == 0) first_loss = mean;
    last_loss = mean;
  }
  cleanup();
  if (stats) {
    stats->first_loss = first_loss;
    stats->final_loss = last_loss;
    stats->steps = steps;
    stats->seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  }
  return rc;
}

int mnhn_encoder_embed(mnhn_encoder *enc, mnhn_sw_ctx *c, float *out) {
  if (!enc || !c || (!out && c->n > 0)) return MNHN_ERR_ARG;
  if (!c->acgt) return MNHN_ERR_ARG;
  if (c->n == 0) return MNHN_OK;
  bf16 *X = nullptr, *E16 = nullptr;
  float *E32 = nullptr;
  int rc = profiles_dev(c, &X, nullptr);
  if (rc) return rc;
  if ((rc = dev_alloc(&E16, (size_t)c->n * D3)) || (rc = dev_alloc(&E32, (size_t)c->n * D3))) {
    dev_free(X); dev_free(E16); dev_free(E32);
    return rc;
  }
  rc = embed_dev(enc, c, X, E16, E32);
  if (rc == MNHN_OK && hipMemcpy(out, E32, (size_t)c->n * D3 * 4, hipMemcpyDeviceToHost) != hipSuccess)
    rc = MNHN_ERR_DEVICE;
  dev_free(X); dev_free(E16); dev_free(E32);
  return rc;
}

// File layout (little-endian): "MNHNENC3", the margin table (int32 n, int32
// upper[n], float margin[n]), the cosine table (same form), then per layer
// int32 in, int32 out, float Wt[out][in], float b[out].  "MNHNENC2" files have
// no cosine table; "MNHNENC1" files hold one float margin in place of both.
static const char kEncMagic[8] = {'M', 'N', 'H', 'N', 'E', 'N', 'C', '3'};
static const char kEncMagicV2[8] = {'M', 'N', 'H', 'N', 'E', 'N', 'C', '2'};
static const char kEncMagicV1[8] = {'M', 'N', 'H', 'N', 'E', 'N', 'C', '1'};

static void write_table(FILE *f, const MarginTable &t) {
  fwrite(&t.n, 4, 1, f);
  fwrite(t.upper, 4, (size_t)t.n, f);
  fwrite(t.margin, 4, (size_t)t.n, f);
}

static bool read_table(FILE *f, MarginTable &t) {
  return fread(&t.n, 4, 1, f) == 1 && t.n >= 1 && t.n <= kMaxMarginBuckets &&
         fread(t.upper, 4, (size_t)t.n, f) == (size_t)t.n && fread(t.margin, 4, (size_t)t.n, f) == (size_t)t.n;
}

int mnhn_encoder_save(const mnhn_encoder *enc, const char *path) {
  if (!enc || !path) return MNHN_ERR_ARG;
  FILE *f = fopen(path, "wb");
  if (!f) return MNHN_ERR_ARG;
  int rc = MNHN_OK;
  fwrite(kEncMagic, 1, 8, f);
  write_table(f, enc->margins);
  write_table(f, enc->cosmin);
  for (const Layer &L : enc->L) {
    std::vector<float> w((size_t)L.in * L.out), b(L.out);
    if (hipMemcpy(w.data(), L.wt, w.size() * 4, hipMemcpyDeviceToHost) != hipSuccess ||
        hipMemcpy(b.data(), L.b, b.size() * 4, hipMemcpyDeviceToHost) != hipSuccess) {
      rc = MNHN_ERR_DEVICE;
      break;
    }
    const int32_t dims[2] = {L.in, L.out};
    fwrite(dims, 4, 2, f);
    fwrite(w.data(), 4, w.size(), f);
    fwrite(b.data(), 4, b.size(), f);
  }
  if (fclose(f) != 0 && rc == MNHN_OK) rc = MNHN_ERR_ARG;
  return rc;
}

int mnhn_encoder_load(mnhn_encoder *enc, const char *path) {
  if (!enc || !path) return MNHN_ERR_ARG;
  FILE *f = fopen(path, "rb");
  if (!f) return MNHN_ERR_ARG;
  char magic[8];
  MarginTable mt{1, {INT32_MAX}, {0.f}}, ct{1, {INT32_MAX}, {-1.f}};
  int rc = MNHN_OK;
  if (fread(magic, 1, 8, f) != 8) {
    rc = MNHN_ERR_ARG;
  } else if (memcmp(magic, kEncMagicV1, 8) == 0) {
    if (fread(&mt.margin[0], 4, 1, f) != 1) rc = MNHN_ERR_ARG;
  } else if (memcmp(magic, kEncMagicV2, 8) == 0) {
    if (!read_table(f, mt)) rc = MNHN_ERR_ARG;
  } else if (memcmp(magic, kEncMagic, 8) == 0) {
    if (!read_table(f, mt) || !read_table(f, ct)) rc = MNHN_ERR_ARG;
  } else {
    rc = MNHN_ERR_ARG;
  }
  for (Layer &L : enc->L) {
    if (rc) break;
    int32_t dims[2];
    if (fread(dims, 4, 2, f) != 2 || dims[0] != L.in || dims[1] != L.out) { rc = MNHN_ERR_ARG; break; }
    std::vector<float> w((size_t)L.in * L.out), b(L.out);
    if (fread(w.data(), 4, w.size(), f) != w.size() || fread(b.data(), 4, b.size(), f) != b.size()) { rc = MNHN_ERR_ARG; break; }
    if (hipMemcpy(L.wt, w.data(), w.size() * 4, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(L.b, b.data(), b.size() * 4, hipMemcpyHostToDevice) != hipSuccess) { rc = MNHN_ERR_DEVICE; break; }
  }
  fclose(f);
  if (rc == MNHN_OK) rc = mnhn_encoder_set_margin_table(enc, mt.upper, mt.margin, mt.n);
  if (rc == MNHN_OK) rc = mnhn_encoder_set_cosine_table(enc, ct.upper, ct.margin, ct.n);
  if (rc == MNHN_OK) {
    rc = refresh(enc);
    if (rc == MNHN_OK && hipStreamSynchronize(enc->stream) != hipSuccess) rc = MNHN_ERR_DEVICE;
  }
  return rc;
}

int mnhn_prefilter_eps_csr(mnhn_encoder *enc, mnhn_sw_ctx *c, int32_t eps_max, mnhn_csr *out,
                           mnhn_prefilter_stats *stats) {
  if (!enc || !c || !out) return MNHN_ERR_ARG;
  if (!c->acgt) return MNHN_ERR_ARG;
  memset(out, 0, sizeof(*out));
  mnhn_prefilter_stats stt{};
  const int n = c->n;
  out->n_rows = n;
  out->eps_max = eps_max;
  out->row_ptr = (int64_t *)calloc((size_t)n + 1, sizeof(int64_t));
  if (!out->row_ptr) return MNHN_ERR_NOMEM;
  if (n == 0) return MNHN_OK;
  hipStream_t st = c->stream;
  hipEvent_t e0, e1, ek0, ek1;
  (void)hipEventCreate(&e0);
  (void)hipEventCreate(&e1);
  (void)hipEventCreate(&ek0);
  (void)hipEventCreate(&ek1);
  auto phase_ms = [&]() {
    float ms = 0.f;
    (void)hipEventRecord(e1, st);
    (void)hipEventSynchronize(e1);
    (void)hipEventElapsedTime(&ms, e0, e1);
    (void)hipEventRecord(e0, st);
    return (double)ms;
  };
  bf16 *X = nullptr, *E16 = nullptr, *Es = nullptr;
  void *Xs = nullptr;  // length-ordered raw 5-mer counts: int8, or bf16 when a count exceeds 127
  bool xs_i8 = true;
  int *d_over = nullptr;
  float *norm = nullptr, *srad = nullptr, *cmin = nullptr, *pnorm = nullptr;
  unsigned long long *cand = nullptr, *d_cnt64 = nullptr;
  // the candidate buffer, kept after the candidates are grouped to become
  // the verification's key buffer (spare_cap keys)
  unsigned long long *spare = nullptr, spare_cap = 0;
  int32_t *rowcnt = nullptr, *gbase = nullptr, *cursor = nullptr, *gtgt = nullptr, *gmax = nullptr,
          *gmin = nullptr, *qids = nullptr, *d_order = nullptr, *lens = nullptr,
          *ibase = nullptr, *seg = nullptr, *sorted = nullptr, *kband = nullptr, *hgs = nullptr, *qids2 = nullptr;
  unsigned long long *tlen = nullptr, *pcells = nullptr;
  int32_t *gparts = nullptr;
  unsigned char *garena = nullptr;  // one allocation for the group tables beside gtgt
  long long *d_rs = nullptr;
  int4 *items = nullptr, *sitems = nullptr;
  int2 *d_tiles = nullptr;
  auto cleanup = [&]() {
    dev_free(X); dev_free(E16); dev_free(Es); dev_free(norm); dev_free(srad); dev_free(cand); dev_free(spare);
    if (Xs) (void)hipFree(Xs);
    Xs = nullptr;
    dev_free(d_over); dev_free(cmin); dev_free(pnorm);
    dev_free(d_cnt64); dev_free(rowcnt); dev_free(gbase); dev_free(cursor); dev_free(gtgt);
    dev_free(qids); dev_free(d_order); dev_free(lens); dev_free(d_tiles);
    dev_free(seg); dev_free(sorted); dev_free(d_rs); dev_free(kband);
    dev_free(garena);  // gmax, gmin, items, ibase, tlen, hgs, qids2, sitems, gparts, pcells
    gmax = gmin = ibase = hgs = qids2 = gparts = nullptr;
    items = sitems = nullptr;
    tlen = pcells = nullptr;
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    (void)hipEventDestroy(ek0);
    (void)hipEventDestroy(ek1);
  };
  auto fail = [&](int code) {
    cleanup();
    mnhn_csr_free(out);
    return code;
  };
  if (eps_max < 0) {  // no pair has a negative distance
    cleanup();
    if (stats) *stats = stt;
    return MNHN_OK;
  }
  int rc = list_bound(c, eps_max, &eps_max);
  if (rc) return fail(rc);
  mnhn_phase_clock clk(st, "prefilter");
  (void)hipEventRecord(e0, st);
  if ((rc = profiles_dev(c, &X, nullptr))) return fail(rc);
  stt.profile_ms = phase_ms();
  if ((rc = dev_alloc(&E16, (size_t)n * D3)) || (rc = embed_dev(enc, c, X, E16, nullptr))) return fail(rc);
  dev_free(X);
  // length-ordered copies for the pairwise stage
  if ((rc = dev_alloc(&Es, (size_t)n * D3)) || (rc = dev_alloc(&norm, (size_t)n)) ||
      (rc = dev_alloc(&srad, (size_t)n)) || (rc = dev_alloc(&cmin, (size_t)n)) ||
      (rc = dev_alloc(&lens, (size_t)n)) || (rc = dev_alloc(&d_order, (size_t)n)))
    return fail(rc);
  bool use_cos = false;
  for (int k = 0; k < enc->cosmin.n; ++k) use_cos |= enc->cosmin.margin[k] >= 0.f;
  if (hipMemcpyAsync(d_order, c->order.data(), (size_t)n * 4, hipMemcpyHostToDevice, st) != hipSuccess)
    return fail(MNHN_ERR_DEVICE);
  const int max_dl = eps_max / c->match;
  hipLaunchKernelGGL(sorted_rows_kernel, dim3(blocks_for(n, 4)), dim3(256), 0, st, E16, d_order, c->d_len, n,
                     (float)eps_max / (float)c->match, enc->margins, enc->cosmin, Es, norm, srad, cmin, lens);
  if (hipGetLastError() != hipSuccess) return fail(MNHN_ERR_DEVICE);
  if (use_cos) {
    // int8 counts for the int8 MFMA stage; bf16 rows (the same products,
    // twice the bytes and MFMA time) only when some count exceeds 127
    int8_t *X8 = nullptr;
    if ((rc = dev_alloc(&X8, (size_t)n * D0)) || (rc = dev_alloc(&pnorm, (size_t)n)) ||
        (rc = dev_alloc(&d_over, 1))) {
      dev_free(X8);
      return fail(rc);
    }
    Xs = X8;
    int over = 0;
    hipError_t e = hipMemsetAsync(d_over, 0, sizeof(int), st);
    if (e == hipSuccess) {
      hipLaunchKernelGGL(kmer_counts_sorted_kernel<int8_t>, dim3((unsigned)std::min(n, 16384)), dim3(256), 0, st,
                         c->d_words, c->d_woff, c->d_len, d_order, n, X8, pnorm, d_over);
      e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipMemcpyAsync(&over, d_over, sizeof(int), hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    if (e != hipSuccess) return fail(MNHN_ERR_DEVICE);
    if (over) {
      (void)hipFree(Xs);
      Xs = nullptr;
      xs_i8 = false;
      bf16 *X16 = nullptr;
      if ((rc = dev_alloc(&X16, (size_t)n * D0))) return fail(rc);
      Xs = X16;
      hipLaunchKernelGGL(kmer_counts_sorted_kernel<bf16>, dim3((unsigned)std::min(n, 16384)), dim3(256), 0, st,
                         c->d_words, c->d_woff, c->d_len, d_order, n, X16, pnorm, d_over);
      if (hipGetLastError() != hipSuccess) return fail(MNHN_ERR_DEVICE);
    }
  }
  dev_free(E16);
  stt.embed_ms = phase_ms();
  clk.mark("profiles, embedding, count rows");
  // ---- exact bucket: the shortest lengths, whose buckets switch both tests
  // off (radius margin >= kMarginOff, no cosine bound), would have every band
  // pair proposed; the all-pairs pass over those sequences lists them
  // directly (positions p0 .. n-1 of the length order) instead of through
  // the candidate lists.  Lists only gain pairs by it, never lose any.
  // MNHN_PREFILTER_EXACT_BUCKET=0 sends those pairs through the candidate
  // lists as well.
  constexpr float kMarginOff = 100.f;
  const char *xb_env = getenv("MNHN_PREFILTER_EXACT_BUCKET");
  const bool exact_bucket = !(xb_env && atoi(xb_env) == 0);
  auto host_bucket = [](const MarginTable &t, int L) {
    float v = t.margin[t.n - 1];
    for (int k = t.n - 1; k >= 0; --k)
      if (L <= t.upper[k]) v = t.margin[k];
    return v;
  };
  int exact_len = -1;
  {
    std::vector<int> edges;
    for (int k = 0; k < enc->margins.n; ++k) edges.push_back(enc->margins.upper[k]);
    for (int k = 0; k < enc->cosmin.n; ++k) edges.push_back(enc->cosmin.upper[k]);
    std::sort(edges.begin(), edges.end());
    edges.erase(std::unique(edges.begin(), edges.end()), edges.end());
    for (int e : edges) {  // both tables are constant between consecutive edges
      if (exact_bucket && host_bucket(enc->margins, e) >= kMarginOff && host_bucket(enc->cosmin, e) < 0.f)
        exact_len = e;
      else break;
    }
  }
  int p0 = n;  // first length-order position with len <= exact_len
  {
    int lo = 0, hi = n;
    while (lo < hi) {
      const int mid = (lo + hi) / 2;
      if (c->len[c->order[mid]] > exact_len) lo = mid + 1;
      else hi = mid;
    }
    p0 = lo;
  }
  // ---- proposals over the length band: row tile ti reaches column tiles
  // ti .. tj_max(ti), the last whose longest length is >= shortest(ti) - max_dl;
  // row tiles whose longest row is in the exact bucket (ti >= t_short) have
  // no pair to propose
  const int tiles = (n + 127) / 128;
  const int t_short = (p0 + 127) / 128;
  std::vector<int32_t> tjmax(tiles);
  long long nblk_ll = 0;
  {
    int tj = 0;
    for (int ti = 0; ti < t_short; ++ti) {
      const long long shortest = c->len[c->order[std::min((ti + 1) * 128, n) - 1]];
      tj = std::max(tj, ti);
      while (tj + 1 < tiles && (long long)c->len[c->order[(tj + 1) * 128]] >= shortest - max_dl) ++tj;
      tjmax[ti] = tj;  // nondecreasing in ti
      nblk_ll += tj - ti + 1;
    }
  }
  if (nblk_ll > 0x7fffffffLL) return fail(MNHN_ERR_RANGE);
  const unsigned nblk = (unsigned)nblk_ll;
  stt.tiles = nblk;
  // the band tiles in blocks of kTileBlock row tiles, by column tile within a block
  {
    std::vector<int2> tl(std::max<size_t>((size_t)nblk, 1));
    size_t o = 0;
    for (int b0 = 0; b0 < t_short; b0 += kTileBlock) {
      const int b1 = std::min(t_short, b0 + kTileBlock);
      for (int tj = b0; tj <= tjmax[b1 - 1]; ++tj)
        for (int ti = b0; ti < b1 && ti <= tj; ++ti)
          if (tj <= tjmax[ti]) tl[o++] = make_int2(ti, tj);
    }
    if (o != (size_t)nblk) return fail(MNHN_ERR_ARG);
    // d_cnt64[0]: candidates, d_cnt64[1]: band tiles that ran the cosine stage
    if ((rc = dev_alloc(&d_tiles, (size_t)nblk)) || (rc = dev_alloc(&d_cnt64, 2))) return fail(rc);
    if (hipMemcpyAsync(d_tiles, tl.data(), (size_t)nblk * sizeof(int2), hipMemcpyHostToDevice, st) != hipSuccess ||
        hipStreamSynchronize(st) != hipSuccess)
      return fail(MNHN_ERR_DEVICE);
  }
  clk.mark("band tile list");
  // candidate buffer: the band's pair count, at most MNHN_PREFILTER_CAND_GB
  // GB (default: 8 GB or 40 % of the free device memory, the larger; an
  // allocation of 100 GB took 0.4 ms on MI355X); grown to the exact count and
  // recomputed once on overflow
  const unsigned long long band_pairs = (unsigned long long)nblk * 128ull * 128ull;
  const char *cg = getenv("MNHN_PREFILTER_CAND_GB");
  double cand_gb = cg ? atof(cg) : 8.0;
  if (!cg) {
    size_t free_b = 0, total_b = 0;
    if (hipMemGetInfo(&free_b, &total_b) == hipSuccess) cand_gb = std::max(cand_gb, 0.4 * (double)free_b / 1e9);
  }
  unsigned long long cap = std::max<unsigned long long>(
      1ull << 20, std::min<unsigned long long>(band_pairs, (unsigned long long)(cand_gb * 1e9 / 8.0))), m = 0;
  for (int attempt = 0; attempt < 2; ++attempt) {
    dev_free(cand);
    if ((rc = dev_alloc(&cand, cap))) break;
    clk.mark("candidate buffer");
    hipError_t e = hipMemsetAsync(d_cnt64, 0, 16, st);
    (void)hipEventRecord(ek0, st);
    if (e == hipSuccess && nblk > 0) {  // no band tile: every pair is the exact pass's
      if (use_cos && xs_i8)
        hipLaunchKernelGGL((pair_candidates_kernel<true, true>), dim3(nblk), dim3(256), 0, st, Es, norm, srad, Xs,
                           pnorm, cmin, lens, n, d_tiles, max_dl, exact_len, cand, d_cnt64, cap, d_cnt64 + 1);
      else if (use_cos)
        hipLaunchKernelGGL((pair_candidates_kernel<true, false>), dim3(nblk), dim3(256), 0, st, Es, norm, srad, Xs,
                           pnorm, cmin, lens, n, d_tiles, max_dl, exact_len, cand, d_cnt64, cap, d_cnt64 + 1);
      else
        hipLaunchKernelGGL((pair_candidates_kernel<false, false>), dim3(nblk), dim3(256), 0, st, Es, norm, srad, Xs,
                           pnorm, cmin, lens, n, d_tiles, max_dl, exact_len, cand, d_cnt64, cap, d_cnt64 + 1);
      e = hipGetLastError();
    }
    (void)hipEventRecord(ek1, st);
    unsigned long long cnt2[2] = {0, 0};
    if (e == hipSuccess) e = hipMemcpyAsync(cnt2, d_cnt64, 16, hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    if (e != hipSuccess) { rc = MNHN_ERR_DEVICE; break; }
    m = cnt2[0];
    stt.stage2_tiles = (int64_t)cnt2[1];
    stt.counts_int8 = use_cos && xs_i8 ? 1 : 0;
    float kms = 0.f;
    (void)hipEventElapsedTime(&kms, ek0, ek1);
    stt.propose_kernel_ms = kms;
    if (m <= cap) break;
    cap = m;  // overflowed: exact size, recompute once
  }
  if (rc || m > cap) return fail(rc ? rc : MNHN_ERR_NOMEM);
  dev_free(Es); dev_free(norm); dev_free(srad); dev_free(lens); dev_free(pnorm); dev_free(cmin);
  if (Xs) (void)hipFree(Xs);
  Xs = nullptr;
  stt.candidates = (int64_t)m;
  stt.propose_ms = phase_ms();
  clk.mark("pairwise kernel, frees");
  // ---- the exact bucket's pass over the shortest sequences: its keys follow
  // the verification's in the key buffer (same attempts); its pairs count as
  // verified candidates
  Work wsub;
  unsigned long long est_sub = 0;
  if (p0 < n - 1) {
    wsub = suffix_work(c, eps_max, p0);
    if (!wsub.items.empty() && (rc = eps_keys_estimate(c, wsub, eps_max, &est_sub))) return fail(rc);
    // the pairs the proposal would have made there: every pair within the
    // length band (positions q > p with len >= len[p] - max_dl)
    long long band = 0;
    for (int p = p0, r = p0; p < n; ++p) {
      r = std::max(r, p);
      while (r + 1 < n && c->len[c->order[r + 1]] >= c->len[c->order[p]] - max_dl) ++r;
      band += r - p;
    }
    stt.candidates += band;
    stt.cells += wsub.cells;
  }
  double kernel_ms = 0, form_cells[3] = {0, 0, 0}, form_ms[3] = {0, 0, 0};
  auto run_sub = [&]() -> int {
    if (wsub.items.empty()) return MNHN_OK;
    const int r = launch_tiles(c, wsub, MODE_EPS, 1, eps_max, nullptr, c->d_keys, c->keys_cap);
    kernel_ms += c->last.kernel_ms;
    for (int f = 0; f < 3; ++f) {
      form_cells[f] += c->last.form_cells[f];
      form_ms[f] += c->last.form_ms[f];
    }
    return r;
  };
  clk.mark("exact bucket work, key estimate");
  // ---- exact verification: candidates grouped per row into targets of 128
  unsigned long long nk = 0;
  if (m > 0) {
    std::vector<int32_t> cnt(n), gb(n + 1, 0);
    if ((rc = dev_alloc(&rowcnt, (size_t)n)) || (rc = dev_alloc(&gbase, (size_t)n + 1)) ||
        (rc = dev_alloc(&cursor, (size_t)n)) || (rc = dev_alloc(&qids, (size_t)n)) ||
        (rc = dev_alloc(&kband, 2 * (size_t)n)))
      return fail(rc);
    hipError_t e = hipMemsetAsync(rowcnt, 0, (size_t)n * 4, st);
    hipLaunchKernelGGL(cand_count_kernel, dim3(stride_blocks((long long)m, 256)), dim3(256), 0, st, cand, (long long)m,
                       d_order, rowcnt);
    if (e == hipSuccess) e = hipMemcpyAsync(cnt.data(), rowcnt, (size_t)n * 4, hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    if (e != hipSuccess) return fail(MNHN_ERR_DEVICE);
    // row segments of the candidates, sorted by target position within rows
    clk.mark("row counts");
    std::vector<long long> rsh((size_t)n + 1, 0);
    for (int i = 0; i < n; ++i) rsh[i + 1] = rsh[i] + cnt[i];
    if ((rc = dev_alloc(&d_rs, (size_t)n + 1)) || (rc = dev_alloc(&seg, (size_t)m))) return fail(rc);
    e = hipMemcpyAsync(d_rs, rsh.data(), ((size_t)n + 1) * sizeof(long long), hipMemcpyHostToDevice, st);
    if (e == hipSuccess) e = hipMemsetAsync(cursor, 0, (size_t)n * 4, st);
    if (e != hipSuccess) return fail(MNHN_ERR_DEVICE);
    hipLaunchKernelGGL(cand_scatter_kernel, dim3(stride_blocks((long long)m, 256)), dim3(256), 0, st, cand,
                       (long long)m, d_order, d_rs, cursor, seg);
    if (hipGetLastError() != hipSuccess) return fail(MNHN_ERR_DEVICE);
    // The candidate buffer is kept as the verification's key buffer (below)
    // when it is larger than the context's: on the 500,000-sequence set that
    // saves freeing ~100 GB and allocating ~100 GB again (150 ms).  Should
    // the next allocations not fit beside it, it is freed first.
    if (cap > c->keys_cap) {
      spare = cand;
      spare_cap = cap;
      cand = nullptr;
    } else {
      dev_free(cand);
    }
    if ((rc = dev_alloc(&sorted, (size_t)m)) && spare) {
      dev_free(spare);
      rc = dev_alloc(&sorted, (size_t)m);
    }
    if (rc) return fail(rc);
    if ((rc = sort_row_segments(st, d_rs, seg, sorted, nullptr, nullptr, n, n))) return fail(rc);
    dev_free(seg);
    clk.mark("scatter, row sort");
    // each row's targets in length bands by the lane format their pairs need
    // (the row's sequence is the longer of each pair, so the target's length
    // bounds the score): match * len beyond 65535 -> 32-bit form, beyond 2048
    // -> packed int16, else exact fp16
    const int t1 = 2048 / c->match, t2 = 65535 / c->match;
    std::vector<int32_t> kb(2 * (size_t)n);
    hipLaunchKernelGGL(band_counts_kernel, dim3(blocks_for(n, 256)), dim3(256), 0, st, sorted, d_rs, d_order,
                       c->d_len, n, t1, t2, kband);
    e = hipGetLastError();
    if (e == hipSuccess) e = hipMemcpyAsync(kb.data(), kband, 2 * (size_t)n * 4, hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    if (e != hipSuccess) return fail(MNHN_ERR_DEVICE);
    clk.mark("length bands");
    // Remainders go to target groups shared by several rows of the same band
    // (band_pieces), grouped longest first: MNHN_VERIFY_SPLIT=0 keeps one row
    // per group, 1 shares half groups only, 2 (default) also quarter groups.
    const char *split_env = getenv("MNHN_VERIFY_SPLIT");
    const int split = split_env ? std::min(std::max(atoi(split_env), 0), 2) : 2;
    // ng[3i + t]: full target groups of row i in band t
    std::vector<int32_t> ng(3 * (size_t)n);
    for (int i = 0; i < n; ++i) {
      const int e0 = kb[2 * (size_t)i], e1 = kb[2 * (size_t)i + 1];
      const int k3[3] = {e0, e1 - e0, cnt[i] - e1};
      int sum = 0;
      for (int t = 0; t < 3; ++t) sum += (ng[3 * (size_t)i + t] = band_pieces(k3[t], t, split).full);
      gb[i + 1] = gb[i] + sum;
    }
    const int Gf = gb[n];
    // rows with an fp16 half / int16 half / fp16 quarter / int16 quarter
    // piece, longest first: the length order breaks ties by id, as a stable
    // sort of the ids by length would
    std::vector<int32_t> hrows[4];
    for (int p = 0; p < n; ++p) {
      const int i = c->order[p];
      const int e0 = kb[2 * (size_t)i], e1 = kb[2 * (size_t)i + 1];
      const Pieces p2 = band_pieces(cnt[i] - e1, 2, split), p1 = band_pieces(e1 - e0, 1, split);
      if (p2.half) hrows[0].push_back(i);
      if (p1.half) hrows[1].push_back(i);
      if (p2.quarter) hrows[2].push_back(i);
      if (p1.quarter) hrows[3].push_back(i);
    }
    // item order by class: fp16 groups of rows short enough for a 32-column
    // strip, the other fp16 groups, the int16 groups, the 32-bit groups
    // (parts 0-3); then the shared groups, by kind (fp16 half, fp16 quarter,
    // int16 half, int16 quarter) and within a kind longest first: fp16 groups
    // whose longest row is longer than 128 run at strip 64 when the context's
    // strip is 64 (parts 5, 8), the other fp16 groups at strip 32 (parts 4,
    // 7), the int16 groups (rows longer than 512) at the context's strip
    // (parts 6, 9)
    const int S = context_strip(c);
    auto band_class = [&](int i, int t) {
      if (t == 0) return 3;
      if (t == 1) return 2;
      return (S == 64 && c->len[i] <= 128) ? 0 : 1;
    };
    constexpr int kParts = 10;
    constexpr int kRowsPer[4] = {2, 2, 4, 4};  // rows per shared group of each kind
    const int kind_order[4] = {0, 2, 1, 3};
    int nshared[4];
    for (int x = 0; x < 4; ++x) nshared[x] = (int)((hrows[x].size() + kRowsPer[x] - 1) / kRowsPer[x]);
    const int np = nshared[0] + nshared[1] + nshared[2] + nshared[3];
    const int G = Gf + np;
    // hg[4i + x]: row i's piece of kind x (shared group << 2 | sub-list)
    std::vector<int32_t> hg(4 * (size_t)n, -1), q2, gpart(np);
    std::vector<int4> sit(np);
    int rng[kParts][2] = {};  // shared parts: their range of sit
    {
      int pi = 0;
      for (int xo = 0; xo < 4; ++xo) {
        const int x = kind_order[xo];
        const int per = kRowsPer[x];
        const std::vector<int32_t> &hr = hrows[x];
        const bool i16 = x == 1 || x == 3;
        const int p_long = x == 0 ? 5 : x == 2 ? 8 : x == 1 ? 6 : 9, p_short = x == 0 ? 4 : x == 2 ? 7 : p_long;
        rng[p_long][0] = rng[p_long][1] = pi;
        for (int gi = 0; gi < nshared[x]; ++gi, ++pi) {
          const int qb = (int)q2.size();
          const size_t k0 = (size_t)gi * per;
          for (int u = 0; u < per; ++u) {
            const size_t k = k0 + u;
            // a short last group repeats its first row, with no targets
            const int r = k < hr.size() ? hr[k] : hr[k0];
            if (k < hr.size()) hg[4 * (size_t)r + x] = ((Gf + pi) << 2) | u;
            q2.push_back(r);
          }
          sit[pi] = make_int4(Gf + pi, qb, qb + per, 0);
          const int part = (i16 || (S == 64 && c->len[hr[k0]] > 128)) ? p_long : p_short;
          gpart[pi] = part;
          if (part == p_short && p_short != p_long && rng[p_short][1] == 0) rng[p_short][0] = pi;
          rng[part][1] = pi + 1;
        }
        if (rng[p_short][1] == 0) rng[p_short][0] = rng[p_short][1] = pi;
      }
    }
    // first item of each (row, band): by class, then row, longest row first
    // (the persistent queue then ends on the cheapest items), then band
    std::vector<int32_t> ib(3 * (size_t)n, 0);
    int part_begin[5] = {0, 0, 0, 0, 0};
    {
      for (int i = 0; i < n; ++i)
        for (int t = 0; t < 3; ++t) part_begin[band_class(i, t) + 1] += ng[3 * (size_t)i + t];
      for (int cls = 0; cls < 4; ++cls) part_begin[cls + 1] += part_begin[cls];
      int off[4] = {part_begin[0], part_begin[1], part_begin[2], part_begin[3]};
      for (int p = 0; p < n; ++p)
        for (int t = 0, i = c->order[p]; t < 3; ++t) {
          const int cls = band_class(i, t);
          ib[3 * (size_t)i + t] = off[cls];
          off[cls] += ng[3 * (size_t)i + t];
        }
    }
    clk.mark("host group layout");
    // gtgt on its own; the other group tables are slices of one allocation,
    // the five that come from the host first (one copy through the
    // context's pinned staging)
    size_t off_arena = 0;
    auto slice = [&](size_t bytes) {
      const size_t o = off_arena;
      off_arena += (bytes + 255) / 256 * 256;
      return o;
    };
    const size_t o_ibase = slice(3 * (size_t)n * 4), o_hgs = slice(4 * (size_t)n * 4),
                 o_qids2 = slice(std::max<size_t>(q2.size(), 1) * 4), o_gparts = slice(std::max(np, 1) * 4ull),
                 o_sitems = slice(std::max(np, 1) * sizeof(int4));
    const size_t up_bytes = off_arena;  // the uploaded part
    const size_t o_gmax = slice((size_t)G * 4), o_gmin = slice((size_t)G * 4),
                 o_items = slice((size_t)std::max(Gf, 1) * sizeof(int4)), o_tlen = slice(7 * (size_t)n * 8),
                 o_pcells = slice((size_t)kParts * 8);
    auto group_buffers = [&]() -> int {
      int r;
      if ((r = dev_alloc(&gtgt, (size_t)G * kGroupSize)) || (r = dev_alloc(&garena, off_arena))) {
        dev_free(gtgt);
        dev_free(garena);
        return r;
      }
      ibase = reinterpret_cast<int32_t *>(garena + o_ibase);
      hgs = reinterpret_cast<int32_t *>(garena + o_hgs);
      qids2 = reinterpret_cast<int32_t *>(garena + o_qids2);
      gparts = reinterpret_cast<int32_t *>(garena + o_gparts);
      sitems = reinterpret_cast<int4 *>(garena + o_sitems);
      gmax = reinterpret_cast<int32_t *>(garena + o_gmax);
      gmin = reinterpret_cast<int32_t *>(garena + o_gmin);
      items = reinterpret_cast<int4 *>(garena + o_items);
      tlen = reinterpret_cast<unsigned long long *>(garena + o_tlen);
      pcells = reinterpret_cast<unsigned long long *>(garena + o_pcells);
      return MNHN_OK;
    };
    if ((rc = group_buffers()) && spare) {
      dev_free(spare);
      rc = group_buffers();
    }
    if (rc) return fail(rc);
    // the uploaded tables, packed as in the arena
    if (c->h_items_bytes < up_bytes) {
      if (c->h_items) (void)hipHostFree(c->h_items);
      c->h_items = nullptr;
      c->h_items_bytes = 0;
      if (hipHostMalloc(&c->h_items, up_bytes, hipHostMallocDefault) == hipSuccess) c->h_items_bytes = up_bytes;
      else c->h_items = nullptr;
    }
    std::vector<unsigned char> up_pageable;
    unsigned char *up = static_cast<unsigned char *>(c->h_items);
    if (!up) {
      up_pageable.resize(up_bytes);
      up = up_pageable.data();
    }
    memcpy(up + o_ibase, ib.data(), 3 * (size_t)n * 4);
    memcpy(up + o_hgs, hg.data(), 4 * (size_t)n * 4);
    if (!q2.empty()) memcpy(up + o_qids2, q2.data(), q2.size() * 4);
    if (np > 0) memcpy(up + o_gparts, gpart.data(), (size_t)np * 4);
    if (np > 0) memcpy(up + o_sitems, sit.data(), (size_t)np * sizeof(int4));
    e = hipMemcpyAsync(gbase, gb.data(), ((size_t)n + 1) * 4, hipMemcpyHostToDevice, st);
    if (e == hipSuccess) e = hipMemcpyAsync(garena, up, up_bytes, hipMemcpyHostToDevice, st);
    hipLaunchKernelGGL(fill_i32_kernel, dim3(stride_blocks((long long)G * kGroupSize, 256)), dim3(256), 0, st, gtgt,
                       (long long)G * kGroupSize, -1);
    if (np > 0) {
      hipLaunchKernelGGL(fill_i32_kernel, dim3(blocks_for(np, 256)), dim3(256), 0, st, gmax + Gf, (long long)np, 0);
      hipLaunchKernelGGL(fill_i32_kernel, dim3(blocks_for(np, 256)), dim3(256), 0, st, gmin + Gf, (long long)np,
                         INT32_MAX);
    }
    hipLaunchKernelGGL(row_groups_kernel, dim3((unsigned)n), dim3(256), 0, st, sorted, d_rs, gbase, kband, hgs,
                       split, d_order, c->d_len, gtgt, gmax, gmin, tlen);
    hipLaunchKernelGGL(cand_items_kernel, dim3(blocks_for(n, 256)), dim3(256), 0, st, gbase, kband, d_rs, ibase, n,
                       split, items);
    hipLaunchKernelGGL(iota_kernel, dim3(blocks_for(n, 256)), dim3(256), 0, st, qids, n);
    if (e == hipSuccess) e = hipGetLastError();
    // cells of each part (band_class; shared pieces by their group's part)
    unsigned long long h_pc[kParts] = {};
    if (e == hipSuccess) e = hipMemsetAsync(pcells, 0, kParts * 8, st);
    hipLaunchKernelGGL(part_cells_kernel, dim3(std::min(blocks_for(n, 256), 1024u)), dim3(256), 0, st, tlen,
                       c->d_len, hgs, gparts, Gf, n, S == 64 ? 128 : -1, pcells);
    if (e == hipSuccess) e = hipGetLastError();
    if (e == hipSuccess) e = hipMemcpyAsync(h_pc, pcells, kParts * 8, hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    if (e != hipSuccess) return fail(MNHN_ERR_DEVICE);
    clk.mark("group buffers, row groups");
    double part_cells[kParts] = {};
    for (int p = 0; p < kParts; ++p) {
      part_cells[p] = (double)h_pc[p];
      stt.cells += part_cells[p];
    }
    dev_free(sorted);
    Work w;
    w.dev = true;
    w.d_gtgt = gtgt;
    w.d_gmax = gmax;
    w.d_gmin = gmin;
    w.max_q_len = c->maxlen;
    // key buffer: both directions of every neighbour, at most two keys per
    // candidate; start from 2m capped at 40 % of the free device memory (at
    // least 2^28 keys) and, if the neighbours need more, grow to the exact
    // count and verify once more
    unsigned long long kroom = std::max(1ull << 28, (unsigned long long)c->keys_cap);
    {
      size_t free_b = 0, total_b = 0;
      if (hipMemGetInfo(&free_b, &total_b) == hipSuccess)
        kroom = std::max<unsigned long long>(kroom, (unsigned long long)(0.4 * (double)free_b / 8.0));
    }
    unsigned long long kcap = std::min(2ull * m, kroom) + est_sub;
    clk.mark("cell counts");
    if (clk.on) {  // lane fill of each verification part: cells / (128 lanes x rows x columns)
      std::vector<int32_t> h_gmax((size_t)G);
      if (hipMemcpy(h_gmax.data(), gmax, (size_t)G * 4, hipMemcpyDeviceToHost) == hipSuccess) {
        double pad[kParts] = {};
        long long nit[kParts] = {};
        for (int i = 0; i < n; ++i) {
          int g = gb[i];
          const double cols = (double)((c->len[i] + 7) / 8 * 8);
          for (int t = 0; t < 3; ++t)
            for (int q = 0; q < ng[3 * (size_t)i + t]; ++q, ++g) {
              pad[band_class(i, t)] += (double)kGroupSize * h_gmax[g] * cols;
              ++nit[band_class(i, t)];
            }
        }
        for (int pi = 0; pi < np; ++pi) {
          int lmax = 0;
          for (int qi = sit[pi].y; qi < sit[pi].z; ++qi) lmax = std::max(lmax, c->len[q2[qi]]);
          const double cols = (double)((lmax + 7) / 8 * 8);
          pad[gpart[pi]] += (double)kGroupSize * h_gmax[Gf + pi] * cols;
          ++nit[gpart[pi]];
        }
        for (int p = 0; p < kParts; ++p)
          if (nit[p])
            fprintf(stderr, "[mnhn] prefilter part %d: %lld items, %.4e cells, lane fill %.3f\n", p, nit[p],
                    part_cells[p], pad[p] > 0 ? part_cells[p] / pad[p] : 0.0);
      }
      clk.mark("lane fill statistics");
    }
    // the work of one verification part (its items, kernel form, strip and queries per group)
    auto part_work = [&](int part) {
      if (part < 4) {
        w.d_items = items + part_begin[part];
        w.n_items = part_begin[part + 1] - part_begin[part];
        w.d_qids = qids;
        w.dev_split = 1;
        w.dev_kernel = part < 2 ? KERNEL_LDS_F16 : part == 2 ? KERNEL_LDS : KERNEL_I32;
        w.dev_f16_ok = part < 2;
        w.dev_strip = part == 0 ? 32 : 0;
      } else {  // shared groups: halves (parts 4, 5 fp16, 6 int16), quarters (7, 8 fp16, 9 int16)
        const bool i16 = part == 6 || part == 9;
        w.d_items = sitems + rng[part][0];
        w.n_items = rng[part][1] - rng[part][0];
        w.d_qids = qids2;
        w.dev_split = part <= 6 ? 3 : 4;
        w.dev_kernel = i16 ? KERNEL_LDS : KERNEL_LDS_F16;
        w.dev_f16_ok = !i16;
        w.dev_strip = (part == 5 || part == 8) ? 64 : (part == 4 || part == 7) ? 32 : (S == 64 ? 64 : 32);
      }
      w.cells = part_cells[part];
    };
    if (spare && spare_cap >= kcap && spare_cap > c->keys_cap) {  // the candidate buffer takes the keys
      dev_free(c->d_keys);
      c->d_keys = spare;
      c->keys_cap = spare_cap;
      spare = nullptr;
    }
    dev_free(spare);
    for (int attempt = 0; attempt < 2; ++attempt) {
      if ((rc = ensure_keys(c, kcap))) return fail(rc);
      clk.mark("key buffer");
      if (hipMemsetAsync(c->d_nkeys, 0, 8, st) != hipSuccess) return fail(MNHN_ERR_DEVICE);
      kernel_ms = 0;
      for (int f = 0; f < 3; ++f) form_cells[f] = form_ms[f] = 0;
      for (int part = 0; part < kParts && rc == MNHN_OK; ++part) {
        part_work(part);
        if (w.n_items == 0) continue;
        rc = launch_tiles(c, w, MODE_EPS, 0, eps_max, nullptr, c->d_keys, c->keys_cap);
        kernel_ms += c->last.kernel_ms;
        if (clk.on)
          fprintf(stderr, "[mnhn] prefilter part %d: %.3f ms, %.0f GCUPS\n", part, c->last.kernel_ms,
                  c->last.kernel_ms > 0 ? part_cells[part] / (c->last.kernel_ms * 1e6) : 0.0);
        for (int f = 0; f < 3; ++f) {
          form_cells[f] += c->last.form_cells[f];
          form_ms[f] += c->last.form_ms[f];
        }
      }
      if (rc == MNHN_OK) rc = run_sub();
      if (rc == MNHN_OK && hipMemcpy(&nk, c->d_nkeys, 8, hipMemcpyDeviceToHost) != hipSuccess) rc = MNHN_ERR_DEVICE;
      if (rc) return fail(rc);
      if (nk <= c->keys_cap) break;
      kcap = nk;
    }
    clk.mark("verification kernels");
    if (nk > c->keys_cap) return fail(MNHN_ERR_NOMEM);
  } else if (!wsub.items.empty()) {  // no candidate: the exact bucket's pass alone
    unsigned long long kcap = std::max<unsigned long long>(est_sub, 1024ull);
    for (int attempt = 0; attempt < 2; ++attempt) {
      if ((rc = ensure_keys(c, kcap))) return fail(rc);
      if (hipMemsetAsync(c->d_nkeys, 0, 8, st) != hipSuccess) return fail(MNHN_ERR_DEVICE);
      kernel_ms = 0;
      for (int f = 0; f < 3; ++f) form_cells[f] = form_ms[f] = 0;
      rc = run_sub();
      if (rc == MNHN_OK && hipMemcpy(&nk, c->d_nkeys, 8, hipMemcpyDeviceToHost) != hipSuccess) rc = MNHN_ERR_DEVICE;
      if (rc) return fail(rc);
      if (nk <= c->keys_cap) break;
      kcap = nk;
    }
    clk.mark("exact bucket kernels");
    if (nk > c->keys_cap) return fail(MNHN_ERR_NOMEM);
  }
  if (m > 0 || !wsub.items.empty()) {
    c->last.kernel_ms = kernel_ms;
    c->last.pairs = (double)m + wsub.pairs;
    c->last.cells = stt.cells;
    for (int f = 0; f < 3; ++f) {
      c->last.form_cells[f] = form_cells[f];
      c->last.form_ms[f] = form_ms[f];
    }
    stt.verify_kernel_ms = kernel_ms;
  }
  stt.verify_ms = phase_ms();
  stt.neighbours = (int64_t)(nk / 2);
  rc = csr_from_keys(c, nk, out);
  stt.csr_ms = phase_ms();
  clk.mark("lists (csr)");
  cleanup();
  clk.mark("release");
  if (stats) *stats = stt;
  return rc;
}

}  // extern "C"
