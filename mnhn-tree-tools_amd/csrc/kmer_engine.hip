// kmer_engine.hip -- gfx950 k-mer frequency tables (include/mnhn_kmer.h), the
// computation behind the reference's fasta2kmer (kmers.c:159-243, 288-376).
//
// The reference scans each sequence once per k-mer of the alphabet (4^k or
// 20^k scans of every sequence).  Here one workgroup owns one sequence at a
// time: every thread hashes the windows i = tid, tid + 256, ... < len - k into
// their column (first base in the low bits for DNA, first residue most
// significant for protein) and counts them with an atomic add in an LDS
// histogram of the row, which is then written out as int16 (the wrap of the
// reference's short counter).  Rows wider than the LDS histogram
// (4^k or 20^k > 32768 columns) count into an int32 row in device memory and
// are narrowed by a second kernel.  Sequences are processed in batches whose
// output fits a fixed device budget; the bytes of all sequences are uploaded
// once.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <vector>

#include "engine_internal.h"
#include "mnhn_kmer.h"

using namespace mnhn_detail;

namespace {

constexpr int kThreads = 256;
constexpr int kLdsColumns = 32768;            // int32 histogram: 128 KB of LDS
constexpr size_t kBatchBytes = size_t(1) << 30;  // device output per batch

struct CodeTable {
  int8_t c[256];
};

CodeTable make_table(int protein) {
  CodeTable t;
  memset(t.c, -1, sizeof(t.c));
  const char *alpha = protein ? "ACDEFGHIKLMNPQRSTVWY" : "ACGT";
  for (int i = 0; alpha[i]; ++i) t.c[(unsigned char)alpha[i]] = (int8_t)i;
  return t;
}

// Column of the window starting at s, or -1 when a byte is outside the
// alphabet.
template <bool PROTEIN>
__device__ __forceinline__ int window_column(const uint8_t *s, int k, const int8_t *code) {
  unsigned idx = 0;
  bool ok = true;
  for (int p = 0; p < k; ++p) {
    const int c = code[s[p]];
    ok = ok && c >= 0;
    if (PROTEIN)
      idx = idx * 20u + (unsigned)(c & 31);
    else
      idx |= (unsigned)(c & 3) << (2 * p);
  }
  return ok ? (int)idx : -1;
}

template <bool PROTEIN>
__global__ __launch_bounds__(kThreads) void kmer_lds_kernel(const uint8_t *__restrict__ bytes,
                                                            const long long *__restrict__ off,
                                                            const int *__restrict__ len, int first, int count,
                                                            int k, int nk, CodeTable tab,
                                                            int16_t *__restrict__ out) {
  extern __shared__ int hist[];
  __shared__ int8_t code[256];
  const int tid = threadIdx.x;
  code[tid] = tab.c[tid];
  for (int s = blockIdx.x; s < count; s += gridDim.x) {
    __syncthreads();
    for (int j = tid; j < nk; j += kThreads) hist[j] = 0;
    __syncthreads();
    const uint8_t *seq = bytes + off[first + s];
    const int windows = len[first + s] - k;
    for (int i = tid; i < windows; i += kThreads) {
      const int col = window_column<PROTEIN>(seq + i, k, code);
      if (col >= 0) atomicAdd(&hist[col], 1);
    }
    __syncthreads();
    int16_t *o = out + (size_t)s * (size_t)nk;
    for (int j = tid; j < nk; j += kThreads) o[j] = (int16_t)hist[j];
  }
}

template <bool PROTEIN>
__global__ __launch_bounds__(kThreads) void kmer_global_kernel(const uint8_t *__restrict__ bytes,
                                                               const long long *__restrict__ off,
                                                               const int *__restrict__ len, int first, int count,
                                                               int k, long long nk, CodeTable tab,
                                                               int *__restrict__ rows) {
  __shared__ int8_t code[256];
  const int tid = threadIdx.x;
  code[tid] = tab.c[tid];
  __syncthreads();
  for (int s = blockIdx.x; s < count; s += gridDim.x) {
    const uint8_t *seq = bytes + off[first + s];
    const int windows = len[first + s] - k;
    int *row = rows + (size_t)s * (size_t)nk;
    for (int i = tid; i < windows; i += kThreads) {
      const int col = window_column<PROTEIN>(seq + i, k, code);
      if (col >= 0) atomicAdd(&row[col], 1);
    }
  }
}

__global__ void narrow_kernel(const int *__restrict__ in, long long total, int16_t *__restrict__ out) {
  for (long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x; t < total;
       t += (long long)gridDim.x * blockDim.x)
    out[t] = (int16_t)in[t];
}

struct Ctx {
  hipStream_t st = nullptr;
  mnhn_detail_stage stage;
  uint8_t *d_bytes = nullptr;
  long long *d_off = nullptr;
  int *d_len = nullptr;
  int16_t *d_out = nullptr;
  int *d_rows = nullptr;
  ~Ctx() {
    if (st) (void)hipStreamSynchronize(st);
    dev_free(d_bytes);
    dev_free(d_off);
    dev_free(d_len);
    dev_free(d_out);
    dev_free(d_rows);
    stage_free(stage);
    if (st) (void)hipStreamDestroy(st);
  }
};

}  // namespace

extern "C" {

size_t mnhn_kmer_columns(int kmer_length, int protein) {
  const int max_k = protein ? 7 : 15;
  if (kmer_length < 1 || kmer_length > max_k) return 0;
  size_t c = 1;
  for (int i = 0; i < kmer_length; ++i) c *= protein ? 20 : 4;
  return c;
}

int mnhn_kmer_frequencies(const char *const *seqs, const size_t *lens, int n, int kmer_length, int protein,
                          int16_t *out) {
  const size_t nk = mnhn_kmer_columns(kmer_length, protein);
  if (nk == 0 || n < 0 || (n > 0 && (!seqs || !lens || !out))) return MNHN_ERR_ARG;
  if (n == 0) return MNHN_OK;
  for (int i = 0; i < n; ++i) {
    if (!seqs[i] && lens[i]) return MNHN_ERR_ARG;
    if (lens[i] > (size_t)0x7fffffff) return MNHN_ERR_RANGE;
  }
  Ctx c;
  if (hipStreamCreateWithFlags(&c.st, hipStreamNonBlocking) != hipSuccess) {
    c.st = nullptr;
    return MNHN_ERR_DEVICE;
  }
  int device = 0, num_cu = 0;
  if (hipGetDevice(&device) != hipSuccess ||
      hipDeviceGetAttribute(&num_cu, hipDeviceAttributeMultiprocessorCount, device) != hipSuccess)
    return MNHN_ERR_DEVICE;

  // all sequence bytes, packed
  std::vector<long long> off(n);
  std::vector<int> len(n);
  size_t total = 0;
  for (int i = 0; i < n; ++i) {
    off[i] = (long long)total;
    len[i] = (int)lens[i];
    total += lens[i];
  }
  std::vector<uint8_t> packed(std::max<size_t>(total, 1));
  for (int i = 0; i < n; ++i)
    if (lens[i]) memcpy(&packed[off[i]], seqs[i], lens[i]);
  int rc;
  if ((rc = dev_alloc(&c.d_bytes, packed.size())) || (rc = dev_alloc(&c.d_off, (size_t)n)) ||
      (rc = dev_alloc(&c.d_len, (size_t)n)))
    return rc;
  if (hipMemcpyAsync(c.d_bytes, packed.data(), packed.size(), hipMemcpyHostToDevice, c.st) != hipSuccess ||
      hipMemcpyAsync(c.d_off, off.data(), (size_t)n * sizeof(long long), hipMemcpyHostToDevice, c.st) !=
          hipSuccess ||
      hipMemcpyAsync(c.d_len, len.data(), (size_t)n * sizeof(int), hipMemcpyHostToDevice, c.st) != hipSuccess)
    return MNHN_ERR_DEVICE;

  const bool lds = nk <= (size_t)kLdsColumns;
  const size_t row_bytes = nk * (lds ? sizeof(int16_t) : sizeof(int16_t) + sizeof(int));
  const int batch = (int)std::max<size_t>(1, std::min<size_t>((size_t)n, kBatchBytes / row_bytes));
  if ((rc = dev_alloc(&c.d_out, (size_t)batch * nk))) return rc;
  if (!lds && (rc = dev_alloc(&c.d_rows, (size_t)batch * nk))) return rc;
  const CodeTable tab = make_table(protein);
  const size_t lds_bytes = nk * sizeof(int);
  if (lds) {
    auto fn = protein ? reinterpret_cast<const void *>(&kmer_lds_kernel<true>)
                      : reinterpret_cast<const void *>(&kmer_lds_kernel<false>);
    if (hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds_bytes) != hipSuccess)
      return MNHN_ERR_DEVICE;
  }
  for (int first = 0; first < n; first += batch) {
    const int count = std::min(batch, n - first);
    const unsigned grid = (unsigned)std::min<long long>(count, (long long)num_cu * 8);
    if (lds) {
      if (protein)
        hipLaunchKernelGGL(kmer_lds_kernel<true>, dim3(grid), dim3(kThreads), lds_bytes, c.st, c.d_bytes, c.d_off,
                           c.d_len, first, count, kmer_length, (int)nk, tab, c.d_out);
      else
        hipLaunchKernelGGL(kmer_lds_kernel<false>, dim3(grid), dim3(kThreads), lds_bytes, c.st, c.d_bytes,
                           c.d_off, c.d_len, first, count, kmer_length, (int)nk, tab, c.d_out);
    } else {
      const long long cells = (long long)count * (long long)nk;
      if (hipMemsetAsync(c.d_rows, 0, (size_t)cells * sizeof(int), c.st) != hipSuccess) return MNHN_ERR_DEVICE;
      if (protein)
        hipLaunchKernelGGL(kmer_global_kernel<true>, dim3(grid), dim3(kThreads), 0, c.st, c.d_bytes, c.d_off,
                           c.d_len, first, count, kmer_length, (long long)nk, tab, c.d_rows);
      else
        hipLaunchKernelGGL(kmer_global_kernel<false>, dim3(grid), dim3(kThreads), 0, c.st, c.d_bytes, c.d_off,
                           c.d_len, first, count, kmer_length, (long long)nk, tab, c.d_rows);
      if (hipGetLastError() != hipSuccess) return MNHN_ERR_DEVICE;
      hipLaunchKernelGGL(narrow_kernel, dim3(stride_blocks(cells, kThreads)), dim3(kThreads), 0, c.st, c.d_rows,
                         cells, c.d_out);
    }
    if (hipGetLastError() != hipSuccess) return MNHN_ERR_DEVICE;
    if (hipStreamSynchronize(c.st) != hipSuccess) return MNHN_ERR_DEVICE;
    rc = copy_d2h(c.st, c.stage, out + (size_t)first * nk, c.d_out, (size_t)count * nk * sizeof(int16_t));
    if (rc) return rc;
  }
  return MNHN_OK;
}

}  // extern "C"
