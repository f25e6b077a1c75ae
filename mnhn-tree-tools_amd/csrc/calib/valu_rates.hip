// valu_rates.hip -- issue-rate calibration of the gfx950 VALU for the op
// kinds the engine's kernels are built from: eight independent chains of one
// op per lane, 8 waves per SIMD, timed with HIP events.  Prints lane
// instructions per clock per CU at the 2.4 GHz peak clock: 128 means a wave64
// op issues every 2 cycles on each of the 4 SIMDs, 64 every 4 cycles.  Used to
// price the rooflines of the Smith-Waterman (packed 16-bit) and vector-scan
// (fp32 sub / add / fma) kernels.
#include <hip/hip_runtime.h>
#include <cstdio>
#include <cstdint>
typedef short s2 __attribute__((ext_vector_type(2)));
typedef unsigned short u2 __attribute__((ext_vector_type(2)));
typedef _Float16 h2 __attribute__((ext_vector_type(2)));

constexpr int ITER = 4096;
#define CHAINS 8

template <int KIND>
__global__ __launch_bounds__(256) void probe(int seed, int *out) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (KIND == 0) {  // fp32 add
    float a[CHAINS]; float b = (float)(seed + t) * 1e-7f;
    for (int c = 0; c < CHAINS; ++c) a[c] = (float)(t + c);
    for (int i = 0; i < ITER; ++i)
#pragma unroll
      for (int c = 0; c < CHAINS; ++c) a[c] = a[c] + b;
    float s = 0; for (int c = 0; c < CHAINS; ++c) s += a[c];
    out[t] = (int)s;
  } else if (KIND == 1) {  // fp32 fma
    float a[CHAINS]; float b = (float)(seed + t) * 1e-7f;
    for (int c = 0; c < CHAINS; ++c) a[c] = (float)(t + c);
    for (int i = 0; i < ITER; ++i)
#pragma unroll
      for (int c = 0; c < CHAINS; ++c) a[c] = __builtin_fmaf(b, b, a[c]);
    float s = 0; for (int c = 0; c < CHAINS; ++c) s += a[c];
    out[t] = (int)s;
  } else if (KIND == 2) {  // packed u16 add
    u2 a[CHAINS]; u2 b = {(unsigned short)(seed + t), (unsigned short)t};
    for (int c = 0; c < CHAINS; ++c) a[c] = (u2){(unsigned short)(t + c), (unsigned short)c};
    for (int i = 0; i < ITER; ++i)
#pragma unroll
      for (int c = 0; c < CHAINS; ++c) asm volatile("v_pk_add_u16 %0, %0, %1" : "+v"(a[c]) : "v"(b));
    int s = 0; for (int c = 0; c < CHAINS; ++c) s += a[c].x + a[c].y;
    out[t] = s;
  } else if (KIND == 3) {  // packed i16 max
    s2 a[CHAINS]; s2 b = {(short)(seed + t), (short)t};
    for (int c = 0; c < CHAINS; ++c) a[c] = (s2){(short)(t + c), (short)c};
    for (int i = 0; i < ITER; ++i)
#pragma unroll
      for (int c = 0; c < CHAINS; ++c) { a[c] = __builtin_elementwise_max(a[c], b); b = b + (s2){1, 1}; }
    int s = 0; for (int c = 0; c < CHAINS; ++c) s += a[c].x + a[c].y;
    out[t] = s;
  } else if (KIND == 4) {  // packed f16 add
    h2 a[CHAINS]; h2 b = {(_Float16)(1e-3f * (seed & 7)), (_Float16)1e-3f};
    for (int c = 0; c < CHAINS; ++c) a[c] = (h2){(_Float16)(float)c, (_Float16)(float)t};
    for (int i = 0; i < ITER; ++i)
#pragma unroll
      for (int c = 0; c < CHAINS; ++c) a[c] = a[c] + b;
    float s = 0; for (int c = 0; c < CHAINS; ++c) s += (float)a[c].x + (float)a[c].y;
    out[t] = (int)s;
  } else if (KIND == 5) {  // int32 add
    int a[CHAINS]; int b = seed + t;
    for (int c = 0; c < CHAINS; ++c) a[c] = t + c;
    for (int i = 0; i < ITER; ++i)
#pragma unroll
      for (int c = 0; c < CHAINS; ++c) asm volatile("v_add_u32 %0, %0, %1" : "+v"(a[c]) : "v"(b));
    int s = 0; for (int c = 0; c < CHAINS; ++c) s += a[c];
    out[t] = s;
  }
}

template <int KIND>
double run(int blocks, int *d_out, hipEvent_t e0, hipEvent_t e1) {
  hipLaunchKernelGGL(probe<KIND>, dim3(blocks), dim3(256), 0, 0, 1, d_out);
  (void)hipDeviceSynchronize();
  (void)hipEventRecord(e0, 0);
  for (int r = 0; r < 5; ++r) hipLaunchKernelGGL(probe<KIND>, dim3(blocks), dim3(256), 0, 0, r, d_out);
  (void)hipEventRecord(e1, 0);
  (void)hipEventSynchronize(e1);
  float ms = 0; (void)hipEventElapsedTime(&ms, e0, e1);
  return ms / 5;
}

int main() {
  int cu = 0; (void)hipDeviceGetAttribute(&cu, hipDeviceAttributeMultiprocessorCount, 0);
  int clk = 0; (void)hipDeviceGetAttribute(&clk, hipDeviceAttributeClockRate, 0);
  const int blocks = cu * 8;  // 32 waves per CU, 8 per SIMD
  int *d_out = nullptr;
  if (hipMalloc(&d_out, (size_t)blocks * 256 * sizeof(int)) != hipSuccess) return 1;
  hipEvent_t e0, e1; (void)hipEventCreate(&e0); (void)hipEventCreate(&e1);
  const double lane_instr = (double)blocks * 256 * ITER * CHAINS;
  const char *names[] = {"fp32 add", "fp32 fma", "pk u16 add", "pk i16 max (+pk add)", "pk f16 add", "i32 add"};
  double ms[6];
  ms[0] = run<0>(blocks, d_out, e0, e1); ms[1] = run<1>(blocks, d_out, e0, e1);
  ms[2] = run<2>(blocks, d_out, e0, e1); ms[3] = run<3>(blocks, d_out, e0, e1);
  ms[4] = run<4>(blocks, d_out, e0, e1); ms[5] = run<5>(blocks, d_out, e0, e1);
  printf("CUs %d, clock attr %d kHz\n", cu, clk);
  for (int k = 0; k < 6; ++k) {
    double instr = lane_instr * (k == 3 ? 2 : 1);
    double per_cu_clk = instr / (ms[k] * 1e-3) / cu / 2.4e9;
    printf("%-22s %8.3f ms  %.3e lane-instr/s  %.1f lane-instr/clk/CU at 2.4 GHz\n", names[k], ms[k],
           instr / (ms[k] * 1e-3), per_cu_clk);
  }
  (void)hipFree(d_out);
  return 0;
}
