// valu_probe.hip -- measures the FP32 VALU issue rate of the device, the
// compute ceiling of the stage kernels (their FP32 arithmetic is not packed:
// the build switches packed FP32 off).
//
// Each kernel runs independent dependency chains of one instruction mix with
// VGPR operands only (an SGPR operand halves the rate of v_fma_f32 on gfx950),
// enough workgroups for 16 waves per SIMD, and no memory traffic inside the
// loop.  The rate is reported in wave instructions per CU per nanosecond, the
// unit of SQ_INSTS_VALU / (CUs * kernel time).
//
// Output: one JSON line {"cus": .., "mixes": {name: rate, ..}, "peak": max}.
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstdlib>

namespace {

constexpr int kChains = 16;
constexpr int kIters = 2048;

// OP 0: fma chains; OP 1: alternating add and mul chains; OP 2: mul, add, fma
template <int OP>
__global__ __launch_bounds__(256) void k_mix(float* out, float b, float c) {
  float a[kChains];
  const float bv = b + threadIdx.x * 1e-9f, cv = c + threadIdx.x * 1e-9f;
#pragma unroll
  for (int j = 0; j < kChains; ++j) a[j] = (float)(threadIdx.x + j) * 1e-3f;
  for (int it = 0; it < kIters; ++it) {
#pragma unroll
    for (int j = 0; j < kChains; ++j) {
      if constexpr (OP == 0) a[j] = __builtin_fmaf(a[j], bv, cv);
      if constexpr (OP == 1) a[j] = (j & 1) ? a[j] * bv : a[j] + cv;
      if constexpr (OP == 2) a[j] = (j % 3 == 0) ? a[j] * bv : ((j % 3 == 1) ? a[j] + cv : __builtin_fmaf(a[j], bv, cv));
    }
  }
  float s = 0.f;
#pragma unroll
  for (int j = 0; j < kChains; ++j) s += a[j];
  out[blockIdx.x * 256 + threadIdx.x] = s;  // keeps the chains live; every index is inside the buffer
}

#define CHECK(x)                                                            \
  do {                                                                      \
    if ((x) != hipSuccess) {                                                \
      fprintf(stderr, "valu_probe: %s failed: %s\n", #x, hipGetErrorString(x)); \
      return 1;                                                             \
    }                                                                       \
  } while (0)

template <int OP>
int time_mix(float* out, int blocks, int cus, double* rate) {
  hipEvent_t e0, e1;
  CHECK(hipEventCreate(&e0));
  CHECK(hipEventCreate(&e1));
  double best = 0;
  for (int rep = 0; rep < 3; ++rep) {
    CHECK(hipEventRecord(e0, nullptr));
    hipLaunchKernelGGL(k_mix<OP>, dim3(blocks), dim3(256), 0, nullptr, out, 0.999f, 1e-4f);
    CHECK(hipGetLastError());
    CHECK(hipEventRecord(e1, nullptr));
    CHECK(hipEventSynchronize(e1));
    float ms = 0;
    CHECK(hipEventElapsedTime(&ms, e0, e1));
    // loop-body wave instructions of the mix (the loop control is scalar)
    const double winstr = (double)blocks * 4.0 * kIters * kChains;
    const double r = winstr / cus / (ms * 1e6);
    if (rep > 0 && r > best) best = r;  // the first launch warms the clocks
  }
  CHECK(hipEventDestroy(e0));
  CHECK(hipEventDestroy(e1));
  *rate = best;
  return 0;
}

}  // namespace

int main() {
  int cus = 0;
  CHECK(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, 0));
  if (cus <= 0) return 1;
  const int blocks = cus * 16;  // 4 waves per workgroup: 16 waves per SIMD
  float* out = nullptr;
  CHECK(hipMalloc(&out, (size_t)blocks * 256 * sizeof(float)));
  double r[3];
  if (time_mix<0>(out, blocks, cus, &r[0]) || time_mix<1>(out, blocks, cus, &r[1]) ||
      time_mix<2>(out, blocks, cus, &r[2]))
    return 1;
  CHECK(hipFree(out));
  double peak = r[0];
  for (double v : r) peak = v > peak ? v : peak;
  printf("{\"cus\": %d, \"unit\": \"wave-instr/CU/ns\", \"mixes\": {\"fma\": %.4f, \"add|mul\": %.4f, "
         "\"mul|add|fma\": %.4f}, \"peak\": %.4f}\n",
         cus, r[0], r[1], r[2], peak);
  return 0;
}
