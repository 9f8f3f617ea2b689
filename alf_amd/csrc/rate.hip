// Issue-rate ceilings of the arithmetic the sweep's kernels are built from, measured on
// the device they run on: FP64 and FP32 MFMA (v_mfma_*_16x16x4), vector FP64 FMA and int8
// MFMA (v_mfma_i32_32x32x32_i8, in tera-operations per second).  Each
// wave runs eight independent accumulation chains for a fixed count, eight workgroups of
// four waves per CU; the result is flops / elapsed time of the launch.  These set the
// attainable peak the GEMMs are compared against (the published MI355X figures are
// 78.6 TF/s FP64, 157.3 TF/s FP32 and 5000 TOP/s int8 dense).
#include <vector>

#include "alfq_common.h"
#include "../../include/alfq.h"

namespace {

typedef double v4d __attribute__((ext_vector_type(4)));
typedef float v4f __attribute__((ext_vector_type(4)));
typedef int v4i __attribute__((ext_vector_type(4)));
typedef int v16i __attribute__((ext_vector_type(16)));
constexpr int RATE_CHAINS = 8;

__global__ __launch_bounds__(256) void rate_mfma64(double* out, int iters) {
  v4d acc[RATE_CHAINS];
  for (int j = 0; j < RATE_CHAINS; ++j) acc[j] = v4d{0, 0, 0, 0};
  const double a = threadIdx.x * 1e-3, b = 1.0 + blockIdx.x * 1e-6;
  for (int i = 0; i < iters; ++i) {
#pragma unroll
    for (int j = 0; j < RATE_CHAINS; ++j) acc[j] = __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, acc[j], 0, 0, 0);
  }
  double s = 0.0;
  for (int j = 0; j < RATE_CHAINS; ++j) s += acc[j][0] + acc[j][1] + acc[j][2] + acc[j][3];
  if (threadIdx.x == 0) out[blockIdx.x] = s;
}

__global__ __launch_bounds__(256) void rate_mfma32(double* out, int iters) {
  v4f acc[RATE_CHAINS];
  for (int j = 0; j < RATE_CHAINS; ++j) acc[j] = v4f{0, 0, 0, 0};
  const float a = threadIdx.x * 1e-3f, b = 1.0f + blockIdx.x * 1e-6f;
  for (int i = 0; i < iters; ++i) {
#pragma unroll
    for (int j = 0; j < RATE_CHAINS; ++j) acc[j] = __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, acc[j], 0, 0, 0);
  }
  float s = 0.f;
  for (int j = 0; j < RATE_CHAINS; ++j) s += acc[j][0] + acc[j][1] + acc[j][2] + acc[j][3];
  if (threadIdx.x == 0) out[blockIdx.x] = s;
}

__global__ __launch_bounds__(256) void rate_mfma_i8(double* out, int iters) {
  v16i acc[RATE_CHAINS];
  for (int j = 0; j < RATE_CHAINS; ++j)
    for (int r = 0; r < 16; ++r) acc[j][r] = 0;
  const int w = (int)threadIdx.x * 0x01010101, u = (int)blockIdx.x | 0x01000100;
  const v4i a = {w, w + 1, w + 2, w + 3}, b = {u, u ^ 1, u ^ 2, u ^ 3};
  for (int i = 0; i < iters; ++i) {
#pragma unroll
    for (int j = 0; j < RATE_CHAINS; ++j) acc[j] = __builtin_amdgcn_mfma_i32_32x32x32_i8(a, b, acc[j], 0, 0, 0);
  }
  long long s = 0;
  for (int j = 0; j < RATE_CHAINS; ++j)
    for (int r = 0; r < 16; ++r) s += acc[j][r];
  if (threadIdx.x == 0) out[blockIdx.x] = (double)s;
}

__global__ __launch_bounds__(256) void rate_vfma64(double* out, int iters) {
  double x[RATE_CHAINS];
  for (int j = 0; j < RATE_CHAINS; ++j) x[j] = threadIdx.x * 1e-3 + j;
  const double a = 1.0 - 1e-9, b = 1e-9 * blockIdx.x;
  for (int i = 0; i < iters; ++i) {
#pragma unroll
    for (int j = 0; j < RATE_CHAINS; ++j) x[j] = fma(x[j], a, b);
  }
  double s = 0.0;
  for (int j = 0; j < RATE_CHAINS; ++j) s += x[j];
  if (threadIdx.x == 0) out[blockIdx.x] = s;
}

}  // namespace

extern "C" int alfq_mfma_rate(int32_t kind, double* tflops) {
  if (!tflops || kind < 0 || kind > 3) { alfq_set_error("mfma_rate: bad arguments"); return ST_GENERIC; }
  NULL_STREAM_LOCK;
  int dev = 0, cus = 0;
  HIP_OK(hipGetDevice(&dev));
  HIP_OK(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
  const int blocks = 8 * (cus > 0 ? cus : 1);
  const int iters = kind == 2 ? 20000 : 2000;
  // operations of one wave per iteration: 8 MFMAs of 16 x 16 x 4 (2048 flops), 8 FMAs x 64
  // lanes, or 8 MFMAs of 32 x 32 x 32 (65536 integer operations)
  const double per_iter = kind == 2 ? RATE_CHAINS * 64.0 * 2.0 : kind == 3 ? RATE_CHAINS * 65536.0 : RATE_CHAINS * 2048.0;
  double* out = nullptr;
  HIP_OK(hipMalloc(&out, blocks * sizeof(double)));
  hipEvent_t e0 = nullptr, e1 = nullptr;
  int st = ST_NONE;
  float best = 0.f;
  if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) st = ST_GENERIC;
  for (int rep = 0; rep < 3 && st == ST_NONE; ++rep) {    // the first launch warms the clocks
    if (hipEventRecord(e0, 0) != hipSuccess) { st = ST_GENERIC; break; }
    if (kind == 0) hipLaunchKernelGGL(rate_mfma64, dim3(blocks), dim3(256), 0, 0, out, iters);
    else if (kind == 1) hipLaunchKernelGGL(rate_mfma32, dim3(blocks), dim3(256), 0, 0, out, iters);
    else if (kind == 3) hipLaunchKernelGGL(rate_mfma_i8, dim3(blocks), dim3(256), 0, 0, out, iters);
    else hipLaunchKernelGGL(rate_vfma64, dim3(blocks), dim3(256), 0, 0, out, iters);
    if (hipGetLastError() != hipSuccess || hipEventRecord(e1, 0) != hipSuccess || hipEventSynchronize(e1) != hipSuccess) {
      st = ST_GENERIC;
      break;
    }
    float ms = 0.f;
    if (hipEventElapsedTime(&ms, e0, e1) != hipSuccess) { st = ST_GENERIC; break; }
    if (rep > 0 && (best == 0.f || ms < best)) best = ms;
  }
  if (e0) hipEventDestroy(e0);
  if (e1) hipEventDestroy(e1);
  hipFree(out);
  if (st) { alfq_set_error("mfma_rate: launch failed"); return st; }
  *tflops = (double)blocks * 4.0 * iters * per_iter / (best * 1e-3) / 1e12;
  return ST_NONE;
}
