// Achievable HBM bandwidth of the device, measured with two streaming
// kernels (16-byte loads and stores, a grid of several workgroups per CU):
// a read-only sum and a copy; and achievable LDS read bandwidth (16-byte LDS
// reads of every thread of four workgroups per CU).  Measurement aids for the
// roofline of the path's kernels; they have no counterpart in the reference.
#include <hip/hip_runtime.h>

#include <vector>

#include "satt_internal.h"

namespace satt {
namespace {

__global__ __launch_bounds__(256) void k_stream_read(const float4 *__restrict__ src, size_t n,
                                                     float *__restrict__ sink) {
  float acc = 0.0f;
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n;
       i += (size_t)gridDim.x * blockDim.x) {
    const float4 v = src[i];
    acc += v.x + v.y + v.z + v.w;
  }
  sink[blockIdx.x * (size_t)blockDim.x + threadIdx.x] = acc;
}

__global__ __launch_bounds__(256) void k_stream_copy(const float4 *__restrict__ src, size_t n,
                                                     float4 *__restrict__ dst) {
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n;
       i += (size_t)gridDim.x * blockDim.x)
    dst[i] = src[i];
}

// LDS read bandwidth: each workgroup fills kLdsVec float4 of LDS, then every
// thread reads 16 bytes per step from a rotating, bank-conflict-free index.
constexpr int kLdsVec = 2048;  // 32 KiB per workgroup
constexpr int kLdsSteps = 4096;
__global__ __launch_bounds__(256) void k_lds_read(float *__restrict__ sink) {
  __shared__ float4 buf[kLdsVec];
  for (int i = threadIdx.x; i < kLdsVec; i += blockDim.x)
    buf[i] = make_float4((float)i, 1.0f, 2.0f, 3.0f);
  __syncthreads();
  float4 acc = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
#pragma unroll 8
  for (int it = 0; it < kLdsSteps; ++it) {
    const float4 v = buf[(threadIdx.x + it * 64) & (kLdsVec - 1)];
    acc.x += v.x; acc.y += v.y; acc.z += v.z; acc.w += v.w;
  }
  sink[blockIdx.x * (size_t)blockDim.x + threadIdx.x] = acc.x + acc.y + acc.z + acc.w;
}

__global__ void k_stream_fill(float4 *p, size_t n) {
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n;
       i += (size_t)gridDim.x * blockDim.x)
    p[i] = make_float4(1.0f, 2.0f, 3.0f, 4.0f);
}

}  // namespace

int stream_probe(hipStream_t s, int cus, size_t nbytes, int reps, double &read_gbs,
                 double &copy_gbs) {
  const size_t n = nbytes / sizeof(float4);
  const unsigned blocks = (unsigned)(cus * 8);
  float4 *a = nullptr, *b = nullptr;
  float *sink = nullptr;
  if (hipMalloc(&a, n * sizeof(float4)) != hipSuccess) return kENoMem;
  if (hipMalloc(&b, n * sizeof(float4)) != hipSuccess) {
    (void)hipFree(a);
    return kENoMem;
  }
  if (hipMalloc(&sink, (size_t)blocks * 256 * sizeof(float)) != hipSuccess) {
    (void)hipFree(a);
    (void)hipFree(b);
    return kENoMem;
  }
  hipLaunchKernelGGL(k_stream_fill, dim3(blocks), dim3(256), 0, s, a, n);
  hipEvent_t e0, e1;
  (void)hipEventCreate(&e0);
  (void)hipEventCreate(&e1);
  float best_r = 0, best_c = 0;
  for (int r = 0; r <= reps; ++r) {  // the first round warms up
    float ms = 0;
    (void)hipEventRecord(e0, s);
    hipLaunchKernelGGL(k_stream_read, dim3(blocks), dim3(256), 0, s, a, n, sink);
    (void)hipEventRecord(e1, s);
    (void)hipEventSynchronize(e1);
    (void)hipEventElapsedTime(&ms, e0, e1);
    if (r > 0 && (best_r == 0 || ms < best_r)) best_r = ms;
    (void)hipEventRecord(e0, s);
    hipLaunchKernelGGL(k_stream_copy, dim3(blocks), dim3(256), 0, s, a, n, b);
    (void)hipEventRecord(e1, s);
    (void)hipEventSynchronize(e1);
    (void)hipEventElapsedTime(&ms, e0, e1);
    if (r > 0 && (best_c == 0 || ms < best_c)) best_c = ms;
  }
  const int st = hipGetLastError() == hipSuccess ? kOk : kEHip;
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  (void)hipFree(a);
  (void)hipFree(b);
  (void)hipFree(sink);
  const double bytes = (double)n * sizeof(float4);
  read_gbs = best_r > 0 ? bytes / (best_r * 1e-3) / 1e9 : 0;
  copy_gbs = best_c > 0 ? 2 * bytes / (best_c * 1e-3) / 1e9 : 0;
  return st;
}

int lds_probe(hipStream_t s, int cus, int reps, double &read_gbs) {
  const unsigned blocks = (unsigned)(cus * 4);
  float *sink = nullptr;
  if (hipMalloc(&sink, (size_t)blocks * 256 * sizeof(float)) != hipSuccess) return kENoMem;
  hipEvent_t e0, e1;
  (void)hipEventCreate(&e0);
  (void)hipEventCreate(&e1);
  float best = 0;
  for (int r = 0; r <= reps; ++r) {  // the first round warms up
    float ms = 0;
    (void)hipEventRecord(e0, s);
    hipLaunchKernelGGL(k_lds_read, dim3(blocks), dim3(256), 0, s, sink);
    (void)hipEventRecord(e1, s);
    (void)hipEventSynchronize(e1);
    (void)hipEventElapsedTime(&ms, e0, e1);
    if (r > 0 && (best == 0 || ms < best)) best = ms;
  }
  const int st = hipGetLastError() == hipSuccess ? kOk : kEHip;
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  (void)hipFree(sink);
  const double bytes = (double)blocks * 256 * kLdsSteps * sizeof(float4);
  read_gbs = best > 0 ? bytes / (best * 1e-3) / 1e9 : 0;
  return st;
}

}  // namespace satt
