// Ceilings of this GPU, measured with the instructions and access patterns the step kernels use:
// dense matrix-core throughput of v_mfma_f32_16x16x4_f32 (FP32 transforms), v_mfma_f64_16x16x4_f64
// (FP64 MFMA) and v_mfma_f32_16x16x16_bf16 (split-BF16x3 transforms), and HBM bandwidth of a
// device copy, a read-only pass and a write-only pass.  The roofline fractions of the step
// kernels are reported against these beside the datasheet figures.
#include "swx_kernels.h"

namespace swx {
namespace {

typedef float v4f __attribute__((ext_vector_type(4)));
typedef double v4d __attribute__((ext_vector_type(4)));
typedef short v4s __attribute__((ext_vector_type(4)));

// Independent accumulator chains per wave: enough in flight to keep the matrix pipe issuing.
constexpr int kAcc = 8;

__global__ void __launch_bounds__(256) k_peak_mfma_f32(float* out, int iters) {
  v4f acc[kAcc];
#pragma unroll
  for (int q = 0; q < kAcc; ++q) acc[q] = v4f{0.f, 0.f, 0.f, 0.f};
  const float a = 1e-3f * (float)(threadIdx.x & 63), b = 1e-3f * (float)(blockIdx.x & 7);
  for (int it = 0; it < iters; ++it) {
#pragma unroll
    for (int q = 0; q < kAcc; ++q) acc[q] = __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, acc[q], 0, 0, 0);
  }
  float s = 0.f;
#pragma unroll
  for (int q = 0; q < kAcc; ++q) s += acc[q][0] + acc[q][1] + acc[q][2] + acc[q][3];
  out[(long long)blockIdx.x * blockDim.x + threadIdx.x] = s;
}

__global__ void __launch_bounds__(256) k_peak_mfma_f64(double* out, int iters) {
  v4d acc[kAcc];
#pragma unroll
  for (int q = 0; q < kAcc; ++q) acc[q] = v4d{0.0, 0.0, 0.0, 0.0};
  const double a = 1e-3 * (double)(threadIdx.x & 63), b = 1e-3 * (double)(blockIdx.x & 7);
  for (int it = 0; it < iters; ++it) {
#pragma unroll
    for (int q = 0; q < kAcc; ++q) acc[q] = __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, acc[q], 0, 0, 0);
  }
  double s = 0.0;
#pragma unroll
  for (int q = 0; q < kAcc; ++q) s += acc[q][0] + acc[q][1] + acc[q][2] + acc[q][3];
  out[(long long)blockIdx.x * blockDim.x + threadIdx.x] = s;
}

__global__ void __launch_bounds__(256) k_peak_mfma_bf16(float* out, int iters) {
  v4f acc[kAcc];
#pragma unroll
  for (int q = 0; q < kAcc; ++q) acc[q] = v4f{0.f, 0.f, 0.f, 0.f};
  const short sa = (short)(0x3c00 + (threadIdx.x & 63)), sb = (short)(0x3c00 + (blockIdx.x & 7));
  const v4s a = {sa, sa, sa, sa}, b = {sb, sb, sb, sb};
  for (int it = 0; it < iters; ++it) {
#pragma unroll
    for (int q = 0; q < kAcc; ++q) acc[q] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, acc[q], 0, 0, 0);
  }
  float s = 0.f;
#pragma unroll
  for (int q = 0; q < kAcc; ++q) s += acc[q][0] + acc[q][1] + acc[q][2] + acc[q][3];
  out[(long long)blockIdx.x * blockDim.x + threadIdx.x] = s;
}

// HBM passes over n 16-byte words, 4 words per thread and iteration in flight.
__global__ void __launch_bounds__(256) k_hbm_copy(uint4* __restrict__ dst, const uint4* __restrict__ src, long long n) {
  const long long stride = (long long)gridDim.x * blockDim.x;
  for (long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x; e < n; e += 4 * stride) {
    uint4 v[4];
#pragma unroll
    for (int u = 0; u < 4; ++u)
      if (e + u * stride < n) v[u] = src[e + u * stride];
#pragma unroll
    for (int u = 0; u < 4; ++u)
      if (e + u * stride < n) dst[e + u * stride] = v[u];
  }
}

__global__ void __launch_bounds__(256) k_hbm_read(const uint4* __restrict__ src, long long n, unsigned* out) {
  const long long stride = (long long)gridDim.x * blockDim.x;
  unsigned x = 0;
  for (long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x; e < n; e += 4 * stride) {
#pragma unroll
    for (int u = 0; u < 4; ++u)
      if (e + u * stride < n) {
        const uint4 v = src[e + u * stride];
        x ^= v.x ^ v.y ^ v.z ^ v.w;
      }
  }
  out[(long long)blockIdx.x * blockDim.x + threadIdx.x] = x;
}

__global__ void __launch_bounds__(256) k_hbm_write(uint4* __restrict__ dst, long long n) {
  const long long stride = (long long)gridDim.x * blockDim.x;
  for (long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x; e < n; e += stride) {
    const unsigned w = (unsigned)e;
    dst[e] = make_uint4(w, w + 1u, w + 2u, w + 3u);
  }
}

// Shortest of 5 timed launches (after 2 untimed ones), in milliseconds; 0 on a HIP error.
template <class F>
float best_ms(F&& launch, hipStream_t s, hipEvent_t e0, hipEvent_t e1) {
  for (int w = 0; w < 2; ++w) launch();
  float best = 0.f;
  for (int r = 0; r < 5; ++r) {
    if (hipEventRecord(e0, s) != hipSuccess) return 0.f;
    launch();
    if (hipEventRecord(e1, s) != hipSuccess || hipEventSynchronize(e1) != hipSuccess) return 0.f;
    float ms = 0.f;
    if (hipEventElapsedTime(&ms, e0, e1) != hipSuccess) return 0.f;
    if (best == 0.f || ms < best) best = ms;
  }
  return hipGetLastError() == hipSuccess ? best : 0.f;
}

}  // namespace

hipError_t measure_ceilings(double* out, size_t bytes, hipStream_t s) {
  int dev = 0, ncu = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e == hipSuccess) e = hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev);
  if (e != hipSuccess) return e;
  const int blocks = 8 * ncu;       // 8 waves per SIMD
  const size_t nres = (size_t)blocks * 256;
  const long long n16 = (long long)(bytes / 16);
  void *res = nullptr, *a = nullptr, *b = nullptr;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  if ((e = hipMalloc(&res, nres * sizeof(double))) != hipSuccess || (e = hipMalloc(&a, bytes)) != hipSuccess ||
      (e = hipMalloc(&b, bytes)) != hipSuccess || (e = hipEventCreate(&e0)) != hipSuccess ||
      (e = hipEventCreate(&e1)) != hipSuccess) {
    (void)hipFree(res);
    (void)hipFree(a);
    (void)hipFree(b);
    if (e0) (void)hipEventDestroy(e0);
    return e;
  }
  const int iters = 1024;
  const double mfma_per_launch = (double)blocks * 4 * iters * kAcc;  // 4 waves per block
  const float t32 = best_ms([&] { hipLaunchKernelGGL(k_peak_mfma_f32, dim3(blocks), dim3(256), 0, s, (float*)res, iters); },
                            s, e0, e1);
  const float t64 = best_ms([&] { hipLaunchKernelGGL(k_peak_mfma_f64, dim3(blocks), dim3(256), 0, s, (double*)res, iters); },
                            s, e0, e1);
  const float tbf = best_ms([&] { hipLaunchKernelGGL(k_peak_mfma_bf16, dim3(blocks), dim3(256), 0, s, (float*)res, iters); },
                            s, e0, e1);
  const int mblocks = 16 * ncu;
  (void)hipMemsetAsync(a, 1, bytes, s);
  const float tw = best_ms([&] { hipLaunchKernelGGL(k_hbm_write, dim3(mblocks), dim3(256), 0, s, (uint4*)a, n16); }, s, e0, e1);
  const float tc = best_ms([&] { hipLaunchKernelGGL(k_hbm_copy, dim3(mblocks), dim3(256), 0, s, (uint4*)b, (const uint4*)a, n16); },
                           s, e0, e1);
  const float tr = best_ms([&] { hipLaunchKernelGGL(k_hbm_read, dim3(mblocks), dim3(256), 0, s, (const uint4*)a, n16, (unsigned*)res); },
                           s, e0, e1);
  e = hipGetLastError();
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  (void)hipFree(res);
  (void)hipFree(a);
  (void)hipFree(b);
  if (e != hipSuccess) return e;
  if (t32 <= 0.f || t64 <= 0.f || tbf <= 0.f || tw <= 0.f || tc <= 0.f || tr <= 0.f) return hipErrorUnknown;
  const double mb = 16.0 * (double)n16;
  out[0] = mfma_per_launch * 2048.0 / (t32 * 1e-3) / 1e12;  // 16x16x4: 2048 flops
  out[1] = mfma_per_launch * 2048.0 / (t64 * 1e-3) / 1e12;
  out[2] = mfma_per_launch * 8192.0 / (tbf * 1e-3) / 1e12;  // 16x16x16: 8192 flops
  out[3] = 2.0 * mb / (tc * 1e-3) / 1e12;                   // bytes read + written
  out[4] = mb / (tr * 1e-3) / 1e12;
  out[5] = mb / (tw * 1e-3) / 1e12;
  return hipSuccess;
}

}  // namespace swx
