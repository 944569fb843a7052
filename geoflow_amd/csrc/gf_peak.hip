// gf_peak.hip -- measured throughput peaks for the roofline report: the
// matrix cores in FP32 and FP64 and the vector units in FP64, each from a
// kernel that keeps every SIMD busy with independent dependency chains.  No
// counterpart in the reference; bench.py reports the operator kernels'
// fractions against these beside the published peaks.
#include <hip/hip_runtime.h>

#include <cstring>
#include <string>

#include "gf_internal.hpp"

namespace {

typedef float v16f __attribute__((ext_vector_type(16)));
typedef double v4d __attribute__((ext_vector_type(4)));

constexpr int kPeakWavesPerSimd = 4;
constexpr int kPeakChains = 8;

// FP32 matrix cores: v_mfma_f32_32x32x2f32, 32 x 32 x 2 multiply-adds (4,096
// flops) per instruction, eight independent accumulators per wave.
__global__ __launch_bounds__(256) void k_peak_mfma_f32(float *out, int iters) {
  v16f acc[kPeakChains];
#pragma unroll
  for (int c = 0; c < kPeakChains; c++)
#pragma unroll
    for (int q = 0; q < 16; q++) acc[c][q] = 0.f;
  const float a = 1e-3f * (float)(threadIdx.x & 63), b = 1.0f - 1e-4f * (float)blockIdx.x;
  for (int it = 0; it < iters; it++)
#pragma unroll
    for (int c = 0; c < kPeakChains; c++)
      acc[c] = __builtin_amdgcn_mfma_f32_32x32x2f32(a, b, acc[c], 0, 0, 0);
  float s = 0.f;
#pragma unroll
  for (int c = 0; c < kPeakChains; c++)
#pragma unroll
    for (int q = 0; q < 16; q++) s += acc[c][q];
  out[blockIdx.x * blockDim.x + threadIdx.x] = s;
}

// FP64 matrix cores: v_mfma_f64_16x16x4f64, 16 x 16 x 4 multiply-adds (2,048
// flops) per instruction.
__global__ __launch_bounds__(256) void k_peak_mfma_f64(double *out, int iters) {
  v4d acc[kPeakChains];
#pragma unroll
  for (int c = 0; c < kPeakChains; c++)
#pragma unroll
    for (int q = 0; q < 4; q++) acc[c][q] = 0.0;
  const double a = 1e-3 * (double)(threadIdx.x & 63), b = 1.0 - 1e-4 * (double)blockIdx.x;
  for (int it = 0; it < iters; it++)
#pragma unroll
    for (int c = 0; c < kPeakChains; c++)
      acc[c] = __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, acc[c], 0, 0, 0);
  double s = 0.0;
#pragma unroll
  for (int c = 0; c < kPeakChains; c++)
#pragma unroll
    for (int q = 0; q < 4; q++) s += acc[c][q];
  out[blockIdx.x * blockDim.x + threadIdx.x] = s;
}

// FP64 vector units: eight independent fma chains per lane (2 flops each).
__global__ __launch_bounds__(256) void k_peak_valu_f64(double *out, int iters) {
  double v[8];
#pragma unroll
  for (int c = 0; c < 8; c++) v[c] = 1e-3 * (double)(threadIdx.x + c);
  const double m = 1.0 - 1e-9 * (double)blockIdx.x, d = 1e-12;
  for (int it = 0; it < iters; it++)
#pragma unroll
    for (int c = 0; c < 8; c++) v[c] = fma(v[c], m, d);
  double s = 0.0;
#pragma unroll
  for (int c = 0; c < 8; c++) s += v[c];
  out[blockIdx.x * blockDim.x + threadIdx.x] = s;
}

}  // namespace

extern "C" int gf_peak_probe(gf_ctx *ctx, const char *which, double *tflops) {
  if (!ctx || !which || !tflops) return GF_ERR_ARG;
  const std::string w(which);
  int dev = 0, cus = 0;
  GF_HIP(ctx, hipGetDevice(&dev));
  GF_HIP(ctx, hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
  // four waves per SIMD on every CU: blocks of 256 threads (4 waves), 4 per CU
  const int nb = cus * kPeakWavesPerSimd, nt = 256;
  double flops_per_wave_iter;
  int iters;
  if (w == "mfma_f32") {
    flops_per_wave_iter = 4096.0 * kPeakChains;
    iters = 4096;
  } else if (w == "mfma_f64") {
    flops_per_wave_iter = 2048.0 * kPeakChains;
    iters = 4096;
  } else if (w == "valu_f64") {
    flops_per_wave_iter = 64.0 * 2.0 * 8.0;
    iters = 16384;
  } else {
    return gf_set_err(ctx, GF_ERR_ARG, "gf_peak_probe: mfma_f32, mfma_f64 or valu_f64");
  }
  void *out = nullptr;
  GF_HIP(ctx, hipMalloc(&out, (size_t)nb * nt * sizeof(double)));
  hipEvent_t e0, e1;
  GF_HIP(ctx, hipEventCreate(&e0));
  GF_HIP(ctx, hipEventCreate(&e1));
  auto launch = [&]() {
    if (w == "mfma_f32")
      hipLaunchKernelGGL(k_peak_mfma_f32, dim3(nb), dim3(nt), 0, ctx->stream, (float *)out, iters);
    else if (w == "mfma_f64")
      hipLaunchKernelGGL(k_peak_mfma_f64, dim3(nb), dim3(nt), 0, ctx->stream, (double *)out, iters);
    else
      hipLaunchKernelGGL(k_peak_valu_f64, dim3(nb), dim3(nt), 0, ctx->stream, (double *)out, iters);
  };
  launch();  // warm-up
  GF_HIP(ctx, hipEventRecord(e0, ctx->stream));
  const int reps = 5;
  for (int r = 0; r < reps; r++) launch();
  GF_HIP(ctx, hipEventRecord(e1, ctx->stream));
  GF_HIP(ctx, hipEventSynchronize(e1));
  float ms = 0.f;
  GF_HIP(ctx, hipEventElapsedTime(&ms, e0, e1));
  hipEventDestroy(e0);
  hipEventDestroy(e1);
  hipFree(out);
  const double waves = (double)nb * (nt / 64);
  *tflops = waves * iters * flops_per_wave_iter * reps / (ms * 1e-3) / 1e12;
  return GF_OK;
}
