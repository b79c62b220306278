// mulaw.hip -- logarithmic ("mu-law") quantisation of raw beam data for
// the disk writer (quantise_power_data / quantise_complex_data,
// src/mdsm-multibeam/multibeam_dedispersion_writer.cpp:21-118), one thread
// per value, HBM-bound (4 bytes in, 1 out).  The arithmetic is the
// reference's, value for value: a 32768-entry log10 table built on the host,
// float products, the quantiser ((int)(d/Q)*Q + Q/2) * 2^bits, and the
// byte conversion through int.  Indices the reference would read out of
// its table (values above the table range, |-32768|) are clamped.
#include <hip/hip_runtime.h>

#include "internal.h"

namespace mdsm {
namespace {

constexpr int MNT = 256;
constexpr int LUT_N = 32768;

__global__ __launch_bounds__(MNT) void max_kernel(const float *__restrict__ x, uint64_t n,
                                                  float *__restrict__ part) {
  __shared__ float s[MNT];
  float m = 0.f;  // the reference starts its maximum at 0
  for (uint64_t i = blockIdx.x * (uint64_t)MNT + threadIdx.x; i < n; i += (uint64_t)gridDim.x * MNT)
    m = x[i] > m ? x[i] : m;
  s[threadIdx.x] = m;
  __syncthreads();
  for (int w = MNT / 2; w >= 1; w >>= 1) {
    if (threadIdx.x < w) s[threadIdx.x] = s[threadIdx.x + w] > s[threadIdx.x] ? s[threadIdx.x + w] : s[threadIdx.x];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = s[0];
}

__global__ __launch_bounds__(MNT) void mulaw_power_kernel(const float *__restrict__ x, uint64_t n,
                                                          const float *__restrict__ lut, float inv_max,
                                                          float log1pmu, float Q, float invQ,
                                                          uint32_t bitrange, uint8_t *__restrict__ out) {
  for (uint64_t i = blockIdx.x * (uint64_t)MNT + threadIdx.x; i < n; i += (uint64_t)gridDim.x * MNT) {
    float d = x[i] * inv_max * (float)LUT_N;
    int k = (int)d;
    k = k < 0 ? 0 : (k >= LUT_N ? LUT_N - 1 : k);
    d = lut[k] * log1pmu;
    const double v = ((float)(int)(d * invQ) * Q + Q * 0.5) * (double)bitrange;
    out[i] = (uint8_t)(int)v;
  }
}

__device__ __forceinline__ uint8_t quant3(float datum, float Q, float invQ, uint32_t bitrange) {
  const double v = ((float)(signed char)(datum * invQ) * Q + Q * 0.5) * (double)bitrange;
  return (uint8_t)(int)v;
}

__global__ __launch_bounds__(MNT) void mulaw_complex_kernel(const int16_t *__restrict__ x, uint64_t n,
                                                            const float *__restrict__ lut, float log1pmu,
                                                            float Q, float invQ, uint32_t bitrange,
                                                            uint8_t *__restrict__ out) {
  for (uint64_t i = blockIdx.x * (uint64_t)MNT + threadIdx.x; i < n; i += (uint64_t)gridDim.x * MNT) {
    const int re = x[2 * i], im = x[2 * i + 1];
    const int ar = re < 0 ? -re : re, ai = im < 0 ? -im : im;
    const uint8_t qr = quant3(lut[ar < LUT_N ? ar : LUT_N - 1] * log1pmu, Q, invQ, bitrange);
    const uint8_t qi = quant3(lut[ai < LUT_N ? ai : LUT_N - 1] * log1pmu, Q, invQ, bitrange);
    out[i] = (uint8_t)((re < 0 ? 0x80 : 0) | ((qr & 0x07) << 4) | (im < 0 ? 0x08 : 0) | (qi & 0x07));
  }
}

inline unsigned grid(uint64_t n) {
  const uint64_t g = (n + MNT - 1) / MNT;
  return (unsigned)(g < 1 ? 1 : (g > 8192 ? 8192 : g));
}

}  // namespace

int launch_max(const float *x, uint64_t n, float *part, uint32_t nparts, hipStream_t s) {
  hipLaunchKernelGGL(max_kernel, dim3(nparts), dim3(MNT), 0, s, x, n, part);
  MDSM_HIP(hipGetLastError());
  return MDSM_OK;
}

int launch_mulaw_power(const float *x, uint64_t n, const float *lut, float inv_max, float log1pmu,
                       float Q, float invQ, uint32_t bitrange, uint8_t *out, hipStream_t s) {
  hipLaunchKernelGGL(mulaw_power_kernel, dim3(grid(n)), dim3(MNT), 0, s, x, n, lut, inv_max, log1pmu, Q,
                     invQ, bitrange, out);
  MDSM_HIP(hipGetLastError());
  return MDSM_OK;
}

int launch_mulaw_complex(const int16_t *x, uint64_t n, const float *lut, float log1pmu, float Q,
                         float invQ, uint32_t bitrange, uint8_t *out, hipStream_t s) {
  hipLaunchKernelGGL(mulaw_complex_kernel, dim3(grid(n)), dim3(MNT), 0, s, x, n, lut, log1pmu, Q, invQ,
                     bitrange, out);
  MDSM_HIP(hipGetLastError());
  return MDSM_OK;
}

}  // namespace mdsm
