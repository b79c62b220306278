// rfi.hip -- bandpass and RFI-clipping kernels on raw [nchans][ld] rows.
//
// Every kernel works on the n samples starting at column col0 of each row
// (the reference's `shift`) and keeps the arithmetic order of the
// reference kernels (multibeam_dedispersion_kernel.cu:302-450, reductions
// over BANDPASS_THREADS = 64 lanes, cache_brute_force.h:16), so results
// are bit-identical to them:
//
//   bandpass_sum  one wave per channel; lane i sums samples i, i+64, ... in
//                 FP64 (16 loads in flight per lane), then the halving tree
//                 32..1 with lane shuffles.  One wave per channel is the
//                 most parallelism the fixed summation order allows; the
//                 stage reads the chunk once and is small next to
//                 dedispersion.
//   channel_flag  one wave per (channel, block): lane-strided FP32 sums,
//                 the same tree, mean > rz(fitted) + thresh flags the block
//                 and its neighbours.
//   channel_clip  flagged blocks take rz(fitted[c]).
//   spectrum_clip one thread per sample, sequential FP32 sum over channels
//                 (coalesced across the wave), out-of-range spectra take the
//                 fitted bandpass.
//   voltage_power (re, im) int16 pairs -> re^2 + im^2.
// All are HBM-bound streaming passes.
#include <hip/hip_runtime.h>

#include "internal.h"

namespace mdsm {
namespace {

constexpr int RFI_NT = 256;

template <typename T>
__device__ __forceinline__ T tree64(T v) {
  // p[i] += p[i + w] for w = 32, 16, ..., 1 (lane 0 holds the result)
  for (int w = 32; w >= 1; w >>= 1) v += __shfl_down(v, w, 64);
  return v;
}

__global__ __launch_bounds__(RFI_NT) void bandpass_sum_kernel(const float *__restrict__ x,
                                                              uint64_t ld, uint32_t nchans,
                                                              uint64_t col0, uint64_t n,
                                                              double *__restrict__ out) {
  const uint32_t c = blockIdx.x * (RFI_NT / 64) + (threadIdx.x >> 6);
  const uint32_t lane = threadIdx.x & 63;
  if (c >= nchans) return;
  const float *row = x + (uint64_t)c * ld + col0;
  double acc = 0.0;
  uint64_t s = lane;
  constexpr int U = 16;
  for (; s + (U - 1) * 64 < n; s += U * 64) {
    float v[U];
#pragma unroll
    for (int u = 0; u < U; ++u) v[u] = __builtin_nontemporal_load(row + s + u * 64);
#pragma unroll
    for (int u = 0; u < U; ++u) acc += (double)v[u];
  }
  for (; s < n; s += 64) acc += (double)row[s];
  acc = tree64(acc);
  if (lane == 0) out[c] = acc / (double)n;
}

__global__ void fill_u8_kernel(uint8_t *p, uint64_t n) {
  for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n;
       i += (uint64_t)gridDim.x * blockDim.x)
    p[i] = 0;
}

__global__ __launch_bounds__(RFI_NT) void channel_flag_kernel(const float *__restrict__ x,
                                                              uint64_t ld, uint32_t nchans,
                                                              uint64_t col0, uint64_t n,
                                                              const double *__restrict__ fitted,
                                                              uint32_t cb, uint32_t nb,
                                                              float thresh,
                                                              uint8_t *__restrict__ flags) {
  const uint64_t item = (uint64_t)blockIdx.x * (RFI_NT / 64) + (threadIdx.x >> 6);
  const uint32_t lane = threadIdx.x & 63;
  if (item >= (uint64_t)nchans * nb) return;
  const uint32_t c = (uint32_t)(item / nb), b = (uint32_t)(item - (uint64_t)c * nb);
  const uint64_t start = (uint64_t)b * cb;
  const uint64_t len = start + cb <= n ? cb : n - start;
  const float *row = x + (uint64_t)c * ld + col0 + start;
  float acc = 0.0f;
  for (uint64_t s = lane; s < len; s += 64) acc += row[s];
  acc = tree64(acc);
  if (lane == 0) {
    const float bp = __double2float_rz(fitted[c]);
    if (acc / (float)len > bp + thresh) {
      uint8_t *f = flags + (uint64_t)c * nb;
      f[b] = 1;
      if (b > 0) f[b - 1] = 1;
      if (b + 2 < nb) f[b + 1] = 1;
    }
  }
}

__global__ __launch_bounds__(RFI_NT) void channel_clip_kernel(float *__restrict__ x, uint64_t ld,
                                                              uint64_t col0, uint64_t n,
                                                              const double *__restrict__ fitted,
                                                              uint32_t cb, uint32_t nb,
                                                              const uint8_t *__restrict__ flags) {
  const uint32_t b = blockIdx.x, c = blockIdx.y;
  if (!flags[(uint64_t)c * nb + b]) return;
  const float bp = __double2float_rz(fitted[c]);
  const uint64_t start = (uint64_t)b * cb;
  const uint64_t end = start + cb <= n ? start + cb : n;
  float *row = x + (uint64_t)c * ld + col0;
  for (uint64_t s = start + threadIdx.x; s < end; s += RFI_NT) row[s] = bp;
}

__global__ __launch_bounds__(RFI_NT) void spectrum_clip_kernel(float *__restrict__ x, uint64_t ld,
                                                               uint32_t nchans, uint64_t col0,
                                                               uint64_t n,
                                                               const double *__restrict__ fitted,
                                                               float hi, float lo,
                                                               uint32_t *__restrict__ count) {
  const uint64_t s = blockIdx.x * (uint64_t)RFI_NT + threadIdx.x;
  if (s >= n) return;
  float *col = x + col0 + s;
  float m = 0.0f;
  uint32_t c = 0;
  constexpr int U = 8;
  for (; c + U <= nchans; c += U) {
    float v[U];
#pragma unroll
    for (int u = 0; u < U; ++u) v[u] = col[(uint64_t)(c + u) * ld];
#pragma unroll
    for (int u = 0; u < U; ++u) m += v[u];
  }
  for (; c < nchans; ++c) m += col[(uint64_t)c * ld];
  m /= (float)nchans;
  if (m > hi || m < lo) {
    for (c = 0; c < nchans; ++c) col[(uint64_t)c * ld] = __double2float_rz(fitted[c]);
    if (count) atomicAdd(count, 1u);
  }
}

__global__ __launch_bounds__(RFI_NT) void voltage_power_kernel(float *__restrict__ x, uint64_t ld,
                                                               uint64_t col0, uint64_t n) {
  const uint64_t s = blockIdx.x * (uint64_t)RFI_NT + threadIdx.x;
  if (s >= n) return;
  float *p = x + (uint64_t)blockIdx.y * ld + col0 + s;
  const uint32_t bits = __float_as_uint(*p);
  const int re = (int16_t)(bits & 0xFFFFu), im = (int16_t)(bits >> 16);
  *p = (float)(re * re + im * im);
}

inline uint32_t grid_for(uint64_t n, uint32_t per) { return (uint32_t)((n + per - 1) / per); }

}  // namespace

int launch_bandpass_sum(const float *x, uint64_t ld, uint32_t nchans, uint64_t col0, uint64_t n,
                        double *out, hipStream_t s) {
  hipLaunchKernelGGL(bandpass_sum_kernel, dim3(grid_for(nchans, RFI_NT / 64)), dim3(RFI_NT), 0, s, x,
                     ld, nchans, col0, n, out);
  MDSM_HIP(hipGetLastError());
  return MDSM_OK;
}

int launch_channel_flag(const float *x, uint64_t ld, uint32_t nchans, uint64_t col0, uint64_t n,
                        const double *fitted, uint32_t cb, float thresh, uint8_t *flags,
                        hipStream_t s) {
  const uint32_t nb = (uint32_t)((n + cb - 1) / cb);
  const uint64_t items = (uint64_t)nchans * nb;
  hipLaunchKernelGGL(fill_u8_kernel, dim3(grid_for(items, RFI_NT) < 4096 ? grid_for(items, RFI_NT) : 4096),
                     dim3(RFI_NT), 0, s, flags, items);
  MDSM_HIP(hipGetLastError());
  hipLaunchKernelGGL(channel_flag_kernel, dim3(grid_for(items, RFI_NT / 64)), dim3(RFI_NT), 0, s, x,
                     ld, nchans, col0, n, fitted, cb, nb, thresh, flags);
  MDSM_HIP(hipGetLastError());
  return MDSM_OK;
}

int launch_channel_clip(float *x, uint64_t ld, uint32_t nchans, uint64_t col0, uint64_t n,
                        const double *fitted, uint32_t cb, const uint8_t *flags, hipStream_t s) {
  const uint32_t nb = (uint32_t)((n + cb - 1) / cb);
  hipLaunchKernelGGL(channel_clip_kernel, dim3(nb, nchans), dim3(RFI_NT), 0, s, x, ld, col0, n,
                     fitted, cb, nb, flags);
  MDSM_HIP(hipGetLastError());
  return MDSM_OK;
}

int launch_spectrum_clip(float *x, uint64_t ld, uint32_t nchans, uint64_t col0, uint64_t n,
                         const double *fitted, float hi, float lo, uint32_t *count, hipStream_t s) {
  hipLaunchKernelGGL(spectrum_clip_kernel, dim3(grid_for(n, RFI_NT)), dim3(RFI_NT), 0, s, x, ld,
                     nchans, col0, n, fitted, hi, lo, count);
  MDSM_HIP(hipGetLastError());
  return MDSM_OK;
}

int launch_voltage_power(float *x, uint64_t ld, uint32_t nchans, uint64_t col0, uint64_t n,
                         hipStream_t s) {
  hipLaunchKernelGGL(voltage_power_kernel, dim3(grid_for(n, RFI_NT), nchans), dim3(RFI_NT), 0, s, x,
                     ld, col0, n);
  MDSM_HIP(hipGetLastError());
  return MDSM_OK;
}

}  // namespace mdsm
