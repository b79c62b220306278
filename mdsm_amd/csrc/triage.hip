// triage.hip -- kernels around the triage network: narrowband RFI
// injection for synthetic training streams, snippet extraction around
// candidate events, batch gather and class probabilities.
#include <hip/hip_runtime.h>

#include "detmath.h"
#include "triage_internal.h"

namespace mdsm {
namespace {

constexpr int TNT = 256;

__device__ __forceinline__ uint16_t bf16_bits(float f) {
  uint32_t u = __float_as_uint(f);
  u += 0x7FFFu + ((u >> 16) & 1u);
  return (uint16_t)(u >> 16);
}

// Philox4x32-10 (Salmon et al. 2011), one round key schedule per round
__device__ __forceinline__ uint32_t philox_u32(uint32_t c0, uint32_t c1, uint32_t c2, uint32_t k0,
                                               uint32_t k1) {
  uint32_t c[4] = {c0, c1, c2, 0u};
#pragma unroll
  for (int r = 0; r < 10; ++r) {
    const uint64_t p0 = (uint64_t)0xD2511F53u * c[0];
    const uint64_t p1 = (uint64_t)0xCD9E8D57u * c[2];
    const uint32_t n0 = (uint32_t)(p1 >> 32) ^ c[1] ^ k0;
    const uint32_t n1 = (uint32_t)p1;
    const uint32_t n2 = (uint32_t)(p0 >> 32) ^ c[3] ^ k1;
    const uint32_t n3 = (uint32_t)p0;
    c[0] = n0; c[1] = n1; c[2] = n2; c[3] = n3;
    k0 += 0x9E3779B9u;
    k1 += 0xBB67AE85u;
  }
  return c[0];
}

// one launch per RFI item (items on one channel add to the same samples, so
// they are applied in order, as the oracle does), threads over its samples
__global__ __launch_bounds__(TNT) void rfi_inject_kernel(const RfiItem *__restrict__ items,
                                                         uint32_t index, const float *__restrict__ gain,
                                                         uint64_t seed, float *__restrict__ raw,
                                                         uint64_t ld, uint64_t n) {
  const RfiItem it = items[index];
  float *row = raw + (uint64_t)it.chan * ld;
  const float g = gain[it.chan];
  const int64_t end = it.start + it.len < (int64_t)n ? it.start + it.len : (int64_t)n;
  for (int64_t t = it.start + blockIdx.x * (int64_t)TNT + threadIdx.x; t < end;
       t += (int64_t)gridDim.x * TNT) {
    if (t < 0) continue;
    float v;
    if (it.kind == RFI_BURST) {
      v = it.amp;
    } else if (it.kind == RFI_MAINS) {
      // sin(2 pi t / period) from IEEE operations only (detmath.h)
      double u = (double)t / it.period;
      u = u - floor(u);
      double sn, cs;
      det_sincos_turn(u, &sn, &cs);
      v = (float)(it.amp * 0.5 * (1.0 + sn));
    } else {  // telegraph: on with probability 1/4 per block of `period` samples
      const uint64_t blk = (uint64_t)t / (uint64_t)it.period;
      const uint32_t r = philox_u32((uint32_t)blk, (uint32_t)(blk >> 32), it.chan + 0x9000u,
                                    (uint32_t)seed, (uint32_t)(seed >> 32));
      v = (r & 3u) == 0 ? it.amp : 0.f;
    }
    row[t] += g * v;
  }
}

// grid (nev, 2): plane 0 = DM-time window of the dedispersed series, plane 1 =
// subband-time window of the normalised input dedispersed at the event DM
__global__ __launch_bounds__(TNT) void snippet_kernel(const float *__restrict__ out, uint32_t ndms,
                                                      uint32_t nsamp, const float *__restrict__ store,
                                                      uint64_t ld, uint32_t nchans,
                                                      const int32_t *__restrict__ table,
                                                      const int32_t *__restrict__ ev_dm,
                                                      const int32_t *__restrict__ ev_t, int SD, int ST,
                                                      int SB, float scale0, float scale1, float clip,
                                                      uint16_t *__restrict__ snip) {
  const uint32_t e = blockIdx.x;
  const int plane = blockIdx.y;
  const int dm0 = ev_dm[e], tc = ev_t[e];
  uint16_t *dst = snip + ((uint64_t)e * 2 + plane) * SD * ST;
  for (int j = threadIdx.x; j < ST; j += TNT) {
    const int64_t t = (int64_t)tc - ST / 2 + j;
    const bool tin = t >= 0 && t < (int64_t)nsamp;
    if (plane == 0) {
      for (int d = 0; d < SD; ++d) {
        const int dm = dm0 - SD / 2 + d;
        float v = (tin && dm >= 0 && dm < (int)ndms) ? out[(uint64_t)dm * nsamp + t] * scale0 : 0.f;
        v = fminf(fmaxf(v, -clip), clip);
        dst[(uint64_t)d * ST + j] = bf16_bits(v);
      }
    } else {
      for (int s = 0; s < SB; ++s) {
        const uint32_t c0 = (uint32_t)((uint64_t)s * nchans / SB);
        const uint32_t c1 = (uint32_t)((uint64_t)(s + 1) * nchans / SB);
        float acc = 0.f;
        if (tin)
          for (uint32_t c = c0; c < c1; ++c) {
            const int64_t i = t + table[(uint64_t)c * ndms + dm0];
            if (i < (int64_t)ld) acc += store[(uint64_t)c * ld + i];
          }
        float v = fminf(fmaxf(acc * scale1, -clip), clip);
        dst[(uint64_t)s * ST + j] = bf16_bits(v);
      }
    }
  }
}

// rows idx[i] of src (row = `row` uint16 elements) -> dst row i; labels too
__global__ __launch_bounds__(TNT) void gather_kernel(const uint16_t *__restrict__ src,
                                                     const int32_t *__restrict__ labels,
                                                     const uint32_t *__restrict__ idx, uint32_t n,
                                                     uint64_t row, uint16_t *__restrict__ dst,
                                                     int32_t *__restrict__ dlab) {
  const uint32_t i = blockIdx.y;
  if (i >= n) return;
  const uint64_t r = idx[i];
  const uint32_t *s = reinterpret_cast<const uint32_t *>(src + r * row);
  uint32_t *d = reinterpret_cast<uint32_t *>(dst + (uint64_t)i * row);
  for (uint64_t k = blockIdx.x * (uint64_t)TNT + threadIdx.x; k < row / 2; k += (uint64_t)gridDim.x * TNT)
    d[k] = s[k];
  if (labels && blockIdx.x == 0 && threadIdx.x == 0) dlab[i] = labels[r];
}

__global__ void prob_kernel(const float *__restrict__ logits, uint32_t n, float *__restrict__ prob) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float d = logits[2 * i + 1] - logits[2 * i];
  prob[i] = 1.f / (1.f + expf(-d));
}

}  // namespace

int launch_rfi_inject(const RfiItem *d_items, uint32_t nitems, int64_t max_len, const float *gain,
                      uint64_t seed, float *raw, uint64_t ld, uint64_t n, hipStream_t s) {
  if (nitems == 0) return MDSM_OK;
  uint64_t g = (uint64_t)(max_len + TNT - 1) / TNT;
  g = g < 1 ? 1 : (g > 1024 ? 1024 : g);
  for (uint32_t i = 0; i < nitems; ++i) {
    hipLaunchKernelGGL(rfi_inject_kernel, dim3((unsigned)g), dim3(TNT), 0, s, d_items, i, gain, seed,
                       raw, ld, n);
    MDSM_HIP(hipGetLastError());
  }
  return MDSM_OK;
}

int launch_snippets(const float *out, uint32_t ndms, uint32_t nsamp, const float *store, uint64_t ld,
                    uint32_t nchans, const int32_t *table, const int32_t *ev_dm, const int32_t *ev_t,
                    uint32_t nev, int SD, int ST, int SB, float scale0, float scale1, float clip,
                    uint16_t *snip, hipStream_t s) {
  if (nev == 0) return MDSM_OK;
  hipLaunchKernelGGL(snippet_kernel, dim3(nev, 2), dim3(TNT), 0, s, out, ndms, nsamp, store, ld, nchans,
                     table, ev_dm, ev_t, SD, ST, SB, scale0, scale1, clip, snip);
  MDSM_HIP(hipGetLastError());
  return MDSM_OK;
}

int launch_gather_rows(const uint16_t *src, const int32_t *labels, const uint32_t *idx, uint32_t n,
                       uint64_t row, uint16_t *dst, int32_t *dlab, hipStream_t s) {
  if (n == 0) return MDSM_OK;
  uint64_t g = (row / 2 + TNT - 1) / TNT;
  g = g < 1 ? 1 : (g > 64 ? 64 : g);
  hipLaunchKernelGGL(gather_kernel, dim3((unsigned)g, n), dim3(TNT), 0, s, src, labels, idx, n, row, dst,
                     dlab);
  MDSM_HIP(hipGetLastError());
  return MDSM_OK;
}

int launch_probabilities(const float *logits, uint32_t n, float *prob, hipStream_t s) {
  hipLaunchKernelGGL(prob_kernel, dim3((n + TNT - 1) / TNT), dim3(TNT), 0, s, logits, n, prob);
  MDSM_HIP(hipGetLastError());
  return MDSM_OK;
}

}  // namespace mdsm
