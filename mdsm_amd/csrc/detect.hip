// detect.hip -- bandpass normalisation, per-block statistics, boxcar
// detection with wavefront compaction, median / detrend filters and the
// device-side synthetic data generator (gfx950).
//
// Reductions use one fixed order everywhere: lane i of a 256-thread
// workgroup sums elements i, i+256, ... sequentially in FP64, then a binary
// tree p[i] += p[i+s] for s = 128..1.  With IEEE FP64 adds and no
// contraction (-ffp-contract=off) the statistics are reproducible bit for bit
// by a CPU restatement of the same order.  These kernels are HBM-bound.
#include <hip/hip_fp16.h>

#include <algorithm>
#include <mutex>

#include "detmath.h"
#include "internal.h"

namespace mdsm {
namespace {

constexpr int NT = 256;

// Sum of the block's 256 per-thread partials (a, q) in one fixed pairing,
// p[i] += p[i + s] for s = 128, 64, ..., 1: the two levels that cross waves
// go through LDS, the six inside the first wave are wavefront shuffles (lane
// l adds lane l + s), so the order of the FP64 adds is the oracle's.  The
// totals land in p1[0] and p2[0], visible to every thread on return.
__device__ __forceinline__ void tree256(double a, double q, double *p1, double *p2) {
  const int i = threadIdx.x;
  if (i >= 128) {
    p1[i] = a;
    p2[i] = q;
  }
  __syncthreads();
  if (i < 128) {
    a = a + p1[i + 128];
    q = q + p2[i + 128];
    if (i >= 64) {
      p1[i] = a;
      p2[i] = q;
    }
  }
  __syncthreads();
  if (i < 64) {
    a = a + p1[i + 64];
    q = q + p2[i + 64];
#pragma unroll
    for (int s = 32; s > 0; s >>= 1) {
      a = a + __shfl_down(a, s);
      q = q + __shfl_down(q, s);
    }
    if (i == 0) {
      p1[0] = a;
      p2[0] = q;
    }
  }
  __syncthreads();
}

__device__ __forceinline__ void moments(double s1, double s2, uint64_t n, double *mean,
                                        double *stdv) {
  const double m = s1 / (double)n;
  const double v = s2 / (double)n - m * m;
  *mean = m;
  *stdv = v > 0.0 ? sqrt(v) : 0.0;
}

// ----------------------------------------------------------- channel stats
__global__ __launch_bounds__(NT) void channel_stats_kernel(const float *__restrict__ raw,
                                                           uint64_t ld, uint64_t n,
                                                           double *__restrict__ mean,
                                                           double *__restrict__ stdv) {
  __shared__ double p1[NT], p2[NT];
  const uint32_t c = blockIdx.x;
  const float *row = raw + (uint64_t)c * ld;
  double a = 0.0, b = 0.0;
  for (uint64_t j = threadIdx.x; j < n; j += NT) {
    const double v = (double)row[j];
    a = a + v;
    b = b + v * v;
  }
  tree256(a, b, p1, p2);
  if (threadIdx.x == 0) moments(p1[0], p2[0], n, &mean[c], &stdv[c]);
}

// ------------------------------------------------------------- normalise
__device__ __forceinline__ uint8_t quant_u8(double v) {
  v = v < -4.0 ? -4.0 : v;
  v = v > 4.0 ? 4.0 : v;
  return (uint8_t)floor((v + 4.0) * 31.875 + 0.5);
}

template <int DT>
__global__ __launch_bounds__(NT) void normalise_kernel(const float *__restrict__ raw, uint64_t ld,
                                                       uint64_t n, const double *__restrict__ mean,
                                                       const double *__restrict__ stdv,
                                                       void *__restrict__ out, uint64_t ldo) {
  const uint32_t c = blockIdx.y;
  const double m = mean[c];
  const double s = stdv[c];
  const double inv = s > 0.0 ? 1.0 / s : 0.0;
  const float *row = raw + (uint64_t)c * ld;
  for (uint64_t t = (uint64_t)blockIdx.x * NT + threadIdx.x; t < n; t += (uint64_t)gridDim.x * NT) {
    const double v = ((double)row[t] - m) * inv;
    if (DT == MDSM_F32) reinterpret_cast<float *>(out)[(uint64_t)c * ldo + t] = (float)v;
    else if (DT == MDSM_F16)
      reinterpret_cast<__half *>(out)[(uint64_t)c * ldo + t] = __float2half_rn((float)v);
    else reinterpret_cast<uint8_t *>(out)[(uint64_t)c * ldo + t] = quant_u8(v);
  }
}

// ------------------------------------------------------------ block stats
__global__ __launch_bounds__(NT) void block_stats_kernel(const float *__restrict__ x,
                                                         uint32_t nsamp, uint32_t L, uint32_t nb,
                                                         double *__restrict__ mean,
                                                         double *__restrict__ stdv) {
  __shared__ double p1[NT], p2[NT];
  const uint32_t b = blockIdx.x, d = blockIdx.y;
  const uint64_t t0 = (uint64_t)b * L;
  const uint32_t n = (uint32_t)min((uint64_t)L, (uint64_t)nsamp - t0);
  const float *row = x + (uint64_t)d * nsamp + t0;
  double a = 0.0, q = 0.0;
  for (uint32_t j = threadIdx.x; j < n; j += NT) {
    const double v = (double)row[j];
    a = a + v;
    q = q + v * v;
  }
  tree256(a, q, p1, p2);
  if (threadIdx.x == 0) moments(p1[0], p2[0], n, &mean[(uint64_t)d * nb + b], &stdv[(uint64_t)d * nb + b]);
}

// ----------------------------------------------------------------- detect
// Tile of DT_T samples of one DM row, staged in LDS as FP64 with a 64-sample
// halo; the dyadic boxcar levels S_k(t) = S_{k-1}(t) + S_{k-1}(t + 2^(k-1))
// are built in place, one barrier pair per level.
constexpr int DT_T = 1024;
constexpr int DT_HALO = 64;
constexpr int DT_Q = DT_T / NT;

__constant__ double kSqrtW[7] = {1.0, 1.4142135623730951, 2.0, 2.8284271247461903,
                                 4.0, 5.656854249492381, 8.0};

__device__ __forceinline__ void emit(bool hit, const mdsm_cand_t &cand, mdsm_cand_t *cands,
                                     uint32_t cap, uint32_t *count) {
  // wave-level compaction: one atomic per wave, lane prefix from the ballot
  const unsigned long long mask = __ballot(hit);
  if (mask == 0ull) return;
  const uint32_t lane = threadIdx.x & 63u;
  const uint32_t leader = (uint32_t)__ffsll((long long)mask) - 1u;
  uint32_t base = 0;
  if (lane == leader) base = atomicAdd(count, (uint32_t)__popcll(mask));
  base = __shfl(base, (int)leader);
  const unsigned long long lt = lane ? (mask & ((1ull << lane) - 1ull)) : 0ull;
  const uint32_t idx = base + (uint32_t)__popcll(lt);
  if (hit && idx < cap) cands[idx] = cand;
}

__global__ __launch_bounds__(NT) void detect_kernel(const float *__restrict__ x, uint32_t nsamp,
                                                    uint32_t L, uint32_t nb,
                                                    const double *__restrict__ mean,
                                                    const double *__restrict__ stdv, int nwidths,
                                                    double thresh, mdsm_cand_t *__restrict__ cands,
                                                    uint32_t cap, uint32_t *__restrict__ count) {
  constexpr int NE = DT_T + DT_HALO;           // staged samples (tile + halo)
  constexpr int NQ = (NE + NT - 1) / NT;       // of them per thread, i = tid + q*NT
  __shared__ double lev[NE];
  const uint32_t d = blockIdx.y;
  const uint32_t t0 = blockIdx.x * DT_T;
  const float *row = x + (uint64_t)d * nsamp;
  // this thread's values of the current level, kept in registers (LDS holds
  // them for the neighbours' t + 2^(k-1) reads)
  double cur[NQ];
#pragma unroll
  for (int q = 0; q < NQ; ++q) {
    const uint32_t i = threadIdx.x + q * NT;
    const uint32_t t = t0 + i;
    cur[q] = (i < NE && t < nsamp) ? (double)row[t] : 0.0;
    if (i < NE) lev[i] = cur[q];
  }
  // block statistics of the thread's own samples, loaded once
  double m[DT_Q], sd[DT_Q];
#pragma unroll
  for (int q = 0; q < DT_Q; ++q) {
    const uint32_t t = t0 + threadIdx.x + q * NT;
    const uint64_t b = (uint64_t)d * nb + (t < nsamp ? t / L : 0u);
    m[q] = mean[b];
    sd[q] = stdv[b];
  }
  // A level can only be the maximum of a candidate if its S/N reaches thresh.
  // a >= lim * den (lim = thresh * (1 - 2^-40), thresh > 0) never rejects
  // such a level: correctly rounded a / den >= thresh needs
  // a >= thresh * den * (1 - 2^-53), above the rounded bound; and a level it
  // rejects has a / den < thresh * (1 - 2^-41), so its quotient rounds below
  // thresh.  The exact FP64 division then runs only for the levels that pass,
  // and best / bk equal those of dividing at every level.  Outside
  // 2^-960 < den < 2^960 every level divides.
  const double lim = thresh > 0.0 ? thresh * (1.0 - 0x1p-40) : -INFINITY;
  double best[DT_Q];
  int bk[DT_Q];
#pragma unroll
  for (int q = 0; q < DT_Q; ++q) { best[q] = -INFINITY; bk[q] = -1; }
  for (int k = 0; k < nwidths; ++k) {
    const uint32_t w = 1u << k;
    if (k > 0) {
      // S_k(i) = S_{k-1}(i) + S_{k-1}(i + 2^(k-1))
      const uint32_t h = w >> 1;
      double nv[NQ];
      __syncthreads();
#pragma unroll
      for (int q = 0; q < NQ; ++q) {
        const uint32_t i = threadIdx.x + q * NT;
        nv[q] = (i + h < (uint32_t)NE) ? cur[q] + lev[i + h] : 0.0;
      }
      __syncthreads();
#pragma unroll
      for (int q = 0; q < NQ; ++q) {
        const uint32_t i = threadIdx.x + q * NT;
        cur[q] = nv[q];
        if (i < (uint32_t)NE) lev[i] = cur[q];
      }
    }
#pragma unroll
    for (int q = 0; q < DT_Q; ++q) {
      const uint32_t t = t0 + threadIdx.x + q * NT;
      if ((uint64_t)t + w <= nsamp && sd[q] > 0.0) {
        const double a = cur[q] - (double)w * m[q];
        const double den = sd[q] * kSqrtW[k];
        // (the bound's relative error holds while lim * den stays normal)
        const bool in_range = den > 0x1p-960 && den < 0x1p960;
        if (!in_range || a >= lim * den) {
          const double snr = a / den;
          if (snr > best[q]) { best[q] = snr; bk[q] = k; }
        }
      }
    }
  }
#pragma unroll
  for (int q = 0; q < DT_Q; ++q) {
    const uint32_t t = t0 + threadIdx.x + q * NT;
    const bool hit = bk[q] >= 0 && best[q] >= thresh;
    mdsm_cand_t cand{(int32_t)d, (int32_t)t, bk[q], (float)best[q]};
    emit(hit, cand, cands, cap, count);
  }
}

// Register-resident detector (block lengths that are multiples of 16): each
// lane holds DW_R = 16 consecutive samples of one DM row and builds every
// boxcar level in place in FP64 registers,
// S_k(i) = S_{k-1}(i) + S_{k-1}(i + h), h = 2^(k-1); the terms past the
// lane's own run come from the next lane (h <= 16) or the lane after it
// (h = 32) by wave shuffles, so a level costs no LDS round trip and no
// barrier.  A wave covers 64 runs; after seven levels the runs of lanes
// 0..56 are complete (a level-k value of lane i depends on lanes up to
// i + 1, 2, 3, 4, 5, 7 for k = 1..6, and lane 63's missing neighbour only
// reaches lanes 57..63), so a wave owns DW_SEG = 57 * 16 samples and reads the
// next 112 again (12 % extra reads, mostly L2 hits).  Each run lies inside
// one statistics block (L % 16 == 0), so a lane forms its block's w * mean
// and threshold bounds once per level instead of once per sample.  Same
// sums, same S/N arithmetic and same pass bound as detect_kernel, so the
// candidates and S/N values are identical.
constexpr int DW_R = 16;
constexpr int DW_LANES = 57;
constexpr int DW_SEG = DW_LANES * DW_R;

// The value of lane + 1 (0 in lane 63): DPP wave_shl:1 on both halves, a
// VALU move instead of a shuffle through the LDS crossbar.
__device__ __forceinline__ double from_next_lane(double x) {
  const uint64_t u = __builtin_bit_cast(uint64_t, x);
  const uint32_t lo = (uint32_t)__builtin_amdgcn_update_dpp(0, (int)(uint32_t)u, 0x130, 0xF, 0xF, true);
  const uint32_t hi = (uint32_t)__builtin_amdgcn_update_dpp(0, (int)(uint32_t)(u >> 32), 0x130, 0xF, 0xF, true);
  return __builtin_bit_cast(double, ((uint64_t)hi << 32) | lo);
}

__global__ __launch_bounds__(NT) void detect_wave_kernel(
    const float *__restrict__ x, uint32_t nsamp, uint32_t L, uint32_t nb,
    const double *__restrict__ mean, const double *__restrict__ stdv, int nwidths, double thresh,
    mdsm_cand_t *__restrict__ cands, uint32_t cap, uint32_t *__restrict__ count) {
  const uint32_t lane = threadIdx.x & 63u;
  const uint32_t d = blockIdx.y;
  const uint64_t seg0 = ((uint64_t)blockIdx.x * (NT / 64) + (threadIdx.x >> 6)) * DW_SEG;
  if (seg0 >= nsamp) return;  // whole wave; this kernel has no barriers
  const uint64_t t0 = seg0 + (uint64_t)DW_R * lane;
  // the row through a buffer resource bounded to nsamp floats: samples past
  // the row end read as 0 (they only feed levels that run past nsamp)
  const float *row = x + (uint64_t)d * nsamp;
  const __amdgpu_buffer_rsrc_t rs =
      __builtin_amdgcn_make_buffer_rsrc(const_cast<float *>(row), (short)0, (int)(4u * nsamp), 0x00020000);
  double v[DW_R];
  const int byte0 = (int)min(t0, (uint64_t)nsamp) * 4;
  // one range-checked dword per sample (no branch on the row end: the
  // hardware returns 0 past num_records)
#pragma unroll
  for (int i = 0; i < DW_R; ++i)
    v[i] = (double)__builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(rs, byte0 + 4 * i, 0, 0));

  // the run's block statistics and per-level bounds (see detect_kernel for
  // the pass bound a >= lim * den)
  const uint32_t b = (uint32_t)(min(t0, (uint64_t)nsamp - 1u) / L);
  const double m = mean[(uint64_t)d * nb + b], sd = stdv[(uint64_t)d * nb + b];
  const bool live = sd > 0.0;
  const double lim = thresh > 0.0 ? thresh * (1.0 - 0x1p-40) : -INFINITY;
  // samples of the run whose level-k window ends inside the row: i < n_in_k
  const int64_t room = (int64_t)nsamp - (int64_t)t0;  // samples from t0 to the row end
  double best[DW_R];
  uint32_t bk = 0xFFFFFFFFu;  // level of best, 4 bits per sample (>= 8: none)
  uint32_t bk2 = 0xFFFFFFFFu;
#pragma unroll
  for (int i = 0; i < DW_R; ++i) best[i] = -INFINITY;
  // Level k: a = S_k - w * mean for all 16 samples and one test of the
  // whole run against the pass bound; only a run with a passing sample takes
  // the (rare) branch that checks the window against the row end, divides
  // and keeps the best level.
  auto test_level = [&](int k) {
    const double w = (double)(1u << k);
    const double wm = w * m;
    const double den = sd * kSqrtW[k];
    const bool in_range = den > 0x1p-960 && den < 0x1p960;
    const double bound = in_range ? lim * den : -INFINITY;
    // One compare per sample: fl(v - wm) >= bound needs v >= wm + bound -
    // 2^-52 |bound| (the subtraction's rounding), and vt lies below that by a
    // margin of 2^-48 (|wm| + |bound|), which covers the roundings of vt's
    // own two operations; so no sample that passes the exact test below is
    // filtered out here.
    const double vt = in_range ? (wm + bound) - (fabs(wm) + fabs(bound)) * 0x1p-48 : -INFINITY;
    bool any = false;
#pragma unroll
    for (int i = 0; i < DW_R; ++i) any |= v[i] >= vt;
    if (live && any) {
      // (a is formed again here, so the common path holds no array of it)
#pragma unroll
      for (int i = 0; i < DW_R; ++i) {
        const double a = v[i] - wm;
        if ((int64_t)i + (int64_t)(1u << k) <= room && a >= bound) {
          const double snr = a / den;
          if (snr > best[i]) {
            best[i] = snr;
            if (i < 8) bk = (bk & ~(15u << (4 * i))) | ((uint32_t)k << (4 * i));
            else bk2 = (bk2 & ~(15u << (4 * (i - 8)))) | ((uint32_t)k << (4 * (i - 8)));
          }
        }
      }
    }
  };
  test_level(0);
#pragma unroll
  for (int k = 1; k < 7; ++k) {
    if (k >= nwidths) break;
    const int h = 1 << (k - 1);
    if (h < DW_R) {
      // the first h values of the next lane, taken before any is updated
      double nx[DW_R / 2];
#pragma unroll
      for (int j = 0; j < DW_R / 2; ++j)
        if (j < h) nx[j] = from_next_lane(v[j]);
#pragma unroll
      for (int i = 0; i < DW_R; ++i) v[i] = v[i] + (i + h < DW_R ? v[i + h] : nx[i + h - DW_R]);
    } else {
      // h = 16: the same element of the next lane; h = 32: of the lane after
      // it (every lane reads the old value before adding)
#pragma unroll
      for (int i = 0; i < DW_R; ++i) {
        const double n1 = from_next_lane(v[i]);
        v[i] = v[i] + (h == DW_R ? n1 : from_next_lane(n1));
      }
    }
    test_level(k);
  }
  const bool own = lane < (uint32_t)DW_LANES;
#pragma unroll
  for (int i = 0; i < DW_R; ++i) {
    const uint32_t k = i < 8 ? (bk >> (4 * i)) & 15u : (bk2 >> (4 * (i - 8))) & 15u;
    const bool hit = own && k < 8u && best[i] >= thresh;
    mdsm_cand_t cand{(int32_t)d, (int32_t)(t0 + (uint64_t)i), (int32_t)k, (float)best[i]};
    emit(hit, cand, cands, cap, count);
  }
}

// ------------------------------------------------- global sums / threshold
// Reference-semantics statistics over the region rows [0, nrows) x columns
// [0, ncols) of a row-major array with row stride ld, in one fixed order:
// chunk g = r * nb + b is columns [b * GS_COLS, (b + 1) * GS_COLS) of row r
// (nb = ceil(ncols / GS_COLS)); workgroup g sums its chunk with the r256
// pattern (lane i: elements i, i + 256, ... in order; then tree256) into
// partial g; global_finish_kernel sums the partials with the same pattern
// (lane i: partials i, i + 256, ...; tree256) and forms
// mean = S1 / n, rms = sqrt(S2 / n - mean^2) as the reference's
// process_brute_clustering does (multibeam_dedispersion_output.cpp:206-215).
__global__ __launch_bounds__(NT) void global_chunk_sums_kernel(const float *__restrict__ x,
                                                               uint64_t ld, uint32_t ncols,
                                                               uint32_t nb,
                                                               double *__restrict__ partials) {
  __shared__ double p1[NT], p2[NT];
  const uint32_t g = blockIdx.x;
  const uint32_t r = g / nb, b = g - r * nb;
  const uint32_t c0 = b * GS_COLS;
  const uint32_t n = min(GS_COLS, ncols - c0);
  const float *row = x + (uint64_t)r * ld + c0;
  double a = 0.0, q = 0.0;
  for (uint32_t j = threadIdx.x; j < n; j += NT) {
    const double v = (double)row[j];
    a = a + v;
    q = q + v * v;
  }
  tree256(a, q, p1, p2);
  if (threadIdx.x == 0) {
    partials[2u * (uint64_t)g] = p1[0];
    partials[2u * (uint64_t)g + 1u] = p2[0];
  }
}

__global__ __launch_bounds__(NT) void global_finish_kernel(const double *__restrict__ partials,
                                                           uint32_t nparts, double count,
                                                           double *__restrict__ stats) {
  __shared__ double p1[NT], p2[NT];
  double a = 0.0, q = 0.0;
  for (uint32_t j = threadIdx.x; j < nparts; j += NT) {
    a = a + partials[2u * (uint64_t)j];
    q = q + partials[2u * (uint64_t)j + 1u];
  }
  tree256(a, q, p1, p2);
  if (threadIdx.x == 0) {
    const double mean = p1[0] / count;
    stats[0] = mean;
    stats[1] = sqrt(p2[0] / count - mean * mean);
  }
}

// Candidate iff x - mean >= rms * thresh in FP64 (thresh a float, as the
// reference's survey field), snr = (x - mean) / rms; one thread per sample
// of the rows [0, nrows) x columns [0, ncols) (stride ld), grid (column
// tiles, rows).  mean and rms come from stats[0..1] on the device, or from
// the arguments when stats is null.
__global__ __launch_bounds__(NT) void threshold_global_kernel(const float *__restrict__ x,
                                                              uint64_t ld, uint32_t ncols,
                                                              const double *__restrict__ stats,
                                                              double mean_arg, double rms_arg,
                                                              float thresh,
                                                              mdsm_cand_t *__restrict__ cands,
                                                              uint32_t cap,
                                                              uint32_t *__restrict__ count) {
  const double mean = stats ? stats[0] : mean_arg;
  const double rms = stats ? stats[1] : rms_arg;
  const double lim = rms * (double)thresh;
  const uint32_t r = blockIdx.y;
  const float *row = x + (uint64_t)r * ld;
  for (uint32_t base = blockIdx.x * NT; base < ncols; base += gridDim.x * NT) {
    const uint32_t j = base + threadIdx.x;
    const bool in = j < ncols;
    const double v = in ? (double)row[j] : 0.0;
    const bool hit = in && (v - mean >= lim);
    mdsm_cand_t cand{(int32_t)r, (int32_t)j, 0, (float)((v - mean) / rms)};
    emit(hit, cand, cands, cap, count);
  }
}

__global__ void zero_u32_kernel(uint32_t *p, uint32_t n) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0u;
}

// ----------------------------------------------------------------- median
__global__ __launch_bounds__(NT) void median5_kernel(const float *__restrict__ x, uint32_t nsamp,
                                                     float *__restrict__ out) {
  const uint32_t d = blockIdx.y;
  const uint32_t t = blockIdx.x * NT + threadIdx.x;
  if (t >= nsamp) return;
  const float *r = x + (uint64_t)d * nsamp;
  float v[5];
#pragma unroll
  for (int k = 0; k < 5; ++k) {
    int i = (int)t + k - 2;
    i = i < 0 ? 0 : (i >= (int)nsamp ? (int)nsamp - 1 : i);
    v[k] = r[i];
  }
#pragma unroll
  for (int i = 1; i < 5; ++i)
#pragma unroll
    for (int j = i; j > 0; --j)
      if (v[j] < v[j - 1]) { const float s = v[j]; v[j] = v[j - 1]; v[j - 1] = s; }
  out[(uint64_t)d * nsamp + t] = v[2];
}

// ------------------------------------------------------------ legacy detector
// process_brute, src/mdsm/dedispersion_output.cpp:135-213.
//
// Statistics: consecutive blocks of block_len values are summed in FP64 in
// index order, one lane per block, because that order fixes the rounding of
// every block total (the host then folds the block totals in block order).
// A block is read with 16-byte loads when its start is 16-byte aligned.
template <bool VEC>
__global__ __launch_bounds__(64) void legacy_block_sums_kernel(const float *__restrict__ x,
                                                               uint32_t block_len,
                                                               uint32_t nblocks,
                                                               double *__restrict__ sums) {
  const uint32_t b = blockIdx.x * 64u + threadIdx.x;
  if (b >= nblocks) return;
  const float *p = x + (uint64_t)b * block_len;
  double t = 0.0, t2 = 0.0;
  if (VEC) {
    const float4 *q = reinterpret_cast<const float4 *>(p);
    for (uint32_t j = 0; j < block_len / 4u; ++j) {
      const float4 v = q[j];
      const double a = (double)v.x, c = (double)v.y, d = (double)v.z, e = (double)v.w;
      t = t + a;
      t2 = t2 + a * a;
      t = t + c;
      t2 = t2 + c * c;
      t = t + d;
      t2 = t2 + d * d;
      t = t + e;
      t2 = t2 + e * e;
    }
  } else {
    for (uint32_t j = 0; j < block_len; ++j) {
      const double a = (double)p[j];
      t = t + a;
      t2 = t2 + a * a;
    }
  }
  sums[2u * b] = t;
  sums[2u * b + 1u] = t2;
}

// The selection network of src/mdsm/dedispersion_output.cpp:14-53: it picks
// one of its five arguments, so the result (signed zeros included) is the one
// that network returns.
__device__ __forceinline__ float median5_network(float n1, float n2, float n3, float n4,
                                                 float n5) {
  float a = n1, b = n2, c = n3, d = n4, t;
  if (b < a) { t = a; a = b; b = t; }
  if (d < c) { t = c; c = d; d = t; }
  if (c < a) { t = b; b = d; d = t; c = a; }
  a = n5;
  if (b < a) { t = a; a = b; b = t; }
  if (a < c) { t = b; b = d; d = t; a = c; }
  return d < a ? d : a;
}

// Median of five over samples l..l+4 of every row, for l < nsamp - 4, then
// the threshold in FP32 as written (no contraction): with per-chunk
// statistics (cmean != nullptr) against cmean/crms[l / spc] with snr =
// (median - mean) / rms, else against the global gmean/grms with snr =
// median / grms; only rows whose DM lowdm + k*dmstep exceeds 1 report.
__global__ __launch_bounds__(NT) void legacy_threshold_kernel(
    const float *__restrict__ x, uint32_t nsamp, float lowdm, float dmstep, float thresh,
    float gmean, float grms, const float *__restrict__ cmean, const float *__restrict__ crms,
    uint32_t spc, mdsm_cand_t *__restrict__ cands, uint32_t cap, uint32_t *__restrict__ count) {
  const uint32_t k = blockIdx.y;
  const uint32_t l = blockIdx.x * NT + threadIdx.x;
  const float *r = x + (uint64_t)k * nsamp;
  const float thisdm = __fadd_rn(lowdm, __fmul_rn((float)k, dmstep));
  bool hit = false;
  float snr = 0.0f;
  if (l + 4u < nsamp && thisdm > 1.0f) {
    const float med = median5_network(r[l], r[l + 1], r[l + 2], r[l + 3], r[l + 4]);
    if (cmean) {
      const uint32_t i = l / spc;
      const float m = cmean[i], sd = crms[i];
      const float dev = __fsub_rn(med, m);
      hit = dev >= __fmul_rn(sd, thresh);
      snr = __fdiv_rn(dev, sd);
    } else {
      hit = __fsub_rn(med, gmean) >= __fmul_rn(grms, thresh);
      snr = __fdiv_rn(med, grms);
    }
  }
  mdsm_cand_t cand{(int32_t)k, (int32_t)l, 0, snr};
  emit(hit, cand, cands, cap, count);
}

// ---------------------------------------------------------------- detrend
__global__ __launch_bounds__(NT) void detrend_kernel(float *__restrict__ x, uint32_t L) {
  __shared__ double p1[NT], p2[NT];
  float *y = x + (uint64_t)blockIdx.x * L;
  const double half = -(double)L / 2.0 + 0.5;
  double sy = 0.0, sxy = 0.0;
  for (uint32_t i = threadIdx.x; i < L; i += NT) {
    const double xi = half + (double)i;
    sy = sy + (double)y[i];
    sxy = sxy + xi * (double)y[i];
  }
  tree256(sy, sxy, p1, p2);
  // sum of x_i^2 for x_i = i - (L-1)/2 is (L^3 - L)/12
  const double sxx = ((double)L * (double)L * (double)L - (double)L) / 12.0;
  const double a = p1[0] / (double)L, b = p2[0] / sxx;
  __syncthreads();
  double ss = 0.0;
  for (uint32_t i = threadIdx.x; i < L; i += NT) {
    const double r = (double)y[i] - (a + b * (half + (double)i));
    ss = ss + r * r;
  }
  tree256(ss, 0.0, p1, p2);
  const double sd = sqrt(p1[0] / (double)L);
  for (uint32_t i = threadIdx.x; i < L; i += NT) {
    const double r = (double)y[i] - (a + b * (half + (double)i));
    y[i] = (float)(r / sd);
  }
}

// -------------------------------------------------------------- generator
__device__ __forceinline__ void philox(uint32_t c[4], uint32_t k0, uint32_t k1) {
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
}

__device__ __forceinline__ double u24(uint32_t x) {
  return ((double)(x >> 8) + 0.5) * (1.0 / 16777216.0);
}

__global__ __launch_bounds__(NT) void generate_kernel(uint64_t seed, int64_t t0, uint64_t n,
                                                      const float *__restrict__ gain,
                                                      const float *__restrict__ offset,
                                                      uint32_t npulses, uint32_t nchans,
                                                      const int64_t *__restrict__ p_start,
                                                      const int32_t *__restrict__ p_width,
                                                      const float *__restrict__ p_amp,
                                                      const int64_t *__restrict__ p_delay,
                                                      float *__restrict__ raw, uint64_t ld) {
  const uint32_t c = blockIdx.y;
  const float g = gain[c], o = offset[c];
  float *row = raw + (uint64_t)c * ld;
  // one thread per quad of samples q = floor(t / 4)
  const int64_t q0 = t0 >> 2;
  const int64_t qlast = (t0 + (int64_t)n - 1) >> 2;
  for (int64_t q = q0 + (int64_t)blockIdx.x * NT + threadIdx.x; q <= qlast;
       q += (int64_t)gridDim.x * NT) {
    uint32_t ctr[4] = {(uint32_t)q, (uint32_t)((uint64_t)q >> 32), c, 0u};
    philox(ctr, (uint32_t)seed, (uint32_t)(seed >> 32));
    // Box-Muller with log and sin/cos from IEEE operations only (detmath.h)
    const double r0 = sqrt(-2.0 * det_log(u24(ctr[0])));
    const double r1 = sqrt(-2.0 * det_log(u24(ctr[2])));
    double s0, c0, s1, c1;
    det_sincos_turn(u24(ctr[1]), &s0, &c0);
    det_sincos_turn(u24(ctr[3]), &s1, &c1);
    float z[4] = {(float)(r0 * c0), (float)(r0 * s0), (float)(r1 * c1), (float)(r1 * s1)};
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      const int64_t t = q * 4 + k;
      if (t < t0 || t >= t0 + (int64_t)n) continue;
      float v = z[k];
      for (uint32_t p = 0; p < npulses; ++p) {
        const int64_t s = p_start[p] + p_delay[(uint64_t)p * nchans + c];
        if (t >= s && t < s + p_width[p]) v = v + p_amp[p];
      }
      row[t - t0] = g * v + o;
    }
  }
}

int grid_for(uint64_t work, int per_block, int cap) {
  uint64_t g = (work + per_block - 1) / per_block;
  if (g < 1) g = 1;
  if (g > (uint64_t)cap) g = cap;
  return (int)g;
}

}  // namespace

int launch_channel_stats(const float *raw, uint64_t ld, uint32_t nchans, uint64_t n, double *mean,
                         double *stdv, hipStream_t s) {
  hipLaunchKernelGGL(channel_stats_kernel, dim3(nchans), dim3(NT), 0, s, raw, ld, n, mean, stdv);
  MDSM_HIP(hipGetLastError());
  return MDSM_OK;
}

int launch_normalise(const float *raw, uint64_t ld, uint32_t nchans, uint64_t n,
                     const double *mean, const double *stdv, void *out, int out_dtype,
                     uint64_t ldo, hipStream_t s) {
  const dim3 grid(grid_for(n, NT * 8, 4096), nchans);
  switch (storage_dtype(out_dtype)) {
    case MDSM_F32:
      hipLaunchKernelGGL(normalise_kernel<MDSM_F32>, grid, dim3(NT), 0, s, raw, ld, n, mean, stdv, out, ldo);
      break;
    case MDSM_F16:
      hipLaunchKernelGGL(normalise_kernel<MDSM_F16>, grid, dim3(NT), 0, s, raw, ld, n, mean, stdv, out, ldo);
      break;
    case MDSM_U8:
      hipLaunchKernelGGL(normalise_kernel<MDSM_U8>, grid, dim3(NT), 0, s, raw, ld, n, mean, stdv, out, ldo);
      break;
    default:
      set_error("unknown output dtype %d", out_dtype);
      return MDSM_ERR_ARG;
  }
  MDSM_HIP(hipGetLastError());
  return MDSM_OK;
}

int launch_block_stats(const float *x, uint32_t nrows, uint32_t nsamp, uint32_t L, double *mean,
                       double *stdv, hipStream_t s) {
  const uint32_t nb = (nsamp + L - 1) / L;
  hipLaunchKernelGGL(block_stats_kernel, dim3(nb, nrows), dim3(NT), 0, s, x, nsamp, L, nb, mean, stdv);
  MDSM_HIP(hipGetLastError());
  return MDSM_OK;
}

int launch_detect(const float *x, uint32_t nrows, uint32_t nsamp, uint32_t L, const double *mean,
                  const double *stdv, int nwidths, double thresh, mdsm_cand_t *cands, uint32_t cap,
                  uint32_t *count, hipStream_t s) {
  const uint32_t nb = (nsamp + L - 1) / L;
  if (L % DW_R == 0 && nsamp < (1u << 29)) {
    // register-resident form: four waves of DW_SEG samples per workgroup
    const uint64_t per_wg = (uint64_t)DW_SEG * (NT / 64);
    const uint32_t nwg = (uint32_t)(((uint64_t)nsamp + per_wg - 1u) / per_wg);
    hipLaunchKernelGGL(detect_wave_kernel, dim3(nwg, nrows), dim3(NT), 0, s, x, nsamp, L, nb, mean,
                       stdv, nwidths, thresh, cands, cap, count);
    MDSM_HIP(hipGetLastError());
    return MDSM_OK;
  }
  const uint32_t ntile = (nsamp + DT_T - 1) / DT_T;
  hipLaunchKernelGGL(detect_kernel, dim3(ntile, nrows), dim3(NT), 0, s, x, nsamp, L, nb, mean, stdv,
                     nwidths, thresh, cands, cap, count);
  MDSM_HIP(hipGetLastError());
  return MDSM_OK;
}

uint32_t global_stats_parts(uint32_t nrows, uint32_t ncols) {
  return nrows * ((ncols + GS_COLS - 1u) / GS_COLS);
}

int launch_global_stats(const float *x, uint64_t ld, uint32_t nrows, uint32_t ncols,
                        double *partials, double *stats, hipStream_t s) {
  const uint32_t nb = (ncols + GS_COLS - 1u) / GS_COLS;
  const uint32_t nparts = nrows * nb;
  hipLaunchKernelGGL(global_chunk_sums_kernel, dim3(nparts), dim3(NT), 0, s, x, ld, ncols, nb,
                     partials);
  MDSM_HIP(hipGetLastError());
  hipLaunchKernelGGL(global_finish_kernel, dim3(1), dim3(NT), 0, s, partials, nparts,
                     (double)nrows * (double)ncols, stats);
  MDSM_HIP(hipGetLastError());
  return MDSM_OK;
}

int launch_threshold_global(const float *x, uint64_t ld, uint32_t nrows, uint32_t ncols,
                            const double *stats, double mean, double rms, float thresh,
                            mdsm_cand_t *cands, uint32_t cap, uint32_t *count, hipStream_t s) {
  const uint32_t gx = std::min<uint32_t>((ncols + NT * 4u - 1u) / (NT * 4u), 4096u);
  hipLaunchKernelGGL(threshold_global_kernel, dim3(gx, nrows), dim3(NT), 0, s, x, ld, ncols, stats,
                     mean, rms, thresh, cands, cap, count);
  MDSM_HIP(hipGetLastError());
  return MDSM_OK;
}

int launch_zero_u32(uint32_t *p, uint32_t n, hipStream_t s) {
  hipLaunchKernelGGL(zero_u32_kernel, dim3((n + 255) / 256), dim3(256), 0, s, p, n);
  MDSM_HIP(hipGetLastError());
  return MDSM_OK;
}

int launch_median5(const float *x, uint32_t nrows, uint32_t nsamp, float *out, hipStream_t s) {
  hipLaunchKernelGGL(median5_kernel, dim3((nsamp + NT - 1) / NT, nrows), dim3(NT), 0, s, x, nsamp, out);
  MDSM_HIP(hipGetLastError());
  return MDSM_OK;
}

int launch_legacy_block_sums(const float *x, uint32_t block_len, uint32_t nblocks, double *sums,
                             hipStream_t s) {
  const bool vec = (block_len % 4u) == 0u && (reinterpret_cast<uintptr_t>(x) & 15u) == 0u;
  const dim3 grid((nblocks + 63u) / 64u);
  if (vec)
    hipLaunchKernelGGL(legacy_block_sums_kernel<true>, grid, dim3(64), 0, s, x, block_len, nblocks,
                       sums);
  else
    hipLaunchKernelGGL(legacy_block_sums_kernel<false>, grid, dim3(64), 0, s, x, block_len,
                       nblocks, sums);
  MDSM_HIP(hipGetLastError());
  return MDSM_OK;
}

int launch_legacy_threshold(const float *x, uint32_t ndms, uint32_t nsamp, float lowdm,
                            float dmstep, float thresh, float gmean, float grms, const float *cmean,
                            const float *crms, uint32_t spc, mdsm_cand_t *cands, uint32_t cap,
                            uint32_t *count, hipStream_t s) {
  hipLaunchKernelGGL(legacy_threshold_kernel, dim3((nsamp - 4u + NT - 1) / NT, ndms), dim3(NT), 0,
                     s, x, nsamp, lowdm, dmstep, thresh, gmean, grms, cmean, crms, spc, cands, cap,
                     count);
  MDSM_HIP(hipGetLastError());
  return MDSM_OK;
}

int launch_detrend(float *x, uint64_t nseg, uint32_t L, hipStream_t s) {
  hipLaunchKernelGGL(detrend_kernel, dim3((uint32_t)nseg), dim3(NT), 0, s, x, L);
  MDSM_HIP(hipGetLastError());
  return MDSM_OK;
}

int launch_generate(uint64_t seed, uint32_t nchans, int64_t t0, uint64_t n, const float *gain,
                    const float *offset, uint32_t npulses, const int64_t *p_start,
                    const int32_t *p_width, const float *p_amp, const int64_t *p_delay, float *raw,
                    uint64_t ld, hipStream_t s) {
  const dim3 grid(grid_for(n / 4 + 2, NT * 4, 4096), nchans);
  hipLaunchKernelGGL(generate_kernel, grid, dim3(NT), 0, s, seed, t0, n, gain, offset, npulses,
                     nchans, p_start, p_width, p_amp, p_delay, raw, ld);
  MDSM_HIP(hipGetLastError());
  return MDSM_OK;
}

}  // namespace mdsm

// ------------------------------------------- allocation-free global detector
// Workspace of mdsm_detect_global_noalloc: a fixed device array, so that the
// call needs no allocation (for callers under graph capture or in loops that
// must not allocate).  Every call on a device uses that device's copy and
// leaves its last contents there; calls that overlap on one device race.
// Room for SCRATCH_PARTS chunk partials (each 65536 samples of one row: at
// most nrows * ceil(stats_nsamp / 65536) = 131072, e.g. 4096 rows of 2^21
// samples), then mean and rms; 2 MiB per device.
namespace {
constexpr uint32_t SCRATCH_PARTS = 131072;
static __device__ __align__(256) double scratch[2 * SCRATCH_PARTS + 2];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
double *g_scratch_ptr[kMaxDevices] = {};

// The device address of this device's scratch, looked up once per device
// with the device current (the array's name in host code is only a host-side
// stand-in, never a device address).
int scratch_address(double **out) {
  int dev = 0;
  MDSM_HIP(hipGetDevice(&dev));
  MDSM_REQUIRE(dev >= 0 && dev < kMaxDevices, MDSM_ERR_ARG, "device %d out of range", dev);
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void *p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) {
      mdsm::set_error("hipGetSymbolAddress(scratch) failed: %s", hipGetErrorString(e));
      return MDSM_ERR_HIP;
    }
    g_scratch_ptr[dev] = static_cast<double *>(p);
  }
  *out = g_scratch_ptr[dev];
  return MDSM_OK;
}
}  // namespace

extern "C" int mdsm_detect_global_noalloc(const float *d_x, uint32_t nrows, uint32_t nsamp,
                                          uint64_t ld, uint32_t stats_nsamp, float thresh,
                                          mdsm_cand_t *d_cands, uint32_t cap, uint32_t *d_count,
                                          double *h_mean, double *h_rms, void *stream) {
  using namespace mdsm;
  int rc = check_global_args(d_x, nrows, nsamp, ld, stats_nsamp, d_cands, d_count);
  if (rc != MDSM_OK) return rc;
  const uint32_t parts = global_stats_parts(nrows, stats_nsamp);
  MDSM_REQUIRE(parts <= SCRATCH_PARTS, MDSM_ERR_ARG,
               "region of %u rows x %u samples needs %u partials; the static workspace holds %u "
               "(rows * ceil(stats_nsamp / 65536) <= %u)",
               nrows, stats_nsamp, parts, SCRATCH_PARTS, SCRATCH_PARTS);
  double *work = nullptr;
  rc = scratch_address(&work);
  if (rc != MDSM_OK) return rc;
  return run_global_detect(d_x, nrows, nsamp, ld, stats_nsamp, thresh, d_cands, cap, d_count,
                           h_mean, h_rms, work, as_stream(stream));
}

