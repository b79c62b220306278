// dedisperse.hip -- brute-force dedispersion kernels for gfx950.
//
// out[d][t] = sum_{c=0}^{C-1} in[c][t + D[c][d]]   (FP32, channel order)
//
// Design (MI355X): a workgroup of 4 wave64s owns a tile of TD = 4*MW DM
// trials x TT = 64*K samples.  Channels are processed in blocks of CB: the
// block's input windows [t0 + D[c][d0], t0 + TT + D[c][d_last]) are staged
// from HBM into LDS with coalesced loads, then every wave adds, for each of
// its MW trials, K samples per lane read from LDS at lane + 64*j + (D[c][d] -
// D[c][d0]).  The delay offsets are wave-uniform and come from the integer
// delay table through scalar loads, so one VALU address add serves K LDS
// reads.  Each staged sample is reused by TD trials (HBM/L2 re-reads divided
// by TD); the kernel is bound by LDS read bandwidth (one 4-byte LDS read per
// add), not by HBM.  No MFMA: the operation is shift-and-add, not a
// contraction.  Accumulation order is channel 0..C-1 for every output, which
// makes the result bit-identical to a sequential FP32 sum.
//
// Reference counterpart: shared_dedispersion
// (src/mdsm-multibeam/multibeam_dedispersion_kernel.cu:44-95) with its
// launcher shared_brute_force (multibeam_dedispersion_thread.cu:204-221),
// which stages one channel per two barriers for a 128 x 32 tile.
#include <hip/hip_fp16.h>

#include <algorithm>

#include "internal.h"

namespace mdsm {
namespace {

// Element loaders: input storage -> FP32 value that enters the sum.  U8 codes
// are accumulated as integers held exactly in FP32 (sum < 2^24 for any
// nchans <= 65793) and rescaled once at the end.
struct LoadF32 {
  using T = float;
  static __device__ __forceinline__ float cvt(T v) { return v; }
};
struct LoadF16 {
  using T = __half;
  static __device__ __forceinline__ float cvt(T v) { return __half2float(v); }
};
struct LoadU8 {
  using T = uint8_t;
  static __device__ __forceinline__ float cvt(T v) { return (float)v; }
};

template <typename L>
__device__ __forceinline__ float finish(float acc, uint32_t nchans);
template <>
__device__ __forceinline__ float finish<LoadF32>(float acc, uint32_t) { return acc; }
template <>
__device__ __forceinline__ float finish<LoadF16>(float acc, uint32_t) { return acc; }
template <>
__device__ __forceinline__ float finish<LoadU8>(float acc, uint32_t nchans) {
  return (float)((double)acc * (8.0 / 255.0) + (double)nchans * -4.0);
}

__device__ __forceinline__ uint32_t uniform(uint32_t v) {
  return __builtin_amdgcn_readfirstlane(v);
}

// One tile.  lds: CB rows of row_words floats.
template <typename L, int MW, int K, int CB>
__device__ __forceinline__ void dedisp_tile(const typename L::T *__restrict__ in, uint64_t ld,
                                            const int32_t *__restrict__ table, uint32_t nchans,
                                            uint32_t nsamp, uint32_t ndms, uint32_t row_words,
                                            uint32_t cbr, float *__restrict__ out, uint32_t tt,
                                            uint32_t dt, float *lds) {
  constexpr int TT = 64 * K;
  constexpr int TD = 4 * MW;
  const uint32_t tid = threadIdx.x;
  const uint32_t lane = tid & 63u;
  const uint32_t wave = uniform(tid >> 6);
  const uint32_t t0 = tt * TT;
  const uint32_t d0 = dt * TD;
  const uint32_t dlast = min(d0 + TD, ndms) - 1u;
  const uint32_t dw = d0 + wave * MW;  // this wave's first trial (may exceed ndms)

  float acc[MW][K];
#pragma unroll
  for (int i = 0; i < MW; ++i)
#pragma unroll
    for (int j = 0; j < K; ++j) acc[i][j] = 0.0f;

  // cbr (<= CB) channels per stage: CB unless the delay spread limits LDS
  for (uint32_t cb = 0; cb < nchans; cb += cbr) {
    const uint32_t nc = min(cbr, nchans - cb);
    // ---- stage the block's windows into LDS (coalesced, zero beyond the row)
    for (uint32_t r = 0; r < nc; ++r) {
      const uint32_t c = cb + r;
      const int32_t *trow = table + (uint64_t)c * ndms;
      const uint32_t base = (uint32_t)trow[d0];
      const uint32_t len = TT + (uint32_t)trow[dlast] - base;
      const uint64_t start = (uint64_t)t0 + base;
      const typename L::T *src = in + (uint64_t)c * ld;
      float *dst = lds + r * row_words;
      for (uint32_t i = tid; i < len; i += 256u) {
        const uint64_t g = start + i;
        dst[i] = g < ld ? L::cvt(src[g]) : 0.0f;
      }
    }
    __syncthreads();
    // ---- accumulate: channel order is preserved for every output
    for (uint32_t r = 0; r < nc; ++r) {
      const int32_t *trow = table + (uint64_t)(cb + r) * ndms;
      const int32_t base = trow[d0];
      const float *row = lds + r * row_words + lane;
#pragma unroll
      for (int i = 0; i < MW; ++i) {
        const uint32_t d = min(dw + (uint32_t)i, ndms - 1u);
        const int32_t rel = trow[d] - base;
        const float *p = row + rel;
#pragma unroll
        for (int j = 0; j < K; ++j) acc[i][j] += p[64 * j];
      }
    }
    __syncthreads();
  }

#pragma unroll
  for (int i = 0; i < MW; ++i) {
    const uint32_t d = dw + (uint32_t)i;
    if (d < ndms) {
      float *o = out + (uint64_t)d * nsamp;
#pragma unroll
      for (int j = 0; j < K; ++j) {
        const uint32_t t = t0 + lane + 64u * j;
        if (t < nsamp) o[t] = finish<L>(acc[i][j], nchans);
      }
    }
  }
}

template <typename L, int MW, int K, int CB>
__global__ __launch_bounds__(256) void dedisp_kernel(const typename L::T *__restrict__ in,
                                                     uint64_t ld,
                                                     const int32_t *__restrict__ table,
                                                     uint32_t nchans, uint32_t nsamp,
                                                     uint32_t ndms, uint32_t row_words,
                                                     uint32_t cbr, uint32_t n_dtiles,
                                                     float *__restrict__ out) {
  extern __shared__ float lds[];
  // DM tiles vary fastest so that the workgroups resident at one time share
  // a narrow time range of the input (reuse in L2 and the 256 MB MALL).
  const uint32_t b = blockIdx.x;
  dedisp_tile<L, MW, K, CB>(in, ld, table, nchans, nsamp, ndms, row_words, cbr, out, b / n_dtiles,
                            b % n_dtiles, lds);
}

// Persistent multi-beam kernel: workgroups pull items from a queue until it
// is exhausted.  Items are ordered by the host (largest beams first); item i
// belongs to the beam with the largest item_begin <= i.
template <typename L, int MW, int K, int CB>
__global__ __launch_bounds__(256) void batch_kernel(const BeamDesc *__restrict__ beams,
                                                    uint32_t nbeams, uint32_t total_items,
                                                    uint32_t cbr,
                                                    uint32_t *__restrict__ counter) {
  extern __shared__ float lds[];
  __shared__ uint32_t item_s;
  while (true) {
    if (threadIdx.x == 0) item_s = atomicAdd(counter, 1u);
    __syncthreads();
    const uint32_t item = item_s;
    __syncthreads();
    if (item >= total_items) break;  // every wave of every workgroup reaches this
    uint32_t lo = 0, hi = nbeams - 1u;
    while (lo < hi) {
      const uint32_t mid = (lo + hi + 1u) >> 1;
      if (beams[mid].item_begin <= item) lo = mid;
      else hi = mid - 1u;
    }
    const BeamDesc &bd = beams[lo];
    const uint32_t local = item - bd.item_begin;
    dedisp_tile<L, MW, K, CB>(reinterpret_cast<const typename L::T *>(bd.in), bd.ld, bd.table,
                              bd.nchans, bd.nsamp, bd.ndms, bd.row_words, cbr, bd.out,
                              local / bd.n_dtiles, local % bd.n_dtiles, lds);
  }
}

// ------------------------------------------------------------ variant table
// Channels staged per barrier: the variant's CB, reduced when CB rows of the
// delay-spread window would not fit in 64 KB of LDS (at most 160 KB per CU).
uint32_t stage_channels(int cb, uint32_t row_words) {
  const uint64_t budget = 64 * 1024 / sizeof(float);
  uint64_t n = (uint64_t)cb;
  while (n > 1 && n * row_words > budget) --n;
  if (n * row_words * sizeof(float) > 160 * 1024) return 0;
  return (uint32_t)n;
}

// Variant 1..N (0 = use the plan's tuned choice).
const TileShape kShapes[] = {
    {8, 4, 8},   // 1: TD 32,  TT 256
    {16, 4, 8},  // 2: TD 64,  TT 256
    {8, 8, 8},   // 3: TD 32,  TT 512
    {16, 8, 4},  // 4: TD 64,  TT 512
    {32, 2, 8},  // 5: TD 128, TT 128
};
constexpr int kNumShapes = sizeof(kShapes) / sizeof(kShapes[0]);

template <typename L>
int launch_one(int variant, const void *in, uint64_t ld, const int32_t *table, uint32_t nchans,
               uint32_t nsamp, uint32_t ndms, uint32_t row_words, float *out, hipStream_t s) {
  const TileShape sh = kShapes[variant - 1];
  const uint32_t n_ttiles = (nsamp + sh.tt() - 1) / sh.tt();
  const uint32_t n_dtiles = (ndms + sh.td() - 1) / sh.td();
  const uint64_t blocks = (uint64_t)n_ttiles * n_dtiles;
  MDSM_REQUIRE(blocks < (1ull << 31), MDSM_ERR_ARG, "dedispersion grid too large");
  const uint32_t cbr = stage_channels(sh.cb, row_words);
  MDSM_REQUIRE(cbr >= 1, MDSM_ERR_ARG,
               "delay spread too large for LDS staging (%u words per channel)", row_words);
  const size_t lds = (size_t)cbr * row_words * sizeof(float);
  const auto *src = reinterpret_cast<const typename L::T *>(in);
#define MDSM_CASE(V, MW, K, CB)                                                                \
  case V:                                                                                      \
    hipLaunchKernelGGL((dedisp_kernel<L, MW, K, CB>), dim3((uint32_t)blocks), dim3(256), lds, s, \
                       src, ld, table, nchans, nsamp, ndms, row_words, cbr, n_dtiles, out);   \
    break;
  switch (variant) {
    MDSM_CASE(1, 8, 4, 8)
    MDSM_CASE(2, 16, 4, 8)
    MDSM_CASE(3, 8, 8, 8)
    MDSM_CASE(4, 16, 8, 4)
    MDSM_CASE(5, 32, 2, 8)
    default:
      set_error("unknown dedispersion variant %d", variant);
      return MDSM_ERR_ARG;
  }
#undef MDSM_CASE
  MDSM_HIP(hipGetLastError());
  return MDSM_OK;
}

template <typename L, int MW, int K, int CB>
int launch_batch_shape(const BeamDesc *beams, uint32_t nbeams, uint32_t total, uint32_t *counter,
                       uint32_t cbr, size_t lds, int grid, hipStream_t s) {
  if (grid <= 0) {
    // persistent: as many workgroups as can be resident at once
    int per_cu = 0, dev = 0, cus = 0;
    MDSM_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, batch_kernel<L, MW, K, CB>,
                                                          256, lds));
    MDSM_HIP(hipGetDevice(&dev));
    MDSM_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
    grid = std::max(1, per_cu) * std::max(1, cus);
  }
  hipLaunchKernelGGL((batch_kernel<L, MW, K, CB>), dim3(grid), dim3(256), lds, s, beams, nbeams,
                     total, cbr, counter);
  MDSM_HIP(hipGetLastError());
  return MDSM_OK;
}

template <typename L>
int launch_batch_one(int variant, const BeamDesc *beams, uint32_t nbeams, uint32_t total,
                     uint32_t *counter, uint32_t cbr, size_t lds, int grid, hipStream_t s) {
#define MDSM_CASE(V, MW, K, CB)                                                                \
  case V:                                                                                      \
    return launch_batch_shape<L, MW, K, CB>(beams, nbeams, total, counter, cbr, lds, grid, s);
  switch (variant) {
    MDSM_CASE(1, 8, 4, 8)
    MDSM_CASE(2, 16, 4, 8)
    MDSM_CASE(3, 8, 8, 8)
    MDSM_CASE(4, 16, 8, 4)
    MDSM_CASE(5, 32, 2, 8)
    default:
      set_error("unknown dedispersion variant %d", variant);
      return MDSM_ERR_ARG;
  }
#undef MDSM_CASE
}

}  // namespace

int num_variants(int) { return kNumShapes; }

TileShape variant_shape(int, int variant) { return kShapes[variant - 1]; }

uint32_t lds_row_words(int, const TileShape &sh, uint32_t spread) {
  const uint32_t w = (uint32_t)sh.tt() + spread;
  return (w + 3u) & ~3u;
}

int launch_dedisperse(int dtype, int variant, const void *in, uint64_t ld, const int32_t *table,
                      uint32_t nchans, uint32_t nsamp, uint32_t ndms, uint32_t row_words,
                      float *out, hipStream_t s) {
  MDSM_REQUIRE(variant >= 1 && variant <= kNumShapes, MDSM_ERR_ARG, "bad variant %d", variant);
  switch (dtype) {
    case MDSM_F32:
      return launch_one<LoadF32>(variant, in, ld, table, nchans, nsamp, ndms, row_words, out, s);
    case MDSM_F16:
      return launch_one<LoadF16>(variant, in, ld, table, nchans, nsamp, ndms, row_words, out, s);
    case MDSM_U8:
      return launch_one<LoadU8>(variant, in, ld, table, nchans, nsamp, ndms, row_words, out, s);
  }
  set_error("unknown input dtype %d", dtype);
  return MDSM_ERR_ARG;
}

size_t batch_lds_bytes(int, int variant, uint32_t max_row_words) {
  return (size_t)stage_channels(kShapes[variant - 1].cb, max_row_words) * max_row_words *
         sizeof(float);
}

int launch_batch_dedisperse(int dtype, int variant, const BeamDesc *d_beams, uint32_t nbeams,
                            const uint32_t *, uint32_t total_items, uint32_t *d_counter,
                            uint32_t max_row_words, int grid, hipStream_t s) {
  MDSM_REQUIRE(variant >= 1 && variant <= kNumShapes, MDSM_ERR_ARG, "bad variant %d", variant);
  const uint32_t cbr = stage_channels(kShapes[variant - 1].cb, max_row_words);
  MDSM_REQUIRE(cbr >= 1, MDSM_ERR_ARG, "delay spread too large for LDS staging");
  const size_t lds = (size_t)cbr * max_row_words * sizeof(float);
  switch (dtype) {
    case MDSM_F32:
      return launch_batch_one<LoadF32>(variant, d_beams, nbeams, total_items, d_counter, cbr, lds, grid, s);
    case MDSM_F16:
      return launch_batch_one<LoadF16>(variant, d_beams, nbeams, total_items, d_counter, cbr, lds, grid, s);
    case MDSM_U8:
      return launch_batch_one<LoadU8>(variant, d_beams, nbeams, total_items, d_counter, cbr, lds, grid, s);
  }
  set_error("unknown input dtype %d", dtype);
  return MDSM_ERR_ARG;
}

}  // namespace mdsm
