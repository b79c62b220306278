// dedisperse.hip -- brute-force dedispersion kernels for gfx950.
//
// out[d][t] = sum_{c=0}^{C-1} in[c][t + D[c][d]]   (FP32 accumulation, channel order)
//
// Design (MI355X).  A workgroup of 4 wave64s owns a tile of TD = 4*MW DM
// trials x TT time samples; every wave owns MW consecutive trials and every
// lane a few samples spaced 64 (FP32) or 128 (packed pairs) apart.  Channels
// are processed in blocks: each block's input windows [t0 + D[c][d0],
// t0 + TT + D[c][d_last]) are staged into one of two LDS buffers while the
// waves add the previous block, so HBM/L2 latency overlaps the adds.  Each
// staged sample is reused by TD trials.  The per-trial offsets inside a
// window are wave-uniform and independent of the time tile: they are
// tabulated once per plan, a block's offsets sit 16 per VGPR, and one trial
// costs one DPP-broadcast address add for K LDS reads (issued in pairs as
// ds_read2st64).  The adds are bound by
// LDS read bandwidth (~128 B/clk/CU), not by HBM, so the storage formats
// trade LDS bytes per add:
//   F32  4 B per add: one float per LDS word, staged by direct buffer->LDS
//        loads (buffer_load_dword ... lds) that use no data VGPRs.
//   F16  2 B per add: LDS words hold pairs of consecutive samples in two
//        copies (even and odd alignment), so any wave-uniform offset reads
//        an aligned word; v_fma_mix_f32 (half operand selected by op_sel,
//        times 1.0, plus the FP32 accumulator) adds one half with one IEEE
//        round-to-nearest-even.  (v_dot2_f32_f16 would add two halves per
//        instruction but is not correctly rounded: it drops low bits of
//        the exact sum in cancellation and wide-exponent cases.)
//   F16_PAIRS  the F16 layout; for each trial the words of channels 2k and
//        2k+1 are added with one v_pk_add_f16 (two samples, each one IEEE
//        FP16 rounding) and the pair sum goes into the FP32 accumulators
//        with v_fma_mix_f32: three VALU instructions per four adds instead
//        of four, at about 1.7x the rounding error of F16 against FP64.
//        The pair-sum tiles build those sums once per delay difference of
//        the two channels instead and read one word per trial (one LDS word
//        and two v_fma_mix_f32 per four adds), where the survey allows.
//   U8   2 B per add: bytes widened to pairs of u16 in the same two-copy
//        layout, summed exactly with v_pk_add_u16 (two adds per
//        instruction); every <= 256 channels their high bytes move into a
//        second u16 pair accumulator (sum = 256 * high + low, exact).
// Results are bit-identical to a sequential FP32 sum in channel order (F32,
// F16 storage), to the FP32 sum of FP16 channel-pair sums (F16_PAIRS) or to
// the exact integer sum (U8).  MFMA is not used: the
// operation is shift-and-add, not a contraction.
//
// Reference counterpart: shared_dedispersion
// (src/mdsm-multibeam/multibeam_dedispersion_kernel.cu:44-95) and its
// launcher shared_brute_force (multibeam_dedispersion_thread.cu:204-221),
// which stages one channel per two barriers for a 128 x 32 tile.
#include <hip/hip_fp16.h>

#include <algorithm>
#include <type_traits>

#include "internal.h"

namespace mdsm {
namespace {

typedef __attribute__((address_space(3))) void lds_void;
typedef _Float16 half2_t __attribute__((ext_vector_type(2)));
typedef unsigned short ushort2_t __attribute__((ext_vector_type(2)));

__device__ __forceinline__ uint32_t uniform(uint32_t v) { return __builtin_amdgcn_readfirstlane(v); }

// s_waitcnt vmcnt(0) with expcnt and lgkmcnt left at their maxima (GFX9 encoding).
__device__ __forceinline__ void wait_vmem() { __builtin_amdgcn_s_waitcnt(0x0F70); }

// One 64-dword chunk from global memory straight into LDS: lane l copies the
// dword at src + 4 l to lds_dst + 4 l (rows are shorter than 2 GiB).  The copy goes through a raw buffer
// resource whose base is src and whose size is `avail` bytes, so lanes at
// or past `avail` read 0 instead of touching memory (the hardware range
// check replaces a per-lane compare), and the only per-lane operand is the
// constant 4*lane (no 64-bit address arithmetic in VGPRs).  Dword 3 of the
// resource (0x00020000: 32-bit data format, raw addressing) is the GFX9
// family value.
__device__ __forceinline__ void chunk_to_lds(const void *src, uint32_t avail, void *lds_dst,
                                             uint32_t lane4) {
  // the operands are wave-uniform; readfirstlane keeps the resource in SGPRs
  const uint64_t a = reinterpret_cast<uint64_t>(src);
  const uint64_t au = ((uint64_t)uniform((uint32_t)(a >> 32)) << 32) | (uint64_t)uniform((uint32_t)a);
  const __amdgpu_buffer_rsrc_t rs = __builtin_amdgcn_make_buffer_rsrc(
      reinterpret_cast<void *>(au), (short)0, (int)uniform(avail), 0x00020000);
  __builtin_amdgcn_raw_ptr_buffer_load_lds(rs, (lds_void *)lds_dst, 4, (int)lane4, 0, 0, 0);
}

// LDS byte offsets of the wave's trials inside the windows of one channel
// block, relative to the block buffer plus the lane's own word.  Lane l of
// register n holds the offset of (r, i) = divmod(16 n + l % 16, MW) (the 16
// values repeat in each row of 16 lanes), i.e. channel cb+r and trial dw+i
// (both clamped into range), from rel = D[c][d] - D[c][d0]:
//   PACK 0 (one sample per word):  word = r*rw + rel
//   PACK 1 (pairs, window starts at element D[c][d0] of an even word
//           origin): q = rel + (D[c][d0] & 1), word = r*rw + q/2 + (q&1)*rw/2
//   PACK 2 (pairs, origin on a 4-sample boundary): as 1 with D[c][d0] & 3
// (copy P is the first half of a row, copy Q the second).  Computed once per
// block; a trial's address is then one v_add_u32 whose offset operand is
// broadcast from lane i of each row by DPP row_newbcast (no v_readlane).
template <int MW, int PACK>
__device__ __forceinline__ int32_t rel_entry(const int32_t *__restrict__ table, uint32_t ndms,
                                             uint32_t nchans, uint32_t cb, uint32_t d0,
                                             uint32_t dw, uint32_t idx, uint32_t rw) {
  const uint32_t r = idx / MW;
  const uint32_t c = min(cb + r, nchans - 1u);
  const uint32_t d = min(dw + idx % MW, ndms - 1u);
  const int32_t *row = table + (uint64_t)c * ndms;
  const int32_t base = row[d0];
  const uint32_t rel = (uint32_t)(row[d] - base);
  uint32_t w;
  if (PACK == 0) {
    w = r * rw + rel;
  } else {
    const uint32_t q = rel + ((uint32_t)base & (PACK == 1 ? 1u : 3u));
    w = r * rw + (q >> 1) + (q & 1u) * (rw >> 1);
  }
  return (int32_t)(4u * w);
}

// The offsets do not depend on the time tile, so they are tabulated once per
// plan and tile shape (trial_offset_kernel) as [DM tile][channel block][wave]
// [register][16] and each block's registers are one load per register from a
// wave-uniform base plus the lane's constant (lane % 16) word.
template <int MW, int CB, int PACK>
struct RelOffsets {
  static constexpr int N = (CB * MW + 15) / 16;
  static constexpr int WORDS = 16 * N;  // table words per (DM tile, block, wave)
  int32_t v[N];
  __device__ __forceinline__ void load(const int32_t *__restrict__ p, uint32_t lane) {
#pragma unroll
    for (int n = 0; n < N; ++n) v[n] = p[16u * n + (lane & 15u)];
  }
  // offset of (r, i) in every lane: DPP row_newbcast of lane idx % 16 (the
  // callers' loops are unrolled, so the switch folds to one case)
  __device__ __forceinline__ int32_t get(int r, int i) const {
    const int idx = r * MW + i;
    const int32_t x = v[idx >> 4];
#define MDSM_BCAST(j) \
  case j: return __builtin_amdgcn_update_dpp(0, x, 0x150 + j, 0xF, 0xF, false);
    switch (idx & 15) {
      MDSM_BCAST(0) MDSM_BCAST(1) MDSM_BCAST(2) MDSM_BCAST(3) MDSM_BCAST(4) MDSM_BCAST(5)
      MDSM_BCAST(6) MDSM_BCAST(7) MDSM_BCAST(8) MDSM_BCAST(9) MDSM_BCAST(10) MDSM_BCAST(11)
      MDSM_BCAST(12) MDSM_BCAST(13) MDSM_BCAST(14) default: return __builtin_amdgcn_update_dpp(0, x, 0x15F, 0xF, 0xF, false);
    }
#undef MDSM_BCAST
  }
};

template <typename T>
__device__ __forceinline__ const T *at_bytes(const T *p, int32_t bytes) {
  return reinterpret_cast<const T *>(reinterpret_cast<const char *>(p) + bytes);
}

// ======================================================================= F32
// BURST > 0: the LDS reads of BURST trials are issued together ahead of
// their adds (a scheduling fence keeps them grouped), so more reads are in
// flight per wave at the cost of BURST*K registers; 0 leaves the order to
// the compiler.
template <int MW, int K, int CB_, int BURST = 0>
struct F32Tile {
  static constexpr bool PAIRS = false;
  static constexpr int TT = 64 * K;
  static constexpr int TD = 4 * MW;
  static constexpr int CB = CB_;
  static constexpr int ALIGN = 1;  // ld multiple, in elements
  using T = float;

  static uint32_t row_words(uint32_t spread) { return ((uint32_t)TT + spread + 63u) & ~63u; }
  static uint32_t lds_words_per_row(uint32_t rw) { return 2u * rw; }
  static bool fits(uint32_t) { return true; }

  static constexpr int PACK = 0;
  using Rel = RelOffsets<MW, CB_, PACK>;
  __device__ __forceinline__ static void run(const float *__restrict__ in, uint64_t ld,
                             const int32_t *__restrict__ table, const int32_t *__restrict__ reltab,
                             uint32_t nchans, uint32_t nsamp, uint32_t ndms, uint32_t rw,
                             uint32_t cbr, float *__restrict__ out, uint32_t tt, uint32_t dt,
                             float *lds) {
    const uint32_t lane = threadIdx.x & 63u;
    const uint32_t wave = uniform(threadIdx.x >> 6);
    const uint32_t t0 = tt * TT, d0 = dt * TD;
    const uint32_t dlast = min(d0 + TD, ndms) - 1u;
    const uint32_t dw = d0 + wave * MW;
    const uint32_t bufw = cbr * rw;  // words per LDS buffer (offsets keep the LDS address space)

    // wave w copies 64-word chunks w, w+4, ... of each row from global
    // memory straight into LDS; words past the end of the row read as 0
    // (they only feed outputs beyond nsamp)
    const uint32_t lane4 = 4u * lane;
    auto stage = [&](uint32_t cb, float *b) {
      const uint32_t nc = min(cbr, nchans - cb);
      for (uint32_t r = 0; r < nc; ++r) {
        const int32_t *trow = table + (uint64_t)(cb + r) * ndms;
        const int32_t base = trow[d0];
        const uint32_t len = (uint32_t)TT + (uint32_t)(trow[dlast] - base);
        const uint64_t g0 = (uint64_t)t0 + (uint32_t)base;
        const float *src = in + (uint64_t)(cb + r) * ld + g0;
        // bytes to the row end (rows are shorter than 2 GiB, so signed 32-bit
        // offsets hold them and the clamps stay scalar)
        const int32_t ld32 = (int32_t)ld;
        const int32_t avail = 4 * (ld32 - min((int32_t)g0, ld32));
        const uint32_t nq = (len + 63u) >> 6;
        for (uint32_t q = wave; q < nq; q += 4u) {
          const int32_t rest = max(avail - 256 * (int32_t)q, 0);
          chunk_to_lds(src + 64u * q, (uint32_t)rest, b + r * rw + 64u * q, lane4);
        }
      }
    };

    float acc[MW][K];
#pragma unroll
    for (int i = 0; i < MW; ++i)
#pragma unroll
      for (int j = 0; j < K; ++j) acc[i][j] = 0.0f;

    Rel rel, rel_next;
    const uint32_t nblocks = (nchans + cbr - 1u) / cbr;
    const int32_t *rt = reltab + ((uint64_t)dt * nblocks * 4u + wave) * Rel::WORDS;
    stage(0, lds);
    rel.load(rt, lane);
    wait_vmem();
    __syncthreads();
    uint32_t k = 0;
    for (uint32_t cb = 0; cb < nchans; cb += cbr, ++k) {
      const float *cur = lds + (k & 1u) * bufw + lane;
      const bool more = cb + cbr < nchans;
      if (more) {
        stage(cb + cbr, lds + ((k + 1u) & 1u) * bufw);
        rel_next.load(rt + (uint64_t)(k + 1u) * 4u * Rel::WORDS, lane);
      }
      const uint32_t nc = min(cbr, nchans - cb);
#pragma unroll
      for (int r = 0; r < CB; ++r) {
        if ((uint32_t)r < nc) {
          if constexpr (BURST > 0) {
#pragma unroll
            for (int i0 = 0; i0 < MW; i0 += BURST) {
              float w[BURST][K];
#pragma unroll
              for (int g = 0; g < BURST; ++g) {
                const float *p = at_bytes(cur, rel.get(r, i0 + g));
#pragma unroll
                for (int j = 0; j < K; ++j) w[g][j] = p[64 * j];
              }
              __builtin_amdgcn_sched_barrier(0);
#pragma unroll
              for (int g = 0; g < BURST; ++g)
#pragma unroll
                for (int j = 0; j < K; ++j) acc[i0 + g][j] += w[g][j];
            }
          } else {
#pragma unroll
            for (int i = 0; i < MW; ++i) {
              const float *p = at_bytes(cur, rel.get(r, i));
#pragma unroll
              for (int j = 0; j < K; ++j) acc[i][j] += p[64 * j];
            }
          }
        }
      }
      wait_vmem();
      __syncthreads();
      if (more) rel = rel_next;
    }

#pragma unroll
    for (int i = 0; i < MW; ++i) {
      const uint32_t d = dw + (uint32_t)i;
      if (d < ndms) {
        float *o = out + (uint64_t)d * nsamp;
#pragma unroll
        for (int j = 0; j < K; ++j) {
          const uint32_t t = t0 + lane + 64u * j;
          if (t < nsamp) o[t] = acc[i][j];
        }
      }
    }
  }
};

// ============================================================ packed (F16, U8)
// LDS row of a channel: copy P (words 0..wp-1) then copy Q (words wp..2wp-1).
// With f the even word origin of the row window (elements 2f, 2f+1, ... of
// the global row), P[w] = (e[2(f+w)], e[2(f+w)+1]) and Q[w] = (e[2(f+w)+1],
// e[2(f+w)+2]) as two 16-bit halves.  For the window start a = g0 + rel of a
// trial (wave-uniform), the lane's pair (a + 2 lane + 128 j, +1) is word
// (a>>1) - f + lane + 64 j of copy (a & 1).
template <int MW, int K2, int CB_>
struct PackedGeom {
  static constexpr int TT = 128 * K2;
  static constexpr int TD = 4 * MW;
  static constexpr int CB = CB_;
  static uint32_t copy_words(uint32_t spread) {
    return ((uint32_t)(64 * K2) + (spread + 4u) / 2u + 2u + 63u) & ~63u;
  }
  static uint32_t row_words(uint32_t spread) { return 2u * copy_words(spread); }
};

// Packed tiles (F16 and U8 storage, rows 4-byte aligned).  Per channel block
// the raw dwords of each window are copied into LDS by direct global->LDS
// loads while the previous block is added; after the barrier one LDS->LDS
// pass builds the pair copies (F16: P is the copy itself and Q its 16-bit
// funnel shift; U8: bytes widened to u16 pairs with v_perm into P and Q).
template <int MODE, int MW, int K2, int CB_, int BURST = 0>
struct PackedTile : PackedGeom<MW, K2, CB_> {
  using G = PackedGeom<MW, K2, CB_>;
  static constexpr int TT = G::TT, TD = G::TD, CB = CB_;
  static constexpr bool IS_U8 = MODE == MDSM_U8;
  // F16_PAIRS: channel blocks hold an even number of rows (stage_channels)
  static constexpr bool PAIRS = MODE == MDSM_F16_PAIRS;
  static_assert(!PAIRS || CB_ % 2 == 0, "F16_PAIRS tiles stage whole channel pairs");
  using T = std::conditional_t<IS_U8, uint8_t, __half>;
  static bool fits(uint32_t) { return true; }
  // raw staging dwords per row (U8): wp/2 + 2 source dwords, padded to 64
  static __host__ __device__ uint32_t raw_words(uint32_t rw) {
    return IS_U8 ? ((rw / 4u + 2u + 63u) & ~63u) : 0u;
  }
  static uint32_t row_words(uint32_t spread) { return G::row_words(spread); }
  static uint32_t lds_words_per_row(uint32_t rw) { return 2u * rw + raw_words(rw); }

  static constexpr int PACK = IS_U8 ? 2 : 1;
  using Rel = RelOffsets<MW, CB_, PACK>;
  __device__ __forceinline__ static void run(const T *__restrict__ in, uint64_t ld,
                                             const int32_t *__restrict__ table,
                                             const int32_t *__restrict__ reltab, uint32_t nchans,
                                             uint32_t nsamp, uint32_t ndms, uint32_t rw,
                                             uint32_t cbr, float *__restrict__ out, uint32_t tt,
                                             uint32_t dt, float *ldsf) {
    uint32_t *lds = reinterpret_cast<uint32_t *>(ldsf);
    const uint32_t tid = threadIdx.x;
    const uint32_t lane = tid & 63u;
    const uint32_t wave = uniform(tid >> 6);
    const uint32_t t0 = tt * TT, d0 = dt * TD;
    const uint32_t dw = d0 + wave * MW;
    const uint32_t wp = rw / 2u;
    const uint32_t rawn = raw_words(rw);
    const uint64_t ldw = IS_U8 ? ld / 4u : ld / 2u;  // source dwords per row
    const uint32_t *src32 = reinterpret_cast<const uint32_t *>(in);
    const uint32_t bufw = cbr * rw;  // words per LDS buffer (offsets keep the LDS address space)
    uint32_t *raw = lds + 2u * cbr * rw;

    auto g0_of = [&](uint32_t c) { return t0 + (uint32_t)table[(uint64_t)c * ndms + d0]; };
    // copy the window dwords of rows [cb, cb+nc) into P of buffer b (F16) or
    // into the raw area (U8); wave w takes 64-dword chunks w, w+4, ...;
    // dwords past the end of the row read as 0
    const uint32_t lane4 = 4u * lane;
    // the window starts of the rows the next stage() takes, loaded one block
    // ahead so that their scalar loads do not delay the copies
    uint32_t g0pre[CB_];
    auto prefetch = [&](uint32_t cb) {
#pragma unroll
      for (int r = 0; r < CB_; ++r)
        if ((uint32_t)r < cbr && cb + (uint32_t)r < nchans) g0pre[r] = g0_of(cb + (uint32_t)r);
    };
    auto stage = [&](uint32_t cb, uint32_t *b) {
      const uint32_t nc = min(cbr, nchans - cb);
#pragma unroll
      for (int rr = 0; rr < CB_; ++rr) {
        if ((uint32_t)rr >= nc) break;
        const uint32_t r = (uint32_t)rr;
        const uint32_t g0 = g0pre[rr];
        const uint64_t w0 = IS_U8 ? (g0 >> 2) : (g0 >> 1);
        const uint32_t *src = src32 + (uint64_t)(cb + r) * ldw + w0;
        const int32_t ldw32 = (int32_t)ldw;  // rows < 2 GiB
        const int32_t avail = 4 * (ldw32 - min((int32_t)w0, ldw32));
        uint32_t *dst = IS_U8 ? raw + r * rawn : b + r * rw;
        const uint32_t n = IS_U8 ? rawn : wp;
        for (uint32_t q = wave; 64u * q < n; q += 4u) {
          const int32_t rest = max(avail - 256 * (int32_t)q, 0);
          chunk_to_lds(src + 64u * q, (uint32_t)rest, dst + 64u * q, lane4);
        }
      }
    };
    // build the pair copies of rows [cb, cb+nc) in buffer b
    // (F16: wave w builds rows w, w+4, ..., four Q words per lane from one
    // 16-byte and one 4-byte LDS read; Q[wp-1], which no trial reads, may
    // take the stale word past P.  U8: also one wave per row.)
    auto build = [&](uint32_t cb, uint32_t *b) {
      const uint32_t nc = min(cbr, nchans - cb);
      if (IS_U8) {
        // wave w widens rows w, w+4, ...: raw dwords n, n+1 (and n+2 for
        // the last Q word) become P words 2n..2n+3 and Q words 2n..2n+3
        // (two 16-byte writes)
        for (uint32_t r = wave; r < nc; r += 4u) {
          uint32_t *P = b + r * rw;
          uint32_t *Q = P + wp;
          const uint32_t *R = raw + r * rawn;
          // two raw dwords per lane (wp is a multiple of 64, so the four
          // P and Q words of a lane are whole 16-byte groups)
          for (uint32_t n = 2u * lane; 2u * n < wp; n += 128u) {
            const uint2 xy = *reinterpret_cast<const uint2 *>(R + n);
            const uint32_t x = xy.x, y = xy.y, z = R[n + 2];
            *reinterpret_cast<uint4 *>(P + 2 * n) =
                make_uint4(__builtin_amdgcn_perm(y, x, 0x0C010C00u), __builtin_amdgcn_perm(y, x, 0x0C030C02u),
                           __builtin_amdgcn_perm(z, y, 0x0C010C00u), __builtin_amdgcn_perm(z, y, 0x0C030C02u));
            *reinterpret_cast<uint4 *>(Q + 2 * n) =
                make_uint4(__builtin_amdgcn_perm(y, x, 0x0C020C01u), __builtin_amdgcn_perm(y, x, 0x0C040C03u),
                           __builtin_amdgcn_perm(z, y, 0x0C020C01u), __builtin_amdgcn_perm(z, y, 0x0C040C03u));
          }
        }
      } else {
        for (uint32_t r = wave; r < nc; r += 4u) {
          uint32_t *P = b + r * rw;
          uint32_t *Q = P + wp;
          // four Q words per lane (wp is a multiple of 64, rows 16-byte aligned)
          for (uint32_t w = 4u * lane; w + 3u < wp; w += 256u) {
            const uint4 a = *reinterpret_cast<const uint4 *>(P + w);
            const uint32_t c = P[w + 4];
            *reinterpret_cast<uint4 *>(Q + w) = make_uint4(__builtin_amdgcn_alignbit(a.y, a.x, 16),
                                                           __builtin_amdgcn_alignbit(a.z, a.y, 16),
                                                           __builtin_amdgcn_alignbit(a.w, a.z, 16),
                                                           __builtin_amdgcn_alignbit(c, a.w, 16));
          }
        }
        // F16_PAIRS, odd block (the last of an odd channel count): row nc,
        // inside the buffer since cbr is even, becomes the zero partner
        if (PAIRS && (nc & 1u)) {
          uint32_t *Z = b + nc * rw;
          for (uint32_t w = 4u * tid; w < rw; w += 1024u)
            *reinterpret_cast<uint4 *>(Z + w) = make_uint4(0u, 0u, 0u, 0u);
        }
      }
    };

    // F16: FP32 accumulators.  U8: two u16 pair accumulators per output
    // pair, the running sum S = 256*uhi + uacc held exactly: a fold moves
    // uacc's high bytes into uhi (three packed instructions per pair) and
    // leaves uacc <= 255, so 256 more channels of bytes fit in 16 bits, and
    // uhi <= S/256 < nchans fits while nchans <= 65536.
    float facc[IS_U8 ? 1 : MW][IS_U8 ? 1 : K2][2];
    ushort2_t uacc[IS_U8 ? MW : 1][IS_U8 ? K2 : 1];
    ushort2_t uhi[IS_U8 ? MW : 1][IS_U8 ? K2 : 1];
#pragma unroll
    for (int i = 0; i < MW; ++i)
#pragma unroll
      for (int j = 0; j < K2; ++j) {
        if (IS_U8) {
          uacc[i][j] = ushort2_t{0, 0};
          uhi[i][j] = ushort2_t{0, 0};
        } else {
          facc[i][j][0] = facc[i][j][1] = 0.0f;
        }
      }
    auto fold = [&]() {
      if (IS_U8) {
#pragma unroll
        for (int i = 0; i < MW; ++i)
#pragma unroll
          for (int j = 0; j < K2; ++j) {
            uhi[i][j] += uacc[i][j] >> (ushort)8;
            uacc[i][j] &= (ushort)0xFF;
          }
      }
    };
    // 1.0 from an opaque scalar move: fma(fpext(h), one, acc) then selects
    // v_fma_mix_f32 instead of being folded into a convert plus an add
    uint32_t one_bits;
    asm("s_mov_b32 %0, 0x3f800000" : "=s"(one_bits));
    const float one = __builtin_bit_cast(float, one_bits);

    Rel rel, rel_next;
    const uint32_t nblocks = (nchans + cbr - 1u) / cbr;
    const int32_t *rt = reltab + ((uint64_t)dt * nblocks * 4u + wave) * Rel::WORDS;
    prefetch(0);
    stage(0, lds);
    if (cbr < nchans) prefetch(cbr);
    rel.load(rt, lane);
    wait_vmem();
    __syncthreads();
    build(0, lds);
    __syncthreads();
    uint32_t k = 0, since_fold = 0;
    for (uint32_t cb = 0; cb < nchans; cb += cbr, ++k) {
      const uint32_t *cur = lds + (k & 1u) * bufw + lane;
      const bool more = cb + cbr < nchans;
      if (more) {
        stage(cb + cbr, lds + ((k + 1u) & 1u) * bufw);
        if (cb + 2u * cbr < nchans) prefetch(cb + 2u * cbr);
        rel_next.load(rt + (uint64_t)(k + 1u) * 4u * Rel::WORDS, lane);
      }
      const uint32_t nc = min(cbr, nchans - cb);
      if (IS_U8) {
        if (since_fold + nc > 256u) {  // 255 + 256 * 255 = 65535 still fits in 16 bits
          fold();
          since_fold = 0;
        }
        since_fold += nc;
      }
      auto add_word = [&](int i, int j, uint32_t word) {
        if (IS_U8) {
          uacc[i][j] += __builtin_bit_cast(ushort2_t, word);
        } else {
          const half2_t h = __builtin_bit_cast(half2_t, word);
          facc[i][j][0] = __builtin_fmaf((float)h.x, one, facc[i][j][0]);
          facc[i][j][1] = __builtin_fmaf((float)h.y, one, facc[i][j][1]);
        }
      };
      if constexpr (PAIRS) {
        // rows r, r+1 of the block are channels 2k, 2k+1 (blocks start at
        // even channels); the pair sum rounds once to FP16, then to FP32.
        // An odd last channel pairs with the zero row build() leaves after
        // it: h + 0 = h exactly, and the accumulators, which start at +0,
        // are never -0, so adding +0 for -0 changes no result.
        auto add_pair = [&](int i, int j, uint32_t a, uint32_t b2) {
          const half2_t s = __builtin_bit_cast(half2_t, a) + __builtin_bit_cast(half2_t, b2);
          add_word(i, j, __builtin_bit_cast(uint32_t, s));
        };
#pragma unroll
        for (int r = 0; r < CB; r += 2) {
          if ((uint32_t)r < nc) {
            if constexpr (BURST > 0) {
#pragma unroll
              for (int i0 = 0; i0 < MW; i0 += BURST) {
                uint32_t w0[BURST][K2], w1[BURST][K2];
#pragma unroll
                for (int g = 0; g < BURST; ++g) {
                  const uint32_t *p0 = at_bytes(cur, rel.get(r, i0 + g));
                  const uint32_t *p1 = at_bytes(cur, rel.get(r + 1, i0 + g));
#pragma unroll
                  for (int j = 0; j < K2; ++j) {
                    w0[g][j] = p0[64 * j];
                    w1[g][j] = p1[64 * j];
                  }
                }
                __builtin_amdgcn_sched_barrier(0);
#pragma unroll
                for (int g = 0; g < BURST; ++g)
#pragma unroll
                  for (int j = 0; j < K2; ++j) add_pair(i0 + g, j, w0[g][j], w1[g][j]);
              }
            } else {
#pragma unroll
              for (int i = 0; i < MW; ++i) {
                const uint32_t *p0 = at_bytes(cur, rel.get(r, i));
                const uint32_t *p1 = at_bytes(cur, rel.get(r + 1, i));
#pragma unroll
                for (int j = 0; j < K2; ++j) add_pair(i, j, p0[64 * j], p1[64 * j]);
              }
            }
          }
        }
      } else {
#pragma unroll
        for (int r = 0; r < CB; ++r) {
          if ((uint32_t)r < nc) {
            if constexpr (BURST > 0) {
#pragma unroll
              for (int i0 = 0; i0 < MW; i0 += BURST) {
                uint32_t w[BURST][K2];
#pragma unroll
                for (int g = 0; g < BURST; ++g) {
                  const uint32_t *p = at_bytes(cur, rel.get(r, i0 + g));
#pragma unroll
                  for (int j = 0; j < K2; ++j) w[g][j] = p[64 * j];
                }
                __builtin_amdgcn_sched_barrier(0);
#pragma unroll
                for (int g = 0; g < BURST; ++g)
#pragma unroll
                  for (int j = 0; j < K2; ++j) add_word(i0 + g, j, w[g][j]);
              }
            } else {
#pragma unroll
              for (int i = 0; i < MW; ++i) {
                const uint32_t *p = at_bytes(cur, rel.get(r, i));
#pragma unroll
                for (int j = 0; j < K2; ++j) add_word(i, j, p[64 * j]);
              }
            }
          }
        }
      }
      wait_vmem();
      __syncthreads();
      if (more) {
        build(cb + cbr, lds + ((k + 1u) & 1u) * bufw);
        __syncthreads();
        rel = rel_next;
      }
    }

    const double base = (double)nchans * -4.0;
#pragma unroll
    for (int i = 0; i < MW; ++i) {
      const uint32_t d = dw + (uint32_t)i;
      if (d < ndms) {
        float *o = out + (uint64_t)d * nsamp;
#pragma unroll
        for (int j = 0; j < K2; ++j) {
          const uint32_t t = t0 + 2u * lane + 128u * j;
          float v0, v1;
          if (IS_U8) {
            const uint32_t s0 = 256u * (uint32_t)uhi[i][j].x + (uint32_t)uacc[i][j].x;
            const uint32_t s1 = 256u * (uint32_t)uhi[i][j].y + (uint32_t)uacc[i][j].y;
            v0 = (float)((double)s0 * (8.0 / 255.0) + base);
            v1 = (float)((double)s1 * (8.0 / 255.0) + base);
          } else {
            v0 = facc[i][j][0];
            v1 = facc[i][j][1];
          }
          if (t < nsamp) o[t] = v0;
          if (t + 1u < nsamp) o[t + 1] = v1;
        }
      }
    }
  }
};

template <int MW, int K2, int CB, int BURST = 0>
using F16Tile = PackedTile<MDSM_F16, MW, K2, CB, BURST>;
template <int MW, int K2, int CB, int BURST = 0>
using U8Tile = PackedTile<MDSM_U8, MW, K2, CB, BURST>;
// The F16_PAIRS tile of an F16 tile shape; grouped reads take half as many
// trials, since each reads two channels.
template <typename F16T>
struct PairsOf;
template <int MW, int K2, int CB, int BURST>
struct PairsOf<PackedTile<MDSM_F16, MW, K2, CB, BURST>> {
  using type = PackedTile<MDSM_F16_PAIRS, MW, K2, CB, BURST / 2>;
};

// ================================================ pair-sum rows (F16, U8)
// F16_PAIRS and U8 tiles that add each channel pair once per distinct delay
// difference instead of once per trial.  For channels c0 = 2k, c1 = 2k+1
// and trial d, the pair sum x0[t + D0(d)] + x1[t + D1(d)] (F16_PAIRS: one
// FP16 rounding; U8: exact in 16 bits) depends on d only through D0(d) and
// delta(d) = D1(d) - D0(d), and inside one DM tile delta takes a few
// consecutive values dmin .. dmin+n-1.  After a channel block is staged (one
// raw window per channel), the build pass adds the pair's windows once per
// value into sum rows S_j[s] = x0[s] + x1[s + dmin + j] (a P and a Q copy,
// as in the F16 layout), and each trial reads one word of S_{delta(d)-dmin}
// at D0(d): one LDS word per four channel adds, then two v_fma_mix_f32
// (F16_PAIRS) or one v_pk_add_u16 (U8), where the packed layouts above read
// two words per trial.  The sums are the same operations on the same
// operands, so the results are identical to those layouts'.  A shape takes
// a survey when n <= NS and delta >= 0 for every DM tile and pair
// (pairsum_accepts); the packed shapes take every survey.
//
// LDS: the raw windows of the block's cbr channels (wp words each, one
// buffer: the build has consumed them before the next block is staged),
// then NS sum rows of 2 wp words (P, Q) per pair.  The trial loop reads only
// the sum rows, so the next block's windows can arrive while it runs.
// Table: the trial offsets (one per pair and trial, relative to the sum
// rows), then per (DM tile, block, pair) the words {dmin, n}.
template <int MODE, int MW, int K2, int CBP, int NS, int BURST = 0>
struct PairSumTile {
  static constexpr int TT = 128 * K2;
  static constexpr int TD = 4 * MW;
  static constexpr int CB = 2 * CBP;
  static constexpr bool PAIRS = true;
  static constexpr bool IS_U8 = MODE == MDSM_U8;
  static constexpr int NSUM = NS;
  // window origins on even samples (F16: one dword = 2 samples) or on
  // multiples of 4 (U8: one dword = 4 samples)
  static constexpr uint32_t ORIGIN = IS_U8 ? 4u : 2u;
  static constexpr int PACK = 1;
  using T = std::conditional_t<IS_U8, uint8_t, __half>;
  using Rel = RelOffsets<MW, CBP, PACK>;
  static_assert(MODE == MDSM_F16_PAIRS || MODE == MDSM_U8, "pair sums of F16 or U8 rows");
  static_assert(NS >= 1 && NS <= 4, "sum rows per pair (shifts up to 4 samples)");
  static_assert(CBP == 1 || CBP == 2 || CBP == 4, "waves per pair in the build: 4 / CBP");
  static bool fits(uint32_t) { return true; }
  // rw: words of one copy (window of TT + spread samples from an even origin)
  static uint32_t row_words(uint32_t spread) { return PackedGeom<MW, K2, CB>::copy_words(spread); }
  // raw window dwords per channel: F16 one copy; U8 the bytes of 2 rw + 8
  // samples, padded to 64
  static __host__ __device__ uint32_t raw_words(uint32_t rw) {
    return IS_U8 ? ((rw / 2u + 2u + 63u) & ~63u) : rw;
  }
  // per staged channel: its raw window and half of its pair's sum rows
  static uint32_t lds_words_per_row(uint32_t rw) { return raw_words(rw) + (uint32_t)NS * rw; }

  // U8: the n sum rows of one pair, exact u16 pairs: row j is P_j[k] =
  // (x0[2k] + x1[2k + SH], x0[2k + 1] + x1[2k + 1 + SH]) with SH = j + F,
  // F = dmin & 3, samples counted from the 4-aligned origins of the two raw
  // windows (bytes widened by v_perm_b32, added by v_pk_add_u16; sums <=
  // 510), and its Q copy.  Lane l builds words k0 + 4 l .. + 3 of every
  // `step` from 3 dwords of R0 and 4 of R1 (reads end at most 2 dwords past
  // word wp / 2 of a row, inside the row's padding).
  template <int F>
  __device__ __forceinline__ static void build_pair_u8(const uint32_t *R0, const uint32_t *R1,
                                                       uint32_t *S, uint32_t wp, uint32_t n,
                                                       bool solo, uint32_t k0, uint32_t step) {
    for (uint32_t k = k0; k < wp; k += step) {
      const uint2 a01 = *reinterpret_cast<const uint2 *>(R0 + k / 2u);
      const uint32_t a2 = R0[k / 2u + 2u];
      // x0 samples 2i, 2i + 1 as a u16 pair, i = 0..4
      const uint32_t av[5] = {__builtin_amdgcn_perm(0u, a01.x, 0x0C010C00u),
                              __builtin_amdgcn_perm(0u, a01.x, 0x0C030C02u),
                              __builtin_amdgcn_perm(0u, a01.y, 0x0C010C00u),
                              __builtin_amdgcn_perm(0u, a01.y, 0x0C030C02u),
                              __builtin_amdgcn_perm(0u, a2, 0x0C010C00u)};
      uint32_t z[4] = {0u, 0u, 0u, 0u};
      if (!solo) {
        const uint2 z01 = *reinterpret_cast<const uint2 *>(R1 + k / 2u);
        const uint2 z23 = *reinterpret_cast<const uint2 *>(R1 + k / 2u + 2u);
        z[0] = z01.x, z[1] = z01.y, z[2] = z23.x, z[3] = z23.y;
      }
#pragma unroll
      for (int j = 0; j < NS; ++j) {
        if ((uint32_t)j >= n) break;
        const int sh = j + F;
        uint32_t sv[5];
#pragma unroll
        for (int i = 0; i < 5; ++i) {
          // x1 bytes b = 2i + sh and b + 1: dword b / 4 (and the next one
          // when b % 4 == 3), selected into the low bytes of the two halves
          const int bb = 2 * i + sh, w = bb >> 2, o = bb & 3;
          const uint32_t lo = z[w], hi = o == 3 ? z[w + 1] : z[w];
          const uint32_t bv = __builtin_amdgcn_perm(hi, lo, 0x0C000C00u | (uint32_t)o | ((uint32_t)(o + 1) << 16));
          const ushort2_t sum = __builtin_bit_cast(ushort2_t, av[i]) + __builtin_bit_cast(ushort2_t, bv);
          sv[i] = __builtin_bit_cast(uint32_t, sum);
        }
        uint32_t *P = S + (uint32_t)j * 2u * wp;
        *reinterpret_cast<uint4 *>(P + k) = make_uint4(sv[0], sv[1], sv[2], sv[3]);
        *reinterpret_cast<uint4 *>(P + wp + k) = make_uint4(
            __builtin_amdgcn_alignbit(sv[1], sv[0], 16), __builtin_amdgcn_alignbit(sv[2], sv[1], 16),
            __builtin_amdgcn_alignbit(sv[3], sv[2], 16), __builtin_amdgcn_alignbit(sv[4], sv[3], 16));
      }
    }
  }

  // F16: the n sum rows of one pair over part of the window: row j is
  // P_j[k] = fl16(R0[k] + R1 shifted by SH = j + F samples) (two halves per
  // word, one v_pk_add_f16) and its Q copy, F = dmin & 1.  Lane l builds
  // words k0 + 4 l .. + 3 of every `step`; R0 and R1 are read once for all
  // rows.  R0 reads run one word past the row and R1 reads up to four (into
  // the next staged row or the sum rows, inside the allocation); those words
  // only reach P and Q words that no trial reads.
  template <int F>
  __device__ __forceinline__ static void build_pair(const uint32_t *R0, const uint32_t *R1, uint32_t *S,
                                                    uint32_t wp, uint32_t n, bool solo, uint32_t k0,
                                                    uint32_t step) {
    for (uint32_t k = k0; k < wp; k += step) {
      const uint4 a = *reinterpret_cast<const uint4 *>(R0 + k);
      const uint32_t av[5] = {a.x, a.y, a.z, a.w, R0[k + 4]};
      uint32_t z[8] = {0u, 0u, 0u, 0u, 0u, 0u, 0u, 0u};
      if (!solo) {
        const uint4 z0 = *reinterpret_cast<const uint4 *>(R1 + k);
        const uint4 z1 = *reinterpret_cast<const uint4 *>(R1 + k + 4);
        z[0] = z0.x, z[1] = z0.y, z[2] = z0.z, z[3] = z0.w;
        z[4] = z1.x, z[5] = z1.y, z[6] = z1.z, z[7] = z1.w;
      }
#pragma unroll
      for (int j = 0; j < NS; ++j) {
        if ((uint32_t)j >= n) break;
        const int sh = j + F, m = sh >> 1;
        const bool e = (sh & 1) != 0;
        uint32_t sv[5];
#pragma unroll
        for (int i = 0; i < 5; ++i) {
          const uint32_t b = e ? __builtin_amdgcn_alignbit(z[m + i + 1], z[m + i], 16) : z[m + i];
          const half2_t h = __builtin_bit_cast(half2_t, av[i]) + __builtin_bit_cast(half2_t, b);
          sv[i] = __builtin_bit_cast(uint32_t, h);
        }
        uint32_t *P = S + (uint32_t)j * 2u * wp;
        *reinterpret_cast<uint4 *>(P + k) = make_uint4(sv[0], sv[1], sv[2], sv[3]);
        *reinterpret_cast<uint4 *>(P + wp + k) = make_uint4(
            __builtin_amdgcn_alignbit(sv[1], sv[0], 16), __builtin_amdgcn_alignbit(sv[2], sv[1], 16),
            __builtin_amdgcn_alignbit(sv[3], sv[2], 16), __builtin_amdgcn_alignbit(sv[4], sv[3], 16));
      }
    }
  }

  __device__ __forceinline__ static void run(const T *__restrict__ in, uint64_t ld,
                                             const int32_t *__restrict__ table,
                                             const int32_t *__restrict__ reltab, uint32_t nchans,
                                             uint32_t nsamp, uint32_t ndms, uint32_t wp,
                                             uint32_t cbr, float *__restrict__ out, uint32_t tt,
                                             uint32_t dt, float *ldsf) {
    uint32_t *lds = reinterpret_cast<uint32_t *>(ldsf);
    const uint32_t tid = threadIdx.x;
    const uint32_t lane = tid & 63u;
    const uint32_t wave = uniform(tid >> 6);
    const uint32_t t0 = tt * TT, d0 = dt * TD;
    const uint32_t dw = d0 + wave * MW;
    const uint64_t ldw = IS_U8 ? ld / 4u : ld / 2u;  // source dwords per row
    const uint32_t *src32 = reinterpret_cast<const uint32_t *>(in);
    const uint32_t rawn = raw_words(wp);
    uint32_t *raw = lds;                 // cbr windows of rawn dwords
    uint32_t *sums = lds + cbr * rawn;   // (cbr / 2) * NS * 2 wp words
    const uint32_t nblocks = (nchans + cbr - 1u) / cbr;
    const uint32_t n_dtiles = (ndms + TD - 1u) / TD;
    const int32_t *dtab =
        reltab + (uint64_t)n_dtiles * nblocks * 4u * Rel::WORDS + (uint64_t)dt * nblocks * CBP * 2u;
    const uint32_t lane4 = 4u * lane;

    // window starts of the next block to stage, and its pairs' {dmin, n}
    uint32_t g0pre[CB];
    int32_t dminpre[CBP];
    uint32_t npre[CBP];
    auto prefetch = [&](uint32_t cb, uint32_t k) {
#pragma unroll
      for (int p = 0; p < CBP; ++p) {
        const uint32_t c0 = cb + 2u * (uint32_t)p;
        if (2u * (uint32_t)p < cbr && c0 < nchans) {
          const uint32_t g = t0 + (uint32_t)table[(uint64_t)c0 * ndms + d0];
          dminpre[p] = dtab[((uint64_t)k * CBP + p) * 2u];
          npre[p] = (uint32_t)dtab[((uint64_t)k * CBP + p) * 2u + 1u];
          g0pre[2 * p] = g;
          g0pre[2 * p + 1] = (g & ~(ORIGIN - 1u)) + (uint32_t)dminpre[p];
        }
      }
    };
    auto stage = [&](uint32_t cb) {
      const uint32_t nc = min(cbr, nchans - cb);
#pragma unroll
      for (int rr = 0; rr < CB; ++rr) {
        if ((uint32_t)rr >= nc) break;
        const uint32_t r = (uint32_t)rr;
        const uint64_t w0 = IS_U8 ? g0pre[rr] >> 2 : g0pre[rr] >> 1;
        const uint32_t *src = src32 + (uint64_t)(cb + r) * ldw + w0;
        const int32_t ldw32 = (int32_t)ldw;  // rows < 2 GiB
        const int32_t avail = 4 * (ldw32 - min((int32_t)w0, ldw32));
        uint32_t *dst = raw + r * rawn;
        for (uint32_t q = wave; 64u * q < rawn; q += 4u) {
          const int32_t rest = max(avail - 256 * (int32_t)q, 0);
          chunk_to_lds(src + 64u * q, (uint32_t)rest, dst + 64u * q, lane4);
        }
      }
    };
    int32_t dmin[CBP];
    uint32_t nrow[CBP];
    auto take = [&]() {
#pragma unroll
      for (int p = 0; p < CBP; ++p) {
        dmin[p] = dminpre[p];
        nrow[p] = npre[p];
      }
    };
    // pair p is built by waves p W .. p W + W - 1 (W = 4 / CBP), wave p W + h
    // taking the 256-word parts h, h + W, ... of the window; all of a pair's
    // sum rows come from one read of its two windows
    auto build = [&](uint32_t cb) {
      const uint32_t nc = min(cbr, nchans - cb);
      constexpr uint32_t W = 4u / CBP;
      const uint32_t p = wave / W, h = wave % W;
      if (2u * p >= nc) return;
      const bool solo = 2u * p + 1u >= nc;
      const uint32_t *R0 = raw + 2u * p * rawn;
      uint32_t *S = sums + p * (uint32_t)NS * 2u * wp;
      const uint32_t k0 = 256u * h + 4u * lane, step = 256u * W;
      // dmin and nrow of pair p, selected with compile-time indices
      int32_t dm = dmin[0];
      uint32_t nr = nrow[0];
#pragma unroll
      for (int q = 1; q < CBP; ++q)
        if (p == (uint32_t)q) dm = dmin[q], nr = nrow[q];
      if constexpr (IS_U8) {
        switch ((uint32_t)dm & 3u) {
          case 0: build_pair_u8<0>(R0, R0 + rawn, S, wp, nr, solo, k0, step); break;
          case 1: build_pair_u8<1>(R0, R0 + rawn, S, wp, nr, solo, k0, step); break;
          case 2: build_pair_u8<2>(R0, R0 + rawn, S, wp, nr, solo, k0, step); break;
          default: build_pair_u8<3>(R0, R0 + rawn, S, wp, nr, solo, k0, step); break;
        }
      } else {
        if ((uint32_t)dm & 1u) build_pair<1>(R0, R0 + rawn, S, wp, nr, solo, k0, step);
        else build_pair<0>(R0, R0 + rawn, S, wp, nr, solo, k0, step);
      }
    };

    // F16: FP32 accumulators.  U8: u16 pair accumulators with the running
    // sum S = 256 * uhi + uacc held exactly, as in the U8 tile: a fold moves
    // uacc's high bytes into uhi and leaves uacc <= 255, so 256 more channels
    // (128 pair sums <= 510) fit in 16 bits.
    float facc[IS_U8 ? 1 : MW][IS_U8 ? 1 : K2][2];
    ushort2_t uacc[IS_U8 ? MW : 1][IS_U8 ? K2 : 1];
    ushort2_t uhi[IS_U8 ? MW : 1][IS_U8 ? K2 : 1];
#pragma unroll
    for (int i = 0; i < MW; ++i)
#pragma unroll
      for (int j = 0; j < K2; ++j) {
        if constexpr (IS_U8) {
          uacc[i][j] = ushort2_t{0, 0};
          uhi[i][j] = ushort2_t{0, 0};
        } else {
          facc[i][j][0] = facc[i][j][1] = 0.0f;
        }
      }
    // 1.0 from an opaque scalar move: fma(fpext(h), one, acc) then selects
    // v_fma_mix_f32 instead of a convert plus an add
    uint32_t one_bits;
    asm("s_mov_b32 %0, 0x3f800000" : "=s"(one_bits));
    const float one = __builtin_bit_cast(float, one_bits);

    Rel rel, rel_next;
    const int32_t *rt = reltab + ((uint64_t)dt * nblocks * 4u + wave) * Rel::WORDS;
    prefetch(0, 0);
    stage(0);
    take();
    if (cbr < nchans) prefetch(cbr, 1);
    rel.load(rt, lane);
    wait_vmem();
    __syncthreads();
    build(0);
    __syncthreads();
    const uint32_t *cur = sums + lane;
    uint32_t k = 0, since_fold = 0;
    for (uint32_t cb = 0; cb < nchans; cb += cbr, ++k) {
      const bool more = cb + cbr < nchans;
      if (more) {
        stage(cb + cbr);
        take();
        if (cb + 2u * cbr < nchans) prefetch(cb + 2u * cbr, k + 2u);
        rel_next.load(rt + (uint64_t)(k + 1u) * 4u * Rel::WORDS, lane);
      }
      const uint32_t nc = min(cbr, nchans - cb);
      if constexpr (IS_U8) {
        if (since_fold + nc > 256u) {  // 255 + 128 * 510 = 65535 still fits in 16 bits
#pragma unroll
          for (int i = 0; i < MW; ++i)
#pragma unroll
            for (int j = 0; j < K2; ++j) {
              uhi[i][j] += uacc[i][j] >> (ushort)8;
              uacc[i][j] &= (ushort)0xFF;
            }
          since_fold = 0;
        }
        since_fold += nc;
      }
      auto add_word = [&](int i, int j, uint32_t w) {
        if constexpr (IS_U8) {
          uacc[i][j] += __builtin_bit_cast(ushort2_t, w);
        } else {
          const half2_t h = __builtin_bit_cast(half2_t, w);
          facc[i][j][0] = __builtin_fmaf((float)h.x, one, facc[i][j][0]);
          facc[i][j][1] = __builtin_fmaf((float)h.y, one, facc[i][j][1]);
        }
      };
#pragma unroll
      for (int p = 0; p < CBP; ++p) {
        if (2u * (uint32_t)p < nc) {
          if constexpr (BURST > 0) {
            // the reads of BURST trials are issued together ahead of their
            // adds (a scheduling fence keeps them grouped)
#pragma unroll
            for (int i0 = 0; i0 < MW; i0 += BURST) {
              uint32_t w[BURST][K2];
#pragma unroll
              for (int g = 0; g < BURST; ++g) {
                const uint32_t *q = at_bytes(cur, rel.get(p, i0 + g));
#pragma unroll
                for (int j = 0; j < K2; ++j) w[g][j] = q[64 * j];
              }
              __builtin_amdgcn_sched_barrier(0);
#pragma unroll
              for (int g = 0; g < BURST; ++g)
#pragma unroll
                for (int j = 0; j < K2; ++j) add_word(i0 + g, j, w[g][j]);
            }
          } else {
#pragma unroll
            for (int i = 0; i < MW; ++i) {
              const uint32_t *q = at_bytes(cur, rel.get(p, i));
#pragma unroll
              for (int j = 0; j < K2; ++j) add_word(i, j, q[64 * j]);
            }
          }
        }
      }
      wait_vmem();
      __syncthreads();
      if (more) {
        build(cb + cbr);
        __syncthreads();
        rel = rel_next;
      }
    }

    const double base = (double)nchans * -4.0;
#pragma unroll
    for (int i = 0; i < MW; ++i) {
      const uint32_t d = dw + (uint32_t)i;
      if (d < ndms) {
        float *o = out + (uint64_t)d * nsamp;
#pragma unroll
        for (int j = 0; j < K2; ++j) {
          const uint32_t t = t0 + 2u * lane + 128u * j;
          float v0, v1;
          if constexpr (IS_U8) {
            const uint32_t s0 = 256u * (uint32_t)uhi[i][j].x + (uint32_t)uacc[i][j].x;
            const uint32_t s1 = 256u * (uint32_t)uhi[i][j].y + (uint32_t)uacc[i][j].y;
            v0 = (float)((double)s0 * (8.0 / 255.0) + base);
            v1 = (float)((double)s1 * (8.0 / 255.0) + base);
          } else {
            v0 = facc[i][j][0];
            v1 = facc[i][j][1];
          }
          if (t < nsamp) o[t] = v0;
          if (t + 1u < nsamp) o[t + 1] = v1;
        }
      }
    }
  }
};

template <typename T>
struct IsPairSum : std::false_type {};
template <int MODE, int MW, int K2, int CBP, int NS, int BURST>
struct IsPairSum<PairSumTile<MODE, MW, K2, CBP, NS, BURST>> : std::true_type {};

// Trial offsets and {dmin, n} of a pair-sum tile (layout above): entry e <
// nrel is ((dt * nblocks + k) * 4 + w) * WORDS + idx with (p, i) =
// divmod(idx, MW); the byte offset of trial dw + i of pair p inside the sum
// rows, from q0 = D0(d) - D0(d0) + (D0(d0) & 1) (sample position from the
// even window origin) and j = delta(d) - dmin: word ((p NS + j) 2 + (q0 & 1))
// wp + q0 / 2.  Entries past nrel hold {dmin, n} per (dt, k, p).  A pair
// without its second channel (odd nchans) sums with zero: dmin 0, n 1.
template <typename Tile>
__global__ __launch_bounds__(256) void pairsum_offset_kernel(const int32_t *__restrict__ table,
                                                             uint32_t nchans, uint32_t ndms,
                                                             uint32_t wp, uint32_t cbr,
                                                             uint32_t nblocks, uint64_t nrel,
                                                             uint64_t total,
                                                             int32_t *__restrict__ reltab) {
  using Rel = typename Tile::Rel;
  constexpr uint32_t MW = Tile::TD / 4, CBP = Tile::CB / 2, NS = Tile::NSUM;
  const uint64_t e = (uint64_t)blockIdx.x * 256u + threadIdx.x;
  if (e >= total) return;
  uint32_t k, dt, p, i = 0, w = 0, field = 0;
  const bool is_rel = e < nrel;
  if (is_rel) {
    const uint32_t idx = (uint32_t)(e % Rel::WORDS);
    const uint64_t g = e / Rel::WORDS;
    w = (uint32_t)(g & 3u);
    const uint64_t blk = g >> 2;
    k = (uint32_t)(blk % nblocks);
    dt = (uint32_t)(blk / nblocks);
    p = min(idx / MW, CBP - 1u);  // padding words repeat the last pair
    i = idx % MW;
  } else {
    const uint64_t q = e - nrel;
    field = (uint32_t)(q & 1u);
    const uint64_t pq = q >> 1;
    p = (uint32_t)(pq % CBP);
    const uint64_t blk = pq / CBP;
    k = (uint32_t)(blk % nblocks);
    dt = (uint32_t)(blk / nblocks);
  }
  const uint32_t d0 = dt * Tile::TD, dl = min(d0 + (uint32_t)Tile::TD, ndms) - 1u;
  const uint32_t c0 = min(k * cbr + 2u * p, nchans - 1u);
  const bool solo = c0 + 1u >= nchans;
  const int32_t *r0 = table + (uint64_t)c0 * ndms;
  const int32_t *r1 = table + (uint64_t)(solo ? c0 : c0 + 1u) * ndms;
  int32_t dmin = 0, dmax = 0;
  if (!solo) {
    dmin = dmax = r1[d0] - r0[d0];
    for (uint32_t d = d0 + 1u; d <= dl; ++d) {
      const int32_t v = r1[d] - r0[d];
      dmin = min(dmin, v);
      dmax = max(dmax, v);
    }
  }
  if (!is_rel) {
    reltab[e] = field ? dmax - dmin + 1 : dmin;
    return;
  }
  const uint32_t d = min(d0 + w * MW + i, ndms - 1u);
  const int32_t base = r0[d0];
  const uint32_t q0 = (uint32_t)(r0[d] - base) + ((uint32_t)base & (Tile::ORIGIN - 1u));
  const uint32_t j = solo ? 0u : (uint32_t)(r1[d] - r0[d] - dmin);
  const uint32_t word = ((p * NS + min(j, NS - 1u)) * 2u + (q0 & 1u)) * wp + (q0 >> 1);
  reltab[e] = (int32_t)(4u * word);
}

// XCD-aware workgroup order for the single-beam launch: off unless
// MDSM_XCD_REMAP=1 (measured at 1024 channels x 4096 trials x 2^20 samples it
// was level for FP32 and 2-6% slower for FP16/U8 than the plain order)
bool xcd_remap_enabled() {
  static const bool on = [] {
    const char *e = getenv("MDSM_XCD_REMAP");
    return e && e[0] == '1';
  }();
  return on;
}

// ================================================================= kernels
template <typename Tile>
__global__ __launch_bounds__(256) void dedisp_kernel(const typename Tile::T *__restrict__ in,
                                                     uint64_t ld,
                                                     const int32_t *__restrict__ table,
                                                     const int32_t *__restrict__ reltab,
                                                     uint32_t nchans, uint32_t nsamp,
                                                     uint32_t ndms, uint32_t rw, uint32_t cbr,
                                                     uint32_t n_dtiles, uint32_t n_ttiles,
                                                     uint32_t xcd_remap, float *__restrict__ out) {
  extern __shared__ __attribute__((aligned(16))) float lds[];
  // DM tiles vary fastest: the workgroups resident at one time then share a
  // narrow time range of the input.  Workgroup IDs are dealt round-robin over
  // the 8 XCDs, each with its own L2; with xcd_remap, XCD x takes the time
  // tiles tt = x (mod 8), so all DM tiles re-reading one time window run
  // behind one L2 (grid: 8 * ceil(n_ttiles / 8) * n_dtiles).
  const uint32_t b = blockIdx.x;
  uint32_t tt, dt;
  if (xcd_remap) {
    const uint32_t x = b & 7u, k = b >> 3;
    tt = ((k / n_dtiles) << 3) + x;
    dt = k % n_dtiles;
    if (tt >= n_ttiles) return;  // whole workgroup, before any barrier
  } else {
    tt = b / n_dtiles;
    dt = b % n_dtiles;
  }
  Tile::run(in, ld, table, reltab, nchans, nsamp, ndms, rw, cbr, out, tt, dt, lds);
}

// Trial-offset table of one tile shape: entry ((dt * nblocks + k) * 4 + w)
// * WORDS + 16 n + l holds the byte offset of (r, i) = divmod(16 n + l, MW)
// for channel block k (cbr channels) and wave w of DM tile dt.
template <typename Tile>
__global__ __launch_bounds__(256) void trial_offset_kernel(const int32_t *__restrict__ table,
                                                           uint32_t nchans, uint32_t ndms,
                                                           uint32_t rw, uint32_t cbr,
                                                           uint32_t nblocks, uint64_t total,
                                                           int32_t *__restrict__ reltab) {
  using Rel = typename Tile::Rel;
  constexpr uint32_t MW = Tile::TD / 4;
  const uint64_t e = (uint64_t)blockIdx.x * 256u + threadIdx.x;
  if (e >= total) return;
  const uint32_t idx = (uint32_t)(e % Rel::WORDS);
  const uint64_t g = e / Rel::WORDS;  // (dt * nblocks + k) * 4 + w
  const uint32_t w = (uint32_t)(g & 3u);
  const uint64_t blk = g >> 2;
  const uint32_t k = (uint32_t)(blk % nblocks), dt = (uint32_t)(blk / nblocks);
  const uint32_t d0 = dt * Tile::TD;
  reltab[e] = rel_entry<MW, Tile::PACK>(table, ndms, nchans, k * cbr, d0, d0 + w * MW, idx, rw);
}

// Multi-beam item numbering (both launch shapes): a beam's items are
// XCD-major, 8 A * ceil(n_ttiles / (8 A)) * n_dtiles of them starting at a
// multiple of 8, A = tgroup; item g belongs to XCD slot s = g % 8, and with
// k = (g - begin) / 8, j = k / (A n_dtiles), r = k % (A n_dtiles) it is DM
// tile r / A of time tile 8 A j + A s + r % A: each XCD takes groups of A
// adjacent time tiles and runs a DM tile for all A of them back to back, so
// their shifted windows overlap in its L2 (A = 1: time tile 8 j + s, DM
// tile r).  Padding items with a time tile past the beam's end do nothing.
struct BeamItem {
  const BeamDesc *bd;
  uint32_t tt, dt;
  bool valid;
};

__device__ __forceinline__ BeamItem decode_item(const BeamDesc *__restrict__ beams, uint32_t nbeams,
                                                uint32_t g) {
  uint32_t lo = 0, hi = nbeams - 1u;
  while (lo < hi) {
    const uint32_t mid = (lo + hi + 1u) >> 1;
    if (beams[mid].item_begin <= g) lo = mid;
    else hi = mid - 1u;
  }
  const BeamDesc *bd = beams + lo;
  const uint32_t local = g - bd->item_begin;
  const uint32_t k = local >> 3, A = bd->tgroup;
  const uint32_t per = A * bd->n_dtiles;
  const uint32_t j = k / per, r = k - j * per;
  BeamItem it;
  it.bd = bd;
  it.dt = r / A;
  it.tt = 8u * A * j + A * (local & 7u) + (r - it.dt * A);
  it.valid = it.tt < bd->n_ttiles;
  return it;
}

// (The dynamic LDS array is declared 16-byte aligned: it follows any static
// LDS of the kernel, and the packed tiles access it with 8-byte operations.)
__device__ __forceinline__ uint64_t uniform64(uint64_t v) {
  return ((uint64_t)uniform((uint32_t)(v >> 32)) << 32) | (uint64_t)uniform((uint32_t)v);
}

// Every tile parameter goes through readfirstlane: inside the persistent
// kernel's item loop the compiler cannot prove the descriptor fields
// wave-uniform and would otherwise keep the per-block address arithmetic in
// VALU registers.
template <typename Tile>
__device__ __forceinline__ void run_item(const BeamDesc *__restrict__ beams, const BeamItem &it,
                                         uint32_t cbr, float *lds) {
  const BeamDesc &bd = *it.bd;
  char *base = reinterpret_cast<char *>(const_cast<BeamDesc *>(beams));
  Tile::run(reinterpret_cast<const typename Tile::T *>(base + (int64_t)uniform64((uint64_t)bd.in_off)),
            uniform64(bd.ld), reinterpret_cast<const int32_t *>(base + (int64_t)uniform64((uint64_t)bd.table_off)),
            reinterpret_cast<const int32_t *>(base + (int64_t)uniform64((uint64_t)bd.rel_off)),
            uniform(bd.nchans), uniform(bd.nsamp), uniform(bd.ndms), uniform(bd.row_words), cbr,
            reinterpret_cast<float *>(base + (int64_t)uniform64((uint64_t)bd.out_off)), uniform(it.tt),
            uniform(it.dt), lds);
}

// Waves per SIMD the persistent kernel is compiled for: its item loop costs
// a few VGPRs over the flat kernel, which drops the 12 x 4 and 10 x 4
// pair-sum tiles from 4 to 3 waves (above 128 VGPRs).  Held at 4 they spill
// a few dwords and still run faster (32-beam batch, persistent launch: FP16
// pair sums 1160 against 1240-1266 ms, U8 1076 against 1270-1285 ms); for
// the 8 x 4 FP16 tile the same hold (6 waves) was slower (1795 against
// 1705-1725 ms), so the other tiles are left to the compiler.
template <typename Tile>
struct PersistWaves {
  static constexpr int value = 1;
};
template <int MODE, int MW, int K2, int CBP, int NS, int BURST>
struct PersistWaves<PairSumTile<MODE, MW, K2, CBP, NS, BURST>> {
  static constexpr int value = 4;
};

// Persistent multi-beam kernel: resident workgroups pull items from eight
// queues, one per XCD slot (workgroup b drains queue b % 8, the XCD the
// dispatcher puts it on), until the queue passes total_items; every wave
// of every workgroup reaches the exit.  Beams are ordered by the host,
// largest first.
template <typename Tile>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(PersistWaves<Tile>::value)))
void batch_kernel(const BeamDesc *__restrict__ beams,
                                                    uint32_t nbeams, uint32_t total_items,
                                                    uint32_t cbr, uint32_t *__restrict__ counter) {
  extern __shared__ __attribute__((aligned(16))) float lds[];
  __shared__ uint32_t item_s;
  const uint32_t slot = blockIdx.x & 7u;
  while (true) {
    if (threadIdx.x == 0) item_s = atomicAdd(counter + slot, 1u) * 8u + slot;
    __syncthreads();
    // wave-uniform: the descriptor lookup and every tile parameter stay scalar
    const uint32_t item = uniform(item_s);
    __syncthreads();
    if (item >= total_items) break;
    const BeamItem it = decode_item(beams, nbeams, item);
    if (it.valid) run_item<Tile>(beams, it, cbr, lds);
    __syncthreads();
  }
}

// Flat multi-beam kernel: one workgroup per item (workgroup ID = item, so
// the dispatcher's round-robin puts item g on XCD g % 8).
template <typename Tile>
__global__ __launch_bounds__(256) void batch_flat_kernel(const BeamDesc *__restrict__ beams,
                                                         uint32_t nbeams, uint32_t cbr) {
  extern __shared__ __attribute__((aligned(16))) float lds[];
  const BeamItem it = decode_item(beams, nbeams, blockIdx.x);
  if (!it.valid) return;  // whole workgroup, before any barrier
  run_item<Tile>(beams, it, cbr, lds);
}

// ------------------------------------------------------------ variant table
// Variant v: F32 tile (MW, K, CB) and packed tiles (MW, K2, CB) with the same
// TD = 4*MW and similar time extent.
template <int V>
struct VariantDef;
template <> struct VariantDef<1> { using F32 = F32Tile<8, 4, 8>;   using F16 = F16Tile<8, 2, 8>;   using U8 = U8Tile<8, 2, 8>; };
template <> struct VariantDef<2> { using F32 = F32Tile<16, 4, 4>;  using F16 = F16Tile<16, 2, 4>;  using U8 = U8Tile<16, 2, 4>; };
template <> struct VariantDef<3> { using F32 = F32Tile<8, 8, 8>;   using F16 = F16Tile<8, 4, 8>;   using U8 = U8Tile<8, 4, 8>; };
template <> struct VariantDef<4> { using F32 = F32Tile<16, 8, 4>;  using F16 = F16Tile<16, 4, 4>;  using U8 = U8Tile<16, 4, 4>; };
template <> struct VariantDef<5> { using F32 = F32Tile<32, 2, 2>;  using F16 = F16Tile<32, 1, 2>;  using U8 = U8Tile<32, 1, 2>; };
// grouped-read forms of 2 and 1
template <> struct VariantDef<6> { using F32 = F32Tile<16, 4, 4, 8>; using F16 = F16Tile<16, 2, 4, 16>; using U8 = U8Tile<16, 2, 4, 16>; };
template <> struct VariantDef<7> { using F32 = F32Tile<8, 4, 8, 8>;  using F16 = F16Tile<8, 2, 8, 8>;   using U8 = U8Tile<8, 2, 8, 8>; };
template <> struct VariantDef<8> { using F32 = F32Tile<16, 4, 4, 4>; using F16 = F16Tile<16, 2, 4, 8>;  using U8 = U8Tile<16, 2, 4, 8>; };
// few trials per wave, many samples per lane: one address add per 8 (F32)
// or 16 (packed) adds, a smaller delay spread per tile, larger windows
template <> struct VariantDef<9> { using F32 = F32Tile<8, 8, 4>;    using F16 = F16Tile<8, 4, 4>;   using U8 = U8Tile<8, 4, 4>; };
template <> struct VariantDef<10> { using F32 = F32Tile<4, 16, 8>;  using F16 = F16Tile<4, 8, 8>;   using U8 = U8Tile<4, 8, 8>; };
template <> struct VariantDef<11> { using F32 = F32Tile<4, 16, 4>;  using F16 = F16Tile<4, 8, 4>;   using U8 = U8Tile<4, 8, 4>; };
// more samples per lane at 12 or 16 trials per wave: one address add per
// 8 or 6 packed adds
template <> struct VariantDef<12> { using F32 = F32Tile<12, 6, 4>;  using F16 = F16Tile<12, 4, 4>;  using U8 = U8Tile<12, 4, 4>; };
template <> struct VariantDef<13> { using F32 = F32Tile<16, 6, 4>;  using F16 = F16Tile<16, 3, 4>;  using U8 = U8Tile<16, 3, 4>; };
// 16 trials per wave over 8-channel blocks: half the barriers per add
template <> struct VariantDef<14> { using F32 = F32Tile<16, 4, 8>;  using F16 = F16Tile<16, 2, 8>;  using U8 = U8Tile<16, 2, 8>; };
// 12 trials per wave, 2 pairs per lane: fewer accumulators, 7 waves per
// SIMD for the packed tiles (68-72 VGPRs)
template <> struct VariantDef<15> { using F32 = F32Tile<12, 4, 4>;  using F16 = F16Tile<12, 2, 4>;  using U8 = U8Tile<12, 2, 4>; };
constexpr int kNumVariants = 15;
// F16_PAIRS and U8 only: pair-sum tiles (MW, K2, pairs per block, sum rows per pair)
template <int V>
struct PairSumDef;
template <> struct PairSumDef<16> { template <int M> using type = PairSumTile<M, 8, 4, 2, 3>; };
template <> struct PairSumDef<17> { template <int M> using type = PairSumTile<M, 12, 4, 2, 3>; };
template <> struct PairSumDef<18> { template <int M> using type = PairSumTile<M, 12, 2, 2, 3>; };
template <> struct PairSumDef<19> { template <int M> using type = PairSumTile<M, 16, 3, 2, 4>; };
// 17 with the reads of 2 or 3 trials issued ahead of their adds
template <> struct PairSumDef<20> { template <int M> using type = PairSumTile<M, 12, 4, 2, 3, 2>; };
template <> struct PairSumDef<21> { template <int M> using type = PairSumTile<M, 12, 4, 2, 3, 3>; };
// 10 trials x 4 pairs with grouped reads: 4 waves per SIMD for U8 (126 VGPRs)
template <> struct PairSumDef<22> { template <int M> using type = PairSumTile<M, 10, 4, 2, 3, 2>; };
constexpr int kNumPairSumVariants = 7;

template <typename V, typename Fn>
int pick_dtype(int dtype, Fn &&fn) {
  switch (dtype) {
    case MDSM_F32: return fn((typename V::F32 *)nullptr);
    case MDSM_F16: return fn((typename V::F16 *)nullptr);
    case MDSM_U8: return fn((typename V::U8 *)nullptr);
    case MDSM_F16_PAIRS: return fn((typename PairsOf<typename V::F16>::type *)nullptr);
  }
  set_error("unknown input dtype %d", dtype);
  return MDSM_ERR_ARG;
}

template <typename Fn>
int with_tile(int dtype, int variant, Fn &&fn) {
  switch (variant) {
    case 1: return pick_dtype<VariantDef<1>>(dtype, fn);
    case 2: return pick_dtype<VariantDef<2>>(dtype, fn);
    case 3: return pick_dtype<VariantDef<3>>(dtype, fn);
    case 4: return pick_dtype<VariantDef<4>>(dtype, fn);
    case 5: return pick_dtype<VariantDef<5>>(dtype, fn);
    case 6: return pick_dtype<VariantDef<6>>(dtype, fn);
    case 7: return pick_dtype<VariantDef<7>>(dtype, fn);
    case 8: return pick_dtype<VariantDef<8>>(dtype, fn);
    case 9: return pick_dtype<VariantDef<9>>(dtype, fn);
    case 10: return pick_dtype<VariantDef<10>>(dtype, fn);
    case 11: return pick_dtype<VariantDef<11>>(dtype, fn);
    case 12: return pick_dtype<VariantDef<12>>(dtype, fn);
    case 13: return pick_dtype<VariantDef<13>>(dtype, fn);
    case 14: return pick_dtype<VariantDef<14>>(dtype, fn);
    case 15: return pick_dtype<VariantDef<15>>(dtype, fn);
  }
  if (dtype == MDSM_F16_PAIRS) {
    switch (variant) {
      case 16: return fn((PairSumDef<16>::type<MDSM_F16_PAIRS> *)nullptr);
      case 17: return fn((PairSumDef<17>::type<MDSM_F16_PAIRS> *)nullptr);
      case 18: return fn((PairSumDef<18>::type<MDSM_F16_PAIRS> *)nullptr);
      case 19: return fn((PairSumDef<19>::type<MDSM_F16_PAIRS> *)nullptr);
      case 20: return fn((PairSumDef<20>::type<MDSM_F16_PAIRS> *)nullptr);
      case 21: return fn((PairSumDef<21>::type<MDSM_F16_PAIRS> *)nullptr);
      case 22: return fn((PairSumDef<22>::type<MDSM_F16_PAIRS> *)nullptr);
    }
  }
  if (dtype == MDSM_U8) {
    switch (variant) {
      case 16: return fn((PairSumDef<16>::type<MDSM_U8> *)nullptr);
      case 17: return fn((PairSumDef<17>::type<MDSM_U8> *)nullptr);
      case 18: return fn((PairSumDef<18>::type<MDSM_U8> *)nullptr);
      case 19: return fn((PairSumDef<19>::type<MDSM_U8> *)nullptr);
      case 20: return fn((PairSumDef<20>::type<MDSM_U8> *)nullptr);
      case 21: return fn((PairSumDef<21>::type<MDSM_U8> *)nullptr);
      case 22: return fn((PairSumDef<22>::type<MDSM_U8> *)nullptr);
    }
  }
  set_error("unknown dedispersion variant %d", variant);
  return MDSM_ERR_ARG;
}

// Channels staged per block: the tile's CB, reduced while the block's LDS
// (two buffers, plus raw staging for U8) exceeds 64 KB (160 KB per CU); even
// for F16_PAIRS, whose channel pairs must not straddle two blocks.
template <typename Tile>
uint32_t stage_channels(uint32_t rw) {
  const uint64_t per = Tile::lds_words_per_row(rw) * sizeof(float);
  const uint64_t step = Tile::PAIRS ? 2 : 1;
  uint64_t n = (uint64_t)Tile::CB;
  while (n > step && n * per > 64 * 1024) n -= step;
  if (n * per > 160 * 1024) return 0;
  return (uint32_t)n;
}

}  // namespace

int num_variants(int dtype) {
  return (dtype == MDSM_F16_PAIRS || dtype == MDSM_U8) ? kNumVariants + kNumPairSumVariants
                                                       : kNumVariants;
}

// Untuned defaults, the fastest shapes measured at 1024 channels x 4096
// trials: F16_PAIRS and U8 prefer the pair-sum shapes (20, 22), which take
// only surveys with few delay differences per channel pair and DM tile, and
// fall back to the fastest shape that takes every survey (15, 2).
int preferred_variant(int dtype) {
  switch (dtype) {
    case MDSM_F32: return 8;
    case MDSM_F16: return 15;
    case MDSM_F16_PAIRS: return 20;
    default: return 22;
  }
}

int fallback_variant(int dtype) {
  switch (dtype) {
    case MDSM_F16_PAIRS: return 15;
    case MDSM_U8: return 2;
    default: return preferred_variant(dtype);
  }
}

TileShape variant_shape(int dtype, int variant) {
  TileShape sh{0, 0, 0};
  with_tile(dtype, variant, [&](auto *tp) {
    using Tile = std::remove_pointer_t<decltype(tp)>;
    sh.mw = Tile::TD / 4;
    sh.k = Tile::TT / 64;
    sh.cb = Tile::CB;
    return MDSM_OK;
  });
  return sh;
}

// Pair-sum tiles: every channel pair's delay difference is >= 0 and takes
// at most ns values inside each DM tile of td trials.
static bool pairsum_accepts(const int32_t *table, uint32_t nchans, uint32_t ndms, uint32_t td,
                            uint32_t ns) {
  for (uint32_t c0 = 0; c0 + 1u < nchans; c0 += 2u) {
    const int32_t *r0 = table + (uint64_t)c0 * ndms, *r1 = r0 + ndms;
    for (uint32_t d0 = 0; d0 < ndms; d0 += td) {
      const uint32_t dl = std::min(d0 + td, ndms);
      int32_t lo = r1[d0] - r0[d0], hi = lo;
      for (uint32_t d = d0 + 1u; d < dl; ++d) {
        const int32_t v = r1[d] - r0[d];
        lo = std::min(lo, v);
        hi = std::max(hi, v);
      }
      if (lo < 0 || (int64_t)hi - lo + 1 > (int64_t)ns) return false;
    }
  }
  return true;
}

uint32_t lds_row_words(int dtype, int variant, uint32_t spread, const int32_t *table,
                       uint32_t nchans, uint32_t ndms) {
  uint32_t rw = 0;
  with_tile(dtype, variant, [&](auto *tp) {
    using Tile = std::remove_pointer_t<decltype(tp)>;
    rw = Tile::fits(spread) ? Tile::row_words(spread) : 0u;
    if constexpr (IsPairSum<Tile>::value) {
      if (!pairsum_accepts(table, nchans, ndms, Tile::TD, Tile::NSUM)) rw = 0;
    }
    return MDSM_OK;
  });
  return rw;
}

uint32_t input_alignment(int dtype) {
  return dtype == MDSM_F32 ? 1u : (storage_dtype(dtype) == MDSM_F16 ? 2u : 4u);
}

uint32_t stage_channels(int dtype, int variant, uint32_t rw) {
  uint32_t cbr = 0;
  with_tile(dtype, variant, [&](auto *tp) {
    using Tile = std::remove_pointer_t<decltype(tp)>;
    cbr = rw ? stage_channels<Tile>(rw) : 0u;
    return MDSM_OK;
  });
  return cbr;
}

uint64_t trial_offset_words(int dtype, int variant, uint32_t nchans, uint32_t ndms, uint32_t cbr) {
  uint64_t n = 0;
  with_tile(dtype, variant, [&](auto *tp) {
    using Tile = std::remove_pointer_t<decltype(tp)>;
    const uint64_t n_dtiles = (ndms + Tile::TD - 1) / Tile::TD;
    const uint64_t nblocks = cbr ? (nchans + cbr - 1) / cbr : 0;
    n = n_dtiles * nblocks * 4u * Tile::Rel::WORDS;
    if constexpr (IsPairSum<Tile>::value) n += n_dtiles * nblocks * (uint64_t)(Tile::CB / 2) * 2u;
    return MDSM_OK;
  });
  return n;
}

int launch_trial_offsets(int dtype, int variant, const int32_t *table, uint32_t nchans,
                         uint32_t ndms, uint32_t rw, uint32_t cbr, int32_t *reltab, hipStream_t s) {
  MDSM_REQUIRE(rw > 0 && cbr > 0, MDSM_ERR_ARG, "tile variant %d cannot stage this survey", variant);
  return with_tile(dtype, variant, [&](auto *tp) -> int {
    using Tile = std::remove_pointer_t<decltype(tp)>;
    const uint32_t nblocks = (nchans + cbr - 1) / cbr;
    const uint64_t total = trial_offset_words(dtype, variant, nchans, ndms, cbr);
    const uint64_t grid = (total + 255u) / 256u;
    MDSM_REQUIRE(grid < (1ull << 31), MDSM_ERR_ARG, "trial offset table too large");
    if constexpr (IsPairSum<Tile>::value) {
      MDSM_REQUIRE(cbr % 2u == 0u, MDSM_ERR_ARG, "pair-sum tiles stage whole channel pairs");
      const uint64_t nrel = (uint64_t)((ndms + Tile::TD - 1) / Tile::TD) * nblocks * 4u * Tile::Rel::WORDS;
      hipLaunchKernelGGL(pairsum_offset_kernel<Tile>, dim3((uint32_t)grid), dim3(256), 0, s, table,
                         nchans, ndms, rw, cbr, nblocks, nrel, total, reltab);
    } else {
      hipLaunchKernelGGL(trial_offset_kernel<Tile>, dim3((uint32_t)grid), dim3(256), 0, s, table,
                         nchans, ndms, rw, cbr, nblocks, total, reltab);
    }
    MDSM_HIP(hipGetLastError());
    return MDSM_OK;
  });
}

int launch_dedisperse(int dtype, int variant, const void *in, uint64_t ld, const int32_t *table,
                      const int32_t *reltab, uint32_t nchans, uint32_t nsamp, uint32_t ndms,
                      uint32_t rw, float *out, hipStream_t s) {
  MDSM_REQUIRE(rw > 0, MDSM_ERR_ARG,
               "delay spread too large for tile variant %d (choose another variant)", variant);
  MDSM_REQUIRE(reltab, MDSM_ERR_STATE, "trial offset table missing for tile variant %d", variant);
  return with_tile(dtype, variant, [&](auto *tp) -> int {
    using Tile = std::remove_pointer_t<decltype(tp)>;
    const uint32_t n_ttiles = (nsamp + Tile::TT - 1) / Tile::TT;
    const uint32_t n_dtiles = (ndms + Tile::TD - 1) / Tile::TD;
    const uint32_t remap = xcd_remap_enabled() ? 1u : 0u;
    const uint64_t blocks =
        (uint64_t)(remap ? (n_ttiles + 7u) / 8u * 8u : n_ttiles) * n_dtiles;
    MDSM_REQUIRE(blocks < (1ull << 31), MDSM_ERR_ARG, "dedispersion grid too large");
    const uint32_t cbr = stage_channels<Tile>(rw);
    MDSM_REQUIRE(cbr >= 1, MDSM_ERR_ARG, "delay spread too large for LDS staging (%u words)", rw);
    const size_t lds = (size_t)cbr * Tile::lds_words_per_row(rw) * sizeof(float);
    hipLaunchKernelGGL(dedisp_kernel<Tile>, dim3((uint32_t)blocks), dim3(256), lds, s,
                       reinterpret_cast<const typename Tile::T *>(in), ld, table, reltab, nchans,
                       nsamp, ndms, rw, cbr, n_dtiles, n_ttiles, remap, out);
    MDSM_HIP(hipGetLastError());
    return MDSM_OK;
  });
}

size_t batch_lds_bytes(int dtype, int variant, uint32_t max_rw) {
  size_t bytes = 0;
  with_tile(dtype, variant, [&](auto *tp) {
    using Tile = std::remove_pointer_t<decltype(tp)>;
    bytes = (size_t)stage_channels<Tile>(max_rw) * Tile::lds_words_per_row(max_rw) * sizeof(float);
    return MDSM_OK;
  });
  return bytes;
}

int launch_batch_flat(int dtype, int variant, const BeamDesc *d_beams, uint32_t nbeams,
                      uint32_t total_items, uint32_t max_rw, hipStream_t s) {
  MDSM_REQUIRE(max_rw > 0, MDSM_ERR_ARG, "delay spread too large for tile variant %d", variant);
  return with_tile(dtype, variant, [&](auto *tp) -> int {
    using Tile = std::remove_pointer_t<decltype(tp)>;
    const uint32_t cbr = stage_channels<Tile>(max_rw);
    MDSM_REQUIRE(cbr >= 1, MDSM_ERR_ARG, "delay spread too large for LDS staging");
    const size_t lds = (size_t)cbr * Tile::lds_words_per_row(max_rw) * sizeof(float);
    hipLaunchKernelGGL(batch_flat_kernel<Tile>, dim3(total_items), dim3(256), lds, s, d_beams,
                       nbeams, cbr);
    MDSM_HIP(hipGetLastError());
    return MDSM_OK;
  });
}

int launch_batch_dedisperse(int dtype, int variant, const BeamDesc *d_beams, uint32_t nbeams,
                            const uint32_t *, uint32_t total_items, uint32_t *d_counter,
                            uint32_t max_rw, int grid, hipStream_t s) {
  MDSM_REQUIRE(max_rw > 0, MDSM_ERR_ARG, "delay spread too large for tile variant %d", variant);
  return with_tile(dtype, variant, [&](auto *tp) -> int {
    using Tile = std::remove_pointer_t<decltype(tp)>;
    const uint32_t cbr = stage_channels<Tile>(max_rw);
    MDSM_REQUIRE(cbr >= 1, MDSM_ERR_ARG, "delay spread too large for LDS staging");
    const size_t lds = (size_t)cbr * Tile::lds_words_per_row(max_rw) * sizeof(float);
    if (grid <= 0) {
      // persistent: as many workgroups as can be resident at once
      int per_cu = 0, dev = 0, cus = 0;
      MDSM_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, batch_kernel<Tile>, 256, lds));
      MDSM_HIP(hipGetDevice(&dev));
      MDSM_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
      grid = std::max(1, per_cu) * std::max(1, cus);
    }
    // every one of the eight queues needs workers: a multiple of 8, at least 8
    grid = std::max(8, (grid + 7) / 8 * 8);
    hipLaunchKernelGGL(batch_kernel<Tile>, dim3(grid), dim3(256), lds, s, d_beams, nbeams,
                       total_items, cbr, d_counter);
    MDSM_HIP(hipGetLastError());
    return MDSM_OK;
  });
}

}  // namespace mdsm
