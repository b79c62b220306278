// internal.h -- shared declarations of libmdsm_amd (not installed).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <vector>

#include "mdsm_abi.h"

namespace mdsm {

void set_error(const char *fmt, ...) __attribute__((format(printf, 1, 2)));

#define MDSM_HIP(call)                                                                   \
  do {                                                                                   \
    hipError_t e_ = (call);                                                              \
    if (e_ != hipSuccess) {                                                              \
      ::mdsm::set_error("%s failed: %s (%s:%d)", #call, hipGetErrorString(e_), __FILE__, \
                        __LINE__);                                                       \
      return MDSM_ERR_HIP;                                                               \
    }                                                                                    \
  } while (0)

#define MDSM_REQUIRE(cond, code, ...) \
  do {                                \
    if (!(cond)) {                    \
      ::mdsm::set_error(__VA_ARGS__); \
      return code;                    \
    }                                 \
  } while (0)

inline hipStream_t as_stream(void *s) { return reinterpret_cast<hipStream_t>(s); }

// Input types: storage element of each (F16_PAIRS rows are F16 rows).
inline bool known_dtype(int dtype) { return dtype >= MDSM_F32 && dtype <= MDSM_F16_PAIRS; }
inline int storage_dtype(int dtype) { return dtype == MDSM_F16_PAIRS ? MDSM_F16 : dtype; }
inline size_t storage_bytes(int dtype) {
  const int st = storage_dtype(dtype);
  return st == MDSM_F32 ? 4u : (st == MDSM_F16 ? 2u : 1u);
}

// ---------------------------------------------------------------- tiles
// A dedispersion tile variant: every wave owns MW consecutive DM trials and
// K time samples per lane (lane + 64*j), a workgroup of 4 waves owns
// TD = 4*MW DM trials x TT = 64*K samples, and CB channels are staged in LDS
// per barrier.
struct TileShape {
  int mw, k, cb;
  int td() const { return 4 * mw; }
  int tt() const { return 64 * k; }
};

int num_variants(int dtype);
TileShape variant_shape(int dtype, int variant);

// Largest delay spread inside any DM tile of width td, over all channels:
// max_c max_{tile} D[c][min(d0+td,M)-1] - D[c][d0].
uint32_t tile_spread(const int32_t *table, uint32_t nchans, uint32_t ndms, int td);

// LDS words per staged channel row of a variant for a delay spread; 0 when
// the variant cannot stage that spread or (pair-sum tiles) cannot take the
// survey's delay table (nchans x ndms).
uint32_t lds_row_words(int dtype, int variant, uint32_t spread, const int32_t *table,
                       uint32_t nchans, uint32_t ndms);
// Required multiple (in elements) of the input row length: 1 (F32), 2 (F16, F16_PAIRS), 4 (U8).
uint32_t input_alignment(int dtype);

// Validates survey parameters and fills the host delay table.
int build_delay_table(const mdsm_survey_t *s, std::vector<int32_t> &table, uint32_t *maxshift);

// ---------------------------------------------------------------- kernels
// Untuned default tile variant per storage type (the fastest measured on a
// 1024-channel, 4096-trial beam).
int preferred_variant(int dtype);
// The shape an untuned plan takes when the preferred one does not take the
// survey (F16_PAIRS, U8: the fastest shape that takes every survey).
int fallback_variant(int dtype);
// Channels staged per LDS block for a tile variant and row width (0: does not fit).
uint32_t stage_channels(int dtype, int variant, uint32_t row_words);
// Words of a variant's trial-offset table, and the kernel that fills it
// from the delay table (the table depends on the row width and the block).
uint64_t trial_offset_words(int dtype, int variant, uint32_t nchans, uint32_t ndms, uint32_t cbr);
int launch_trial_offsets(int dtype, int variant, const int32_t *table, uint32_t nchans,
                         uint32_t ndms, uint32_t row_words, uint32_t cbr, int32_t *reltab,
                         hipStream_t stream);
int launch_dedisperse(int dtype, int variant, const void *in, uint64_t ld, const int32_t *table,
                      const int32_t *reltab, uint32_t nchans, uint32_t nsamp, uint32_t ndms,
                      uint32_t row_words, float *out, hipStream_t stream);

// Per-beam descriptor of the persistent batched dedisperser.  Buffers are
// given as byte offsets from the descriptor array itself: pointers derived
// from a kernel argument keep the global address space (global_load rather
// than flat_load, which would also count against the LDS wait counter).
struct BeamDesc {
  int64_t in_off;
  int64_t out_off;
  int64_t table_off;
  int64_t rel_off;  // trial-offset table of the bound variant
  uint64_t ld;
  uint32_t nchans, nsamp, ndms, row_words;
  uint32_t n_ttiles, n_dtiles;
  uint32_t item_begin;  // first work item of this beam in the queue order
  uint32_t tgroup;      // adjacent time tiles per XCD group (1: every 8th time tile)
};

int launch_batch_flat(int dtype, int variant, const BeamDesc *d_beams, uint32_t nbeams,
                      uint32_t total_items, uint32_t max_rw, hipStream_t s);
int launch_batch_dedisperse(int dtype, int variant, const BeamDesc *d_beams, uint32_t nbeams,
                            const uint32_t *d_order, uint32_t total_items, uint32_t *d_counter,
                            uint32_t max_row_words, int grid, hipStream_t stream);
size_t batch_lds_bytes(int dtype, int variant, uint32_t max_row_words);

int launch_channel_stats(const float *raw, uint64_t ld, uint32_t nchans, uint64_t n, double *mean,
                         double *stdv, hipStream_t s);
int launch_normalise(const float *raw, uint64_t ld, uint32_t nchans, uint64_t n,
                     const double *mean, const double *stdv, void *out, int out_dtype,
                     uint64_t ldo, hipStream_t s);
int launch_block_stats(const float *x, uint32_t nrows, uint32_t nsamp, uint32_t L, double *mean,
                       double *stdv, hipStream_t s);
int launch_detect(const float *x, uint32_t nrows, uint32_t nsamp, uint32_t L, const double *mean,
                  const double *stdv, int nwidths, double thresh, mdsm_cand_t *cands, uint32_t cap,
                  uint32_t *count, hipStream_t s);
// Reference-semantics global statistics: columns per chunk of the fixed
// reduction order, number of chunk partials for a region, and the two
// kernels (chunk sums into partials[2 * parts], then mean and rms into
// stats[2]); the threshold kernel reads mean and rms from stats.
constexpr uint32_t GS_COLS = 65536;
uint32_t global_stats_parts(uint32_t nrows, uint32_t ncols);
int launch_global_stats(const float *x, uint64_t ld, uint32_t nrows, uint32_t ncols,
                        double *partials, double *stats, hipStream_t s);
int launch_threshold_global(const float *x, uint64_t ld, uint32_t nrows, uint32_t ncols,
                            const double *stats, double mean, double rms, float thresh,
                            mdsm_cand_t *cands, uint32_t cap, uint32_t *count, hipStream_t s);
int launch_zero_u32(uint32_t *p, uint32_t n, hipStream_t s);
// The global detector's argument checks and kernels on a given workspace
// (partials at work[0 .. 2*parts), mean and rms after them).
int check_global_args(const float *d_x, uint32_t nrows, uint32_t nsamp, uint64_t ld,
                      uint32_t stats_nsamp, const mdsm_cand_t *d_cands, const uint32_t *d_count);
int run_global_detect(const float *d_x, uint32_t nrows, uint32_t nsamp, uint64_t ld,
                      uint32_t stats_nsamp, float thresh, mdsm_cand_t *d_cands, uint32_t cap,
                      uint32_t *d_count, double *h_mean, double *h_rms, double *work,
                      hipStream_t s);
int plan_time_variants(mdsm_plan plan, const void *d_in, uint64_t ld, float *d_out,
                       hipStream_t s, std::vector<float> &t);
int pick_variant(int dtype, const std::vector<float> &t);
int launch_median5(const float *x, uint32_t nrows, uint32_t nsamp, float *out, hipStream_t s);
int launch_legacy_block_sums(const float *x, uint32_t block_len, uint32_t nblocks, double *sums,
                             hipStream_t s);
int launch_legacy_threshold(const float *x, uint32_t ndms, uint32_t nsamp, float lowdm,
                            float dmstep, float thresh, float gmean, float grms, const float *cmean,
                            const float *crms, uint32_t spc, mdsm_cand_t *cands, uint32_t cap,
                            uint32_t *count, hipStream_t s);
int launch_detrend(float *x, uint64_t nseg, uint32_t L, hipStream_t s);
int launch_unpack_corner_turn(const void *in, int nbits, uint32_t nspec, uint32_t nchans,
                              uint32_t nbeams, float *out, uint64_t ld, uint64_t beam_stride,
                              uint64_t col0, hipStream_t s);
int launch_bandpass_sum(const float *x, uint64_t ld, uint32_t nchans, uint64_t col0, uint64_t n,
                        double *out, hipStream_t s);
int launch_channel_flag(const float *x, uint64_t ld, uint32_t nchans, uint64_t col0, uint64_t n,
                        const double *fitted, uint32_t cb, float thresh, uint8_t *flags,
                        hipStream_t s);
int launch_channel_clip(float *x, uint64_t ld, uint32_t nchans, uint64_t col0, uint64_t n,
                        const double *fitted, uint32_t cb, const uint8_t *flags, hipStream_t s);
int launch_spectrum_clip(float *x, uint64_t ld, uint32_t nchans, uint64_t col0, uint64_t n,
                         const double *fitted, float hi, float lo, uint32_t *count, hipStream_t s);
int launch_voltage_power(float *x, uint64_t ld, uint32_t nchans, uint64_t col0, uint64_t n,
                         hipStream_t s);
int launch_generate(uint64_t seed, uint32_t nchans, int64_t t0, uint64_t n, const float *gain,
                    const float *offset, uint32_t npulses, const int64_t *p_start,
                    const int32_t *p_width, const float *p_amp, const int64_t *p_delay, float *raw,
                    uint64_t ld, hipStream_t s);

}  // namespace mdsm
