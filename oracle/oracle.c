/*
 * oracle.c -- CPU restatement of MDSM's brute-force dedispersion and
 * single-pulse detection path.  TEST INFRASTRUCTURE ONLY.
 *
 * This file is the oracle the tests compare with.  Only the tests, the
 * smoke check and the CPU baseline leg of the benchmark may load it; the
 * product library never links
 * or calls it.  Everything here is plain C (gcc, OpenMP), written for clarity
 * and for bit-exact reproducibility, not for speed (except
 * oracle_dedisperse_f32seq, which doubles as the multithreaded CPU baseline).
 *
 * Build: oracle/Makefile (gcc -O2 -fopenmp -ffp-contract=off).  Floating
 * point contraction is disabled so that every expression rounds exactly as
 * written; the product kernels are compiled the same way for the
 * expressions that must match bit for bit (delay table, statistics, boxcar
 * tree, thresholds).
 *
 * Reference functions restated (paths relative to the MDSM tree):
 *   delay constant / dm_shifts  src/mdsm-multibeam/multibeam_dedispersion_manager.cpp:325-326,366-369
 *   maxshift / nsamp            src/mdsm-multibeam/multibeam_dedispersion_manager.cpp:329-347,371-387
 *   integer delay table         src/mdsm-multibeam/multibeam_dedispersion_thread.cu:546-559
 *   LDS halo (shared_shift)     src/mdsm-multibeam/multibeam_dedispersion_thread.cu:542-543
 *   brute-force sum             src/mdsm-multibeam/multibeam_dedispersion_kernel.cu:44-95
 *   global mean/rms + threshold src/mdsm-multibeam/multibeam_dedispersion_output.cpp:148-185,188-248
 *   DBSCAN (3-D box variant)    src/mdsm-multibeam/dbscan.cpp:194-276,390-415
 *   cluster classification      src/mdsm-multibeam/dbscan.cpp:25-171
 *   mu-law raw-data quantiser   src/mdsm-multibeam/multibeam_dedispersion_writer.cpp:21-118
 *   median of five (legacy)     src/mdsm/dedispersion_output.cpp:14-53,135-213
 *   detrend + normalise         src/mdsm-multibeam/multibeam_dedispersion_kernel.cu:207-295
 *   SIGPROC sample unpack       src/mdsm-multibeam/file_handler.cpp:193-287
 *   corner turn into beams      src/MDSM.cpp:71-104
 *   bandpass / RFI stage        src/mdsm-multibeam/multibeam_dedispersion_kernel.cu:302-450,
 *                               src/mdsm-multibeam/multibeam_dedispersion_thread.cu:224-390
 * Extensions beyond the reference restated from their own definitions (no reference
 * code exists for them): FP64 round() delay table with 4148.808, per-block
 * FP64 mean/RMS, dyadic boxcar detection, 2-D Chebyshev DBSCAN, FP16/uint8
 * input storage, Philox4x32-10 synthetic data.
 *
 * Pinning (what checks each function; see tests/test_oracle_*.py):
 *   philox4x32_10      Random123 known-answer vectors (Salmon et al. 2011)
 *   delay tables       closed-form special cases (DM=0, top channel, exact
 *                      integer configurations), monotonicity, +-0.5 bound
 *   dedisperse_*       known answers (in[c][t]=c, constant, impulse on a DM
 *                      track, DM 0 = channel sum), FP64 vs FP32 agreement,
 *                      numpy brute force on tiny inputs
 *   stats              numpy mean/std (tolerance) and exact sums on
 *                      integer-valued data (order independent)
 *   boxcar detect      numpy cumulative-sum boxcar (tolerance), exact on
 *                      integer data, injected pulse recovered at its DM/width
 *   dbscan2d           scikit-learn DBSCAN(metric="chebyshev")
 *   dbscan_ref         hand-derived cases from the reference algorithm's
 *                      text, scikit-learn DBSCAN where the neighbour graph
 *                      is a union of cliques, and on chained random-walk
 *                      neighbourhoods a closed form: each seed with enough
 *                      unvisited neighbours takes its connected component
 *                      among the unvisited points (scipy graph components)
 *                      (the reference needs Qt headers: unbuildable here)
 *   classify_ref       hand-derived cases (RFI and width rules, the
 *                      histogram smoothing and median), and the whole score
 *                      on clusters whose smoothed DM-S/N curve is the
 *                      Cordes & McLaughlin (2003) smearing response: 0 when
 *                      only the peak bin passes the 0.9 cut, (1/m-1)^2
 *                      sum(f^2)/N otherwise (the reference source needs
 *                      PGPLOT and Qt headers: not compiled here)
 *   median5            sorted median over random and exhaustive-order inputs
 *   legacy detector    the block-count quirk (last block left out), median
 *                      spikes and plateaus, DM > 1 rule, hand-derived
 *                      per-chunk statistics cases
 *   f16 conversion     numpy float16 (round to nearest even)
 *   dedisperse_f16_pairs  a numpy restatement that rounds the exact float64
 *                      pair sum straight to float16 (no float32 step),
 *                      ties-to-even known answers (2048 + 1 -> 2048,
 *                      1 + 2^-11 -> 1), odd channel counts, exact integer
 *                      sums
 *   bandpass / RFI     exact sums on integer data, float64 means within
 *                      float rounding, hand-built bursts and spikes whose
 *                      flagged blocks and clipped spectra are known; the
 *                      least-squares fit itself is numpy's SVD lstsq (the
 *                      method of the reference's gsl_multifit_linear)
 *   mu-law quantiser   closed forms (0 -> Q/2 level, the table's end -> top
 *                      level), monotonicity, sign bits of complex pairs
 *   sigproc unpack     the reference's read_block compiled from its own
 *                      source (oracle/_ref/libref_file.so) and the values
 *                      the fixture writer packed (tests/golden)
 */
#include <math.h>
#include <omp.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#define FLT_MAX_D 3.40282346638528859812e+38
#define FLT_MIN_D 1.17549435082228750797e-38
#define FLT_MAX_F 3.40282346638528859812e+38F
#define FLT_MIN_F 1.17549435082228750797e-38F
#define M_PI_D 3.14159265358979323846
#define KDM_FP64_ROUND 4148.808
#define KDM_REFERENCE 4148.741601
#define R256 256

/* ------------------------------------------------------------------ */
/* Philox4x32-10 (Salmon, Moraes, Dror, Shaw, SC'11), counter-based RNG */
/* ------------------------------------------------------------------ */
static inline void mulhilo32(uint32_t a, uint32_t b, uint32_t *hi, uint32_t *lo) {
  uint64_t p = (uint64_t)a * (uint64_t)b;
  *hi = (uint32_t)(p >> 32);
  *lo = (uint32_t)p;
}

void oracle_philox4x32_10(const uint32_t ctr_in[4], const uint32_t key_in[2], uint32_t out[4]) {
  uint32_t c0 = ctr_in[0], c1 = ctr_in[1], c2 = ctr_in[2], c3 = ctr_in[3];
  uint32_t k0 = key_in[0], k1 = key_in[1];
  for (int r = 0; r < 10; ++r) {
    uint32_t hi0, lo0, hi1, lo1;
    mulhilo32(0xD2511F53u, c0, &hi0, &lo0);
    mulhilo32(0xCD9E8D57u, c2, &hi1, &lo1);
    uint32_t n0 = hi1 ^ c1 ^ k0;
    uint32_t n1 = lo1;
    uint32_t n2 = hi0 ^ c3 ^ k1;
    uint32_t n3 = lo0;
    c0 = n0; c1 = n1; c2 = n2; c3 = n3;
    k0 += 0x9E3779B9u;
    k1 += 0xBB67AE85u;
  }
  out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

static inline double u24(uint32_t x) { return ((double)(x >> 8) + 0.5) * (1.0 / 16777216.0); }

/* ln x for normal x > 0 from IEEE +,-,*,/ only (the device generator's
 * det_log): x = m 2^e, m in [sqrt(1/2), sqrt(2)), ln x = e ln2 + 2 atanh(s),
 * s = (m-1)/(m+1), odd series of atanh to s^23, Horner from the top term. */
double oracle_det_log(double x) {
  uint64_t b;
  memcpy(&b, &x, 8);
  int e = (int)((b >> 52) & 0x7FFu) - 1023;
  uint64_t mb = (b & 0xFFFFFFFFFFFFFull) | 0x3FF0000000000000ull;
  double m;
  memcpy(&m, &mb, 8);
  if (m > 1.4142135623730951) { m = m * 0.5; e = e + 1; }
  double s = (m - 1.0) / (m + 1.0), s2 = s * s;
  double p = 1.0 / 23.0;
  for (int k = 21; k >= 3; k -= 2) p = p * s2 + 1.0 / (double)k;
  p = p * s2 + 1.0;
  return (double)e * 0.6931471805599453 + 2.0 * (s * p);
}

/* sin and cos of 2 pi u, 0 <= u <= 1 (the device's det_sincos_turn): the
 * quarter turn q = (int)(4u + 0.5) is split off, x = 2 pi (u - q/4), Taylor
 * series of sin to x^17 and cos to x^18 by Horner, then the quadrant. */
void oracle_det_sincos_turn(double u, double *sn, double *cs) {
  static const double fact[19] = {1.0, 1.0, 2.0, 6.0, 24.0, 120.0, 720.0, 5040.0, 40320.0,
                                  362880.0, 3628800.0, 39916800.0, 479001600.0, 6227020800.0,
                                  87178291200.0, 1307674368000.0, 20922789888000.0,
                                  355687428096000.0, 6402373705728000.0};
  int q = (int)(u * 4.0 + 0.5);
  double r = u - (double)q * 0.25;
  double x = r * 6.283185307179586, x2 = x * x;
  /* ps = 1/5! - x^2/7! + ... + x^12/17! */
  double ps = 1.0 / fact[17];
  for (int k = 15; k >= 5; k -= 2) ps = ((k / 2) % 2 == 0) ? ps * x2 + 1.0 / fact[k] : ps * x2 - 1.0 / fact[k];
  double s = (x - (x * x2) * (1.0 / 6.0)) + ((x * x2) * x2) * ps;
  /* pc = -1/6! + x^2/8! - ... - x^12/18! */
  double pc = -1.0 / fact[18];
  for (int k = 16; k >= 6; k -= 2) pc = ((k / 2) % 2 == 0) ? pc * x2 + 1.0 / fact[k] : pc * x2 - 1.0 / fact[k];
  double c = ((1.0 - x2 * 0.5) + (x2 * x2) * (1.0 / 24.0)) + ((x2 * x2) * x2) * pc;
  switch (q & 3) {
    case 0: *sn = s; *cs = c; break;
    case 1: *sn = c; *cs = -s; break;
    case 2: *sn = -s; *cs = -c; break;
    default: *sn = -c; *cs = s; break;
  }
}

/* Four N(0,1) draws for channel c, sample quad q (samples 4q..4q+3). */
static void gauss4(uint64_t seed, uint32_t c, uint64_t q, float z[4]) {
  uint32_t ctr[4] = {(uint32_t)q, (uint32_t)(q >> 32), c, 0u};
  uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
  uint32_t o[4];
  oracle_philox4x32_10(ctr, key, o);
  double r0 = sqrt(-2.0 * oracle_det_log(u24(o[0])));
  double r1 = sqrt(-2.0 * oracle_det_log(u24(o[2])));
  double s0, c0, s1, c1;
  oracle_det_sincos_turn(u24(o[1]), &s0, &c0);
  oracle_det_sincos_turn(u24(o[3]), &s1, &c1);
  z[0] = (float)(r0 * c0);
  z[1] = (float)(r0 * s0);
  z[2] = (float)(r1 * c1);
  z[3] = (float)(r1 * s1);
}

/* Per-channel bandpass gain and offset, drawn from stream 1. */
void oracle_channel_bandpass(uint64_t seed, int nchans, double g_lo, double g_hi, double o_lo,
                             double o_hi, float *gain, float *offset) {
  uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
  for (int c = 0; c < nchans; ++c) {
    uint32_t ctr[4] = {(uint32_t)c, 0u, 0u, 1u}, o[4];
    oracle_philox4x32_10(ctr, key, o);
    gain[c] = (float)(g_lo + (g_hi - g_lo) * u24(o[0]));
    offset[c] = (float)(o_lo + (o_hi - o_lo) * u24(o[1]));
  }
}

/* Delay of channel c for an arbitrary DM (4148.808, FP64, round). */
static inline int64_t ns_delay(double fch1, double foff, int c, double tsamp, double dm) {
  double f = fch1 + foff * (double)c;
  double k = 1.0 / (f * f) - 1.0 / (fch1 * fch1);
  return (int64_t)round(((KDM_FP64_ROUND * dm) * k) / tsamp);
}

/*
 * Synthetic filterbank block, layout [c][ld], samples t0..t0+n-1 of a stream.
 * raw = gain_c * (z + sum of pulse amplitudes covering (c,t)) + offset_c.
 * Pulses: arrays of dm, start sample, width, amplitude (normalised units);
 * a pulse covers [start + delay_c(dm), start + delay_c(dm) + width).
 */
void oracle_generate(uint64_t seed, int nchans, double fch1, double foff, double tsamp,
                     int64_t t0, int64_t n, const float *gain, const float *offset, int npulses,
                     const double *p_dm, const int64_t *p_start, const int32_t *p_width,
                     const float *p_amp, float *out, int64_t ld) {
#pragma omp parallel for schedule(static)
  for (int c = 0; c < nchans; ++c) {
    float *row = out + (int64_t)c * ld;
    float z[4];
    int64_t q_cached = -1;
    for (int64_t i = 0; i < n; ++i) {
      int64_t t = t0 + i;
      int64_t q = t >> 2;
      if (q != q_cached) { gauss4(seed, (uint32_t)c, (uint64_t)q, z); q_cached = q; }
      float v = z[t & 3];
      for (int p = 0; p < npulses; ++p) {
        int64_t s = p_start[p] + ns_delay(fch1, foff, c, tsamp, p_dm[p]);
        if (t >= s && t < s + p_width[p]) v = v + p_amp[p];
      }
      row[i] = gain[c] * v + offset[c];
    }
  }
}

/* ------------------------------------------------------------------ */
/* Delay tables                                                        */
/* ------------------------------------------------------------------ */
/* FP64 rounded table: D[c][d] = round(4148.808*DM_d*(f_c^-2 - f_top^-2)/tsamp), FP64,
 * f_c = fch1 + foff*c (channel 0 = top), DM_d = lowdm + d*dmstep. */
void oracle_delay_table(double fch1, double foff, int nchans, double tsamp, double lowdm,
                        double dmstep, int ndms, int32_t *table) {
  for (int c = 0; c < nchans; ++c) {
    double f = fch1 + foff * (double)c;
    double k = 1.0 / (f * f) - 1.0 / (fch1 * fch1);
    for (int d = 0; d < ndms; ++d) {
      double dm = lowdm + dmstep * (double)d;
      table[(int64_t)c * ndms + d] = (int32_t)round(((KDM_FP64_ROUND * dm) * k) / tsamp);
    }
  }
}

/* Reference dm_shifts (multibeam_dedispersion_manager.cpp:325-326,366-369):
 * float F1 = fch1 + foff*j (float arithmetic), dmdelay evaluated in double,
 * stored as float. */
void oracle_dm_shifts_ref(float fch1, float foff, int nchans, float *shifts) {
  for (int j = 0; j < nchans; ++j) {
    float F1 = fch1 + foff * (float)j;
    float F2 = fch1;
    shifts[j] = (float)(KDM_REFERENCE * ((1.0 / F1 / F1) - (1.0 / F2 / F2)));
  }
}

/* Reference delay table (multibeam_dedispersion_thread.cu:551-553):
 * (int)(dm_shifts[c] * (lowdm + d*dmstep) / tsamp), float arithmetic, truncation. */
void oracle_delay_table_ref(float fch1, float foff, int nchans, float tsamp, float lowdm,
                            float dmstep, int ndms, int32_t *table) {
  float *s = (float *)malloc(sizeof(float) * (size_t)nchans);
  oracle_dm_shifts_ref(fch1, foff, nchans, s);
  for (int c = 0; c < nchans; ++c)
    for (int d = 0; d < ndms; ++d) {
      float dm = lowdm + (float)d * dmstep;
      float v = s[c] * dm;
      v = v / tsamp;
      table[(int64_t)c * ndms + d] = (int32_t)v;
    }
  free(s);
}

/* Legacy single-beam brute-force delays (src/mdsm/dedispersion_kernel.cu:55,68):
 * shift_temp = (startdm + d*dmstep)/tsamp, delay = floor(dm_shifts[c]*shift_temp),
 * float throughout, with the same dm_shifts (src/mdsm/dedispersion_manager.cpp:262-265). */
void oracle_delay_table_legacy(float fch1, float foff, int nchans, float tsamp, float lowdm,
                               float dmstep, int ndms, int32_t *table) {
  float *s = (float *)malloc(sizeof(float) * (size_t)nchans);
  oracle_dm_shifts_ref(fch1, foff, nchans, s);
  for (int d = 0; d < ndms; ++d) {
    float shift_temp = (lowdm + (float)d * dmstep) / tsamp;
    for (int c = 0; c < nchans; ++c) table[(int64_t)c * ndms + d] = (int32_t)floorf(s[c] * shift_temp);
  }
  free(s);
}

/* Reference maxshift for a list of beams (manager.cpp:362-380), including the
 * round-up to a multiple of 256 that is applied once per beam inside the loop. */
uint32_t oracle_maxshift_ref(const float *fch1, const float *foff, int nbeams, int nchans,
                             float tsamp, float lowdm, float dmstep, int ndms) {
  uint32_t greatest = 0;
  float *s = (float *)malloc(sizeof(float) * (size_t)nchans);
  for (int b = 0; b < nbeams; ++b) {
    oracle_dm_shifts_ref(fch1[b], foff[b], nchans, s);
    float high_dm = lowdm + dmstep * (float)(ndms - 1);
    uint32_t ms = (uint32_t)(s[nchans - 1] * high_dm / tsamp);
    greatest = ms > greatest ? ms : greatest;
    greatest = (greatest / 256) * 256 + 256;
  }
  free(s);
  return greatest;
}

/* Reference LDS halo size (thread.cu:542-543), round() of float products. */
uint32_t oracle_shared_shift_ref(float bottom_shift, float lowdm, float dmstep, int ndms,
                                 float tsamp, int dedisp_dms) {
  double a = round(bottom_shift * (lowdm + (float)ndms * dmstep) / tsamp);
  double b = round(bottom_shift * (lowdm + (float)(ndms - dedisp_dms) * dmstep) / tsamp);
  return (uint32_t)(a - b);
}

/* Reference nsamp sizing (manager.cpp:329-347). */
uint32_t oracle_calculate_nsamp_ref(uint32_t nsamp, uint32_t maxshift, uint32_t nchans,
                                    uint32_t ndms, uint64_t memory_mb) {
  if (nsamp == 0)
    nsamp = (uint32_t)(((double)memory_mb * 1000 * 0.99 / sizeof(float) -
                        (double)maxshift * nchans) / (double)(nchans + ndms));
  if (maxshift > nsamp) nsamp = (uint32_t)pow(2, ceil(log((double)maxshift) / log(2)));
  return nsamp;
}

/* ------------------------------------------------------------------ */
/* Brute-force dedispersion                                            */
/* out[d][t] = sum_c in[c][t + D[c][d]]  (multibeam_dedispersion_kernel.cu:44-95) */
/* ------------------------------------------------------------------ */
void oracle_dedisperse_f64(const float *in, int64_t ld, int nchans, int nsamp, int ndms,
                           const int32_t *table, double *out) {
#pragma omp parallel for schedule(dynamic, 1)
  for (int d = 0; d < ndms; ++d) {
    double *o = out + (int64_t)d * nsamp;
    for (int t = 0; t < nsamp; ++t) o[t] = 0.0;
    for (int c = 0; c < nchans; ++c) {
      const float *r = in + (int64_t)c * ld + table[(int64_t)c * ndms + d];
      for (int t = 0; t < nsamp; ++t) o[t] += (double)r[t];
    }
  }
}

/* FP32 accumulation in channel order 0..C-1, one rounding per add: the
 * arithmetic of the reference kernel.  Multithreaded over (DM, time tile) so
 * it also serves as the CPU baseline. */
void oracle_dedisperse_f32seq(const float *in, int64_t ld, int nchans, int nsamp, int ndms,
                              const int32_t *table, float *out) {
  const int TT = 2048;
  int ntiles = (nsamp + TT - 1) / TT;
  int64_t work = (int64_t)ndms * ntiles;
#pragma omp parallel
  {
    float acc[2048] __attribute__((aligned(64)));
#pragma omp for schedule(dynamic, 4)
    for (int64_t w = 0; w < work; ++w) {
      int d = (int)(w / ntiles);
      int t0 = (int)(w % ntiles) * TT;
      int n = nsamp - t0 < TT ? nsamp - t0 : TT;
      for (int t = 0; t < n; ++t) acc[t] = 0.0f;
      for (int c = 0; c < nchans; ++c) {
        const float *r = in + (int64_t)c * ld + table[(int64_t)c * ndms + d] + t0;
#pragma omp simd
        for (int t = 0; t < n; ++t) acc[t] += r[t];
      }
      memcpy(out + (int64_t)d * nsamp + t0, acc, sizeof(float) * (size_t)n);
    }
  }
}

/* ------------------------------------------------------------------ */
/* IEEE binary16 conversion (round to nearest, ties to even)           */
/* ------------------------------------------------------------------ */
uint16_t oracle_f32_to_f16(float f) {
  uint32_t x;
  memcpy(&x, &f, 4);
  uint32_t sign = (x >> 16) & 0x8000u;
  uint32_t ax = x & 0x7FFFFFFFu;
  if (ax >= 0x7F800000u) return (uint16_t)(sign | (ax > 0x7F800000u ? 0x7E00u : 0x7C00u));
  if (ax >= 0x477FF000u) return (uint16_t)(sign | 0x7C00u); /* rounds to >= 65520: inf */
  if (ax < 0x38800000u) {                                   /* subnormal half or zero */
    if (ax < 0x33000000u) return (uint16_t)sign;            /* < 2^-25: rounds to 0 */
    uint32_t e = ax >> 23;
    uint32_t m = (ax & 0x7FFFFFu) | 0x800000u;
    /* value = m * 2^(e-150); in half subnormal units of 2^-24 that is m >> (126-e) */
    uint32_t shift = 126u - e;
    uint32_t rem = m & ((1u << shift) - 1u);
    uint32_t half = 1u << (shift - 1u);
    uint32_t q = m >> shift;
    if (rem > half || (rem == half && (q & 1u))) q += 1u;
    return (uint16_t)(sign | q);
  }
  uint32_t e = (ax >> 23) - 112u; /* rebias 127 -> 15 */
  uint32_t m = ax & 0x7FFFFFu;
  uint32_t h = (e << 10) | (m >> 13);
  uint32_t rem = m & 0x1FFFu;
  if (rem > 0x1000u || (rem == 0x1000u && (h & 1u))) h += 1u;
  return (uint16_t)(sign | h);
}

float oracle_f16_to_f32(uint16_t h) {
  uint32_t sign = ((uint32_t)h & 0x8000u) << 16;
  uint32_t e = (h >> 10) & 0x1Fu, m = h & 0x3FFu, x;
  if (e == 0) {
    if (m == 0) x = sign;
    else {
      float v = (float)m * (1.0f / 16777216.0f);
      memcpy(&x, &v, 4);
      x |= sign;
    }
  } else if (e == 31) x = sign | 0x7F800000u | (m << 13);
  else x = sign | ((e + 112u) << 23) | (m << 13);
  float f;
  memcpy(&f, &x, 4);
  return f;
}

void oracle_f32_to_f16_array(const float *in, uint16_t *out, int64_t n) {
  for (int64_t i = 0; i < n; ++i) out[i] = oracle_f32_to_f16(in[i]);
}

void oracle_f16_to_f32_array(const uint16_t *in, float *out, int64_t n) {
  for (int64_t i = 0; i < n; ++i) out[i] = oracle_f16_to_f32(in[i]);
}

/* FP16-stored input, FP32 accumulation in channel order. */
void oracle_dedisperse_f16(const uint16_t *in, int64_t ld, int nchans, int nsamp, int ndms,
                           const int32_t *table, float *out) {
#pragma omp parallel for schedule(dynamic, 1)
  for (int d = 0; d < ndms; ++d) {
    float *o = out + (int64_t)d * nsamp;
    for (int t = 0; t < nsamp; ++t) o[t] = 0.0f;
    for (int c = 0; c < nchans; ++c) {
      const uint16_t *r = in + (int64_t)c * ld + table[(int64_t)c * ndms + d];
      for (int t = 0; t < nsamp; ++t) o[t] += oracle_f16_to_f32(r[t]);
    }
  }
}

/* FP16-stored input summed in channel pairs (the library's MDSM_F16_PAIRS;
 * no reference counterpart: a storage-precision extension of the same sum,
 * multibeam_dedispersion_kernel.cu:44-95).  For k = 0, 1, ...:
 * s = fl16(h[2k] + h[2k+1]) (no partner past the last channel: s = h[2k]),
 * out = fl32(out + s).  The float sum of two halves rounds to 24 bits, and
 * 24 >= 2*11 + 2, so rounding it again to FP16 gives the correctly rounded
 * FP16 sum of the exact value (double rounding is innocuous at that
 * precision; sums in the FP16 subnormal range are exact in both). */
void oracle_dedisperse_f16_pairs(const uint16_t *in, int64_t ld, int nchans, int nsamp, int ndms,
                                 const int32_t *table, float *out) {
#pragma omp parallel for schedule(dynamic, 1)
  for (int d = 0; d < ndms; ++d) {
    float *o = out + (int64_t)d * nsamp;
    for (int t = 0; t < nsamp; ++t) o[t] = 0.0f;
    for (int c = 0; c < nchans; c += 2) {
      const uint16_t *r0 = in + (int64_t)c * ld + table[(int64_t)c * ndms + d];
      if (c + 1 < nchans) {
        const uint16_t *r1 = in + (int64_t)(c + 1) * ld + table[(int64_t)(c + 1) * ndms + d];
        for (int t = 0; t < nsamp; ++t) {
          const float f = oracle_f16_to_f32(r0[t]) + oracle_f16_to_f32(r1[t]);
          o[t] += oracle_f16_to_f32(oracle_f32_to_f16(f));
        }
      } else {
        for (int t = 0; t < nsamp; ++t) o[t] += oracle_f16_to_f32(r0[t]);
      }
    }
  }
}

/* uint8 input: value = u * (8/255) - 4 in normalised units.  The sum over
 * channels is an exact integer S; out = (float)(S * (8/255) - 4*C). */
#define U8_SCALE (8.0 / 255.0)
#define U8_OFFSET (-4.0)
void oracle_dedisperse_u8(const uint8_t *in, int64_t ld, int nchans, int nsamp, int ndms,
                          const int32_t *table, float *out) {
#pragma omp parallel for schedule(dynamic, 1)
  for (int d = 0; d < ndms; ++d) {
    uint32_t *s = (uint32_t *)calloc((size_t)nsamp, sizeof(uint32_t));
    for (int c = 0; c < nchans; ++c) {
      const uint8_t *r = in + (int64_t)c * ld + table[(int64_t)c * ndms + d];
      for (int t = 0; t < nsamp; ++t) s[t] += r[t];
    }
    for (int t = 0; t < nsamp; ++t)
      out[(int64_t)d * nsamp + t] = (float)((double)s[t] * U8_SCALE + (double)nchans * U8_OFFSET);
    free(s);
  }
}

/* ------------------------------------------------------------------ */
/* Reductions with a fixed order (the order the GPU kernels use):       */
/* lane i of 256 sums elements i, i+256, ... sequentially; then a binary */
/* tree p[i] += p[i+s] for s = 128, 64, ..., 1.                          */
/* ------------------------------------------------------------------ */
static void reduce_r256_f(const float *x, int64_t n, double *s1, double *s2) {
  double p1[R256], p2[R256];
  for (int i = 0; i < R256; ++i) {
    double a = 0.0, b = 0.0;
    for (int64_t j = i; j < n; j += R256) {
      double v = (double)x[j];
      a = a + v;
      b = b + v * v;
    }
    p1[i] = a;
    p2[i] = b;
  }
  for (int s = R256 / 2; s > 0; s >>= 1)
    for (int i = 0; i < s; ++i) {
      p1[i] = p1[i] + p1[i + s];
      p2[i] = p2[i] + p2[i + s];
    }
  *s1 = p1[0];
  *s2 = p2[0];
}

static void reduce_r256_d(const double *x, int64_t n, double *s1, double *s2) {
  double p1[R256], p2[R256];
  for (int i = 0; i < R256; ++i) {
    double a = 0.0, b = 0.0;
    for (int64_t j = i; j < n; j += R256) {
      a = a + x[j];
      b = b + x[j] * x[j];
    }
    p1[i] = a;
    p2[i] = b;
  }
  for (int s = R256 / 2; s > 0; s >>= 1)
    for (int i = 0; i < s; ++i) {
      p1[i] = p1[i] + p1[i + s];
      p2[i] = p2[i] + p2[i + s];
    }
  *s1 = p1[0];
  *s2 = p2[0];
}

static inline void moments(double s1, double s2, int64_t n, double *mean, double *stdv) {
  double m = s1 / (double)n;
  double v = s2 / (double)n - m * m;
  *mean = m;
  *stdv = v > 0.0 ? sqrt(v) : 0.0;
}

/* ------------------------------------------------------------------ */
/* Bandpass normalisation: per-channel FP64 mean/std over n samples,     */
/* then v = (x - mean) * (1/std) in FP64, stored as f32, f16 or u8.      */
/* ------------------------------------------------------------------ */
void oracle_channel_stats(const float *raw, int64_t ld, int nchans, int64_t n, double *mean,
                          double *stdv) {
#pragma omp parallel for schedule(dynamic, 1)
  for (int c = 0; c < nchans; ++c) {
    double s1, s2;
    reduce_r256_f(raw + (int64_t)c * ld, n, &s1, &s2);
    moments(s1, s2, n, &mean[c], &stdv[c]);
  }
}

static inline double norm_value(float x, double mean, double inv) {
  return ((double)x - mean) * inv;
}

static inline double inv_std(double s) { return s > 0.0 ? 1.0 / s : 0.0; }

void oracle_normalise_f32(const float *raw, int64_t ld, int nchans, int64_t n, const double *mean,
                          const double *stdv, float *out, int64_t ldo) {
#pragma omp parallel for schedule(static)
  for (int c = 0; c < nchans; ++c) {
    double inv = inv_std(stdv[c]);
    for (int64_t t = 0; t < n; ++t)
      out[(int64_t)c * ldo + t] = (float)norm_value(raw[(int64_t)c * ld + t], mean[c], inv);
  }
}

void oracle_normalise_f16(const float *raw, int64_t ld, int nchans, int64_t n, const double *mean,
                          const double *stdv, uint16_t *out, int64_t ldo) {
#pragma omp parallel for schedule(static)
  for (int c = 0; c < nchans; ++c) {
    double inv = inv_std(stdv[c]);
    for (int64_t t = 0; t < n; ++t)
      out[(int64_t)c * ldo + t] =
          oracle_f32_to_f16((float)norm_value(raw[(int64_t)c * ld + t], mean[c], inv));
  }
}

static inline uint8_t quant_u8(double v) {
  if (v < -4.0) v = -4.0;
  if (v > 4.0) v = 4.0;
  double q = floor((v + 4.0) * 31.875 + 0.5);
  return (uint8_t)q;
}

void oracle_normalise_u8(const float *raw, int64_t ld, int nchans, int64_t n, const double *mean,
                         const double *stdv, uint8_t *out, int64_t ldo) {
#pragma omp parallel for schedule(static)
  for (int c = 0; c < nchans; ++c) {
    double inv = inv_std(stdv[c]);
    for (int64_t t = 0; t < n; ++t)
      out[(int64_t)c * ldo + t] = quant_u8(norm_value(raw[(int64_t)c * ld + t], mean[c], inv));
  }
}

/* ------------------------------------------------------------------ */
/* Detection (own definition): per-block FP64 mean/RMS, dyadic boxcars, */
/* best width per sample, threshold.                                     */
/* ------------------------------------------------------------------ */
/* Block (d, b) covers samples [b*L, min((b+1)*L, nsamp)) of row d. */
void oracle_block_stats(const double *x, int nrows, int nsamp, int L, double *mean, double *stdv) {
  int nb = (nsamp + L - 1) / L;
#pragma omp parallel for schedule(dynamic, 1)
  for (int d = 0; d < nrows; ++d)
    for (int b = 0; b < nb; ++b) {
      int64_t t0 = (int64_t)b * L;
      int64_t n = nsamp - t0 < L ? nsamp - t0 : L;
      double s1, s2;
      reduce_r256_d(x + (int64_t)d * nsamp + t0, n, &s1, &s2);
      moments(s1, s2, n, &mean[(int64_t)d * nb + b], &stdv[(int64_t)d * nb + b]);
    }
}

/* sqrt(2^k) for k = 0..6, correctly rounded. */
static const double SQRT_W[7] = {1.0, 1.4142135623730951, 2.0, 2.8284271247461903,
                                 4.0, 5.656854249492381, 8.0};

/*
 * For each row d and sample t: S_0(t) = x(t), S_{k+1}(t) = S_k(t) + S_k(t + 2^k)
 * (FP64, this exact pairing), snr_k(t) = (S_k(t) - 2^k*mean)/(std*sqrt(2^k)) with
 * the stats of the block holding t, for every k with t + 2^k <= nsamp and
 * k < nwidths.  Best width = first k reaching the maximum.  A sample is a
 * candidate when the best snr >= thresh.  Candidates are written in (d, t)
 * order; returns the total count (entries beyond cap are not written).
 */
int64_t oracle_detect(const double *x, int nrows, int nsamp, int L, const double *mean,
                      const double *stdv, int nwidths, double thresh, int64_t cap, int32_t *cd,
                      int32_t *ct, int32_t *ck, float *csnr) {
  int nb = (nsamp + L - 1) / L;
  int64_t total = 0;
  double *lev = (double *)malloc(sizeof(double) * (size_t)nsamp);
  double *best = (double *)malloc(sizeof(double) * (size_t)nsamp);
  int *bestk = (int *)malloc(sizeof(int) * (size_t)nsamp);
  for (int d = 0; d < nrows; ++d) {
    const double *r = x + (int64_t)d * nsamp;
    for (int t = 0; t < nsamp; ++t) { lev[t] = r[t]; best[t] = -INFINITY; bestk[t] = -1; }
    for (int k = 0; k < nwidths; ++k) {
      int w = 1 << k;
      if (k > 0) {
        int h = w >> 1;
        for (int t = 0; t + w <= nsamp; ++t) lev[t] = lev[t] + lev[t + h];
      }
      for (int t = 0; t + w <= nsamp; ++t) {
        int b = t / L;
        double m = mean[(int64_t)d * nb + b], s = stdv[(int64_t)d * nb + b];
        if (!(s > 0.0)) continue;
        double snr = (lev[t] - (double)w * m) / (s * SQRT_W[k]);
        if (snr > best[t]) { best[t] = snr; bestk[t] = k; }
      }
    }
    for (int t = 0; t < nsamp; ++t)
      if (bestk[t] >= 0 && best[t] >= thresh) {
        if (total < cap) {
          cd[total] = d; ct[total] = t; ck[total] = bestk[t]; csnr[total] = (float)best[t];
        }
        total++;
      }
  }
  free(lev); free(best); free(bestk);
  return total;
}

/* ------------------------------------------------------------------ */
/* Reference global statistics + threshold (process_brute and            */
/* process_brute_clustering, multibeam_dedispersion_output.cpp:148-185,   */
/* 188-248): mean and rms in double over the first sc samples of each of  */
/* the nrows rows (sc = all samples except for the reference's last      */
/* buffer, :216-229), rms = sqrt(S2/n - mean^2); candidate if             */
/* x - mean >= rms*thresh over all samples, snr = (x - mean)/rms.         */
/* ------------------------------------------------------------------ */

/* The reference's own order: one sequential loop over the values
 * (row-major), S2 += pow(x, 2) (exact square of a float in double). */
void oracle_global_stats_seq(const float *x, int64_t ld, int nrows, int sc, double *mean,
                             double *rms) {
  double m = 0.0, s = 0.0;
  for (int r = 0; r < nrows; ++r)
    for (int j = 0; j < sc; ++j) {
      double v = (double)x[(int64_t)r * ld + j];
      m = m + v;
      s = s + v * v;
    }
  double n = (double)nrows * (double)sc;
  m = m / n;
  *mean = m;
  *rms = sqrt(s / n - m * m);
}

/* The product's fixed order (mdsm_detect_global): chunk g = r*nb + b holds
 * samples [b*65536, min((b+1)*65536, sc)) of row r; each chunk is reduced
 * with the r256 pattern (lane i sums elements i, i+256, ... in order, then
 * the binary tree s = 128..1), and the chunk partials, in chunk order, are
 * reduced with the same pattern. */
#define GS_COLS 65536
static void r256_pair(const double *a, const double *b, int64_t n, int64_t stride, double *s1,
                      double *s2) {
  double p1[R256], p2[R256];
  for (int i = 0; i < R256; ++i) {
    double x = 0.0, y = 0.0;
    for (int64_t j = i; j < n; j += R256) {
      x = x + a[j * stride];
      y = y + b[j * stride];
    }
    p1[i] = x;
    p2[i] = y;
  }
  for (int s = R256 / 2; s > 0; s >>= 1)
    for (int i = 0; i < s; ++i) {
      p1[i] = p1[i] + p1[i + s];
      p2[i] = p2[i] + p2[i + s];
    }
  *s1 = p1[0];
  *s2 = p2[0];
}

void oracle_global_stats(const float *x, int64_t ld, int nrows, int sc, double *mean,
                         double *rms) {
  int64_t nb = ((int64_t)sc + GS_COLS - 1) / GS_COLS, np = (int64_t)nrows * nb;
  double *part = (double *)malloc(sizeof(double) * 2 * (size_t)np);
#pragma omp parallel for schedule(dynamic, 16)
  for (int64_t g = 0; g < np; ++g) {
    int64_t r = g / nb, b = g % nb, c0 = b * GS_COLS;
    int64_t n = sc - c0 < GS_COLS ? sc - c0 : GS_COLS;
    reduce_r256_f(x + r * ld + c0, n, &part[2 * g], &part[2 * g + 1]);
  }
  double s1, s2;
  r256_pair(part, part + 1, np, 2, &s1, &s2);
  free(part);
  double n = (double)nrows * (double)sc;
  double m = s1 / n;
  *mean = m;
  *rms = sqrt(s2 / n - m * m);
}

int64_t oracle_threshold_ref(const float *x, int64_t ld, int nrows, int nsamp, double mean,
                             double rms, float thresh, int64_t cap, int32_t *cd, int32_t *ct,
                             float *csnr) {
  int64_t total = 0;
  for (int k = 0; k < nrows; ++k)
    for (int l = 0; l < nsamp; ++l) {
      float v = x[(int64_t)k * ld + l];
      if (v - mean >= rms * thresh) {
        if (total < cap) { cd[total] = k; ct[total] = l; csnr[total] = (float)((v - mean) / rms); }
        total++;
      }
    }
  return total;
}

/* ------------------------------------------------------------------ */
/* DBSCAN, 2-D variant: points (d, t), Chebyshev distance,              */
/* neighbour iff |dd| <= eps and |dt| <= eps, core iff the neighbourhood  */
/* (self included) has >= minpts points.  Clusters numbered 0.. in order  */
/* of their first core point; a border point joins the lowest-numbered    */
/* cluster with a core point in reach; noise = -1.                        */
/* ------------------------------------------------------------------ */
int oracle_dbscan2d(const int32_t *d, const int32_t *t, int n, int eps, int minpts, int32_t *label) {
  char *core = (char *)calloc((size_t)n, 1);
  int *queue = (int *)malloc(sizeof(int) * (size_t)(n > 0 ? n : 1));
  for (int i = 0; i < n; ++i) {
    int cnt = 0;
    for (int j = 0; j < n; ++j)
      if (abs(d[i] - d[j]) <= eps && abs(t[i] - t[j]) <= eps) cnt++;
    core[i] = cnt >= minpts;
    label[i] = -1;
  }
  int nclusters = 0;
  for (int i = 0; i < n; ++i) {
    if (!core[i] || label[i] != -1) continue;
    int id = nclusters++;
    int qh = 0, qt = 0;
    label[i] = id;
    queue[qt++] = i;
    while (qh < qt) {
      int p = queue[qh++];
      for (int j = 0; j < n; ++j) {
        if (abs(d[p] - d[j]) > eps || abs(t[p] - t[j]) > eps) continue;
        if (label[j] != -1) continue;
        label[j] = id;
        if (core[j]) queue[qt++] = j;
      }
    }
  }
  free(core);
  free(queue);
  return nclusters;
}

/* ------------------------------------------------------------------ */
/* DBSCAN, reference 3-D box variant (dbscan.cpp:194-276, 390-415).      */
/* cluster: in/out, 0 = unassigned on entry; ids from 1; noise -1.        */
/* ------------------------------------------------------------------ */
static unsigned ref_neighbours(const double *tm, const double *dm, const double *snr,
                               const int32_t *cluster, unsigned n, unsigned index, char *nb,
                               const char *visited, float min_time, float min_dm, float min_snr) {
  unsigned count = 0;
  for (unsigned i = 0; i < n; ++i) {
    if (index != i && cluster[i] <= 0) {
      float x = (float)fabs(tm[index] - tm[i]);
      float y = (float)fabs(dm[index] - dm[i]);
      float z = (float)fabs(snr[index] - snr[i]);
      if (x < min_time && y < min_dm && z < min_snr && nb[i] == 0 && visited[i] == 0) {
        nb[i] = 1;
        count++;
      }
    }
  }
  return count;
}

int oracle_dbscan_ref(const double *tm, const double *dm, const double *snr, unsigned n,
                      float min_time, float min_dm, float min_snr, unsigned min_points,
                      int32_t *cluster) {
  char *nb = (char *)calloc(n ? n : 1, 1);
  char *visited = (char *)calloc(n ? n : 1, 1);
  int nclusters = 0;
  for (unsigned i = 0; i < n; ++i) cluster[i] = 0;
  for (unsigned i = 0; i < n; ++i) {
    if (visited[i]) continue;
    visited[i] = 1;
    unsigned found = ref_neighbours(tm, dm, snr, cluster, n, i, nb, visited, min_time, min_dm, min_snr);
    if (found < min_points) {
      cluster[i] = -1;
      memset(nb, 0, n);
      continue;
    }
    int id = ++nclusters;
    cluster[i] = id;
    while (found != 0) {
      unsigned j = 0;
      while (nb[j] == 0) {
        j++;
        if (j >= n) j = 0;
      }
      if (!visited[j]) {
        visited[j] = 1;
        found += ref_neighbours(tm, dm, snr, cluster, n, j, nb, visited, min_time, min_dm, min_snr);
      }
      if (cluster[j] == 0) cluster[j] = id;
      nb[j] = 0;
      found--;
    }
  }
  free(nb);
  free(visited);
  return nclusters;
}

/* ------------------------------------------------------------------ */
/* Median of five, restating the selection network of                    */
/* src/mdsm/dedispersion_output.cpp:14-53 operation for operation.        */
/* ------------------------------------------------------------------ */
float oracle_median5(float n1, float n2, float n3, float n4, float n5) {
  float *a = &n1, *b = &n2, *c = &n3, *d = &n4, *e = &n5, *tmp;
  if (*b < *a) { tmp = a; a = b; b = tmp; }
  if (*d < *c) { tmp = c; c = d; d = tmp; }
  if (*c < *a) { tmp = b; b = d; d = tmp; c = a; }
  a = e;
  if (*b < *a) { tmp = a; a = b; b = tmp; }
  if (*a < *c) { tmp = b; b = d; d = tmp; a = c; }
  if (*d < *a) return *d;
  return *a;
}

/* 5-tap running median per DM row, out-of-place, edge samples replicated
 * (the intent of median_filter, multibeam_dedispersion_kernel.cu:98-157). */
void oracle_median_filter5(const float *x, int nrows, int nsamp, float *out) {
  for (int d = 0; d < nrows; ++d) {
    const float *r = x + (int64_t)d * nsamp;
    for (int t = 0; t < nsamp; ++t) {
      float v[5];
      for (int k = -2; k <= 2; ++k) {
        int i = t + k;
        i = i < 0 ? 0 : (i >= nsamp ? nsamp - 1 : i);
        v[k + 2] = r[i];
      }
      /* insertion sort, take the middle */
      for (int i = 1; i < 5; ++i)
        for (int j = i; j > 0 && v[j] < v[j - 1]; --j) { float s = v[j]; v[j] = v[j - 1]; v[j - 1] = s; }
      out[(int64_t)d * nsamp + t] = v[2];
    }
  }
}

/* Legacy detector (dedispersion_output.cpp:135-213, mean == NULL branch):
 * mean/rms from 32K-value blocks, median of five, threshold with thisdm > 1. */
int64_t oracle_process_brute_legacy(const float *x, int ndms, int nsamp, float lowdm, float dmstep,
                                    float thresh, int64_t nvals, int64_t cap, int32_t *cd,
                                    int32_t *ct, float *csnr, float *out_mean, float *out_rms) {
  /* vals = read_nsamp * tdms (:147): the first nvals values of the row-major
   * array enter the statistics (all of them when nvals is 0) */
  unsigned vals = nvals ? (unsigned)nvals : (unsigned)nsamp * (unsigned)ndms;
  unsigned mod_factor = vals < 32 * 1024 ? vals : 32 * 1024;
  unsigned iters = 0, j;
  double mean2 = 0, total, total2;
  float localmean = 0, localrms;
  while (1) {
    total = 0;
    total2 = 0;
    for (j = 0; j < mod_factor; j++) {
      total += x[iters * mod_factor + j];
      total2 += pow(x[iters * mod_factor + j], 2);
    }
    localmean = (float)((double)localmean + total / j);
    mean2 += (total2 / j);
    iters++;
    if (iters * mod_factor + j >= vals) break;
  }
  localmean /= (float)iters;
  float lm2 = localmean * localmean;
  localrms = (float)sqrt(mean2 / iters - (double)lm2);
  *out_mean = localmean;
  *out_rms = localrms;
  int64_t total_c = 0;
  for (int k = 0; k < ndms; ++k) {
    const float *r = x + (int64_t)k * nsamp;
    for (int l = 0; l < nsamp - 4; ++l) {
      float med = oracle_median5(r[l], r[l + 1], r[l + 2], r[l + 3], r[l + 4]);
      float thisdm = lowdm + k * dmstep;
      if (med - localmean >= localrms * thresh && thisdm > 1.0) {
        if (total_c < cap) { cd[total_c] = k; ct[total_c] = l; csnr[total_c] = med / localrms; }
        total_c++;
      }
    }
  }
  return total_c;
}

/* Legacy detector, branch with caller-supplied statistics
 * (dedispersion_output.cpp:198-209, mean != NULL): sample l of every row is
 * judged against mean[l / spc] and rms[l / spc] (per-chunk noise estimates),
 * snr = (median - mean) / rms.  Same median of five and DM > 1 rule. */
int64_t oracle_process_brute_legacy_chunks(const float *x, int ndms, int nsamp, float lowdm,
                                           float dmstep, float thresh, const float *mean,
                                           const float *rms, int spc, int64_t cap, int32_t *cd,
                                           int32_t *ct, float *csnr) {
  int64_t total_c = 0;
  for (int k = 0; k < ndms; ++k) {
    const float *r = x + (int64_t)k * nsamp;
    for (int l = 0; l < nsamp - 4; ++l) {
      float med = oracle_median5(r[l], r[l + 1], r[l + 2], r[l + 3], r[l + 4]);
      float thisdm = lowdm + k * dmstep;
      float m = mean[l / spc], s = rms[l / spc];
      if (med - m >= s * thresh && thisdm > 1.0) {
        if (total_c < cap) { cd[total_c] = k; ct[total_c] = l; csnr[total_c] = (med - m) / s; }
        total_c++;
      }
    }
  }
  return total_c;
}

/* Linear detrend and normalisation of consecutive segments of length L
 * (detrend_normalise, multibeam_dedispersion_kernel.cu:207-295), FP64
 * accumulation: x_i = -L/2 + 0.5 + i; fit y = a + b x by least squares
 * (a = mean(y), b = sum(xy)/sum(xx)); subtract; divide by sqrt(mean(r^2)). */
void oracle_detrend_normalise(const float *x, int64_t nseg, int L, float *out) {
  for (int64_t s = 0; s < nseg; ++s) {
    const float *y = x + s * L;
    double sy = 0, sxy = 0, sxx = 0;
    for (int i = 0; i < L; ++i) {
      double xi = -L / 2.0 + 0.5 + i;
      sy += y[i];
      sxy += xi * y[i];
      sxx += xi * xi;
    }
    double a = sy / L, b = sxy / sxx, ss = 0;
    for (int i = 0; i < L; ++i) {
      double xi = -L / 2.0 + 0.5 + i;
      double r = (double)y[i] - (a + b * xi);
      ss += r * r;
    }
    double sd = sqrt(ss / L);
    for (int i = 0; i < L; ++i) {
      double xi = -L / 2.0 + 0.5 + i;
      out[s * L + i] = (float)(((double)y[i] - (a + b * xi)) / sd);
    }
  }
}

int oracle_num_threads(void) { return omp_get_max_threads(); }

/* SIGPROC sample unpack (read_block, file_handler.cpp:193-287) over a
 * contiguous stream of n samples: 1, 2 and 4-bit fields are taken from the
 * least significant bits of each byte upward, 8 and 16-bit samples are
 * unsigned, 32-bit samples are IEEE floats. */
void oracle_sigproc_unpack(const uint8_t *bytes, int nbits, int64_t n, float *out) {
  for (int64_t i = 0; i < n; ++i) {
    if (nbits < 8) {
      int64_t bit = i * nbits;
      unsigned v = (bytes[bit >> 3] >> (bit & 7)) & ((1u << nbits) - 1u);
      out[i] = (float)v;
    } else if (nbits == 8) {
      out[i] = (float)bytes[i];
    } else if (nbits == 16) {
      uint16_t v;
      memcpy(&v, bytes + 2 * i, 2);
      out[i] = (float)v;
    } else {
      memcpy(out + i, bytes + 4 * i, 4);
    }
  }
}

/* Corner turn of time-major spectra [t][nbeams*nchans] into beam rows
 * (readBinaryData, src/MDSM.cpp:71-104): out[b*beam_stride + c*ld + col0 + t]
 * = spectra[t*nbeams*nchans + b*nchans + c]. */
void oracle_corner_turn(const float *spectra, int64_t nspectra, int nchans, int nbeams, float *out,
                        int64_t ld, int64_t beam_stride, int64_t col0) {
  const int64_t ct = (int64_t)nchans * nbeams;
  for (int64_t t = 0; t < nspectra; ++t)
    for (int b = 0; b < nbeams; ++b)
      for (int c = 0; c < nchans; ++c)
        out[b * beam_stride + c * ld + col0 + t] = spectra[t * ct + (int64_t)b * nchans + c];
}

/* ------------------------------------------------------------------ */
/* Bandpass and RFI stage.  Rows are [nchans][ld]; every function works on
 * the n samples starting at column col0 of each row (the reference's
 * `shift`, with `total` = ld).  The reduction orders are those of the
 * reference kernels with BANDPASS_THREADS = 64 (cache_brute_force.h:16). */
#define BP_T 64

/* Per-channel mean: lane i sums samples i, i+64, ... in double, then the
 * halving tree i = 32..1 (bandpass_power_sum, kernel.cu:317-350). */
void oracle_bandpass_sum(const float *x, int64_t ld, int nchans, int64_t col0, int64_t n,
                         double *bandpass) {
  for (int c = 0; c < nchans; ++c) {
    double p[BP_T];
    for (int i = 0; i < BP_T; ++i) {
      p[i] = 0.0;
      for (int64_t s = i; s < n; s += BP_T) p[i] += x[c * ld + col0 + s];
    }
    for (int w = BP_T / 2; w >= 1; w /= 2)
      for (int i = 0; i < w; ++i) p[i] += p[i + w];
    bandpass[c] = p[0] / (double)n;
  }
}

/* Channel masks (thread.cu:245-258): channels from..to take the mean of
 * the two neighbouring values, computed in double and stored through a
 * float. */
void oracle_bandpass_mask(double *bp, int nchans, const int *from, const int *to, int nmasks) {
  for (int m = 0; m < nmasks; ++m) {
    int a = from[m], b = to[m];
    float value;
    if (a == 0)
      value = (float)((bp[b + 1] + bp[b + 2]) * 0.5);
    else if (b >= nchans - 1)
      value = (float)((bp[a - 1] + bp[a - 2]) * 0.5);
    else
      value = (float)((bp[a - 1] + bp[b + 1]) * 0.5);
    for (int i = a; i <= b; ++i) bp[i] = value;
  }
}

/* Polynomial evaluation at x_i = i/nchans (thread.cu:241,264-270):
 * fitted[i] = sum_j coeffs[j] * pow(x_i, j), accumulated in j order. */
void oracle_poly_eval(const double *coeffs, int ncoeffs, int nchans, double *fitted) {
  for (int i = 0; i < nchans; ++i) {
    double xi = 0 + i / (1.0 * nchans), acc = 0.0;
    for (int j = 0; j < ncoeffs; ++j) acc += coeffs[j] * pow(xi, j);
    fitted[i] = acc;
  }
}

/* Bandpass statistics and RFI thresholds (thread.cu:276-296, 344-346):
 * float accumulators receiving double terms, as the reference's float
 * SURVEY fields do.  out = {mean, std, rmse, channel_thresh,
 * spectrum_hi, spectrum_lo}. */
void oracle_bandpass_stats(const double *bp, const double *fitted, int nchans,
                           float channel_thresh, float spectrum_thresh, float *out) {
  float mean = 0, stdv = 0, rmse = 0;
  for (int i = 0; i < nchans; ++i) mean = (float)(mean + bp[i] / nchans);
  for (int i = 0; i < nchans; ++i) stdv = (float)(stdv + (bp[i] - mean) * (bp[i] - mean));
  stdv = sqrtf(stdv / (float)nchans);
  for (int i = 0; i < nchans; ++i) rmse = (float)(rmse + pow(bp[i] - fitted[i], 2));
  rmse = sqrtf(rmse / (float)nchans);
  const float prod_hi = spectrum_thresh * stdv; /* rounded before the add */
  out[0] = mean;
  out[1] = stdv;
  out[2] = rmse;
  out[3] = channel_thresh * rmse;
  out[4] = mean + prod_hi;
  out[5] = mean - prod_hi;
}

/* float(double) rounded toward zero (__double2float_rz). */
static float f64_to_f32_rz(double d) {
  float f = (float)d;
  if (fabs((double)f) > fabs(d)) f = nextafterf(f, 0.0f);
  return f;
}

float oracle_f64_to_f32_rz(double d) { return f64_to_f32_rz(d); }

/* Channel-block flags (channel_flagger, kernel.cu:353-398): the mean of
 * block b of channel c (lane-strided float sums, halving tree, / cb) above
 * rz(fitted[c]) + thresh flags blocks b, b-1 (b > 0) and b+1 (b + 2 <
 * nblocks, the reference's `b < num_blocks - 2`).  Blocks cover
 * [b*cb, min((b+1)*cb, n)) and a short last block is averaged over its own
 * samples (the reference reads past the chunk there).  flags:
 * [nchans][nblocks], cleared first. */
void oracle_channel_flagger(const float *x, int64_t ld, int nchans, int64_t col0, int64_t n,
                            const double *fitted, int cb, float thresh, uint8_t *flags) {
  const int64_t nb = (n + cb - 1) / cb;
  memset(flags, 0, (size_t)(nchans * nb));
  for (int c = 0; c < nchans; ++c) {
    const float bp = f64_to_f32_rz(fitted[c]);
    for (int64_t b = 0; b < nb; ++b) {
      float p[BP_T];
      const int64_t len = (b + 1) * cb <= n ? cb : n - b * cb;
      for (int i = 0; i < BP_T; ++i) {
        p[i] = 0.0f;
        for (int64_t s = i; s < len; s += BP_T) p[i] += x[c * ld + col0 + b * cb + s];
      }
      for (int w = BP_T / 2; w >= 1; w /= 2)
        for (int i = 0; i < w; ++i) p[i] += p[i + w];
      if (p[0] / (float)len > bp + thresh) {
        flags[c * nb + b] = 1;
        if (b > 0) flags[c * nb + b - 1] = 1;
        if (b + 2 < nb) flags[c * nb + b + 1] = 1;
      }
    }
  }
}

/* Flagged blocks take rz(fitted[c]) (channel_clipper, kernel.cu:402-423). */
void oracle_channel_clipper(float *x, int64_t ld, int nchans, int64_t col0, int64_t n,
                            const double *fitted, int cb, const uint8_t *flags) {
  const int64_t nb = (n + cb - 1) / cb;
  for (int c = 0; c < nchans; ++c) {
    const float bp = f64_to_f32_rz(fitted[c]);
    for (int64_t b = 0; b < nb; ++b) {
      if (!flags[c * nb + b]) continue;
      for (int64_t s = b * cb; s < (b + 1) * cb && s < n; ++s) x[c * ld + col0 + s] = bp;
    }
  }
}

/* Spectra whose channel mean (sequential float sum / nchans) lies above hi
 * or below lo are replaced by rz(fitted[c]) (spectrum_clipper,
 * kernel.cu:427-450).  Returns the number of clipped spectra. */
int64_t oracle_spectrum_clipper(float *x, int64_t ld, int nchans, int64_t col0, int64_t n,
                                const double *fitted, float hi, float lo) {
  int64_t clipped = 0;
  for (int64_t s = 0; s < n; ++s) {
    float m = 0.0f;
    for (int c = 0; c < nchans; ++c) m += x[c * ld + col0 + s];
    m /= (float)nchans;
    if (m > hi || m < lo) {
      for (int c = 0; c < nchans; ++c) x[c * ld + col0 + s] = f64_to_f32_rz(fitted[c]);
      ++clipped;
    }
  }
  return clipped;
}

/* Complex 16-bit voltages (re, im) stored in each 32-bit slot become their
 * power re^2 + im^2 (voltage_to_power, kernel.cu:302-312). */
void oracle_voltage_to_power(float *x, int64_t ld, int nchans, int64_t col0, int64_t n) {
  for (int c = 0; c < nchans; ++c)
    for (int64_t s = 0; s < n; ++s) {
      int16_t v[2];
      memcpy(v, x + c * ld + col0 + s, 4);
      x[c * ld + col0 + s] = (float)(v[0] * v[0] + v[1] * v[1]);
    }
}

/* Cluster classification (Cluster::computeTransientProbability,
 * dbscan.cpp:25-171) over the n member points of one cluster.  The float /
 * double mix of the reference is kept: a float DM-S/N histogram, a 3-tap
 * smoothing that reads (h[i] + h[i] + h[i+1]) / 3, the median DM of the
 * bins above 0.9 of the maximum, the width from the time extent of points
 * within one DM step (initial extremes FLT_MAX / FLT_MIN), and the mean
 * squared difference from the model sqrt(pi)/2 erf(y)/y.  Returns 1 for a
 * cluster rejected by a rule (DM <= 2, bad width, model out of range),
 * else the MSE.  info = {maxDM, maxSNR, width, position}. */
float oracle_classify_ref(const double *tm, const double *dm, const double *snr, int n, float minDm,
                          float dmStep, unsigned numDMs, float min_pulse_width, float fch1,
                          float foff, unsigned nchans, double *info) {
  float *h = (float *)calloc(numDMs, sizeof(float));
  float *tmp = (float *)calloc(numDMs, sizeof(float));
  float *fit = (float *)calloc(numDMs, sizeof(float));
  int *lst = (int *)malloc(sizeof(int) * (numDMs ? numDMs : 1));
  float result = 1.0f, maxDM = 0, maxSNR = 0, width = 0;
  double position = 0;
  for (int i = 0; i < n; ++i) h[(int)round((dm[i] - minDm) / dmStep)] = (float)(h[(int)round((dm[i] - minDm) / dmStep)] + snr[i]);
  for (unsigned i = 1; i + 1 < numDMs; ++i) tmp[i] = (float)((h[i] + h[i] + h[i + 1]) / 3.0);
  for (unsigned i = 1; i + 1 < numDMs; ++i) h[i] = tmp[i];
  float mx = 0;
  for (unsigned i = 0; i < numDMs; ++i) mx = h[i] > mx ? h[i] : mx;
  const float inv = (float)(1.0 / mx);
  int size = 0;
  mx = 0;
  for (unsigned i = 0; i < numDMs; ++i)
    if (h[i] * inv > 0.9) {
      lst[size++] = (int)i;
      mx += h[i];
    }
  mx /= (float)size;
  if (size % 2 == 0)
    maxDM = (float)((lst[size / 2 - 1] + lst[size / 2]) * 0.5 * dmStep + minDm);
  else
    maxDM = (float)lst[size / 2] * dmStep + minDm;
  if (maxDM <= 2.0) goto done;
  for (int i = 0; i < n; ++i)
    if (fabs(dm[i] - maxDM) < dmStep * 0.5 && snr[i] > maxSNR) maxSNR = (float)snr[i];
  {
    double minT = FLT_MAX_D, maxT = FLT_MIN_D;
    for (int i = 0; i < n; ++i)
      if (fabs(dm[i] - maxDM) < dmStep) {
        minT = tm[i] < minT ? tm[i] : minT;
        maxT = tm[i] > maxT ? tm[i] : maxT;
      }
    width = (float)((maxT - minT) * 1e3 * 0.4);
    position = minT + (maxT - minT) / 2;
  }
  if (!(width <= FLT_MAX_F && width >= FLT_MIN_F && width > min_pulse_width)) goto done;
  {
    const float band = fabsf(foff * (float)nchans);
    const float freq = (float)pow(1e-3 * fch1 - (band * 1e-3) * 0.5, 3);
    float fitMax = 0;
    for (unsigned i = 0; i < numDMs; ++i) {
      const float x = (float)i * dmStep + minDm - maxDM;
      const float y = (float)(6.91e-3 * x * (band / (width * freq)));
      fit[i] = (float)(sqrt(M_PI_D) * 0.5 * (1.0 / y) * erff(y));
      fitMax = fitMax > fit[i] ? fitMax : fit[i];
    }
    fit[(int)roundf((maxDM - minDm) / dmStep)] = 1;
    if ((fitMax - 1.0) > 0.1) goto done;
    for (unsigned i = 0; i < numDMs; ++i) h[i] /= mx;
    float mse = 0;
    for (unsigned i = 0; i < numDMs; ++i) mse += (h[i] - fit[i]) * (h[i] - fit[i]);
    mse /= (float)numDMs;
    result = mse;
  }
done:
  if (info) {
    info[0] = maxDM;
    info[1] = maxSNR;
    info[2] = width;
    info[3] = position;
  }
  free(h);
  free(tmp);
  free(fit);
  free(lst);
  return result;
}

/* Mu-law quantiser (quantise_power_data / quantise_complex_data,
 * multibeam_dedispersion_writer.cpp:21-118, LOG_LOOKUP_LENGTH 32768):
 * table log10(1 + i*mu/32768) in float; power values scaled by 32768/max to
 * a table index (clamped to the table; the reference reads past it),
 * quantised as ((int)(d/Q)*Q + Q/2) * 2^bits and stored through int to a
 * byte; complex pairs as [sign_re, 3 bits, sign_im, 3 bits]. */
static void mulaw_table(unsigned mu, float *lut) {
  for (unsigned i = 0; i < 32768; ++i) lut[i] = log10f(1.0f + (float)(i * mu) / 32768.0f);
}

void oracle_mulaw_power(const float *x, int64_t n, unsigned bits, unsigned mu, float max_value,
                        uint8_t *out) {
  float *lut = (float *)malloc(32768 * sizeof(float));
  mulaw_table(mu, lut);
  const float Q = (float)(1.0 / pow(2, bits)), invQ = (float)(1.0 / Q);
  const float l1pmu = (float)(1.0 / log10(1 + mu)), inv_max = (float)(1.0 / max_value);
  const unsigned bitrange = (unsigned)pow(2, bits);
  for (int64_t i = 0; i < n; ++i) {
    float d = x[i] * inv_max * 32768;
    int k = (int)d;
    k = k < 0 ? 0 : (k > 32767 ? 32767 : k);
    d = lut[k] * l1pmu;
    out[i] = (uint8_t)(int)(((float)(int)(d * invQ) * Q + Q * 0.5) * bitrange);
  }
  free(lut);
}

void oracle_mulaw_complex(const int16_t *x, int64_t n, unsigned bits, unsigned mu, uint8_t *out) {
  float *lut = (float *)malloc(32768 * sizeof(float));
  mulaw_table(mu, lut);
  const unsigned bitrange = (unsigned)pow(2, bits);
  const float Q = (float)(1.0 / bitrange), invQ = (float)(1.0 / Q);
  const float l1pmu = (float)(1.0 / log10(1 + mu));
  for (int64_t i = 0; i < n; ++i) {
    const int re = x[2 * i], im = x[2 * i + 1];
    int ar = abs(re), ai = abs(im);
    ar = ar > 32767 ? 32767 : ar;
    ai = ai > 32767 ? 32767 : ai;
    const float dr = lut[ar] * l1pmu, di = lut[ai] * l1pmu;
    const uint8_t qr = (uint8_t)(int)(((float)(signed char)(dr * invQ) * Q + Q * 0.5) * bitrange);
    const uint8_t qi = (uint8_t)(int)(((float)(signed char)(di * invQ) * Q + Q * 0.5) * bitrange);
    const signed char rs = re < 0 ? -1 : 1, is = im < 0 ? -1 : 1;
    out[i] = (uint8_t)((rs & 0x80) | ((qr & 0x07) << 4) | ((is & 0x80) >> 4) | (qi & 0x07));
  }
  free(lut);
}

