/*
 * triage.c -- CPU restatement of the triage-data kernels (narrowband RFI
 * injection, event snippets).  TEST INFRASTRUCTURE ONLY.  These kernels
 * have no counterpart in the reference; the definitions are the product's
 * own (include/mdsm_abi.h, mdsm_inject_rfi and mdsm_triage_sim_run).
 * Pinned by hand-computed cases (tests/test_oracle_triage.py).
 */
#include <math.h>
#include <stdint.h>
#include <string.h>

typedef struct {
  uint32_t chan;
  int32_t kind;
  int64_t start, len;
  float amp, period;
} rfi_item_t;

void oracle_philox4x32_10(const uint32_t ctr_in[4], const uint32_t key_in[2], uint32_t out[4]);
void oracle_det_sincos_turn(double u, double *sn, double *cs);

static float bf16_value(float f) {
  uint32_t u;
  memcpy(&u, &f, 4);
  u = (u + 0x7FFFu + ((u >> 16) & 1u)) & 0xFFFF0000u;
  memcpy(&f, &u, 4);
  return f;
}

/* raw[chan][t] += gain[chan] * v(t) for t in [start, min(start+len, n)) */
void oracle_rfi_inject(const rfi_item_t *items, int nitems, const float *gain, uint64_t seed,
                       float *raw, int64_t ld, int64_t n) {
  for (int i = 0; i < nitems; ++i) {
    const rfi_item_t *it = items + i;
    const int64_t end = it->start + it->len < n ? it->start + it->len : n;
    for (int64_t t = it->start < 0 ? 0 : it->start; t < end; ++t) {
      float v;
      if (it->kind == 0) {
        v = it->amp;
      } else if (it->kind == 1) {
        double u = (double)t / it->period, sn, cs;
        u = u - floor(u);
        oracle_det_sincos_turn(u, &sn, &cs);
        v = (float)(it->amp * 0.5 * (1.0 + sn));
      } else {
        const uint64_t blk = (uint64_t)t / (uint64_t)it->period;
        const uint32_t ctr[4] = {(uint32_t)blk, (uint32_t)(blk >> 32), it->chan + 0x9000u, 0u};
        const uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
        uint32_t r[4];
        oracle_philox4x32_10(ctr, key, r);
        v = (r[0] & 3u) == 0 ? it->amp : 0.f;
      }
      raw[it->chan * ld + t] += gain[it->chan] * v;
    }
  }
}

/* Snippets of one event (BF16 values as floats): plane 0 [SD][ST] =
 * out[dm0 - SD/2 + d][tc - ST/2 + j] * scale0; plane 1 [SB][ST] = sum over the
 * channels of subband s of store[c][t + table[c][dm0]] * scale1; zero
 * outside the arrays; clipped to +-clip; sums in channel order (FP32). */
void oracle_snippet(const float *out, int ndms, int nsamp, const float *store, int64_t ld, int nchans,
                    const int32_t *table, int dm0, int tc, int SD, int ST, int SB, float scale0,
                    float scale1, float clip, float *snip) {
  for (int j = 0; j < ST; ++j) {
    const int64_t t = (int64_t)tc - ST / 2 + j;
    const int tin = t >= 0 && t < nsamp;
    for (int d = 0; d < SD; ++d) {
      const int dm = dm0 - SD / 2 + d;
      float v = (tin && dm >= 0 && dm < ndms) ? out[(int64_t)dm * nsamp + t] * scale0 : 0.f;
      v = fminf(fmaxf(v, -clip), clip);
      snip[(int64_t)d * ST + j] = bf16_value(v);
    }
    for (int s = 0; s < SB; ++s) {
      const int c0 = (int)((int64_t)s * nchans / SB), c1 = (int)((int64_t)(s + 1) * nchans / SB);
      float acc = 0.f;
      if (tin)
        for (int c = c0; c < c1; ++c) {
          const int64_t i = t + table[(int64_t)c * ndms + dm0];
          if (i < ld) acc += store[(int64_t)c * ld + i];
        }
      float v = fminf(fmaxf(acc * scale1, -clip), clip);
      snip[(int64_t)(SD + s) * ST + j] = bf16_value(v);
    }
  }
}
