/*
 * cnn.c -- CPU oracle of the event-triage CNN.  TEST INFRASTRUCTURE ONLY.
 *
 * The network has no counterpart in the reference (its triage is
 * DBSCAN + computeTransientProbability, dbscan.cpp:25-276); it is defined
 * here and by the product kernels together:
 *
 *   input  x[B][2][H][W]   (DM-time and subband-time snippets)
 *   conv k (k = 1..3): 3x3, stride 1, zero padding 1, channels 2->16->32->64,
 *          + bias, ReLU, 2x2 max pool (first maximum in the order
 *          (0,0),(0,1),(1,0),(1,1) on ties)
 *   fc1    64*(H/8)*(W/8) -> 64, + bias, ReLU
 *   fc2    64 -> 2 logits, + bias
 *   loss   mean softmax cross-entropy over the batch
 *
 * Parameters are one flat FP64 array in the order w1 b1 w2 b2 w3 b3 fw1 fb1
 * fw2 fb2 with w_k[co][ci][3][3] and fw[out][in] (in = c*(H/8)*(W/8) +
 * y*(W/8) + x).  Everything accumulates in FP64.  With `bf16` set, values
 * are rounded (through float, to nearest even) at the points where the
 * product kernels store BF16 operands: the input, the weights (biases are
 * FP32), the input of every convolution and linear layer, and the output
 * gradients entering the backward products (bias gradients are sums of
 * those rounded gradients).
 *
 * Pinned by central finite differences of the loss (tests/test_oracle_cnn.py)
 * and by closed-form single-layer cases.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

static double bf16r(double v) {
  float f = (float)v;
  uint32_t u;
  memcpy(&u, &f, 4);
  if ((u & 0x7F800000u) != 0x7F800000u) u = (u + 0x7FFFu + ((u >> 16) & 1u)) & 0xFFFF0000u;
  memcpy(&f, &u, 4);
  return (double)f;
}

static void round_array(double *a, int64_t n, int on) {
  if (!on) return;
  for (int64_t i = 0; i < n; ++i) a[i] = bf16r(a[i]);
}

double oracle_bf16_round(double v) { return bf16r(v); }

static const int CH[4] = {2, 16, 32, 64};
#define FC_HIDDEN 64
#define NCLASS 2

int64_t oracle_cnn_param_count(int H, int W) {
  int64_t n = 0;
  for (int k = 0; k < 3; ++k) n += (int64_t)CH[k + 1] * CH[k] * 9 + CH[k + 1];
  const int64_t fin = (int64_t)64 * (H / 8) * (W / 8);
  n += fin * FC_HIDDEN + FC_HIDDEN + FC_HIDDEN * NCLASS + NCLASS;
  return n;
}

/* y[b][co][h][w] = bias + sum w[co][ci][ky][kx] * x[b][ci][h+ky-1][w+kx-1] */
static void conv_fwd(const double *x, int B, int Ci, int Co, int H, int W, const double *w,
                     const double *bias, double *y) {
  for (int b = 0; b < B; ++b)
    for (int co = 0; co < Co; ++co)
      for (int h = 0; h < H; ++h)
        for (int v = 0; v < W; ++v) {
          double acc = bias[co];
          for (int ci = 0; ci < Ci; ++ci)
            for (int ky = 0; ky < 3; ++ky) {
              const int hh = h + ky - 1;
              if (hh < 0 || hh >= H) continue;
              for (int kx = 0; kx < 3; ++kx) {
                const int vv = v + kx - 1;
                if (vv < 0 || vv >= W) continue;
                acc += w[((co * Ci + ci) * 3 + ky) * 3 + kx] *
                       x[(((int64_t)b * Ci + ci) * H + hh) * W + vv];
              }
            }
          y[(((int64_t)b * Co + co) * H + h) * W + v] = acc;
        }
}

static void conv_bwd(const double *x, const double *dy, int B, int Ci, int Co, int H, int W,
                     const double *w, double *dx, double *dw, double *db) {
  memset(dw, 0, sizeof(double) * Co * Ci * 9);
  memset(db, 0, sizeof(double) * Co);
  if (dx) memset(dx, 0, sizeof(double) * (int64_t)B * Ci * H * W);
  for (int b = 0; b < B; ++b)
    for (int co = 0; co < Co; ++co)
      for (int h = 0; h < H; ++h)
        for (int v = 0; v < W; ++v) {
          const double g = dy[(((int64_t)b * Co + co) * H + h) * W + v];
          if (g == 0.0) continue;
          db[co] += g;
          for (int ci = 0; ci < Ci; ++ci)
            for (int ky = 0; ky < 3; ++ky) {
              const int hh = h + ky - 1;
              if (hh < 0 || hh >= H) continue;
              for (int kx = 0; kx < 3; ++kx) {
                const int vv = v + kx - 1;
                if (vv < 0 || vv >= W) continue;
                const int64_t xi = (((int64_t)b * Ci + ci) * H + hh) * W + vv;
                dw[((co * Ci + ci) * 3 + ky) * 3 + kx] += g * x[xi];
                if (dx) dx[xi] += g * w[((co * Ci + ci) * 3 + ky) * 3 + kx];
              }
            }
        }
}

/* ReLU then 2x2 max pool; idx records the argmax (0..3) */
static void relu_pool(const double *z, int B, int C, int H, int W, double *p, uint8_t *idx) {
  const int Ho = H / 2, Wo = W / 2;
  for (int64_t bc = 0; bc < (int64_t)B * C; ++bc)
    for (int h = 0; h < Ho; ++h)
      for (int v = 0; v < Wo; ++v) {
        double best = -1.0;
        int arg = 0;
        for (int k = 0; k < 4; ++k) {
          double a = z[(bc * H + 2 * h + (k >> 1)) * W + 2 * v + (k & 1)];
          a = a > 0 ? a : 0.0;
          if (a > best) {
            best = a;
            arg = k;
          }
        }
        p[(bc * Ho + h) * Wo + v] = best;
        idx[(bc * Ho + h) * Wo + v] = (uint8_t)arg;
      }
}

static void relu_pool_bwd(const double *dp, const double *z, const uint8_t *idx, int B, int C,
                          int H, int W, double *dz) {
  const int Ho = H / 2, Wo = W / 2;
  memset(dz, 0, sizeof(double) * (int64_t)B * C * H * W);
  for (int64_t bc = 0; bc < (int64_t)B * C; ++bc)
    for (int h = 0; h < Ho; ++h)
      for (int v = 0; v < Wo; ++v) {
        const int k = idx[(bc * Ho + h) * Wo + v];
        const int64_t zi = (bc * H + 2 * h + (k >> 1)) * W + 2 * v + (k & 1);
        if (z[zi] > 0) dz[zi] = dp[(bc * Ho + h) * Wo + v];
      }
}

/* Loss and gradients of the whole network.  x: [B][2][H][W]; labels in
 * {0,1}; grads may be NULL (forward only); logits (B x 2) may be NULL.
 * Returns the mean cross-entropy. */
double oracle_cnn_loss_grad(const double *x_in, const int32_t *labels, int B, int H, int W,
                            const double *params_in, double *grads, double *logits_out, int bf16) {
  const int64_t np = oracle_cnn_param_count(H, W);
  double *params = malloc(sizeof(double) * np);
  memcpy(params, params_in, sizeof(double) * np);
  /* weights are used as BF16 copies, biases stay FP32 */
  const double *w[3], *bs[3];
  int64_t off = 0;
  for (int k = 0; k < 3; ++k) {
    w[k] = params + off;
    round_array(params + off, (int64_t)CH[k + 1] * CH[k] * 9, bf16);
    off += (int64_t)CH[k + 1] * CH[k] * 9;
    bs[k] = params + off;
    for (int i = 0; i < CH[k + 1]; ++i) params[off + i] = bf16 ? (double)(float)params[off + i] : params[off + i];
    off += CH[k + 1];
  }
  const int fin = 64 * (H / 8) * (W / 8);
  double *fw1 = params + off, *fb1 = fw1 + (int64_t)fin * FC_HIDDEN;
  double *fw2 = fb1 + FC_HIDDEN, *fb2 = fw2 + FC_HIDDEN * NCLASS;
  round_array(fw1, (int64_t)fin * FC_HIDDEN, bf16);
  round_array(fw2, FC_HIDDEN * NCLASS, bf16);
  if (bf16)
    for (int i = 0; i < FC_HIDDEN + NCLASS; ++i) {
      double *bp = i < FC_HIDDEN ? fb1 + i : fb2 + (i - FC_HIDDEN);
      *bp = (double)(float)*bp;
    }
  /* forward */
  double *a[4], *z[3];
  uint8_t *idx[3];
  int h = H, v = W;
  a[0] = malloc(sizeof(double) * (int64_t)B * 2 * H * W);
  memcpy(a[0], x_in, sizeof(double) * (int64_t)B * 2 * H * W);
  round_array(a[0], (int64_t)B * 2 * H * W, bf16);
  for (int k = 0; k < 3; ++k) {
    const int64_t nz = (int64_t)B * CH[k + 1] * h * v;
    z[k] = malloc(sizeof(double) * nz);
    conv_fwd(a[k], B, CH[k], CH[k + 1], h, v, w[k], bs[k], z[k]);
    a[k + 1] = malloc(sizeof(double) * nz / 4);
    idx[k] = malloc(nz / 4);
    relu_pool(z[k], B, CH[k + 1], h, v, a[k + 1], idx[k]);
    round_array(a[k + 1], nz / 4, bf16);
    h /= 2;
    v /= 2;
  }
  double *z4 = malloc(sizeof(double) * B * FC_HIDDEN), *h4 = malloc(sizeof(double) * B * FC_HIDDEN);
  for (int b = 0; b < B; ++b)
    for (int o = 0; o < FC_HIDDEN; ++o) {
      double acc = fb1[o];
      for (int i = 0; i < fin; ++i) acc += fw1[(int64_t)o * fin + i] * a[3][(int64_t)b * fin + i];
      z4[b * FC_HIDDEN + o] = acc;
      h4[b * FC_HIDDEN + o] = acc > 0 ? acc : 0.0;
    }
  round_array(h4, (int64_t)B * FC_HIDDEN, bf16);
  double *lg = malloc(sizeof(double) * B * NCLASS), *dl = malloc(sizeof(double) * B * NCLASS);
  double loss = 0;
  for (int b = 0; b < B; ++b) {
    for (int o = 0; o < NCLASS; ++o) {
      double acc = fb2[o];
      for (int i = 0; i < FC_HIDDEN; ++i) acc += fw2[o * FC_HIDDEN + i] * h4[b * FC_HIDDEN + i];
      lg[b * NCLASS + o] = acc;
    }
    const double m = fmax(lg[b * 2], lg[b * 2 + 1]);
    const double e0 = exp(lg[b * 2] - m), e1 = exp(lg[b * 2 + 1] - m), s = e0 + e1;
    loss += -(lg[b * 2 + labels[b]] - m - log(s));
    dl[b * 2] = (e0 / s - (labels[b] == 0)) / B;
    dl[b * 2 + 1] = (e1 / s - (labels[b] == 1)) / B;
  }
  loss /= B;
  if (logits_out) memcpy(logits_out, lg, sizeof(double) * B * NCLASS);
  if (grads) {
    double *g = grads;
    int64_t goff = 0;
    double *gw[3], *gb[3];
    for (int k = 0; k < 3; ++k) {
      gw[k] = g + goff;
      goff += (int64_t)CH[k + 1] * CH[k] * 9;
      gb[k] = g + goff;
      goff += CH[k + 1];
    }
    double *gfw1 = g + goff, *gfb1 = gfw1 + (int64_t)fin * FC_HIDDEN;
    double *gfw2 = gfb1 + FC_HIDDEN, *gfb2 = gfw2 + FC_HIDDEN * NCLASS;
    round_array(dl, (int64_t)B * NCLASS, bf16);
    /* fc2 */
    double *dh = calloc((size_t)B * FC_HIDDEN, sizeof(double));
    for (int o = 0; o < NCLASS; ++o) {
      double sb = 0;
      for (int b = 0; b < B; ++b) sb += dl[b * 2 + o];
      gfb2[o] = sb;
      for (int i = 0; i < FC_HIDDEN; ++i) {
        double s = 0;
        for (int b = 0; b < B; ++b) s += dl[b * 2 + o] * h4[b * FC_HIDDEN + i];
        gfw2[o * FC_HIDDEN + i] = s;
      }
    }
    for (int b = 0; b < B; ++b)
      for (int i = 0; i < FC_HIDDEN; ++i) {
        double s = 0;
        for (int o = 0; o < NCLASS; ++o) s += dl[b * 2 + o] * fw2[o * FC_HIDDEN + i];
        dh[b * FC_HIDDEN + i] = z4[b * FC_HIDDEN + i] > 0 ? s : 0.0;
      }
    round_array(dh, (int64_t)B * FC_HIDDEN, bf16);
    /* fc1 */
    double *da = calloc((size_t)B * fin, sizeof(double));
    for (int o = 0; o < FC_HIDDEN; ++o) {
      double sb = 0;
      for (int b = 0; b < B; ++b) sb += dh[b * FC_HIDDEN + o];
      gfb1[o] = sb;
      for (int i = 0; i < fin; ++i) {
        double s = 0;
        for (int b = 0; b < B; ++b) s += dh[b * FC_HIDDEN + o] * a[3][(int64_t)b * fin + i];
        gfw1[(int64_t)o * fin + i] = s;
      }
    }
    for (int b = 0; b < B; ++b)
      for (int i = 0; i < fin; ++i) {
        double s = 0;
        for (int o = 0; o < FC_HIDDEN; ++o) s += dh[b * FC_HIDDEN + o] * fw1[(int64_t)o * fin + i];
        da[(int64_t)b * fin + i] = s;
      }
    /* conv stack, top down */
    for (int k = 2; k >= 0; --k) {
      const int hk = H >> k, vk = W >> k;
      const int64_t nz = (int64_t)B * CH[k + 1] * hk * vk;
      double *dz = malloc(sizeof(double) * nz);
      relu_pool_bwd(da, z[k], idx[k], B, CH[k + 1], hk, vk, dz);
      round_array(dz, nz, bf16);
      double *dprev = k > 0 ? malloc(sizeof(double) * (int64_t)B * CH[k] * hk * vk) : NULL;
      conv_bwd(a[k], dz, B, CH[k], CH[k + 1], hk, vk, w[k], dprev, gw[k], gb[k]);
      free(dz);
      free(da);
      da = dprev;
    }
    free(dh);
  }
  for (int k = 0; k < 4; ++k) free(a[k]);
  for (int k = 0; k < 3; ++k) {
    free(z[k]);
    free(idx[k]);
  }
  free(z4);
  free(h4);
  free(lg);
  free(dl);
  free(params);
  return loss;
}
