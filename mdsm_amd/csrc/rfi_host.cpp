// rfi_host.cpp -- bandpass fit and RFI clipping entry points.
//
// The sequence is the reference's bandpass_fitting + rfi_clipping
// (multibeam_dedispersion_thread.cu:224-390): per-channel means on the
// device, channel masks and a least-squares polynomial fit on the host
// (x_i = i/nchans, ncoeffs terms), the fitted bandpass back on the device,
// then channel-block flagging/clipping and spectrum clipping against
// thresholds derived from the bandpass statistics.  The reference fits
// with GSL's SVD least squares; this file solves the same problem with
// Householder QR on column-equilibrated FP64 (backward stable; agrees with
// an SVD solution to ~1e-12 relative on the fitted values).
#include <hip/hip_runtime.h>

#include <cmath>
#include <vector>

#include "internal.h"

using namespace mdsm;

namespace {

// least squares min ||X c - y|| for X[i][j] = x_i^j, Householder QR
void poly_lstsq(const std::vector<double> &xs, const std::vector<double> &y, int p,
                std::vector<double> &coef) {
  const int m = (int)xs.size();
  std::vector<double> A((size_t)m * p), b(y);
  for (int i = 0; i < m; ++i)
    for (int j = 0; j < p; ++j) A[(size_t)i * p + j] = std::pow(xs[i], j);
  std::vector<double> scale(p, 1.0);
  for (int j = 0; j < p; ++j) {
    double s = 0;
    for (int i = 0; i < m; ++i) s += A[(size_t)i * p + j] * A[(size_t)i * p + j];
    s = std::sqrt(s);
    if (s > 0) {
      scale[j] = s;
      for (int i = 0; i < m; ++i) A[(size_t)i * p + j] /= s;
    }
  }
  std::vector<double> diag(p);
  for (int k = 0; k < p; ++k) {
    double norm = 0;
    for (int i = k; i < m; ++i) norm += A[(size_t)i * p + k] * A[(size_t)i * p + k];
    norm = std::sqrt(norm);
    const double akk = A[(size_t)k * p + k];
    const double alpha = akk > 0 ? -norm : norm;
    diag[k] = alpha;
    if (norm == 0) continue;
    // v = a_k - alpha e_k stored in place; H = I - 2 v v^T / (v^T v)
    A[(size_t)k * p + k] = akk - alpha;
    double vtv = 0;
    for (int i = k; i < m; ++i) vtv += A[(size_t)i * p + k] * A[(size_t)i * p + k];
    for (int j = k + 1; j < p; ++j) {
      double d = 0;
      for (int i = k; i < m; ++i) d += A[(size_t)i * p + k] * A[(size_t)i * p + j];
      const double f = 2 * d / vtv;
      for (int i = k; i < m; ++i) A[(size_t)i * p + j] -= f * A[(size_t)i * p + k];
    }
    double d = 0;
    for (int i = k; i < m; ++i) d += A[(size_t)i * p + k] * b[i];
    const double f = 2 * d / vtv;
    for (int i = k; i < m; ++i) b[i] -= f * A[(size_t)i * p + k];
  }
  coef.assign(p, 0.0);
  for (int k = p - 1; k >= 0; --k) {
    double s = b[k];
    for (int j = k + 1; j < p; ++j) s -= A[(size_t)k * p + j] * coef[j];
    coef[k] = diag[k] != 0 ? s / diag[k] : 0.0;
  }
  for (int j = 0; j < p; ++j) coef[j] /= scale[j];
}

int check_rows(const void *x, uint64_t ld, uint32_t nchans, uint64_t col0, uint64_t n) {
  MDSM_REQUIRE(x, MDSM_ERR_ARG, "null input");
  MDSM_REQUIRE(nchans >= 1 && nchans <= 65535, MDSM_ERR_ARG, "nchans must be in 1..65535");
  MDSM_REQUIRE(n >= 1 && col0 + n <= ld, MDSM_ERR_ARG, "col0 + n exceeds the row length");
  return MDSM_OK;
}

}  // namespace

extern "C" int mdsm_bandpass_sum(const float *d_x, uint64_t ld, uint32_t nchans, uint64_t col0,
                                 uint64_t n, double *d_bandpass, void *stream) {
  int rc = check_rows(d_x, ld, nchans, col0, n);
  if (rc != MDSM_OK) return rc;
  MDSM_REQUIRE(d_bandpass, MDSM_ERR_ARG, "null output");
  return launch_bandpass_sum(d_x, ld, nchans, col0, n, d_bandpass, as_stream(stream));
}

extern "C" int mdsm_bandpass_fit(const double *bandpass, uint32_t nchans, const mdsm_rfi_params_t *p,
                                 double *fitted, mdsm_bandpass_stats_t *st) {
  MDSM_REQUIRE(bandpass && p && fitted && st, MDSM_ERR_ARG, "null argument");
  MDSM_REQUIRE(p->ncoeffs >= 1 && p->ncoeffs <= 32 && p->ncoeffs <= nchans, MDSM_ERR_ARG,
               "ncoeffs must be in 1..min(32, nchans)");
  MDSM_REQUIRE(p->nmasks == 0 || (p->mask_from && p->mask_to), MDSM_ERR_ARG, "null mask arrays");
  std::vector<double> bp(bandpass, bandpass + nchans);
  // channel masks take the mean of their neighbours, stored through a float
  for (uint32_t m = 0; m < p->nmasks; ++m) {
    const int64_t a = p->mask_from[m], b = p->mask_to[m];
    MDSM_REQUIRE(a >= 0 && a <= b && b < (int64_t)nchans, MDSM_ERR_ARG, "mask %u out of range", m);
    float value;
    if (a == 0) {
      MDSM_REQUIRE(b + 2 < (int64_t)nchans, MDSM_ERR_ARG, "mask %u leaves no neighbours", m);
      value = (float)((bp[b + 1] + bp[b + 2]) * 0.5);
    } else if (b >= (int64_t)nchans - 1) {
      MDSM_REQUIRE(a >= 2, MDSM_ERR_ARG, "mask %u leaves no neighbours", m);
      value = (float)((bp[a - 1] + bp[a - 2]) * 0.5);
    } else {
      value = (float)((bp[a - 1] + bp[b + 1]) * 0.5);
    }
    for (int64_t i = a; i <= b; ++i) bp[i] = value;
  }
  std::vector<double> xs(nchans);
  for (uint32_t i = 0; i < nchans; ++i) xs[i] = 0 + i / (1.0 * nchans);
  std::vector<double> coef;
  poly_lstsq(xs, bp, (int)p->ncoeffs, coef);
  for (uint32_t i = 0; i < nchans; ++i) {
    double acc = 0.0;
    for (uint32_t j = 0; j < p->ncoeffs; ++j) acc += coef[j] * std::pow(xs[i], (double)j);
    fitted[i] = acc;
  }
  // statistics in float accumulators, as the reference's SURVEY fields
  float mean = 0, stdv = 0, rmse = 0;
  for (uint32_t i = 0; i < nchans; ++i) mean = (float)(mean + bp[i] / nchans);
  for (uint32_t i = 0; i < nchans; ++i) stdv = (float)(stdv + (bp[i] - mean) * (bp[i] - mean));
  stdv = std::sqrt(stdv / (float)nchans);
  for (uint32_t i = 0; i < nchans; ++i) rmse = (float)(rmse + std::pow(bp[i] - fitted[i], 2));
  rmse = std::sqrt(rmse / (float)nchans);
  const float spread = p->spectrum_thresh * stdv;
  st->mean = mean;
  st->std = stdv;
  st->rmse = rmse;
  st->channel_thresh = p->channel_thresh * rmse;
  st->spectrum_hi = mean + spread;
  st->spectrum_lo = mean - spread;
  return MDSM_OK;
}

extern "C" int mdsm_rfi_channel_flag(const float *d_x, uint64_t ld, uint32_t nchans, uint64_t col0,
                                     uint64_t n, const double *d_fitted, uint32_t channel_block,
                                     float thresh, uint8_t *d_flags, void *stream) {
  int rc = check_rows(d_x, ld, nchans, col0, n);
  if (rc != MDSM_OK) return rc;
  MDSM_REQUIRE(d_fitted && d_flags && channel_block >= 1, MDSM_ERR_ARG, "bad arguments");
  return launch_channel_flag(d_x, ld, nchans, col0, n, d_fitted, channel_block, thresh, d_flags,
                             as_stream(stream));
}

extern "C" int mdsm_rfi_channel_clip(float *d_x, uint64_t ld, uint32_t nchans, uint64_t col0,
                                     uint64_t n, const double *d_fitted, uint32_t channel_block,
                                     const uint8_t *d_flags, void *stream) {
  int rc = check_rows(d_x, ld, nchans, col0, n);
  if (rc != MDSM_OK) return rc;
  MDSM_REQUIRE(d_fitted && d_flags && channel_block >= 1, MDSM_ERR_ARG, "bad arguments");
  MDSM_REQUIRE((n + channel_block - 1) / channel_block <= 0x7FFFFFFFull, MDSM_ERR_ARG,
               "too many channel blocks");
  return launch_channel_clip(d_x, ld, nchans, col0, n, d_fitted, channel_block, d_flags,
                             as_stream(stream));
}

extern "C" int mdsm_rfi_spectrum_clip(float *d_x, uint64_t ld, uint32_t nchans, uint64_t col0,
                                      uint64_t n, const double *d_fitted, float hi, float lo,
                                      uint32_t *d_count, void *stream) {
  int rc = check_rows(d_x, ld, nchans, col0, n);
  if (rc != MDSM_OK) return rc;
  MDSM_REQUIRE(d_fitted, MDSM_ERR_ARG, "null bandpass");
  MDSM_REQUIRE(n <= 0xFFFFFFFFull * 256, MDSM_ERR_ARG, "n too large");
  return launch_spectrum_clip(d_x, ld, nchans, col0, n, d_fitted, hi, lo, d_count,
                              as_stream(stream));
}

extern "C" int mdsm_voltage_to_power(float *d_x, uint64_t ld, uint32_t nchans, uint64_t col0,
                                     uint64_t n, void *stream) {
  int rc = check_rows(d_x, ld, nchans, col0, n);
  if (rc != MDSM_OK) return rc;
  return launch_voltage_power(d_x, ld, nchans, col0, n, as_stream(stream));
}

extern "C" int mdsm_rfi_clean(float *d_x, uint64_t ld, uint32_t nchans, uint64_t col0, uint64_t n,
                              const mdsm_rfi_params_t *p, double *d_fitted, uint8_t *d_flags,
                              mdsm_bandpass_stats_t *stats, void *stream) {
  int rc = check_rows(d_x, ld, nchans, col0, n);
  if (rc != MDSM_OK) return rc;
  MDSM_REQUIRE(p && d_fitted && p->channel_block >= 1, MDSM_ERR_ARG, "bad arguments");
  hipStream_t s = as_stream(stream);
  rc = launch_bandpass_sum(d_x, ld, nchans, col0, n, d_fitted, s);
  if (rc != MDSM_OK) return rc;
  std::vector<double> bp(nchans), fitted(nchans);
  MDSM_HIP(hipMemcpyAsync(bp.data(), d_fitted, nchans * 8, hipMemcpyDeviceToHost, s));
  MDSM_HIP(hipStreamSynchronize(s));
  mdsm_bandpass_stats_t st;
  rc = mdsm_bandpass_fit(bp.data(), nchans, p, fitted.data(), &st);
  if (rc != MDSM_OK) return rc;
  MDSM_HIP(hipMemcpyAsync(d_fitted, fitted.data(), nchans * 8, hipMemcpyHostToDevice, s));
  if (d_flags) {
    rc = launch_channel_flag(d_x, ld, nchans, col0, n, d_fitted, p->channel_block,
                             st.channel_thresh, d_flags, s);
    if (rc == MDSM_OK && p->apply_channel_clip)
      rc = launch_channel_clip(d_x, ld, nchans, col0, n, d_fitted, p->channel_block, d_flags, s);
    if (rc != MDSM_OK) return rc;
  }
  rc = launch_spectrum_clip(d_x, ld, nchans, col0, n, d_fitted, st.spectrum_hi, st.spectrum_lo,
                            nullptr, s);
  if (rc != MDSM_OK) return rc;
  MDSM_HIP(hipStreamSynchronize(s));  // the fitted vector on the host stack was copied from
  if (stats) *stats = st;
  return MDSM_OK;
}
