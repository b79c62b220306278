// mulaw_host.cpp -- entry points of the logarithmic raw-data quantiser
// (multibeam_dedispersion_writer.cpp:21-118): table, constants and the
// first-buffer maximum on the host side, kernels in mulaw.hip.
#include <hip/hip_runtime.h>

#include <cmath>
#include <mutex>
#include <vector>

#include "internal.h"

using namespace mdsm;

namespace mdsm {
int launch_max(const float *x, uint64_t n, float *part, uint32_t nparts, hipStream_t s);
int launch_mulaw_power(const float *x, uint64_t n, const float *lut, float inv_max, float log1pmu,
                       float Q, float invQ, uint32_t bitrange, uint8_t *out, hipStream_t s);
int launch_mulaw_complex(const int16_t *x, uint64_t n, const float *lut, float log1pmu, float Q,
                         float invQ, uint32_t bitrange, uint8_t *out, hipStream_t s);
}  // namespace mdsm

namespace {

constexpr uint32_t LUT_N = 32768;

// log10(1 + i*mu/32768) in float, as the reference builds it; the device
// copy is cached per compression value
struct LutCache {
  std::mutex mu;
  float *d = nullptr;
  uint32_t compression = 0;
  int device = -1;
} g_lut;

int device_lut(uint32_t compression, const float **out) {
  std::lock_guard<std::mutex> lock(g_lut.mu);
  int dev = 0;
  MDSM_HIP(hipGetDevice(&dev));
  if (!g_lut.d || g_lut.compression != compression || g_lut.device != dev) {
    std::vector<float> lut(LUT_N);
    for (uint32_t i = 0; i < LUT_N; ++i)
      lut[i] = std::log10(1.0f + (float)(i * compression) / (float)LUT_N);
    if (g_lut.d && g_lut.device == dev) (void)hipFree(g_lut.d);
    g_lut.d = nullptr;
    MDSM_HIP(hipMalloc(&g_lut.d, LUT_N * sizeof(float)));
    MDSM_HIP(hipMemcpy(g_lut.d, lut.data(), LUT_N * sizeof(float), hipMemcpyHostToDevice));
    g_lut.compression = compression;
    g_lut.device = dev;
  }
  *out = g_lut.d;
  return MDSM_OK;
}

}  // namespace

extern "C" int mdsm_mulaw_encode_power(const float *d_in, uint64_t n, mdsm_mulaw_state_t *st,
                                       uint8_t *d_out, void *stream) {
  MDSM_REQUIRE(d_in && st && d_out && n, MDSM_ERR_ARG, "null argument");
  MDSM_REQUIRE(st->bits >= 1 && st->bits <= 8 && st->compression >= 1, MDSM_ERR_ARG,
               "bits must be 1..8 and compression >= 1");
  hipStream_t s = as_stream(stream);
  if (st->max_value == 0.f) {  // first buffer: twice its maximum
    constexpr uint32_t P = 1024;
    float *d_part = nullptr;
    MDSM_HIP(hipMalloc(&d_part, P * sizeof(float)));
    std::vector<float> part(P);
    int rc = launch_max(d_in, n, d_part, P, s);
    if (rc == MDSM_OK && (hipMemcpyAsync(part.data(), d_part, P * 4, hipMemcpyDeviceToHost, s) != hipSuccess ||
                          hipStreamSynchronize(s) != hipSuccess))
      rc = MDSM_ERR_HIP;
    (void)hipFree(d_part);
    if (rc != MDSM_OK) return rc;
    float m = 0.f;
    for (float v : part) m = v > m ? v : m;
    st->max_value = m * 2;
    MDSM_REQUIRE(st->max_value > 0.f, MDSM_ERR_ARG, "first buffer has no positive value");
  }
  const float *lut = nullptr;
  int rc = device_lut(st->compression, &lut);
  if (rc != MDSM_OK) return rc;
  const float Q = (float)(1.0 / std::pow(2, st->bits));
  const float invQ = (float)(1.0 / Q);
  const float log1pmu = (float)(1.0 / std::log10(1 + st->compression));
  const float inv_max = (float)(1.0 / st->max_value);
  const uint32_t bitrange = (uint32_t)std::pow(2, st->bits);
  return launch_mulaw_power(d_in, n, lut, inv_max, log1pmu, Q, invQ, bitrange, d_out, s);
}

extern "C" int mdsm_mulaw_encode_complex(const int16_t *d_in, uint64_t n, uint32_t bits,
                                         uint32_t compression, uint8_t *d_out, void *stream) {
  MDSM_REQUIRE(d_in && d_out && n, MDSM_ERR_ARG, "null argument");
  MDSM_REQUIRE(bits >= 1 && bits <= 3 && compression >= 1, MDSM_ERR_ARG,
               "bits must be 1..3 (3 bits + sign per part) and compression >= 1");
  const float *lut = nullptr;
  int rc = device_lut(compression, &lut);
  if (rc != MDSM_OK) return rc;
  const uint32_t bitrange = (uint32_t)std::pow(2, bits);
  const float Q = (float)(1.0 / bitrange);
  const float invQ = (float)(1.0 / Q);
  const float log1pmu = (float)(1.0 / std::log10(1 + compression));
  return launch_mulaw_complex(d_in, n, lut, log1pmu, Q, invQ, bitrange, d_out, as_stream(stream));
}
