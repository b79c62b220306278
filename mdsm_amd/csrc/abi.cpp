// abi.cpp -- host side of libmdsm_amd: the C ABI declared in mdsm_abi.h.
//
// Replaces the seam src/mdsm-multibeam/multibeam_dedispersion_wrapper.cu:4-23
// and the host halves of multibeam_dedispersion_thread.cu (delay table
// :546-559, launch :204-221) and multibeam_dedispersion_manager.cpp
// (dmdelay :325-326, maxshift :371-387).  Errors are status codes with a
// thread-local message; nothing here calls exit().
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <numeric>
#include <string>
#include <vector>

#include "internal.h"

namespace mdsm {

static thread_local char g_err[512] = "";

void set_error(const char *fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_err, sizeof(g_err), fmt, ap);
  va_end(ap);
}

// FP64 rounded delay: round(4148.808*DM*(f^-2 - ftop^-2)/tsamp) in FP64, the
// expression evaluated in exactly this order (built with -ffp-contract=off).
static int32_t delay_ns(double fch1, double foff, uint32_t c, double tsamp, double dm) {
  const double f = fch1 + foff * (double)c;
  const double k = 1.0 / (f * f) - 1.0 / (fch1 * fch1);
  return (int32_t)std::round(((4148.808 * dm) * k) / tsamp);
}

int build_delay_table(const mdsm_survey_t *s, std::vector<int32_t> &table, uint32_t *maxshift) {
  MDSM_REQUIRE(s != nullptr, MDSM_ERR_ARG, "null survey");
  MDSM_REQUIRE(s->nchans > 0 && s->ndms > 0 && s->nsamp > 0, MDSM_ERR_ARG,
               "nchans, ndms and nsamp must be positive");
  MDSM_REQUIRE(s->tsamp > 0.0 && std::isfinite(s->tsamp), MDSM_ERR_ARG, "tsamp must be > 0");
  MDSM_REQUIRE(s->foff < 0.0, MDSM_ERR_ARG,
               "foff must be negative (channel 0 is the top of the band)");
  MDSM_REQUIRE(s->fch1 + s->foff * (double)(s->nchans - 1) > 0.0, MDSM_ERR_ARG,
               "lowest channel frequency must be positive");
  MDSM_REQUIRE(s->lowdm >= 0.0 && s->dmstep >= 0.0, MDSM_ERR_ARG, "DMs must be non-negative");
  const uint64_t n = (uint64_t)s->nchans * s->ndms;
  table.resize(n);
  if (s->delay_mode == MDSM_DELAY_FP64_ROUND) {
    for (uint32_t c = 0; c < s->nchans; ++c)
      for (uint32_t d = 0; d < s->ndms; ++d)
        table[(uint64_t)c * s->ndms + d] =
            delay_ns(s->fch1, s->foff, c, s->tsamp, s->lowdm + s->dmstep * (double)d);
  } else if (s->delay_mode == MDSM_DELAY_REFERENCE || s->delay_mode == MDSM_DELAY_LEGACY) {
    // float dm_shifts (manager.cpp:366-369), float arithmetic throughout.
    // REFERENCE: the truncated (shift * dm) / tsamp of thread.cu:553;
    // LEGACY: floor(shift * (dm / tsamp)) of src/mdsm/dedispersion_kernel.cu:55,68.
    const float fch1 = (float)s->fch1, foff = (float)s->foff, tsamp = (float)s->tsamp;
    const float lowdm = (float)s->lowdm, dmstep = (float)s->dmstep;
    const bool legacy = s->delay_mode == MDSM_DELAY_LEGACY;
    for (uint32_t c = 0; c < s->nchans; ++c) {
      const float F1 = fch1 + foff * (float)c;
      const float sh = (float)(4148.741601 * ((1.0 / F1 / F1) - (1.0 / fch1 / fch1)));
      for (uint32_t d = 0; d < s->ndms; ++d) {
        const float dm = lowdm + (float)d * dmstep;
        if (legacy) {
          const float per_sample = dm / tsamp;
          table[(uint64_t)c * s->ndms + d] = (int32_t)std::floor(sh * per_sample);
        } else {
          float v = sh * dm;
          v = v / tsamp;
          table[(uint64_t)c * s->ndms + d] = (int32_t)v;
        }
      }
    }
  } else {
    set_error("unknown delay_mode %d", s->delay_mode);
    return MDSM_ERR_ARG;
  }
  // the kernels rely on non-negative delays, non-decreasing in DM
  int32_t mx = 0;
  for (uint32_t c = 0; c < s->nchans; ++c) {
    const int32_t *r = &table[(uint64_t)c * s->ndms];
    MDSM_REQUIRE(r[0] >= 0, MDSM_ERR_ARG, "negative delay at channel %u", c);
    for (uint32_t d = 1; d < s->ndms; ++d)
      MDSM_REQUIRE(r[d] >= r[d - 1], MDSM_ERR_ARG, "delays not monotone in DM at channel %u", c);
    mx = std::max(mx, r[s->ndms - 1]);
  }
  MDSM_REQUIRE((uint64_t)s->nsamp + (uint64_t)mx < (1ull << 31), MDSM_ERR_ARG,
               "nsamp + maxshift must be < 2^31");
  if (maxshift) *maxshift = (uint32_t)mx;
  return MDSM_OK;
}

uint32_t tile_spread(const int32_t *table, uint32_t nchans, uint32_t ndms, int td) {
  int32_t best = 0;
  for (uint32_t c = 0; c < nchans; ++c) {
    const int32_t *r = table + (uint64_t)c * ndms;
    for (uint32_t d0 = 0; d0 < ndms; d0 += (uint32_t)td) {
      const uint32_t dl = std::min(d0 + (uint32_t)td, ndms) - 1u;
      best = std::max(best, r[dl] - r[d0]);
    }
  }
  return (uint32_t)best;
}

}  // namespace mdsm

using namespace mdsm;

// ======================================================================
extern "C" const char *mdsm_last_error(void) { return g_err; }
extern "C" int mdsm_version(void) { return 10100; }

extern "C" int mdsm_device_count(int *count) {
  MDSM_REQUIRE(count, MDSM_ERR_ARG, "null count");
  int n = 0;
  hipError_t e = hipGetDeviceCount(&n);
  *count = (e == hipSuccess) ? n : 0;
  return MDSM_OK;
}

extern "C" int mdsm_set_device(int device) {
  MDSM_HIP(hipSetDevice(device));
  return MDSM_OK;
}

extern "C" int mdsm_malloc(void **ptr, size_t bytes) {
  MDSM_REQUIRE(ptr, MDSM_ERR_ARG, "null ptr");
  hipError_t e = hipMalloc(ptr, bytes ? bytes : 1);
  if (e != hipSuccess) {
    set_error("hipMalloc(%zu) failed: %s", bytes, hipGetErrorString(e));
    *ptr = nullptr;
    return MDSM_ERR_NOMEM;
  }
  return MDSM_OK;
}

extern "C" int mdsm_free(void *ptr) {
  if (ptr) MDSM_HIP(hipFree(ptr));
  return MDSM_OK;
}

extern "C" int mdsm_host_malloc(void **ptr, size_t bytes) {
  MDSM_REQUIRE(ptr, MDSM_ERR_ARG, "null ptr");
  hipError_t e = hipHostMalloc(ptr, bytes ? bytes : 1, hipHostMallocDefault);
  if (e != hipSuccess) {
    set_error("hipHostMalloc(%zu) failed: %s", bytes, hipGetErrorString(e));
    *ptr = nullptr;
    return MDSM_ERR_NOMEM;
  }
  return MDSM_OK;
}

extern "C" int mdsm_host_free(void *ptr) {
  if (ptr) MDSM_HIP(hipHostFree(ptr));
  return MDSM_OK;
}

extern "C" int mdsm_memcpy_htod(void *dst, const void *src, size_t bytes, void *stream) {
  MDSM_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, as_stream(stream)));
  MDSM_HIP(hipStreamSynchronize(as_stream(stream)));
  return MDSM_OK;
}

extern "C" int mdsm_memcpy_dtoh(void *dst, const void *src, size_t bytes, void *stream) {
  MDSM_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, as_stream(stream)));
  MDSM_HIP(hipStreamSynchronize(as_stream(stream)));
  return MDSM_OK;
}

extern "C" int mdsm_memcpy_dtod(void *dst, const void *src, size_t bytes, void *stream) {
  MDSM_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, as_stream(stream)));
  return MDSM_OK;
}

extern "C" int mdsm_memset(void *ptr, int value, size_t bytes, void *stream) {
  MDSM_HIP(hipMemsetAsync(ptr, value, bytes, as_stream(stream)));
  return MDSM_OK;
}

extern "C" int mdsm_stream_create(void **stream) {
  MDSM_REQUIRE(stream, MDSM_ERR_ARG, "null stream");
  hipStream_t s;
  MDSM_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
  *stream = s;
  return MDSM_OK;
}

extern "C" int mdsm_stream_destroy(void *stream) {
  if (stream) MDSM_HIP(hipStreamDestroy(as_stream(stream)));
  return MDSM_OK;
}

extern "C" int mdsm_stream_synchronize(void *stream) {
  MDSM_HIP(hipStreamSynchronize(as_stream(stream)));
  return MDSM_OK;
}

extern "C" int mdsm_device_synchronize(void) {
  MDSM_HIP(hipDeviceSynchronize());
  return MDSM_OK;
}

extern "C" int mdsm_event_create(void **event) {
  MDSM_REQUIRE(event, MDSM_ERR_ARG, "null event");
  hipEvent_t e;
  MDSM_HIP(hipEventCreate(&e));
  *event = e;
  return MDSM_OK;
}

extern "C" int mdsm_event_destroy(void *event) {
  if (event) MDSM_HIP(hipEventDestroy(reinterpret_cast<hipEvent_t>(event)));
  return MDSM_OK;
}

extern "C" int mdsm_event_record(void *event, void *stream) {
  MDSM_HIP(hipEventRecord(reinterpret_cast<hipEvent_t>(event), as_stream(stream)));
  return MDSM_OK;
}

extern "C" int mdsm_event_elapsed(void *start, void *stop, float *ms) {
  MDSM_HIP(hipEventSynchronize(reinterpret_cast<hipEvent_t>(stop)));
  MDSM_HIP(hipEventElapsedTime(ms, reinterpret_cast<hipEvent_t>(start),
                               reinterpret_cast<hipEvent_t>(stop)));
  return MDSM_OK;
}

extern "C" int mdsm_delay_table(const mdsm_survey_t *s, int32_t *table, uint32_t *maxshift) {
  std::vector<int32_t> t;
  int rc = build_delay_table(s, t, maxshift);
  if (rc != MDSM_OK) return rc;
  if (table) std::memcpy(table, t.data(), t.size() * sizeof(int32_t));
  return MDSM_OK;
}

// ======================================================================
// plans
struct mdsm_plan_s {
  mdsm_survey_t s;
  int dtype;
  uint32_t maxshift;
  std::vector<int32_t> h_table;
  int32_t *d_table = nullptr;
  int variant = 0;  // 0 until set or tuned; then 1..num_variants
  std::vector<uint32_t> row_words;  // per variant (index 1..)
  std::vector<int32_t *> d_rel;     // trial-offset table per variant, built on first use
};

// Build the trial-offset table of variant v on stream s if it is missing
// (one allocation and one kernel, the first time the variant is used; the
// stream is then synchronised so that later calls on any stream see it).
static int ensure_trial_offsets(mdsm_plan_s *p, int v, hipStream_t s) {
  if (p->d_rel[v]) return MDSM_OK;
  const uint32_t rw = p->row_words[v];
  const uint32_t cbr = stage_channels(p->dtype, v, rw);
  MDSM_REQUIRE(rw > 0 && cbr > 0, MDSM_ERR_ARG,
               "delay spread too large for tile variant %d (choose another variant)", v);
  const uint64_t words = trial_offset_words(p->dtype, v, p->s.nchans, p->s.ndms, cbr);
  int32_t *d = nullptr;
  hipError_t e = hipMalloc(&d, words * sizeof(int32_t));
  if (e != hipSuccess) {
    set_error("hipMalloc(trial offsets) failed: %s", hipGetErrorString(e));
    return MDSM_ERR_NOMEM;
  }
  int rc = launch_trial_offsets(p->dtype, v, p->d_table, p->s.nchans, p->s.ndms, rw, cbr, d, s);
  if (rc == MDSM_OK && hipStreamSynchronize(s) != hipSuccess) {
    set_error("trial offset table build failed");
    rc = MDSM_ERR_HIP;
  }
  if (rc != MDSM_OK) {
    (void)hipFree(d);
    return rc;
  }
  p->d_rel[v] = d;
  return MDSM_OK;
}

// A tile shape can stage this survey when its delay spread has an LDS row
// and the LDS holds at least one channel of that row.
static bool variant_fits(const mdsm_plan_s *p, int v) {
  return p->row_words[v] > 0 && stage_channels(p->dtype, v, p->row_words[v]) > 0;
}

// Untuned default: the preferred tile shape if it can stage this survey's
// delay spread, else the storage type's fallback shape, else the first one
// that can.
static int default_variant(const mdsm_plan_s *p) {
  const int pref = preferred_variant(p->dtype);
  if (variant_fits(p, pref)) return pref;
  const int fb = fallback_variant(p->dtype);
  if (variant_fits(p, fb)) return fb;
  for (int v = 1; v < (int)p->row_words.size(); ++v)
    if (variant_fits(p, v)) return v;
  return pref;
}

static int plan_variant(const mdsm_plan_s *p) { return p->variant ? p->variant : default_variant(p); }

extern "C" int mdsm_plan_create(const mdsm_survey_t *s, int in_dtype, mdsm_plan *plan) {
  MDSM_REQUIRE(plan, MDSM_ERR_ARG, "null plan");
  *plan = nullptr;
  MDSM_REQUIRE(known_dtype(in_dtype), MDSM_ERR_ARG,
               "unknown input dtype %d", in_dtype);
  MDSM_REQUIRE(s && s->nchans <= 65536, MDSM_ERR_ARG, "nchans must be <= 65536");
  auto *p = new mdsm_plan_s();
  p->s = *s;
  p->dtype = in_dtype;
  int rc = build_delay_table(s, p->h_table, &p->maxshift);
  if (rc != MDSM_OK) {
    delete p;
    return rc;
  }
  const int nv = num_variants(in_dtype);
  p->row_words.assign(nv + 1, 0);
  p->d_rel.assign(nv + 1, nullptr);
  for (int v = 1; v <= nv; ++v) {
    const TileShape sh = variant_shape(in_dtype, v);
    p->row_words[v] =
        lds_row_words(in_dtype, v, tile_spread(p->h_table.data(), s->nchans, s->ndms, sh.td()),
                      p->h_table.data(), s->nchans, s->ndms);
  }
  const size_t bytes = p->h_table.size() * sizeof(int32_t);
  hipError_t e = hipMalloc(&p->d_table, bytes);
  if (e != hipSuccess) {
    set_error("hipMalloc(delay table) failed: %s", hipGetErrorString(e));
    delete p;
    return MDSM_ERR_NOMEM;
  }
  e = hipMemcpy(p->d_table, p->h_table.data(), bytes, hipMemcpyHostToDevice);
  if (e != hipSuccess) {
    set_error("delay table upload failed: %s", hipGetErrorString(e));
    (void)hipFree(p->d_table);
    delete p;
    return MDSM_ERR_HIP;
  }
  *plan = p;
  return MDSM_OK;
}

extern "C" int mdsm_plan_destroy(mdsm_plan plan) {
  if (!plan) return MDSM_OK;
  for (int32_t *d : plan->d_rel)
    if (d) MDSM_HIP(hipFree(d));
  if (plan->d_table) MDSM_HIP(hipFree(plan->d_table));
  delete plan;
  return MDSM_OK;
}

extern "C" int mdsm_plan_info(mdsm_plan plan, uint32_t *maxshift, uint64_t *min_ld) {
  MDSM_REQUIRE(plan, MDSM_ERR_ARG, "null plan");
  if (maxshift) *maxshift = plan->maxshift;
  if (min_ld) {
    const uint64_t a = input_alignment(plan->dtype);
    *min_ld = ((uint64_t)plan->s.nsamp + plan->maxshift + a - 1) / a * a;
  }
  return MDSM_OK;
}

extern "C" int mdsm_plan_device_table(mdsm_plan plan, const int32_t **table) {
  MDSM_REQUIRE(plan && table, MDSM_ERR_ARG, "null argument");
  *table = plan->d_table;
  return MDSM_OK;
}

extern "C" int mdsm_tile_variants(int in_dtype, int *count) {
  MDSM_REQUIRE(count, MDSM_ERR_ARG, "null count");
  *count = num_variants(in_dtype);
  return MDSM_OK;
}

extern "C" int mdsm_plan_set_variant(mdsm_plan plan, int variant) {
  MDSM_REQUIRE(plan, MDSM_ERR_ARG, "null plan");
  MDSM_REQUIRE(variant >= 0 && variant <= num_variants(plan->dtype), MDSM_ERR_ARG,
               "variant %d out of range", variant);
  plan->variant = variant;
  // build the variant's trial-offset table now, so that dedisperse calls
  // (and graphs captured from them) allocate nothing
  const int v = plan_variant(plan);
  if (variant_fits(plan, v)) {
    const int rc = ensure_trial_offsets(plan, v, nullptr);
    if (rc != MDSM_OK) return rc;
  }
  return MDSM_OK;
}

extern "C" int mdsm_tile_preferred(int in_dtype, int *variant) {
  MDSM_REQUIRE(variant, MDSM_ERR_ARG, "null variant");
  MDSM_REQUIRE(known_dtype(in_dtype), MDSM_ERR_ARG,
               "unknown input dtype %d", in_dtype);
  *variant = preferred_variant(in_dtype);
  return MDSM_OK;
}

extern "C" int mdsm_plan_variant_channels(mdsm_plan plan, int variant, uint32_t *channels) {
  MDSM_REQUIRE(plan && channels, MDSM_ERR_ARG, "null argument");
  MDSM_REQUIRE(variant >= 1 && variant <= num_variants(plan->dtype), MDSM_ERR_ARG,
               "variant %d out of range", variant);
  *channels = variant_fits(plan, variant)
                  ? stage_channels(plan->dtype, variant, plan->row_words[variant])
                  : 0u;
  return MDSM_OK;
}

extern "C" int mdsm_plan_get_variant(mdsm_plan plan, int *variant) {
  MDSM_REQUIRE(plan && variant, MDSM_ERR_ARG, "null argument");
  *variant = plan_variant(plan);
  return MDSM_OK;
}

extern "C" int mdsm_dedisperse(mdsm_plan plan, const void *d_in, uint64_t ld, float *d_out,
                               void *stream) {
  MDSM_REQUIRE(plan && d_in && d_out, MDSM_ERR_ARG, "null argument");
  MDSM_REQUIRE(ld >= (uint64_t)plan->s.nsamp + plan->maxshift, MDSM_ERR_ARG,
               "input row length %llu < nsamp + maxshift = %llu", (unsigned long long)ld,
               (unsigned long long)plan->s.nsamp + plan->maxshift);
  MDSM_REQUIRE(ld % input_alignment(plan->dtype) == 0 && ((uintptr_t)d_in & 3u) == 0, MDSM_ERR_ARG,
               "input rows must start on 4-byte boundaries (row length a multiple of %u)",
               input_alignment(plan->dtype));
  MDSM_REQUIRE(ld / input_alignment(plan->dtype) < (1ull << 29), MDSM_ERR_ARG,
               "input rows must be shorter than 2 GiB");
  const int v = plan_variant(plan);
  const int rc = ensure_trial_offsets(plan, v, as_stream(stream));
  if (rc != MDSM_OK) return rc;
  return launch_dedisperse(plan->dtype, v, d_in, ld, plan->d_table, plan->d_rel[v], plan->s.nchans,
                           plan->s.nsamp, plan->s.ndms, plan->row_words[v], d_out,
                           as_stream(stream));
}

namespace mdsm {

// Time every tile variant of a plan on the given buffers: one warm-up launch,
// then the best of three timed launches; t[v] < 0 where a variant does not fit.
int plan_time_variants(mdsm_plan plan, const void *d_in, uint64_t ld, float *d_out,
                       hipStream_t s, std::vector<float> &t) {
  hipEvent_t e0, e1;
  MDSM_HIP(hipEventCreate(&e0));
  MDSM_HIP(hipEventCreate(&e1));
  const int nv = num_variants(plan->dtype);
  const int keep = plan->variant;
  t.assign(nv + 1, -1.0f);
  for (int v = 1; v <= nv; ++v) {
    plan->variant = v;
    if (mdsm_dedisperse(plan, d_in, ld, d_out, s) != MDSM_OK) continue;  // warm-up / LDS fit
    float vms = -1.0f;
    for (int rep = 0; rep < 3; ++rep) {
      float ms = 0.0f;
      MDSM_HIP(hipEventRecord(e0, s));
      int rc = mdsm_dedisperse(plan, d_in, ld, d_out, s);
      MDSM_HIP(hipEventRecord(e1, s));
      MDSM_HIP(hipEventSynchronize(e1));
      if (rc != MDSM_OK) break;
      MDSM_HIP(hipEventElapsedTime(&ms, e0, e1));
      if (vms < 0.0f || ms < vms) vms = ms;
    }
    t[v] = vms;
  }
  plan->variant = keep;
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  return MDSM_OK;
}

// The fastest entry of t (index 1..), or the storage type's default when
// that is within 2 % of it, so that near-ties do not flip between runs.
int pick_variant(int dtype, const std::vector<float> &t) {
  int best = 0;
  for (int v = 1; v < (int)t.size(); ++v)
    if (t[v] >= 0.0f && (!best || t[v] < t[best])) best = v;
  const int pref = preferred_variant(dtype);
  if (best && pref != best && pref < (int)t.size() && t[pref] >= 0.0f && t[pref] <= 1.02f * t[best])
    best = pref;
  return best;
}

}  // namespace mdsm

extern "C" int mdsm_plan_autotune(mdsm_plan plan, const void *d_in, uint64_t ld, float *d_out,
                                  void *stream, int *best_variant, float *best_ms) {
  MDSM_REQUIRE(plan && d_in && d_out, MDSM_ERR_ARG, "null argument");
  std::vector<float> t;
  int rc = plan_time_variants(plan, d_in, ld, d_out, as_stream(stream), t);
  if (rc != MDSM_OK) return rc;
  const int best = pick_variant(plan->dtype, t);
  MDSM_REQUIRE(best > 0, MDSM_ERR_ARG, "no tile variant fits this survey");
  plan->variant = best;
  if (best_variant) *best_variant = best;
  if (best_ms) *best_ms = t[best];
  return MDSM_OK;
}

// ======================================================================
// normalisation / detection wrappers
extern "C" int mdsm_channel_stats(const float *d_raw, uint64_t ld, uint32_t nchans, uint64_t n,
                                  double *d_mean, double *d_std, void *stream) {
  MDSM_REQUIRE(d_raw && d_mean && d_std && n > 0 && n <= ld, MDSM_ERR_ARG, "bad arguments");
  return launch_channel_stats(d_raw, ld, nchans, n, d_mean, d_std, as_stream(stream));
}

extern "C" int mdsm_normalise(const float *d_raw, uint64_t ld, uint32_t nchans, uint64_t n,
                              const double *d_mean, const double *d_std, void *d_out, int out_dtype,
                              uint64_t ldo, void *stream) {
  MDSM_REQUIRE(d_raw && d_mean && d_std && d_out && n <= ld && n <= ldo, MDSM_ERR_ARG,
               "bad arguments");
  return launch_normalise(d_raw, ld, nchans, n, d_mean, d_std, d_out, out_dtype, ldo,
                          as_stream(stream));
}

extern "C" int mdsm_block_stats(const float *d_x, uint32_t nrows, uint32_t nsamp, uint32_t L,
                                double *d_mean, double *d_std, void *stream) {
  MDSM_REQUIRE(d_x && d_mean && d_std && L > 0 && nsamp > 0 && nrows > 0 && nrows <= 65535,
               MDSM_ERR_ARG, "bad arguments");
  return launch_block_stats(d_x, nrows, nsamp, L, d_mean, d_std, as_stream(stream));
}

extern "C" int mdsm_detect(const float *d_x, uint32_t nrows, uint32_t nsamp, uint32_t L,
                           const double *d_mean, const double *d_std, int nwidths, double thresh,
                           mdsm_cand_t *d_cands, uint32_t cap, uint32_t *d_count, void *stream) {
  MDSM_REQUIRE(d_x && d_mean && d_std && d_cands && d_count && L > 0 && nrows > 0 &&
                   nrows <= 65535 && nsamp > 0,
               MDSM_ERR_ARG, "bad arguments");
  MDSM_REQUIRE(nwidths >= 1 && nwidths <= 7, MDSM_ERR_ARG, "nwidths must be 1..7");
  int rc = launch_zero_u32(d_count, 1, as_stream(stream));
  if (rc != MDSM_OK) return rc;
  return launch_detect(d_x, nrows, nsamp, L, d_mean, d_std, nwidths, thresh, d_cands, cap, d_count,
                       as_stream(stream));
}

namespace mdsm {

// Device workspace of the global detector, one per (device, stream): a call
// enqueues its kernels on its stream, so calls on one stream reuse the
// buffer in order and calls on different streams never share one.  Buffers
// come from the stream-ordered allocator and grow (free + allocate on the
// same stream) when a larger region arrives.
namespace {
struct StreamWorkspace {
  int device;
  hipStream_t stream;
  void *ptr;
  size_t bytes;
};
std::mutex g_ws_mu;
std::vector<StreamWorkspace> g_ws;
}  // namespace

int stream_workspace(hipStream_t s, size_t bytes, void **out) {
  int dev = 0;
  MDSM_HIP(hipGetDevice(&dev));
  std::lock_guard<std::mutex> lock(g_ws_mu);
  for (auto &w : g_ws)
    if (w.device == dev && w.stream == s) {
      if (w.bytes < bytes) {
        MDSM_HIP(hipFreeAsync(w.ptr, s));
        w.ptr = nullptr;
        w.bytes = 0;
        MDSM_HIP(hipMallocAsync(&w.ptr, bytes, s));
        w.bytes = bytes;
      }
      *out = w.ptr;
      return MDSM_OK;
    }
  void *p = nullptr;
  MDSM_HIP(hipMallocAsync(&p, bytes, s));
  g_ws.push_back({dev, s, p, bytes});
  *out = p;
  return MDSM_OK;
}

// Checks shared by the global detector's entry points.
int check_global_args(const float *d_x, uint32_t nrows, uint32_t nsamp, uint64_t ld,
                      uint32_t stats_nsamp, const mdsm_cand_t *d_cands, const uint32_t *d_count) {
  MDSM_REQUIRE(d_x && d_cands && d_count, MDSM_ERR_ARG, "null argument");
  MDSM_REQUIRE(nrows > 0 && nrows <= 65535 && nsamp > 0, MDSM_ERR_ARG,
               "need 1 <= nrows <= 65535 and nsamp > 0");
  MDSM_REQUIRE(ld >= nsamp, MDSM_ERR_ARG, "row stride %llu < nsamp %u", (unsigned long long)ld,
               nsamp);
  MDSM_REQUIRE(stats_nsamp > 0 && stats_nsamp <= nsamp, MDSM_ERR_ARG,
               "stats_nsamp must be in 1..nsamp");
  MDSM_REQUIRE(nsamp < (1u << 31), MDSM_ERR_ARG, "nsamp must be < 2^31");
  return MDSM_OK;
}

// The global detector's kernels on a given workspace: chunk partials at
// work[0 .. 2*parts), mean and rms at work[2*parts], work[2*parts + 1].
int run_global_detect(const float *d_x, uint32_t nrows, uint32_t nsamp, uint64_t ld,
                      uint32_t stats_nsamp, float thresh, mdsm_cand_t *d_cands, uint32_t cap,
                      uint32_t *d_count, double *h_mean, double *h_rms, double *work,
                      hipStream_t s) {
  const uint32_t parts = global_stats_parts(nrows, stats_nsamp);
  double *stats = work + 2u * (uint64_t)parts;
  int rc = launch_global_stats(d_x, ld, nrows, stats_nsamp, work, stats, s);
  if (rc == MDSM_OK) rc = launch_zero_u32(d_count, 1, s);
  if (rc == MDSM_OK)
    rc = launch_threshold_global(d_x, ld, nrows, nsamp, stats, 0.0, 0.0, thresh, d_cands, cap,
                                 d_count, s);
  if (rc != MDSM_OK || (!h_mean && !h_rms)) return rc;
  double h[2];
  MDSM_HIP(hipMemcpyAsync(h, stats, sizeof(h), hipMemcpyDeviceToHost, s));
  MDSM_HIP(hipStreamSynchronize(s));
  if (h_mean) *h_mean = h[0];
  if (h_rms) *h_rms = h[1];
  return MDSM_OK;
}

}  // namespace mdsm

extern "C" int mdsm_detect_global(const float *d_x, uint32_t nrows, uint32_t nsamp, uint64_t ld,
                                  uint32_t stats_nsamp, float thresh, mdsm_cand_t *d_cands,
                                  uint32_t cap, uint32_t *d_count, double *h_mean, double *h_rms,
                                  void *stream) {
  int rc = check_global_args(d_x, nrows, nsamp, ld, stats_nsamp, d_cands, d_count);
  if (rc != MDSM_OK) return rc;
  hipStream_t s = as_stream(stream);
  const uint64_t words = 2ull * global_stats_parts(nrows, stats_nsamp) + 2ull;
  void *work = nullptr;
  rc = stream_workspace(s, words * sizeof(double), &work);
  if (rc != MDSM_OK) return rc;
  return run_global_detect(d_x, nrows, nsamp, ld, stats_nsamp, thresh, d_cands, cap, d_count,
                           h_mean, h_rms, static_cast<double *>(work), s);
}

extern "C" int mdsm_threshold_global(const float *d_x, uint32_t nrows, uint32_t nsamp,
                                     uint64_t ld, double mean, double rms, float thresh,
                                     mdsm_cand_t *d_cands, uint32_t cap, uint32_t *d_count,
                                     void *stream) {
  int rc = check_global_args(d_x, nrows, nsamp, ld, nsamp, d_cands, d_count);
  if (rc != MDSM_OK) return rc;
  hipStream_t s = as_stream(stream);
  rc = launch_zero_u32(d_count, 1, s);
  if (rc != MDSM_OK) return rc;
  return launch_threshold_global(d_x, ld, nrows, nsamp, nullptr, mean, rms, thresh, d_cands, cap,
                                 d_count, s);
}

extern "C" int mdsm_detect_legacy(const float *d_x, uint32_t ndms, uint32_t nsamp, float lowdm,
                                  float dmstep, float thresh, uint64_t nvals, const float *d_mean,
                                  const float *d_rms, uint32_t samples_per_chunk, uint32_t nstats,
                                  mdsm_cand_t *d_cands, uint32_t cap, uint32_t *d_count,
                                  float *h_mean, float *h_rms, void *stream) {
  MDSM_REQUIRE(d_x && d_cands && d_count && ndms > 0 && nsamp > 0 && ndms <= 65535, MDSM_ERR_ARG,
               "bad arguments");
  // the reference counts values in an unsigned int
  MDSM_REQUIRE((uint64_t)ndms * nsamp < (1ull << 32), MDSM_ERR_ARG,
               "ndms * nsamp must stay below 2^32");
  MDSM_REQUIRE(!d_mean == !d_rms, MDSM_ERR_ARG, "give both per-chunk mean and rms, or neither");
  if (d_mean) {
    MDSM_REQUIRE(samples_per_chunk > 0, MDSM_ERR_ARG, "samples_per_chunk must be positive");
    const uint32_t last = nsamp > 4 ? nsamp - 5 : 0;
    MDSM_REQUIRE((uint64_t)nstats >= (uint64_t)(last / samples_per_chunk) + 1, MDSM_ERR_ARG,
                 "too few per-chunk statistics for nsamp");
  }
  hipStream_t s = as_stream(stream);
  float gmean = 0.0f, grms = 0.0f;
  if (!d_mean) {
    // dedispersion_output.cpp:145-166: blocks of mod_factor values; the loop
    // leaves after block n once (n + 1) * mod_factor >= vals, so with more
    // than one block the last block (full or partial) never enters the mean
    MDSM_REQUIRE(nvals <= (uint64_t)ndms * nsamp, MDSM_ERR_ARG, "nvals exceeds ndms * nsamp");
    const uint32_t vals = nvals ? (uint32_t)nvals : ndms * nsamp;
    const uint32_t mod = vals < 32u * 1024u ? vals : 32u * 1024u;
    const uint32_t nfull = (uint32_t)(((uint64_t)vals + mod - 1) / mod);
    const uint32_t nblocks = nfull > 1 ? nfull - 1 : 1;
    double *d_part = nullptr;
    MDSM_HIP(hipMalloc((void **)&d_part, sizeof(double) * 2 * nblocks));
    int rc = launch_legacy_block_sums(d_x, mod, nblocks, d_part, s);
    std::vector<double> h(2 * (size_t)nblocks);
    if (rc == MDSM_OK) {
      MDSM_HIP(hipMemcpyAsync(h.data(), d_part, sizeof(double) * 2 * nblocks,
                              hipMemcpyDeviceToHost, s));
    }
    MDSM_HIP(hipStreamSynchronize(s));
    MDSM_HIP(hipFree(d_part));
    if (rc != MDSM_OK) return rc;
    // FP32 running mean of FP64 block means, FP64 mean of squares, as the
    // reference's variable types round them
    float localmean = 0.0f;
    double mean2 = 0.0;
    for (uint32_t i = 0; i < nblocks; ++i) {
      localmean = (float)((double)localmean + h[2 * i] / (double)mod);
      mean2 += h[2 * i + 1] / (double)mod;
    }
    localmean = localmean / (float)nblocks;
    const float lm2 = localmean * localmean;
    gmean = localmean;
    grms = (float)std::sqrt(mean2 / (double)nblocks - (double)lm2);
    if (h_mean) *h_mean = gmean;
    if (h_rms) *h_rms = grms;
  }
  int rc = launch_zero_u32(d_count, 1, s);
  if (rc != MDSM_OK || nsamp < 5) return rc;
  return launch_legacy_threshold(d_x, ndms, nsamp, lowdm, dmstep, thresh, gmean, grms, d_mean,
                                 d_rms, samples_per_chunk, d_cands, cap, d_count, s);
}

extern "C" int mdsm_median_filter5(const float *d_x, uint32_t nrows, uint32_t nsamp, float *d_out,
                                   void *stream) {
  MDSM_REQUIRE(d_x && d_out && d_x != d_out && nrows <= 65535, MDSM_ERR_ARG,
               "bad arguments (out of place only)");
  return launch_median5(d_x, nrows, nsamp, d_out, as_stream(stream));
}

extern "C" int mdsm_detrend_normalise(float *d_x, uint64_t nseg, uint32_t L, void *stream) {
  MDSM_REQUIRE(d_x && L > 1 && nseg > 0 && nseg < (1ull << 31), MDSM_ERR_ARG, "bad arguments");
  return launch_detrend(d_x, nseg, L, as_stream(stream));
}

// ======================================================================
// synthetic generator
extern "C" int mdsm_generate(uint64_t seed, uint32_t nchans, int64_t t0, uint64_t n,
                             const float *gain, const float *offset, uint32_t npulses,
                             const int64_t *p_start, const int32_t *p_width, const float *p_amp,
                             const int64_t *p_delay, float *d_raw, uint64_t ld, void *stream) {
  MDSM_REQUIRE(gain && offset && d_raw && n <= ld && t0 >= 0, MDSM_ERR_ARG, "bad arguments");
  hipStream_t s = as_stream(stream);
  // pack the host parameter arrays into one device buffer
  const size_t n_g = nchans, n_p = npulses ? npulses : 1;
  std::vector<char> blob;
  auto put = [&](const void *src, size_t bytes) {
    size_t off = (blob.size() + 15) & ~size_t(15);
    blob.resize(off + bytes);
    if (src) std::memcpy(blob.data() + off, src, bytes);
    return off;
  };
  const size_t o_gain = put(gain, n_g * 4), o_off = put(offset, n_g * 4);
  const size_t o_st = put(npulses ? p_start : nullptr, n_p * 8);
  const size_t o_w = put(npulses ? p_width : nullptr, n_p * 4);
  const size_t o_a = put(npulses ? p_amp : nullptr, n_p * 4);
  const size_t o_dl = put(npulses ? p_delay : nullptr, n_p * n_g * 8);
  char *d_blob = nullptr;
  MDSM_HIP(hipMalloc((void **)&d_blob, blob.size()));
  int rc = MDSM_OK;
  hipError_t e = hipMemcpyAsync(d_blob, blob.data(), blob.size(), hipMemcpyHostToDevice, s);
  if (e != hipSuccess) {
    set_error("generator upload failed: %s", hipGetErrorString(e));
    rc = MDSM_ERR_HIP;
  }
  if (rc == MDSM_OK)
    rc = launch_generate(seed, nchans, t0, n, (const float *)(d_blob + o_gain),
                         (const float *)(d_blob + o_off), npulses, (const int64_t *)(d_blob + o_st),
                         (const int32_t *)(d_blob + o_w), (const float *)(d_blob + o_a),
                         (const int64_t *)(d_blob + o_dl), d_raw, ld, s);
  e = hipStreamSynchronize(s);
  (void)hipFree(d_blob);
  if (rc == MDSM_OK && e != hipSuccess) {
    set_error("generator failed: %s", hipGetErrorString(e));
    rc = MDSM_ERR_HIP;
  }
  return rc;
}
