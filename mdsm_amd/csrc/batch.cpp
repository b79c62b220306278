// batch.cpp -- multi-beam batching: many beams of uneven nchans / nsamp /
// ndms dedispersed by one launch: a persistent grid draining eight device
// work queues (one per XCD), or a flat grid of one workgroup per tile.
//
// The reference runs one host pthread per beam, each launching
// shared_dedispersion on its own slice of device memory
// (src/mdsm-multibeam/multibeam_dedispersion_thread.cu:464-839,
// multibeam_dedispersion_manager.cpp:495-512).  Here every beam's tiles are
// items: a beam descriptor table in device memory, beams ordered largest
// first, items numbered so that the DM tiles sharing a time window land on
// one XCD (and its L2), eight atomic counters.  mdsm_batch_launch issues
// only a counter-clear kernel and the persistent kernel, so it can be
// captured in a hipGraph and replayed per chunk.
#include <algorithm>
#include <cstdlib>
#include <cstring>
#include <numeric>
#include <vector>

#include "internal.h"

using namespace mdsm;

struct mdsm_batch_s {
  int dtype = MDSM_F32;
  int variant = 0;
  int mode = MDSM_BATCH_FLAT;  // the faster shape measured on the 32-beam batch
  uint32_t bound_mode = MDSM_BATCH_FLAT;
  std::vector<mdsm_survey_t> beams;
  std::vector<uint32_t> maxshift;
  std::vector<std::vector<int32_t>> h_tables;
  std::vector<int32_t *> d_tables;
  std::vector<std::vector<uint32_t>> row_words;  // [beam][variant]
  std::vector<std::vector<int32_t *>> d_rel;     // [beam][variant] trial-offset tables
  std::vector<uint32_t> order;                   // processing order of beams
  BeamDesc *d_desc = nullptr;
  uint32_t *d_counter = nullptr;
  bool bound = false;
  uint32_t bound_variant = 0;
  uint32_t total_items = 0;
  uint32_t max_row_words = 0;
  std::vector<BeamDesc> h_desc;
};

// Adjacent time tiles per XCD group (the item order is described at
// decode_item in dedisperse.hip).  Measured on the 32-beam batch: 4 cuts the
// input fetched from beyond L2 from 35.7x to 19.9x the input bytes (2:
// 23.3x; 8 overflows L2: 34.6x) and runs 0-2 % faster than 1; MDSM_TGROUP
// (1..16) overrides it.
static uint32_t batch_tgroup() {
  const char *e = getenv("MDSM_TGROUP");
  const int a = e ? atoi(e) : 4;
  return a >= 1 && a <= 16 ? (uint32_t)a : 4u;
}

// A tile shape can run the batch when every beam's delay spread has an LDS
// row (row_words > 0) and the LDS holds at least one staged channel of the
// widest row (the launch stages every beam with the widest beam's block).
static bool batch_fits(const mdsm_batch_s *b, int v) {
  uint32_t maxrw = 0;
  for (const auto &rw : b->row_words) {
    if (!rw[v]) return false;
    maxrw = std::max(maxrw, rw[v]);
  }
  return stage_channels(b->dtype, v, maxrw) > 0;
}

// Untuned default: the preferred tile shape when it fits every beam, else the
// storage type's fallback shape, else the first shape that does.
static int batch_variant(const mdsm_batch_s *b) {
  if (b->variant) return b->variant;
  auto fits = [&](int v) { return batch_fits(b, v); };
  const int pref = preferred_variant(b->dtype);
  if (fits(pref)) return pref;
  const int fb = fallback_variant(b->dtype);
  if (fits(fb)) return fb;
  for (int v = 1; v <= num_variants(b->dtype); ++v)
    if (fits(v)) return v;
  return pref;
}

static void free_batch(mdsm_batch_s *b) {
  for (auto *p : b->d_tables)
    if (p) (void)hipFree(p);
  for (auto &per : b->d_rel)
    for (auto *p : per)
      if (p) (void)hipFree(p);
  if (b->d_desc) (void)hipFree(b->d_desc);
  if (b->d_counter) (void)hipFree(b->d_counter);
  delete b;
}

extern "C" int mdsm_batch_create(const mdsm_survey_t *beams, uint32_t nbeams, int in_dtype,
                                 mdsm_batch *batch) {
  MDSM_REQUIRE(batch && beams && nbeams > 0, MDSM_ERR_ARG, "bad arguments");
  MDSM_REQUIRE(known_dtype(in_dtype), MDSM_ERR_ARG,
               "unknown input dtype %d", in_dtype);
  for (uint32_t i = 0; i < nbeams; ++i)
    MDSM_REQUIRE(beams[i].nchans <= 65536, MDSM_ERR_ARG, "beam %u: nchans must be <= 65536", i);
  *batch = nullptr;
  auto *b = new mdsm_batch_s();
  b->dtype = in_dtype;
  b->beams.assign(beams, beams + nbeams);
  b->maxshift.resize(nbeams);
  b->h_tables.resize(nbeams);
  b->d_tables.assign(nbeams, nullptr);
  b->row_words.resize(nbeams);
  const int nv = num_variants(in_dtype);
  b->d_rel.assign(nbeams, std::vector<int32_t *>(nv + 1, nullptr));
  for (uint32_t i = 0; i < nbeams; ++i) {
    int rc = build_delay_table(&beams[i], b->h_tables[i], &b->maxshift[i]);
    if (rc != MDSM_OK) {
      free_batch(b);
      return rc;
    }
    b->row_words[i].assign(nv + 1, 0);
    for (int v = 1; v <= nv; ++v) {
      const TileShape sh = variant_shape(in_dtype, v);
      b->row_words[i][v] = lds_row_words(
          in_dtype, v, tile_spread(b->h_tables[i].data(), beams[i].nchans, beams[i].ndms, sh.td()),
          b->h_tables[i].data(), beams[i].nchans, beams[i].ndms);
    }
    const size_t bytes = b->h_tables[i].size() * sizeof(int32_t);
    if (hipMalloc(&b->d_tables[i], bytes) != hipSuccess ||
        hipMemcpy(b->d_tables[i], b->h_tables[i].data(), bytes, hipMemcpyHostToDevice) != hipSuccess) {
      set_error("delay table upload failed for beam %u", i);
      free_batch(b);
      return MDSM_ERR_NOMEM;
    }
  }
  // largest beams (most channel-sample-trial adds) first: the queue tail
  // then holds the cheapest items, which evens out the finish
  b->order.resize(nbeams);
  std::iota(b->order.begin(), b->order.end(), 0u);
  std::stable_sort(b->order.begin(), b->order.end(), [&](uint32_t x, uint32_t y) {
    const auto &a = beams[x], &c = beams[y];
    return (double)a.nchans * a.ndms * a.nsamp > (double)c.nchans * c.ndms * c.nsamp;
  });
  if (hipMalloc(&b->d_desc, sizeof(BeamDesc) * nbeams) != hipSuccess ||
      hipMalloc(&b->d_counter, 8 * sizeof(uint32_t)) != hipSuccess) {
    set_error("batch descriptor allocation failed");
    free_batch(b);
    return MDSM_ERR_NOMEM;
  }
  *batch = b;
  return MDSM_OK;
}

extern "C" int mdsm_batch_destroy(mdsm_batch batch) {
  if (batch) free_batch(batch);
  return MDSM_OK;
}

extern "C" int mdsm_batch_info(mdsm_batch batch, uint32_t beam, uint32_t *maxshift,
                               uint64_t *min_ld) {
  MDSM_REQUIRE(batch && beam < batch->beams.size(), MDSM_ERR_ARG, "bad beam index");
  if (maxshift) *maxshift = batch->maxshift[beam];
  if (min_ld) {
    const uint64_t a = input_alignment(batch->dtype);
    *min_ld = ((uint64_t)batch->beams[beam].nsamp + batch->maxshift[beam] + a - 1) / a * a;
  }
  return MDSM_OK;
}

extern "C" int mdsm_batch_set_variant(mdsm_batch batch, int variant) {
  MDSM_REQUIRE(batch, MDSM_ERR_ARG, "null batch");
  MDSM_REQUIRE(variant >= 0 && variant <= num_variants(batch->dtype), MDSM_ERR_ARG,
               "variant %d out of range", variant);
  batch->variant = variant;
  batch->bound = false;
  return MDSM_OK;
}

// Tile autotuner for a batch, in two stages: every beam, truncated to 2^16
// samples, is dedispersed with every tile shape (warm-up, best of three)
// from the caller's input into a scratch output allocated here, and a
// shape's batch time is estimated as the sum of its beams' times scaled to
// their full length; the three least estimates among the shapes that fit
// every beam are then timed as one whole launch each, and the fastest kept.
extern "C" int mdsm_batch_autotune(mdsm_batch batch, const void *const *d_in, const uint64_t *ld,
                                   float *const *d_out_full, void *stream, int *best_variant,
                                   float *best_ms) {
  MDSM_REQUIRE(batch && d_in && ld && d_out_full, MDSM_ERR_ARG, "null argument");
  const uint32_t nb = (uint32_t)batch->beams.size();
  const int nv = num_variants(batch->dtype);
  const uint32_t probe_n = 1u << 16;
  size_t out_words = 0;
  for (uint32_t i = 0; i < nb; ++i) {
    MDSM_REQUIRE(d_in[i], MDSM_ERR_ARG, "null input for beam %u", i);
    out_words = std::max(out_words, (size_t)batch->beams[i].ndms *
                                        std::min<uint32_t>(batch->beams[i].nsamp, probe_n));
  }
  float *d_out = nullptr;
  if (hipMalloc((void **)&d_out, sizeof(float) * out_words) != hipSuccess) {
    set_error("autotune: scratch output allocation failed");
    return MDSM_ERR_NOMEM;
  }
  std::vector<double> est(nv + 1, 0.0);
  std::vector<bool> ok(nv + 1, true);
  ok[0] = false;
  int rc = MDSM_OK;
  for (uint32_t i = 0; i < nb && rc == MDSM_OK; ++i) {
    mdsm_survey_t probe = batch->beams[i];
    probe.nsamp = std::min<uint32_t>(probe.nsamp, probe_n);
    mdsm_plan p = nullptr;
    rc = mdsm_plan_create(&probe, batch->dtype, &p);
    if (rc != MDSM_OK) break;
    std::vector<float> t;
    rc = plan_time_variants(p, d_in[i], ld[i], d_out, as_stream(stream), t);
    mdsm_plan_destroy(p);
    const double scale = (double)batch->beams[i].nsamp / (double)probe.nsamp;
    for (int v = 1; v <= nv; ++v) {
      if (t.size() <= (size_t)v || t[v] < 0.0f || !batch_fits(batch, v)) ok[v] = false;
      else est[v] += scale * t[v];
    }
  }
  (void)hipFree(d_out);
  if (rc != MDSM_OK) return rc;
  // the three least estimates are then timed as whole launches on the
  // caller's buffers (the probes miss the launch's shared L2 and XCD
  // placement), and the fastest of them is kept
  std::vector<int> cand;
  for (int u = 1; u <= nv; ++u)
    if (ok[u]) cand.push_back(u);
  std::sort(cand.begin(), cand.end(), [&](int a, int b) { return est[a] < est[b]; });
  if (cand.size() > 3) cand.resize(3);
  const int keep_variant = batch->variant;
  hipStream_t s = as_stream(stream);
  hipEvent_t e0, e1;
  MDSM_HIP(hipEventCreate(&e0));
  MDSM_HIP(hipEventCreate(&e1));
  int v = 0;
  float vms = 0.0f;
  for (int u : cand) {
    batch->variant = u;
    rc = mdsm_batch_bind(batch, d_in, ld, d_out_full);
    if (rc == MDSM_OK) rc = mdsm_batch_launch(batch, stream);  // warm-up
    if (rc != MDSM_OK) break;
    float ms = 0.0f;
    MDSM_HIP(hipEventRecord(e0, s));
    rc = mdsm_batch_launch(batch, stream);
    MDSM_HIP(hipEventRecord(e1, s));
    MDSM_HIP(hipEventSynchronize(e1));
    if (rc != MDSM_OK) break;
    MDSM_HIP(hipEventElapsedTime(&ms, e0, e1));
    if (!v || ms < vms) {
      v = u;
      vms = ms;
    }
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  batch->variant = v > 0 ? v : keep_variant;
  batch->bound = false;
  if (rc != MDSM_OK) return rc;
  if (best_variant) *best_variant = batch_variant(batch);
  if (best_ms) *best_ms = vms;
  return MDSM_OK;
}

extern "C" int mdsm_batch_bind(mdsm_batch batch, const void *const *d_in, const uint64_t *ld,
                               float *const *d_out) {
  MDSM_REQUIRE(batch && d_in && ld && d_out, MDSM_ERR_ARG, "null argument");
  const uint32_t nb = (uint32_t)batch->beams.size();
  const int v = batch_variant(batch);
  const TileShape sh = variant_shape(batch->dtype, v);
  batch->h_desc.resize(nb);
  uint64_t items = 0;
  uint32_t maxrw = 0;
  for (uint32_t i = 0; i < nb; ++i) maxrw = std::max(maxrw, batch->row_words[i][v]);
  // one launch stages every beam with the same channel block (from the
  // widest row), so each beam's trial-offset table is built for that block
  const uint32_t cbr = stage_channels(batch->dtype, v, maxrw);
  for (uint32_t i = 0; i < nb; ++i) {
    const auto &s = batch->beams[i];
    MDSM_REQUIRE(batch->row_words[i][v] > 0 && cbr > 0, MDSM_ERR_ARG,
                 "beam %u: delay spread too large for tile variant %d", i, v);
    if (batch->d_rel[i][v]) continue;
    const uint64_t words = trial_offset_words(batch->dtype, v, s.nchans, s.ndms, cbr);
    int32_t *d = nullptr;
    MDSM_HIP(hipMalloc(&d, words * sizeof(int32_t)));
    const int rc = launch_trial_offsets(batch->dtype, v, batch->d_tables[i], s.nchans, s.ndms,
                                        batch->row_words[i][v], cbr, d, nullptr);
    if (rc != MDSM_OK) {
      (void)hipFree(d);
      return rc;
    }
    batch->d_rel[i][v] = d;
  }
  MDSM_HIP(hipStreamSynchronize(nullptr));
  for (uint32_t k = 0; k < nb; ++k) {
    const uint32_t i = batch->order[k];
    const auto &s = batch->beams[i];
    MDSM_REQUIRE(d_in[i] && d_out[i], MDSM_ERR_ARG, "null buffer for beam %u", i);
    MDSM_REQUIRE(ld[i] >= (uint64_t)s.nsamp + batch->maxshift[i], MDSM_ERR_ARG,
                 "beam %u: input row length below nsamp + maxshift", i);
    MDSM_REQUIRE(ld[i] % input_alignment(batch->dtype) == 0 && ((uintptr_t)d_in[i] & 3u) == 0,
                 MDSM_ERR_ARG, "beam %u: input rows must start on 4-byte boundaries", i);
    MDSM_REQUIRE(ld[i] / input_alignment(batch->dtype) < (1ull << 29), MDSM_ERR_ARG,
                 "beam %u: input rows must be shorter than 2 GiB", i);
    MDSM_REQUIRE(batch->row_words[i][v] > 0, MDSM_ERR_ARG,
                 "beam %u: delay spread too large for tile variant %d", i, v);
    BeamDesc &d = batch->h_desc[k];
    const intptr_t base = reinterpret_cast<intptr_t>(batch->d_desc);
    d.in_off = reinterpret_cast<intptr_t>(d_in[i]) - base;
    d.out_off = reinterpret_cast<intptr_t>(d_out[i]) - base;
    d.table_off = reinterpret_cast<intptr_t>(batch->d_tables[i]) - base;
    d.rel_off = reinterpret_cast<intptr_t>(batch->d_rel[i][v]) - base;
    d.ld = ld[i];
    d.nchans = s.nchans;
    d.nsamp = s.nsamp;
    d.ndms = s.ndms;
    d.row_words = batch->row_words[i][v];
    d.n_ttiles = (s.nsamp + sh.tt() - 1) / sh.tt();
    d.n_dtiles = (s.ndms + sh.td() - 1) / sh.td();
    d.item_begin = (uint32_t)items;
    d.tgroup = batch_tgroup();
    // items are numbered XCD-major: 8 slots per group of 8 * tgroup time tiles
    const uint32_t span = 8u * d.tgroup;
    items += (uint64_t)(d.n_ttiles + span - 1u) / span * span * d.n_dtiles;
  }
  MDSM_REQUIRE(items < (1ull << 31), MDSM_ERR_ARG, "too many work items");
  MDSM_HIP(hipMemcpy(batch->d_desc, batch->h_desc.data(), sizeof(BeamDesc) * nb,
                     hipMemcpyHostToDevice));
  batch->total_items = (uint32_t)items;
  batch->max_row_words = maxrw;
  batch->bound = true;
  batch->bound_variant = (uint32_t)v;
  batch->bound_mode = (uint32_t)batch->mode;
  return MDSM_OK;
}

extern "C" int mdsm_batch_set_mode(mdsm_batch batch, int mode) {
  MDSM_REQUIRE(batch && (mode == MDSM_BATCH_PERSISTENT || mode == MDSM_BATCH_FLAT), MDSM_ERR_ARG,
               "mode must be MDSM_BATCH_PERSISTENT or MDSM_BATCH_FLAT");
  batch->mode = mode;
  batch->bound = false;
  return MDSM_OK;
}

extern "C" int mdsm_batch_launch(mdsm_batch batch, void *stream) {
  MDSM_REQUIRE(batch && batch->bound, MDSM_ERR_STATE, "batch not bound to buffers");
  hipStream_t s = as_stream(stream);
  if (batch->bound_mode == MDSM_BATCH_FLAT)
    return launch_batch_flat(batch->dtype, (int)batch->bound_variant, batch->d_desc,
                             (uint32_t)batch->beams.size(), batch->total_items,
                             batch->max_row_words, s);
  int rc = launch_zero_u32(batch->d_counter, 8, s);  // one queue per XCD slot
  if (rc != MDSM_OK) return rc;
  return launch_batch_dedisperse(batch->dtype, (int)batch->bound_variant, batch->d_desc,
                                 (uint32_t)batch->beams.size(), nullptr, batch->total_items,
                                 batch->d_counter, batch->max_row_words, 0, s);
}

extern "C" int mdsm_batch_dedisperse(mdsm_batch batch, const void *const *d_in,
                                     const uint64_t *ld, float *const *d_out, void *stream) {
  int rc = mdsm_batch_bind(batch, d_in, ld, d_out);
  if (rc != MDSM_OK) return rc;
  return mdsm_batch_launch(batch, stream);
}
