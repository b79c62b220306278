// pipeline.cpp -- the per-chunk multi-beam search sequence (normalise ->
// batched dedispersion -> block statistics -> detection) as one unit that
// is captured into a hipGraph and replayed per chunk.
//
// The reference spreads these stages over host threads and barriers
// (multibeam_dedispersion_thread.cu:464-839 for the GPU stages,
// multibeam_dedispersion_output.cpp:325-452 for detection on the CPU).
// Here every stage is a kernel on one stream; the launch sequence has no
// allocation, host copy or memset (counters are zeroed by a kernel), so it
// records into a graph whose replay costs one launch call per chunk.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <vector>

#include "internal.h"

using namespace mdsm;

struct mdsm_pipeline_s {
  std::vector<mdsm_survey_t> beams;
  int dtype = MDSM_F16;
  mdsm_detect_params_t dp{};
  mdsm_batch batch = nullptr;
  std::vector<void *> d_store;     // normalised input per beam
  std::vector<uint64_t> ld_store;  // its row length (samples)
  std::vector<float *> d_out;
  std::vector<double *> d_cmean, d_cstd;  // per-channel statistics
  std::vector<double *> d_bmean, d_bstd;  // per-block statistics
  std::vector<mdsm_cand_t *> d_cands;
  uint32_t *d_counts = nullptr;  // one counter per beam
  std::vector<const float *> raw;
  std::vector<uint64_t> ld_raw;
  bool bound = false;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  hipStream_t graph_stream = nullptr;
};


static void drop_graph(mdsm_pipeline_s *p) {
  if (p->exec) (void)hipGraphExecDestroy(p->exec);
  if (p->graph) (void)hipGraphDestroy(p->graph);
  p->exec = nullptr;
  p->graph = nullptr;
}

static void free_pipeline(mdsm_pipeline_s *p) {
  drop_graph(p);
  if (p->graph_stream) (void)hipStreamDestroy(p->graph_stream);
  if (p->batch) mdsm_batch_destroy(p->batch);
  auto fr = [](void *q) { if (q) (void)hipFree(q); };
  for (auto *q : p->d_store) fr(q);
  for (auto *q : p->d_out) fr(q);
  for (auto *q : p->d_cmean) fr(q);
  for (auto *q : p->d_cstd) fr(q);
  for (auto *q : p->d_bmean) fr(q);
  for (auto *q : p->d_bstd) fr(q);
  for (auto *q : p->d_cands) fr(q);
  fr(p->d_counts);
  delete p;
}

extern "C" int mdsm_pipeline_create(const mdsm_survey_t *beams, uint32_t nbeams, int dtype,
                                    const mdsm_detect_params_t *dp, mdsm_pipeline *out) {
  MDSM_REQUIRE(beams && nbeams && dp && out, MDSM_ERR_ARG, "null argument");
  MDSM_REQUIRE(dp->nwidths >= 1 && dp->nwidths <= 7 && dp->stats_block > 0 && dp->cand_cap > 0,
               MDSM_ERR_ARG, "bad detection parameters");
  *out = nullptr;
  auto *p = new mdsm_pipeline_s();
  p->beams.assign(beams, beams + nbeams);
  p->dtype = dtype;
  p->dp = *dp;
  int rc = mdsm_batch_create(beams, nbeams, dtype, &p->batch);
  if (rc != MDSM_OK) {
    free_pipeline(p);
    return rc;
  }
  p->d_store.assign(nbeams, nullptr);
  p->ld_store.assign(nbeams, 0);
  p->d_out.assign(nbeams, nullptr);
  p->d_cmean.assign(nbeams, nullptr);
  p->d_cstd.assign(nbeams, nullptr);
  p->d_bmean.assign(nbeams, nullptr);
  p->d_bstd.assign(nbeams, nullptr);
  p->d_cands.assign(nbeams, nullptr);
  bool ok = hipMalloc(&p->d_counts, sizeof(uint32_t) * nbeams) == hipSuccess;
  for (uint32_t b = 0; ok && b < nbeams; ++b) {
    const auto &s = beams[b];
    uint64_t min_ld = 0;
    mdsm_batch_info(p->batch, b, nullptr, &min_ld);
    p->ld_store[b] = min_ld;
    const size_t nb = (s.nsamp + dp->stats_block - 1) / dp->stats_block;
    ok = hipMalloc(&p->d_store[b], s.nchans * min_ld * storage_bytes(dtype)) == hipSuccess &&
         hipMalloc((void **)&p->d_out[b], (size_t)s.ndms * s.nsamp * 4) == hipSuccess &&
         hipMalloc((void **)&p->d_cmean[b], s.nchans * 8) == hipSuccess &&
         hipMalloc((void **)&p->d_cstd[b], s.nchans * 8) == hipSuccess &&
         hipMalloc((void **)&p->d_bmean[b], s.ndms * nb * 8) == hipSuccess &&
         hipMalloc((void **)&p->d_bstd[b], s.ndms * nb * 8) == hipSuccess &&
         hipMalloc((void **)&p->d_cands[b], (size_t)dp->cand_cap * sizeof(mdsm_cand_t)) == hipSuccess;
    if (ok) ok = hipMemset(p->d_store[b], 0, s.nchans * min_ld * storage_bytes(dtype)) == hipSuccess;
  }
  if (!ok) {
    set_error("pipeline allocation failed");
    free_pipeline(p);
    return MDSM_ERR_NOMEM;
  }
  std::vector<const void *> ins(p->d_store.begin(), p->d_store.end());
  std::vector<float *> outs(p->d_out.begin(), p->d_out.end());
  rc = mdsm_batch_bind(p->batch, ins.data(), p->ld_store.data(), outs.data());
  if (rc != MDSM_OK) {
    free_pipeline(p);
    return rc;
  }
  *out = p;
  return MDSM_OK;
}

extern "C" int mdsm_pipeline_destroy(mdsm_pipeline p) {
  if (p) free_pipeline(p);
  return MDSM_OK;
}

extern "C" int mdsm_pipeline_bind(mdsm_pipeline p, const float *const *d_raw, const uint64_t *ld_raw) {
  MDSM_REQUIRE(p && d_raw && ld_raw, MDSM_ERR_ARG, "null argument");
  const uint32_t nb = (uint32_t)p->beams.size();
  for (uint32_t b = 0; b < nb; ++b) {
    uint32_t maxshift = 0;
    mdsm_batch_info(p->batch, b, &maxshift, nullptr);
    MDSM_REQUIRE(d_raw[b] && ld_raw[b] >= (uint64_t)p->beams[b].nsamp + maxshift, MDSM_ERR_ARG,
                 "beam %u: raw rows shorter than nsamp + maxshift", b);
  }
  p->raw.assign(d_raw, d_raw + nb);
  p->ld_raw.assign(ld_raw, ld_raw + nb);
  p->bound = true;
  drop_graph(p);  // pointers changed: the next graph run records again
  return MDSM_OK;
}

extern "C" int mdsm_pipeline_set_variant(mdsm_pipeline p, int variant) {
  MDSM_REQUIRE(p, MDSM_ERR_ARG, "null pipeline");
  int rc = mdsm_batch_set_variant(p->batch, variant);
  if (rc != MDSM_OK) return rc;
  std::vector<const void *> ins(p->d_store.begin(), p->d_store.end());
  std::vector<float *> outs(p->d_out.begin(), p->d_out.end());
  drop_graph(p);
  return mdsm_batch_bind(p->batch, ins.data(), p->ld_store.data(), outs.data());
}

// the launch sequence of one chunk (no allocation, copy or memset)
static int enqueue(mdsm_pipeline_s *p, hipStream_t s) {
  const uint32_t nb = (uint32_t)p->beams.size();
  for (uint32_t b = 0; b < nb; ++b) {
    const auto &sv = p->beams[b];
    // statistics over the samples that reach the store (nsamp + maxshift,
    // rounded up to the storage alignment, within the raw rows)
    const uint64_t n = std::min<uint64_t>(p->ld_store[b], p->ld_raw[b]);
    int rc = launch_channel_stats(p->raw[b], p->ld_raw[b], sv.nchans, n, p->d_cmean[b], p->d_cstd[b], s);
    if (rc == MDSM_OK)
      rc = launch_normalise(p->raw[b], p->ld_raw[b], sv.nchans, n, p->d_cmean[b], p->d_cstd[b],
                            p->d_store[b], p->dtype, p->ld_store[b], s);
    if (rc != MDSM_OK) return rc;
  }
  int rc = mdsm_batch_launch(p->batch, s);
  if (rc != MDSM_OK) return rc;
  rc = launch_zero_u32(p->d_counts, nb, s);
  if (rc != MDSM_OK) return rc;
  for (uint32_t b = 0; b < nb; ++b) {
    const auto &sv = p->beams[b];
    rc = launch_block_stats(p->d_out[b], sv.ndms, sv.nsamp, p->dp.stats_block, p->d_bmean[b],
                            p->d_bstd[b], s);
    if (rc == MDSM_OK)
      rc = launch_detect(p->d_out[b], sv.ndms, sv.nsamp, p->dp.stats_block, p->d_bmean[b],
                         p->d_bstd[b], p->dp.nwidths, p->dp.thresh, p->d_cands[b], p->dp.cand_cap,
                         p->d_counts + b, s);
    if (rc != MDSM_OK) return rc;
  }
  return MDSM_OK;
}

extern "C" int mdsm_pipeline_run(mdsm_pipeline p, void *stream, int use_graph) {
  MDSM_REQUIRE(p && p->bound, MDSM_ERR_STATE, "pipeline not bound to raw inputs");
  hipStream_t s = as_stream(stream);
  if (!use_graph) return enqueue(p, s);
  if (!p->exec) {
    // record on a private stream (capture is not allowed on the null stream)
    if (!p->graph_stream) MDSM_HIP(hipStreamCreateWithFlags(&p->graph_stream, hipStreamNonBlocking));
    MDSM_HIP(hipStreamBeginCapture(p->graph_stream, hipStreamCaptureModeThreadLocal));
    int rc = enqueue(p, p->graph_stream);
    hipGraph_t g = nullptr;
    hipError_t e = hipStreamEndCapture(p->graph_stream, &g);
    if (rc != MDSM_OK) {
      if (g) (void)hipGraphDestroy(g);
      return rc;
    }
    MDSM_HIP(e);
    p->graph = g;
    MDSM_HIP(hipGraphInstantiate(&p->exec, p->graph, nullptr, nullptr, 0));
  }
  MDSM_HIP(hipGraphLaunch(p->exec, s));
  return MDSM_OK;
}

extern "C" int mdsm_pipeline_candidates(mdsm_pipeline p, uint32_t beam, mdsm_cand_t *host,
                                        uint32_t cap, uint32_t *n) {
  MDSM_REQUIRE(p && beam < p->beams.size() && n, MDSM_ERR_ARG, "bad arguments");
  uint32_t cnt = 0;
  MDSM_HIP(hipMemcpy(&cnt, p->d_counts + beam, 4, hipMemcpyDeviceToHost));
  *n = cnt;
  const uint32_t kept = std::min(cnt, p->dp.cand_cap);
  const uint32_t k = std::min(kept, cap);
  if (k) {
    std::vector<mdsm_cand_t> all(kept);
    MDSM_HIP(hipMemcpy(all.data(), p->d_cands[beam], kept * sizeof(mdsm_cand_t), hipMemcpyDeviceToHost));
    mdsm_sort_candidates(all.data(), kept);
    std::copy(all.begin(), all.begin() + k, host);
  }
  if (cnt > p->dp.cand_cap || cnt > cap) {
    set_error("beam %u: %u candidates exceed the capacity", beam, cnt);
    return MDSM_ERR_OVERFLOW;
  }
  return MDSM_OK;
}

extern "C" int mdsm_pipeline_output(mdsm_pipeline p, uint32_t beam, float **d_out) {
  MDSM_REQUIRE(p && beam < p->beams.size() && d_out, MDSM_ERR_ARG, "bad arguments");
  *d_out = p->d_out[beam];
  return MDSM_OK;
}
