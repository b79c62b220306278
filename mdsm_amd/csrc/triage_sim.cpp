// triage_sim.cpp -- synthetic triage streams (the plain search over seeded
// injections and RFI) and the training / inference loops of the triage
// network.
//
// Each chunk: host draws pulses (DM, width, S/N, arrival), zero-DM spikes
// and narrowband RFI items; the device generates the raw rows (Philox
// noise, bandpass gain/offset, dispersed pulses), adds the RFI, normalises
// per channel, dedisperses (FP32 plan, FP64 rounded delays), computes
// per-block statistics and boxcar candidates; the host sorts, clusters
// (2-D DBSCAN), takes cluster peaks, labels each peak by matching it to the
// injected pulses, and the device cuts the two snippets of every event.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <numeric>
#include <random>
#include <vector>

#include "triage_internal.h"

using namespace mdsm;

struct mdsm_triage_sim_s {
  mdsm_triage_config_t cfg{};
  uint64_t seed = 0, chunk = 0;
  mdsm_plan plan = nullptr;
  uint32_t maxshift = 0;
  uint64_t ld = 0;
  const int32_t *d_table = nullptr;
  float *d_raw = nullptr, *d_store = nullptr, *d_out = nullptr, *d_gain = nullptr, *d_off = nullptr;
  double *d_cmean = nullptr, *d_cstd = nullptr, *d_bmean = nullptr, *d_bstd = nullptr;
  mdsm_cand_t *d_cands = nullptr;
  uint32_t *d_count = nullptr, cap = 1u << 22;
  int64_t *d_pstart = nullptr, *d_pdelay = nullptr;
  int32_t *d_pwidth = nullptr, *d_evdm = nullptr, *d_evt = nullptr;
  float *d_pamp = nullptr;
  RfiItem *d_rfi = nullptr;
  uint32_t max_ev_chunk = 1u << 16;
  hipStream_t stream = nullptr;
  std::vector<void *> allocs;
};

namespace {

template <typename T>
bool dalloc(mdsm_triage_sim_s *s, T **p, size_t count) {
  void *q = nullptr;
  if (hipMalloc(&q, std::max<size_t>(count, 1) * sizeof(T)) != hipSuccess) return false;
  s->allocs.push_back(q);
  *p = static_cast<T *>(q);
  return true;
}

void free_sim(mdsm_triage_sim_s *s) {
  if (s->stream) (void)hipStreamSynchronize(s->stream);
  for (void *q : s->allocs) (void)hipFree(q);
  if (s->plan) mdsm_plan_destroy(s->plan);
  if (s->stream) (void)hipStreamDestroy(s->stream);
  delete s;
}

constexpr double KDM = 4148.808;

struct Pulse {
  double dm;
  int64_t start;
  int32_t width;
  double snr;
};

}  // namespace

extern "C" int mdsm_triage_sim_create(const mdsm_triage_config_t *cfg, uint64_t seed,
                                      mdsm_triage_sim *out) {
  MDSM_REQUIRE(cfg && out, MDSM_ERR_ARG, "null argument");
  MDSM_REQUIRE(cfg->snip_dms > 0 && cfg->snip_samples > 0 && cfg->snip_subbands == cfg->snip_dms &&
                   cfg->snip_dms % 8 == 0 && cfg->snip_samples % 8 == 0 &&
                   (uint32_t)cfg->snip_subbands <= cfg->survey.nchans,
               MDSM_ERR_ARG, "snippet shape: rows (DMs = subbands) and samples multiples of 8");
  MDSM_REQUIRE(cfg->width_lo >= 1 && cfg->width_hi >= cfg->width_lo && cfg->dm_hi >= cfg->dm_lo &&
                   cfg->dm_lo >= 0 && cfg->survey.nsamp > 2 * cfg->width_hi,
               MDSM_ERR_ARG, "bad pulse ranges");
  const double dm_max = cfg->survey.lowdm + cfg->survey.dmstep * (cfg->survey.ndms - 1);
  MDSM_REQUIRE(cfg->dm_hi <= dm_max, MDSM_ERR_ARG, "pulse DMs beyond the trial range");
  *out = nullptr;
  auto *s = new mdsm_triage_sim_s();
  s->cfg = *cfg;
  s->seed = seed;
  mdsm_survey_t sv = cfg->survey;
  sv.delay_mode = MDSM_DELAY_FP64_ROUND;
  s->cfg.survey = sv;
  int rc = mdsm_plan_create(&sv, MDSM_F32, &s->plan);
  if (rc != MDSM_OK) {
    free_sim(s);
    return rc;
  }
  mdsm_plan_info(s->plan, &s->maxshift, &s->ld);
  mdsm_plan_device_table(s->plan, &s->d_table);
  const uint32_t C = sv.nchans, M = sv.ndms, N = sv.nsamp;
  const size_t nb = (N + cfg->stats_block - 1) / cfg->stats_block;
  const uint32_t maxp = cfg->npulses + cfg->n_spikes;
  const uint32_t maxr = cfg->n_bursts + cfg->n_mains + cfg->n_telegraph;
  bool ok = hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking) == hipSuccess &&
            dalloc(s, &s->d_raw, (size_t)C * s->ld) && dalloc(s, &s->d_store, (size_t)C * s->ld) &&
            dalloc(s, &s->d_out, (size_t)M * N) && dalloc(s, &s->d_gain, C) && dalloc(s, &s->d_off, C) &&
            dalloc(s, &s->d_cmean, C) && dalloc(s, &s->d_cstd, C) && dalloc(s, &s->d_bmean, M * nb) &&
            dalloc(s, &s->d_bstd, M * nb) && dalloc(s, &s->d_cands, s->cap) && dalloc(s, &s->d_count, 1) &&
            dalloc(s, &s->d_pstart, maxp) && dalloc(s, &s->d_pdelay, (size_t)maxp * C) &&
            dalloc(s, &s->d_pwidth, maxp) && dalloc(s, &s->d_pamp, maxp) && dalloc(s, &s->d_rfi, maxr) &&
            dalloc(s, &s->d_evdm, s->max_ev_chunk) && dalloc(s, &s->d_evt, s->max_ev_chunk);
  if (!ok) {
    free_sim(s);
    set_error("triage stream allocation failed");
    return MDSM_ERR_NOMEM;
  }
  *out = s;
  return MDSM_OK;
}

extern "C" int mdsm_triage_sim_destroy(mdsm_triage_sim s) {
  if (s) free_sim(s);
  return MDSM_OK;
}

extern "C" int mdsm_inject_rfi(const mdsm_rfi_item_t *items, uint32_t nitems, const float *d_gain,
                               uint64_t seed, float *d_raw, uint64_t ld, uint64_t n, void *stream) {
  MDSM_REQUIRE((items || !nitems) && d_gain && d_raw, MDSM_ERR_ARG, "null argument");
  if (!nitems) return MDSM_OK;
  int64_t max_len = 0;
  for (uint32_t i = 0; i < nitems; ++i) {
    MDSM_REQUIRE(items[i].len >= 0 && items[i].kind >= 0 && items[i].kind <= 2 &&
                     (items[i].kind == MDSM_RFI_BURST || items[i].period >= 1.f),
                 MDSM_ERR_ARG, "bad RFI item %u", i);
    max_len = std::max(max_len, items[i].len);
  }
  RfiItem *d = nullptr;
  MDSM_HIP(hipMalloc(&d, sizeof(RfiItem) * nitems));
  hipStream_t s = as_stream(stream);
  int rc = MDSM_OK;
  if (hipMemcpyAsync(d, items, sizeof(RfiItem) * nitems, hipMemcpyHostToDevice, s) != hipSuccess)
    rc = MDSM_ERR_HIP;
  if (rc == MDSM_OK) rc = launch_rfi_inject(d, nitems, max_len, d_gain, seed, d_raw, ld, n, s);
  (void)hipStreamSynchronize(s);
  (void)hipFree(d);
  return rc;
}

extern "C" int mdsm_triage_snippets(const float *d_out, uint32_t ndms, uint32_t nsamp,
                                    const float *d_store, uint64_t ld, uint32_t nchans,
                                    const int32_t *d_table, const int32_t *ev_dm, const int32_t *ev_t,
                                    uint32_t nev, int SD, int ST, int SB, float scale0, float scale1,
                                    float clip, uint16_t *d_snip, void *stream) {
  MDSM_REQUIRE(d_out && d_store && d_table && d_snip && (nev == 0 || (ev_dm && ev_t)), MDSM_ERR_ARG,
               "null argument");
  MDSM_REQUIRE(SD > 0 && ST > 0 && SB == SD && (uint32_t)SB <= nchans && ld >= nsamp, MDSM_ERR_ARG,
               "bad snippet shape");
  for (uint32_t i = 0; i < nev; ++i)
    MDSM_REQUIRE(ev_dm[i] >= 0 && (uint32_t)ev_dm[i] < ndms, MDSM_ERR_ARG, "event %u: DM index out of range", i);
  if (!nev) return MDSM_OK;
  int32_t *d = nullptr;
  MDSM_HIP(hipMalloc(&d, (size_t)nev * 8));
  hipStream_t s = as_stream(stream);
  int rc = MDSM_OK;
  if (hipMemcpyAsync(d, ev_dm, nev * 4, hipMemcpyHostToDevice, s) != hipSuccess ||
      hipMemcpyAsync(d + nev, ev_t, nev * 4, hipMemcpyHostToDevice, s) != hipSuccess)
    rc = MDSM_ERR_HIP;
  if (rc == MDSM_OK)
    rc = launch_snippets(d_out, ndms, nsamp, d_store, ld, nchans, d_table, d, d + nev, nev, SD, ST, SB,
                         scale0, scale1, clip, d_snip, s);
  (void)hipStreamSynchronize(s);
  (void)hipFree(d);
  return rc;
}

extern "C" int mdsm_triage_sim_run(mdsm_triage_sim s, uint32_t nchunks, uint16_t *d_snip,
                                   mdsm_event_t *events, uint32_t max_events, uint32_t *n_events,
                                   mdsm_triage_stats_t *stats) {
  MDSM_REQUIRE(s && d_snip && events && n_events && stats, MDSM_ERR_ARG, "null argument");
  const mdsm_triage_config_t &cf = s->cfg;
  const mdsm_survey_t &sv = cf.survey;
  const uint32_t C = sv.nchans, M = sv.ndms, N = sv.nsamp;
  const uint64_t nraw = (uint64_t)N + s->maxshift;
  const uint64_t snip_elems = 2ull * cf.snip_dms * cf.snip_samples;
  hipStream_t st = s->stream;
  std::vector<double> finv2(C);
  for (uint32_t c = 0; c < C; ++c) {
    const double f = sv.fch1 + sv.foff * c;
    finv2[c] = 1.0 / (f * f) - 1.0 / (sv.fch1 * sv.fch1);
  }
  std::vector<mdsm_cand_t> cands;
  std::vector<int32_t> labels;
  for (uint32_t k = 0; k < nchunks && *n_events < max_events; ++k, ++s->chunk) {
    std::mt19937_64 rng(s->seed * 0x9E3779B97F4A7C15ull + s->chunk * 0xD1B54A32D192ED03ull + 1);
    std::uniform_real_distribution<double> U(0.0, 1.0);
    auto logu = [&](double lo, double hi) { return std::exp(std::log(lo) + U(rng) * (std::log(hi) - std::log(lo))); };
    // ---- injections
    std::vector<Pulse> pulses;
    for (uint32_t p = 0; p < cf.npulses; ++p) {
      Pulse q;
      q.dm = cf.dm_lo + U(rng) * (cf.dm_hi - cf.dm_lo);
      q.width = std::max(1, (int32_t)std::lround(logu(cf.width_lo, cf.width_hi)));
      q.snr = cf.snr_lo + U(rng) * (cf.snr_hi - cf.snr_lo);
      q.start = (int64_t)(U(rng) * (double)(N - q.width));
      pulses.push_back(q);
    }
    const uint32_t np = cf.npulses + cf.n_spikes;
    std::vector<int64_t> pstart(np), pdelay((size_t)np * C, 0);
    std::vector<int32_t> pwidth(np);
    std::vector<float> pamp(np);
    for (uint32_t p = 0; p < np; ++p) {
      if (p < cf.npulses) {
        const Pulse &q = pulses[p];
        pstart[p] = q.start;
        pwidth[p] = q.width;
        pamp[p] = (float)(q.snr / std::sqrt((double)C * q.width));
        for (uint32_t c = 0; c < C; ++c)
          pdelay[(size_t)p * C + c] = (int64_t)std::llround(KDM * q.dm * finv2[c] / sv.tsamp);
      } else {  // zero-DM broadband spike
        const int32_t w = std::max(1, (int32_t)std::lround(logu(1.0, 8.0)));
        pwidth[p] = w;
        pstart[p] = (int64_t)(U(rng) * (double)(nraw - w));
        pamp[p] = (float)((cf.rfi_lo + U(rng) * (cf.rfi_hi - cf.rfi_lo)) / std::sqrt((double)C * w));
      }
    }
    std::vector<RfiItem> rfi;
    auto add_rfi = [&](int kind, uint32_t count) {
      for (uint32_t i = 0; i < count; ++i) {
        RfiItem it{};
        it.chan = (uint32_t)(U(rng) * C) % C;
        it.kind = kind;
        it.amp = (float)(cf.rfi_lo + U(rng) * (cf.rfi_hi - cf.rfi_lo));
        if (kind == RFI_BURST) {
          it.len = (int64_t)logu(16.0, 4096.0);
          it.start = (int64_t)(U(rng) * (double)(nraw - it.len));
          it.period = 1.f;
        } else {
          it.start = 0;
          it.len = (int64_t)nraw;
          it.period = kind == RFI_MAINS ? (float)(1.0 / (50.0 * sv.tsamp)) : 64.f;
        }
        rfi.push_back(it);
      }
    };
    add_rfi(RFI_BURST, cf.n_bursts);
    add_rfi(RFI_MAINS, cf.n_mains);
    add_rfi(RFI_TELEGRAPH, cf.n_telegraph);
    std::vector<float> gain(C), off(C);
    for (uint32_t c = 0; c < C; ++c) {
      gain[c] = (float)(0.5 + 1.5 * U(rng));
      off[c] = (float)(10.0 * U(rng));
    }
    const uint64_t gseed = rng();
    // ---- device: generate, inject, normalise, dedisperse, detect
    MDSM_HIP(hipMemcpyAsync(s->d_gain, gain.data(), C * 4, hipMemcpyHostToDevice, st));
    MDSM_HIP(hipMemcpyAsync(s->d_off, off.data(), C * 4, hipMemcpyHostToDevice, st));
    MDSM_HIP(hipMemcpyAsync(s->d_pstart, pstart.data(), np * 8, hipMemcpyHostToDevice, st));
    MDSM_HIP(hipMemcpyAsync(s->d_pwidth, pwidth.data(), np * 4, hipMemcpyHostToDevice, st));
    MDSM_HIP(hipMemcpyAsync(s->d_pamp, pamp.data(), np * 4, hipMemcpyHostToDevice, st));
    MDSM_HIP(hipMemcpyAsync(s->d_pdelay, pdelay.data(), pdelay.size() * 8, hipMemcpyHostToDevice, st));
    int64_t max_len = 0;
    for (auto &it : rfi) max_len = std::max(max_len, it.len);
    if (!rfi.empty())
      MDSM_HIP(hipMemcpyAsync(s->d_rfi, rfi.data(), rfi.size() * sizeof(RfiItem), hipMemcpyHostToDevice, st));
    int rc = launch_generate(gseed, C, 0, nraw, s->d_gain, s->d_off, np, s->d_pstart, s->d_pwidth,
                             s->d_pamp, s->d_pdelay, s->d_raw, s->ld, st);
    if (rc == MDSM_OK)
      rc = launch_rfi_inject(s->d_rfi, (uint32_t)rfi.size(), max_len, s->d_gain, gseed ^ 0x5A5A5A5Aull,
                             s->d_raw, s->ld, nraw, st);
    if (rc == MDSM_OK) rc = launch_channel_stats(s->d_raw, s->ld, C, nraw, s->d_cmean, s->d_cstd, st);
    if (rc == MDSM_OK)
      rc = launch_normalise(s->d_raw, s->ld, C, nraw, s->d_cmean, s->d_cstd, s->d_store, MDSM_F32, s->ld, st);
    if (rc == MDSM_OK) rc = mdsm_dedisperse(s->plan, s->d_store, s->ld, s->d_out, st);
    if (rc == MDSM_OK) rc = launch_block_stats(s->d_out, M, N, cf.stats_block, s->d_bmean, s->d_bstd, st);
    if (rc == MDSM_OK) rc = launch_zero_u32(s->d_count, 1, st);
    if (rc == MDSM_OK)
      rc = launch_detect(s->d_out, M, N, cf.stats_block, s->d_bmean, s->d_bstd, cf.nwidths, cf.thresh,
                         s->d_cands, s->cap, s->d_count, st);
    if (rc != MDSM_OK) return rc;
    uint32_t ncand = 0;
    MDSM_HIP(hipMemcpyAsync(&ncand, s->d_count, 4, hipMemcpyDeviceToHost, st));
    MDSM_HIP(hipStreamSynchronize(st));
    ncand = std::min(ncand, s->cap);
    cands.resize(ncand);
    if (ncand) MDSM_HIP(hipMemcpy(cands.data(), s->d_cands, ncand * sizeof(mdsm_cand_t), hipMemcpyDeviceToHost));
    mdsm_sort_candidates(cands.data(), ncand);
    labels.assign(ncand, -1);
    uint32_t ncl = 0;
    if (ncand) mdsm_dbscan2d(cands.data(), ncand, cf.eps, cf.minpts, labels.data(), &ncl);
    std::vector<mdsm_cand_t> peaks(ncl);
    std::vector<uint32_t> sizes(ncl);
    if (ncl) mdsm_cluster_peaks(cands.data(), labels.data(), ncand, ncl, peaks.data(), sizes.data());
    // ---- label the cluster peaks against the injected pulses
    std::vector<char> found(cf.npulses, 0);
    const uint32_t room = std::min<uint32_t>(max_events - *n_events, s->max_ev_chunk);
    std::vector<int32_t> evdm, evt;
    uint64_t false_cl = 0;
    for (uint32_t i = 0; i < ncl; ++i) {
      const mdsm_cand_t &pk = peaks[i];
      const double dm = sv.lowdm + sv.dmstep * pk.dm_idx;
      const double tc = pk.sample + 0.5 * (1 << pk.width_log2);
      int32_t match = -1;
      for (uint32_t p = 0; p < cf.npulses && match < 0; ++p) {
        const Pulse &q = pulses[p];
        const double tol_dm = std::max(5.0 * sv.dmstep, 0.1 * q.dm);
        const double dt = std::fabs(tc - (q.start + 0.5 * q.width));
        if (std::fabs(dm - q.dm) <= tol_dm && dt <= q.width + (1 << pk.width_log2) + 16) match = (int32_t)p;
      }
      if (match >= 0) found[match] = 1;
      else ++false_cl;
      if (evdm.size() < room) {
        mdsm_event_t &e = events[*n_events + evdm.size()];
        e.dm_idx = pk.dm_idx;
        e.sample = pk.sample;
        e.width_log2 = pk.width_log2;
        e.snr = pk.snr;
        e.cluster_size = (int32_t)sizes[i];
        e.label = match >= 0 ? 1 : 0;
        e.pulse = match;
        e.chunk = (int32_t)s->chunk;
        evdm.push_back(pk.dm_idx);
        evt.push_back((int32_t)std::lround(tc));
      }
    }
    stats->chunks += 1;
    stats->injected += cf.npulses;
    stats->injected_found += (uint64_t)std::count(found.begin(), found.end(), 1);
    stats->clusters += ncl;
    stats->false_clusters += false_cl;
    stats->candidates += ncand;
    // ---- snippets
    const uint32_t ne = (uint32_t)evdm.size();
    if (ne) {
      MDSM_HIP(hipMemcpyAsync(s->d_evdm, evdm.data(), ne * 4, hipMemcpyHostToDevice, st));
      MDSM_HIP(hipMemcpyAsync(s->d_evt, evt.data(), ne * 4, hipMemcpyHostToDevice, st));
      const float scale0 = (float)(1.0 / std::sqrt((double)C));
      const float scale1 = (float)(1.0 / std::sqrt((double)C / cf.snip_subbands));
      rc = launch_snippets(s->d_out, M, N, s->d_store, s->ld, C, s->d_table, s->d_evdm, s->d_evt, ne,
                           cf.snip_dms, cf.snip_samples, cf.snip_subbands, scale0, scale1, 20.f,
                           d_snip + (uint64_t)*n_events * snip_elems, st);
      if (rc != MDSM_OK) return rc;
      MDSM_HIP(hipStreamSynchronize(st));
      *n_events += ne;
    }
  }
  MDSM_HIP(hipStreamSynchronize(st));
  return MDSM_OK;
}

// ---------------------------------------------------------------- training
extern "C" int mdsm_cnn_train_epoch(mdsm_cnn net, const uint16_t *d_snip, const int32_t *d_labels,
                                    uint32_t n, uint32_t batch, float lr, uint64_t seed,
                                    float *mean_loss, void *stream) {
  MDSM_REQUIRE(net && d_snip && d_labels && n && batch && mean_loss, MDSM_ERR_ARG, "bad arguments");
  hipStream_t st = as_stream(stream);
  int H = 0, W = 0, bmax = 0;
  mdsm_cnn_shape(net, &H, &W, &bmax);
  MDSM_REQUIRE((int)batch <= bmax, MDSM_ERR_ARG, "batch exceeds the network's max_batch");
  const uint64_t row = 2ull * H * W;
  std::vector<uint32_t> perm(n);
  std::iota(perm.begin(), perm.end(), 0u);
  std::mt19937_64 rng(seed);
  std::shuffle(perm.begin(), perm.end(), rng);
  const uint32_t nb = (n + batch - 1) / batch;
  uint32_t *d_idx = nullptr;
  uint16_t *d_x = nullptr;
  int32_t *d_y = nullptr;
  float *d_losses = nullptr;
  int rc = MDSM_OK;
  if (hipMalloc(&d_idx, (size_t)n * 4) != hipSuccess || hipMalloc(&d_x, (size_t)batch * row * 2) != hipSuccess ||
      hipMalloc(&d_y, (size_t)batch * 4) != hipSuccess || hipMalloc(&d_losses, (size_t)nb * 4) != hipSuccess)
    rc = MDSM_ERR_NOMEM;
  if (rc == MDSM_OK && hipMemcpyAsync(d_idx, perm.data(), (size_t)n * 4, hipMemcpyHostToDevice, st) != hipSuccess)
    rc = MDSM_ERR_HIP;
  for (uint32_t b = 0; b < nb && rc == MDSM_OK; ++b) {
    const uint32_t i0 = b * batch, cnt = std::min(batch, n - i0);
    rc = launch_gather_rows(d_snip, d_labels, d_idx + i0, cnt, row, d_x, d_y, st);
    if (rc == MDSM_OK) rc = mdsm_cnn_train_step(net, d_x, d_y, cnt, lr, d_losses + b, st);
  }
  std::vector<float> losses(nb, 0.f);
  if (rc == MDSM_OK && hipMemcpyAsync(losses.data(), d_losses, (size_t)nb * 4, hipMemcpyDeviceToHost, st) != hipSuccess)
    rc = MDSM_ERR_HIP;
  (void)hipStreamSynchronize(st);
  (void)hipFree(d_idx);
  (void)hipFree(d_x);
  (void)hipFree(d_y);
  (void)hipFree(d_losses);
  if (rc == MDSM_ERR_NOMEM) set_error("training buffers could not be allocated");
  if (rc != MDSM_OK) return rc;
  double sum = 0;
  for (float l : losses) sum += l;
  *mean_loss = (float)(sum / nb);
  return MDSM_OK;
}

extern "C" int mdsm_cnn_predict(mdsm_cnn net, const uint16_t *d_snip, uint32_t n, float *d_prob,
                                void *stream) {
  MDSM_REQUIRE(net && d_snip && d_prob, MDSM_ERR_ARG, "null argument");
  hipStream_t st = as_stream(stream);
  int H = 0, W = 0, bmax = 0;
  mdsm_cnn_shape(net, &H, &W, &bmax);
  const uint64_t row = 2ull * H * W;
  float *d_logits = nullptr;
  MDSM_HIP(hipMalloc(&d_logits, (size_t)bmax * 8));
  int rc = MDSM_OK;
  for (uint32_t i0 = 0; i0 < n && rc == MDSM_OK; i0 += bmax) {
    const uint32_t cnt = std::min<uint32_t>(bmax, n - i0);
    rc = mdsm_cnn_forward(net, d_snip + (uint64_t)i0 * row, cnt, d_logits, st);
    if (rc == MDSM_OK) rc = launch_probabilities(d_logits, cnt, d_prob + i0, st);
  }
  (void)hipStreamSynchronize(st);
  (void)hipFree(d_logits);
  return rc;
}
