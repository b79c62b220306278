// compat.cpp -- the MDSM multibeam manager API (initialiseMDSM,
// get_buffer_pointer, next_chunk_multibeam, start_processing, tearDownMDSM)
// on top of the C ABI, for drivers written against
// src/mdsm-multibeam/multibeam_dedispersion_manager.h:9-28.
//
// The reference overlaps read(k+1), GPU(k) and detect(k-1) with one pthread
// per beam, an output thread and two barriers
// (multibeam_dedispersion_manager.cpp:467-590,
// multibeam_dedispersion_thread.cu:464-839,
// multibeam_dedispersion_output.cpp:325-452).  On one GPU the same overlap
// comes from three HIP streams ordered by events, driven from the caller's
// thread:
//   s_in   chunk k: overlap-save tail copy, upload of the new samples of
//          every beam, voltage power / RFI clipping, and for FP16 / U8
//          storage the per-chunk bandpass normalisation;
//   s_dd   chunk k: one batched dedispersion launch over all beams, then
//          median filter / detrending;
//   s_out  chunk k: the reference detector (global FP64 mean/rms and
//          threshold) on the device, then the download of the dedispersed
//          series into the pinned output ring (skipped, with no ring, when
//          SURVEY::series_on_device is set).
// Device inputs and outputs are double-buffered by chunk parity, host
// outputs form a ring of three (a returned buffer stays valid until the
// next next_chunk_multibeam call).  start_processing(k) enqueues chunk k and
// then finishes chunk k-1 on the host (candidate download, reference
// DBSCAN / classification, CSV lines) while the GPU runs chunk k.
#include <hip/hip_runtime.h>
#include <unistd.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <ctime>
#include <string>
#include <vector>

#include "internal.h"
#include "mdsm_compat.h"

extern "C" const char *mdsm_compat_xml_error(void);

namespace {

constexpr int NRING = 3;  // host output ring (chunk j in slot j % 3)

struct BeamState {
  mdsm_survey_t s{};
  FILE *fp = nullptr;
  unsigned clusters_written = 0;
  mdsm_cand_t *d_cands[2] = {nullptr, nullptr};  // per chunk parity: cap entries, then the counter
  uint32_t cap[2] = {0, 0};
  std::vector<double> ctime, cdm, csnr;
  std::vector<int> ccl;
};

struct ChunkInfo {
  unsigned data_read = 0;
  double timestamp = 0, blockRate = 0;
};

struct State {
  SURVEY *survey = nullptr;
  unsigned nsamp = 0, maxshift = 0, nchans = 0, tdms = 0, nbeams = 0;
  uint64_t ld = 0;         // input row length (>= nsamp + maxshift, multiple of 4)
  int storage = MDSM_F32;  // dedisperser input storage
  std::vector<BeamState> beams;
  float *d_raw[2] = {nullptr, nullptr};   // [beam][chan][ld] float rows, per chunk parity
  void *d_store[2] = {nullptr, nullptr};  // storage rows (the float rows for FP32)
  float *d_out[2] = {nullptr, nullptr};   // [beam][tdms][nsamp]
  float *d_tmp = nullptr;                 // median filter output [tdms][nsamp]
  double *d_fitted = nullptr;             // fitted bandpass of the beam being cleaned
  double *d_cmean = nullptr, *d_cstd = nullptr;  // channel statistics for normalisation
  uint32_t *h_count = nullptr;            // pinned [2][nbeams] candidate counts
  mdsm_batch batch[2] = {nullptr, nullptr};
  float *input[MDSM_STAGES] = {nullptr, nullptr, nullptr};  // pinned [beam][chan][nsamp]
  float **out_ring[NRING] = {nullptr, nullptr, nullptr};    // per beam [tdms][nsamp] (pinned)
  hipStream_t s_in = nullptr, s_dd = nullptr, s_out = nullptr, s_aux = nullptr;
  hipEvent_t ev_in[2] = {}, ev_dd[2] = {}, ev_det[2] = {}, ev_out[2] = {};
  hipEvent_t ev_h2d[MDSM_STAGES] = {}, ev_ring[NRING] = {};
  ChunkInfo chunk[NRING];
  std::vector<int32_t> mask_from, mask_to;
  int loop_counter = 0;
  bool out_switch = true;
  unsigned pnsamp = 0, ppnsamp = 0;
  double timestamp = 0, blockRate = 0;
  int pending = 0;  // chunk whose host-side detection is outstanding (0: none)
  double written_samples = 0;
  bool ok = false;
} g;

thread_local std::string g_err;

void fail(const char *fmt, ...) {
  char b[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(b, sizeof(b), fmt, ap);
  va_end(ap);
  g_err = b;
  fprintf(stderr, "mdsm: %s\n", b);
}

bool hip_ok(hipError_t e, const char *what) {
  if (e == hipSuccess) return true;
  fail("%s: %s", what, hipGetErrorString(e));
  return false;
}

bool abi_ok(int rc, const char *what, unsigned beam) {
  if (rc == MDSM_OK) return true;
  fail("beam %u: %s: %s", beam, what, mdsm_last_error());
  return false;
}

// reference dmdelay (multibeam_dedispersion_manager.cpp:325-326)
inline float dmdelay(float F1, float F2) {
  return (float)(4148.741601 * ((1.0 / F1 / F1) - (1.0 / F2 / F2)));
}

// Candidate files (create_files, multibeam_dedispersion_output.cpp:21-86):
// single-file mode opens <basedir>/<prefix>[_beam_<b>].dat once; multi-file
// mode opens <basedir>/<prefix>_beam_<b>_<time>.dat when a file period
// starts, timestamped with the PC clock or the chunk's timestamp.
bool open_files(double timestamp) {
  SURVEY *s = g.survey;
  for (unsigned b = 0; b < g.nbeams; ++b) {
    char path[512];
    if (s->single_file_mode) {
      if (s->nbeams > 1) snprintf(path, sizeof(path), "%s/%s_beam_%u.dat", s->basedir, s->fileprefix, b);
      else snprintf(path, sizeof(path), "%s/%s.dat", s->basedir, s->fileprefix);
    } else {
      time_t now = s->use_pc_time ? time(nullptr) : (time_t)timestamp;
      char ts[32];
      strftime(ts, sizeof(ts), "%F_%T", localtime(&now));
      snprintf(path, sizeof(path), "%s/%s_beam_%u_%s.dat", s->basedir, s->fileprefix, b, ts);
    }
    g.beams[b].fp = fopen(path, "w");
    if (!g.beams[b].fp) {
      fail("Invalid output file path: %s", path);
      return false;
    }
  }
  return true;
}

void close_files() {
  for (auto &b : g.beams)
    if (b.fp) {
      fclose(b.fp);
      b.fp = nullptr;
    }
}

// Candidate lines of one beam's chunk in the reference's formats
// (multibeam_dedispersion_output.cpp:179 and :262-312).
void write_candidates(unsigned b, const ChunkInfo &ck, const std::vector<mdsm_cand_t> &c) {
  SURVEY *s = g.survey;
  BeamState &bs = g.beams[b];
  const uint32_t n = (uint32_t)c.size();
  bs.ctime.resize(n);
  bs.cdm.resize(n);
  bs.csnr.resize(n);
  bs.ccl.assign(n, 0);
  for (uint32_t i = 0; i < n; ++i) {
    bs.ctime[i] = ck.timestamp + c[i].sample * ck.blockRate;
    bs.cdm[i] = (double)(s->lowdm + c[i].dm_idx * s->dmstep);
    bs.csnr[i] = c[i].snr;
  }
  if (!bs.fp) return;
  if (!s->apply_clustering) {
    for (uint32_t i = 0; i < n; ++i)
      fprintf(bs.fp, "%lf, %f, %f\n", bs.ctime[i], (float)bs.cdm[i], (float)bs.csnr[i]);
  } else {
    uint32_t nc = 0;
    mdsm_dbscan_ref(bs.ctime.data(), bs.cdm.data(), bs.csnr.data(), n,
                    (float)(s->dbscan_time_range * ck.blockRate), s->dbscan_dm_range,
                    s->dbscan_snr_range, s->dbscan_min_points, bs.ccl.data(), &nc);
    // clusters in id order, then noise; with classification a cluster
    // scoring >= 0.1 is written with id -2
    std::vector<double> ct, cd, cs;
    for (uint32_t id = 1; id <= nc; ++id) {
      int written_id = (int)id + (int)bs.clusters_written;
      if (s->apply_classification) {
        ct.clear();
        cd.clear();
        cs.clear();
        for (uint32_t i = 0; i < n; ++i)
          if (bs.ccl[i] == (int)id) {
            ct.push_back(bs.ctime[i]);
            cd.push_back(bs.cdm[i]);
            cs.push_back(bs.csnr[i]);
          }
        float score = 1.f;
        if (mdsm_classify_cluster_ref(ct.data(), cd.data(), cs.data(), (uint32_t)ct.size(), s->lowdm,
                                      s->dmstep, s->tdms, s->min_pulse_width, s->beams[b].fch1,
                                      s->beams[b].foff, s->nchans, &score, nullptr) != MDSM_OK)
          score = 1.f;
        if (!(score < 0.1)) written_id = -2;
      }
      for (uint32_t i = 0; i < n; ++i)
        if (bs.ccl[i] == (int)id)
          fprintf(bs.fp, "%lf,%f,%f,%d,%d\n", bs.ctime[i], (float)bs.cdm[i], (float)bs.csnr[i],
                  written_id, 0);
    }
    for (uint32_t i = 0; i < n; ++i)
      if (bs.ccl[i] == -1)
        fprintf(bs.fp, "%lf,%f,%f,%d,%d\n", bs.ctime[i], (float)bs.cdm[i], (float)bs.csnr[i], -1, 0);
    bs.clusters_written += nc;
  }
  fflush(bs.fp);
}

float *out_ptr(int parity, unsigned b) { return g.d_out[parity] + (size_t)b * g.tdms * g.nsamp; }

void *store_ptr(int parity, unsigned b) {
  return static_cast<char *>(g.d_store[parity]) +
         (size_t)b * g.nchans * g.ld * mdsm::storage_bytes(g.storage);
}

float *raw_ptr(int parity, unsigned b) { return g.d_raw[parity] + (size_t)b * g.nchans * g.ld; }

// Candidate buffer of a parity and beam: cap candidates, then the counter.
bool alloc_cands(BeamState &bs, int parity, uint32_t cap) {
  if (bs.d_cands[parity]) (void)hipFree(bs.d_cands[parity]);
  bs.d_cands[parity] = nullptr;
  bs.cap[parity] = 0;
  if (!hip_ok(hipMalloc((void **)&bs.d_cands[parity], sizeof(mdsm_cand_t) * ((size_t)cap + 1)),
              "candidate buffer"))
    return false;
  bs.cap[parity] = cap;
  return true;
}

// The reference detector on one beam's chunk, enqueued on stream st:
// columns [0, ncols) of the [tdms][nsamp] series, statistics over the first
// stats_cols of them; the count is copied into the pinned h_count.
// Detrended series are thresholded at mean 0 and rms 1 (process_brute,
// multibeam_dedispersion_output.cpp:165-166; the clustering variant's 0 / 0
// at :201-205 would divide by zero and is not kept).
bool enqueue_detect(int parity, unsigned b, unsigned ncols, unsigned stats_cols, hipStream_t st) {
  BeamState &bs = g.beams[b];
  uint32_t *cnt = g.h_count + parity * g.nbeams + b;
  uint32_t *d_count = reinterpret_cast<uint32_t *>(bs.d_cands[parity] + bs.cap[parity]);
  int rc;
  if (g.survey->apply_detrending)
    rc = mdsm_threshold_global(out_ptr(parity, b), g.tdms, ncols, g.nsamp, 0.0, 1.0,
                               g.survey->detection_threshold, bs.d_cands[parity], bs.cap[parity],
                               d_count, st);
  else
    rc = mdsm_detect_global(out_ptr(parity, b), g.tdms, ncols, g.nsamp, stats_cols,
                            g.survey->detection_threshold, bs.d_cands[parity], bs.cap[parity],
                            d_count, nullptr, nullptr, st);
  return abi_ok(rc, "detection", b) &&
         hip_ok(hipMemcpyAsync(cnt, d_count, 4, hipMemcpyDeviceToHost, st), "count download");
}

// Host side of chunk j: candidate download, clustering and file output.
// last: the final chunk of the stream; when the caller set last_nsamp below
// the chunk's length, the statistics are taken over the first last_nsamp
// samples of each row, as the reference does for its last buffer
// (multibeam_dedispersion_output.cpp:216-229).
bool finish_chunk(int j, bool last) {
  SURVEY *s = g.survey;
  const int p = j & 1;
  const ChunkInfo ck = g.chunk[j % NRING];
  if (!hip_ok(hipEventSynchronize(g.ev_det[p]), "detection")) return false;
  const bool short_stats =
      last && s->last_nsamp > 0 && s->last_nsamp < ck.data_read && !s->apply_detrending;
  const unsigned stats_cols = short_stats ? s->last_nsamp : ck.data_read;
  if (!s->single_file_mode && g.written_samples == 0 && !open_files(ck.timestamp)) return false;
  for (unsigned b = 0; b < g.nbeams; ++b) {
    BeamState &bs = g.beams[b];
    uint32_t n = g.h_count[p * g.nbeams + b];
    // run again on the auxiliary stream when the statistics change or the
    // buffer overflowed (the chunk's series stay on the device until chunk
    // j + 2 is enqueued), with a buffer that holds every candidate
    for (int pass = 0; pass < 2 && (n > bs.cap[p] || (short_stats && pass == 0)); ++pass) {
      if (n > bs.cap[p] && !alloc_cands(bs, p, std::max(n, 2 * bs.cap[p]))) return false;
      if (!enqueue_detect(p, b, ck.data_read, stats_cols, g.s_aux) ||
          !hip_ok(hipStreamSynchronize(g.s_aux), "detection"))
        return false;
      n = g.h_count[p * g.nbeams + b];
    }
    if (n > bs.cap[p]) {
      fail("beam %u: candidate buffer overflow (%u)", b, n);
      return false;
    }
    std::vector<mdsm_cand_t> c(n);
    if (n && (!hip_ok(hipMemcpyAsync(c.data(), bs.d_cands[p], n * sizeof(mdsm_cand_t),
                                     hipMemcpyDeviceToHost, g.s_aux), "candidates") ||
              !hip_ok(hipStreamSynchronize(g.s_aux), "candidates")))
      return false;
    mdsm_sort_candidates(c.data(), n);
    write_candidates(b, ck, c);
  }
  // file rotation (multibeam_dedispersion_output.cpp:403-406)
  g.written_samples += ck.data_read;
  if (!s->single_file_mode && g.written_samples * ck.blockRate > s->secs_per_file) {
    g.written_samples = 0;
    close_files();
  }
  return true;
}

// Input conditioning of one beam's rows over [col0, col0 + n): complex
// voltages to power, then bandpass fit and RFI clipping
// (multibeam_dedispersion_thread.cu:700-752).  The reference clears the
// channel flags before channel_clipper runs (:364), so only spectrum
// clipping changes the data; the flags are not computed here.
bool clean_input(float *rows, unsigned b, uint64_t col0, uint64_t n) {
  SURVEY *s = g.survey;
  if (n == 0) return true;
  if (s->voltage &&
      !abi_ok(mdsm_voltage_to_power(rows, g.ld, g.nchans, col0, n, g.s_in), "voltage to power", b))
    return false;
  if (!s->apply_rfi_clipper) return true;
  mdsm_rfi_params_t prm{};
  prm.ncoeffs = s->ncoeffs;
  prm.channel_block = s->channel_block;
  prm.channel_thresh = s->channel_thresh;
  prm.spectrum_thresh = s->spectrum_thresh;
  prm.nmasks = (uint32_t)g.mask_from.size();
  prm.mask_from = g.mask_from.data();
  prm.mask_to = g.mask_to.data();
  prm.apply_channel_clip = 0;
  mdsm_bandpass_stats_t st{};
  if (!abi_ok(mdsm_rfi_clean(rows, g.ld, g.nchans, col0, n, &prm, g.d_fitted, nullptr, &st, g.s_in),
              "RFI clipping", b))
    return false;
  s->bandpass_mean = st.mean;
  s->bandpass_std = st.std;
  s->bandpass_rmse = st.rmse;
  return true;
}

// Median filter and detrending of the dedispersed series (thread.cu:766-772).
bool filter_output(int parity, unsigned b) {
  SURVEY *s = g.survey;
  float *y = out_ptr(parity, b);
  if (s->apply_median_filter) {
    if (!abi_ok(mdsm_median_filter5(y, g.tdms, g.nsamp, g.d_tmp, g.s_dd), "median filter", b) ||
        !hip_ok(hipMemcpyAsync(y, g.d_tmp, (size_t)g.tdms * g.nsamp * 4, hipMemcpyDeviceToDevice,
                               g.s_dd), "median copy"))
      return false;
  }
  if (s->apply_detrending &&
      !abi_ok(mdsm_detrend_normalise(y, g.tdms, g.nsamp, g.s_dd), "detrending", b))
    return false;
  return true;
}

// Binds the two parities' batches to their buffers; the tile shape is timed
// once on the (zeroed) buffers unless MDSM_COMPAT_NO_TUNE is set.
bool bind_batches() {
  std::vector<const void *> in(g.nbeams);
  std::vector<uint64_t> lds(g.nbeams, g.ld);
  std::vector<float *> out(g.nbeams);
  for (int p = 0; p < 2; ++p) {
    for (unsigned b = 0; b < g.nbeams; ++b) {
      in[b] = store_ptr(p, b);
      out[b] = out_ptr(p, b);
    }
    if (p == 0 && !getenv("MDSM_COMPAT_NO_TUNE")) {
      int v = 0;
      float ms = 0;
      if (!abi_ok(mdsm_batch_autotune(g.batch[0], in.data(), lds.data(), out.data(), g.s_dd, &v, &ms),
                  "autotune", 0) ||
          !abi_ok(mdsm_batch_set_variant(g.batch[1], v), "variant", 0))
        return false;
    }
    if (!abi_ok(mdsm_batch_bind(g.batch[p], in.data(), lds.data(), out.data()), "bind", 0))
      return false;
  }
  return true;
}

}  // namespace

extern "C" const char *mdsm_compat_last_error(void) { return g_err.c_str(); }

extern "C" unsigned mdsm_compat_candidates(unsigned beam, const double **time, const double **dm,
                                           const double **snr, const int **cluster) {
  if (beam >= g.beams.size()) return 0;
  BeamState &b = g.beams[beam];
  if (time) *time = b.ctime.data();
  if (dm) *dm = b.cdm.data();
  if (snr) *snr = b.csnr.data();
  if (cluster) *cluster = b.ccl.data();
  return (unsigned)b.ctime.size();
}

extern "C" int mdsm_compat_storage(void) { return g.storage; }

void initialiseMDSM(SURVEY *survey) {
  g = State();
  g.survey = survey;
  if (!survey || survey->nbeams == 0 || survey->nchans == 0 || survey->tdms == 0) {
    fail("invalid survey (no beams, channels or DMs)");
    return;
  }
  const unsigned st = survey->input_storage;
  if (!mdsm::known_dtype((int)st)) {
    fail("unknown input storage %u", st);
    return;
  }
  g.storage = (int)st;
  g.nchans = survey->nchans;
  g.tdms = survey->tdms;
  g.nbeams = survey->nbeams;
  // the output directory must exist and be writable (the reference exits
  // when a candidate file cannot be created, multibeam_dedispersion_output.cpp:41-45)
  if (access(survey->basedir, W_OK | X_OK) != 0) {
    fail("Invalid output directory: %s", survey->basedir);
    return;
  }
  // per-beam dm_shifts and the common maxshift, rounded up to a multiple of
  // 256 once per beam as the reference does (manager.cpp:362-387)
  unsigned greatest = 0;
  for (unsigned i = 0; i < survey->nbeams; ++i) {
    BEAM *beam = &survey->beams[i];
    beam->dm_shifts = (float *)malloc(survey->nchans * sizeof(float));
    for (unsigned j = 0; j < survey->nchans; ++j)
      beam->dm_shifts[j] = dmdelay(beam->fch1 + beam->foff * (float)j, beam->fch1);
    const float high_dm = survey->lowdm + survey->dmstep * (float)(survey->tdms - 1);
    const unsigned ms = (unsigned)(beam->dm_shifts[survey->nchans - 1] * high_dm / survey->tsamp);
    greatest = std::max(ms, greatest);
    greatest = (greatest / 256) * 256 + 256;
  }
  for (unsigned i = 0; i < survey->nbeams; ++i) survey->beams[i].maxshift = greatest;
  g.maxshift = greatest;
  // nsamp: as given, else sized from device memory (two chunks in flight:
  // float and storage input rows and outputs per beam), capped at 2^20
  const double sb = (double)mdsm::storage_bytes(g.storage);
  if (survey->nsamp == 0) {
    size_t fr = 0, tot = 0;
    if (hipMemGetInfo(&fr, &tot) != hipSuccess) tot = (size_t)8 << 30;
    const double in_per = (4.0 + (g.storage == MDSM_F32 ? 0.0 : sb)) * survey->nchans;
    const double per = (double)survey->nbeams * 2.0 * (in_per + 4.0 * survey->tdms) +
                       (survey->apply_median_filter ? 4.0 * survey->tdms : 0.0);
    const double n = ((double)tot * 0.8 - 2.0 * greatest * in_per * survey->nbeams) / per;
    survey->nsamp = (unsigned)std::max(0.0, std::min(n, (double)(1u << 20)));
    survey->nsamp -= survey->nsamp % DEDISP_THREADS;
  }
  if (greatest > survey->nsamp) survey->nsamp = 1u << (unsigned)std::ceil(std::log2((double)greatest));
  g.nsamp = survey->nsamp;
  g.ld = ((uint64_t)g.nsamp + g.maxshift + 3u) & ~(uint64_t)3u;
  survey->num_threads = survey->nbeams;
  for (hipStream_t *s : {&g.s_in, &g.s_dd, &g.s_out, &g.s_aux})
    if (!hip_ok(hipStreamCreateWithFlags(s, hipStreamNonBlocking), "stream")) return;
  for (int p = 0; p < 2; ++p)
    for (hipEvent_t *e : {&g.ev_in[p], &g.ev_dd[p], &g.ev_det[p], &g.ev_out[p]})
      if (!hip_ok(hipEventCreateWithFlags(e, hipEventDisableTiming), "event")) return;
  for (int k = 0; k < NRING; ++k)
    for (hipEvent_t *e : {&g.ev_h2d[k], &g.ev_ring[k]})
      if (!hip_ok(hipEventCreateWithFlags(e, hipEventDisableTiming), "event")) return;
  g.beams.resize(survey->nbeams);
  std::vector<mdsm_survey_t> sv(survey->nbeams);
  for (unsigned i = 0; i < survey->nbeams; ++i) {
    mdsm_survey_t &b = g.beams[i].s;
    b.fch1 = survey->beams[i].fch1;
    b.foff = survey->beams[i].foff;
    b.tsamp = survey->tsamp;
    b.lowdm = survey->lowdm;
    b.dmstep = survey->dmstep;
    b.nchans = survey->nchans;
    b.ndms = survey->tdms;
    b.nsamp = g.nsamp;
    b.delay_mode = MDSM_DELAY_REFERENCE;  // the reference's float dm_shifts and truncation
    sv[i] = b;
  }
  for (int p = 0; p < 2; ++p)
    if (mdsm_batch_create(sv.data(), g.nbeams, g.storage, &g.batch[p]) != MDSM_OK) {
      fail("%s", mdsm_last_error());
      return;
    }
  for (unsigned i = 0; i < g.nbeams; ++i) {
    uint64_t min_ld = 0;
    if (mdsm_batch_info(g.batch[0], i, nullptr, &min_ld) != MDSM_OK || min_ld > g.ld) {
      fail("beam %u: input rows too short for its delays", i);
      return;
    }
  }
  const size_t in_bytes = (size_t)g.nbeams * g.nchans * g.ld * 4;
  const size_t st_bytes = (size_t)g.nbeams * g.nchans * g.ld * mdsm::storage_bytes(g.storage);
  const size_t out_bytes = (size_t)g.nbeams * g.tdms * g.nsamp * 4;
  for (int p = 0; p < 2; ++p) {
    if (!hip_ok(hipMalloc((void **)&g.d_raw[p], in_bytes), "input buffer") ||
        !hip_ok(hipMemset(g.d_raw[p], 0, in_bytes), "input clear") ||
        !hip_ok(hipMalloc((void **)&g.d_out[p], out_bytes), "output buffer"))
      return;
    if (g.storage == MDSM_F32) {
      g.d_store[p] = g.d_raw[p];
    } else if (!hip_ok(hipMalloc(&g.d_store[p], st_bytes), "storage buffer") ||
               !hip_ok(hipMemset(g.d_store[p], 0, st_bytes), "storage clear")) {
      return;
    }
  }
  // candidates per beam and chunk: 2^20 to start (MDSM_COMPAT_CAND_CAP
  // overrides), grown when a chunk has more
  uint32_t cap = 1u << 20;
  if (const char *e = getenv("MDSM_COMPAT_CAND_CAP")) cap = std::max(1u, (uint32_t)strtoul(e, nullptr, 10));
  for (unsigned i = 0; i < g.nbeams; ++i)
    for (int p = 0; p < 2; ++p)
      if (!alloc_cands(g.beams[i], p, cap)) return;
  if (!hip_ok(hipHostMalloc((void **)&g.h_count, sizeof(uint32_t) * 2 * g.nbeams), "counts")) return;
  memset(g.h_count, 0, sizeof(uint32_t) * 2 * g.nbeams);
  for (int k = 0; k < MDSM_STAGES; ++k)
    if (!hip_ok(hipHostMalloc((void **)&g.input[k],
                              (size_t)g.nbeams * g.nsamp * g.nchans * sizeof(float)),
                "pinned input"))
      return;
  for (int k = 0; k < NRING && !survey->series_on_device; ++k) {
    g.out_ring[k] = (float **)calloc(g.nbeams, sizeof(float *));
    for (unsigned i = 0; i < g.nbeams; ++i) {
      if (!hip_ok(hipHostMalloc((void **)&g.out_ring[k][i], (size_t)g.tdms * g.nsamp * sizeof(float)),
                  "pinned output"))
        return;
      memset(g.out_ring[k][i], 0, (size_t)g.tdms * g.nsamp * sizeof(float));
    }
  }
  if (!hip_ok(hipMalloc(&g.d_fitted, sizeof(double) * g.nchans), "bandpass") ||
      !hip_ok(hipMalloc(&g.d_cmean, sizeof(double) * g.nchans), "channel means") ||
      !hip_ok(hipMalloc(&g.d_cstd, sizeof(double) * g.nchans), "channel deviations"))
    return;
  if (survey->apply_median_filter &&
      !hip_ok(hipMalloc(&g.d_tmp, (size_t)g.tdms * g.nsamp * sizeof(float)), "median buffer"))
    return;
  for (unsigned m = 0; m < survey->num_masks; ++m) {
    g.mask_from.push_back((int32_t)survey->channel_mask[m].from);
    g.mask_to.push_back((int32_t)survey->channel_mask[m].to);
  }
  if (!bind_batches() || !hip_ok(hipStreamSynchronize(g.s_dd), "autotune")) return;
  if (survey->single_file_mode && !open_files(0.0)) return;
  printf("Observation Params: ndms = %u, maxDM = %f\n", survey->tdms,
         survey->lowdm + survey->dmstep * (survey->tdms - 1));
  printf("Observation Params: nchans = %u, nsamp = %u, tsamp = %f, maxshift = %u\n", g.nchans,
         g.nsamp, survey->tsamp, g.maxshift);
  g.ok = true;
}

float *get_buffer_pointer() { return g.ok ? g.input[g.loop_counter % MDSM_STAGES] : nullptr; }

unsigned char *get_antenna_pointer() { return nullptr; }  // beamforming is not part of this library

float **next_chunk_multibeam(unsigned int data_read, unsigned &samples, double timestamp,
                             double blockRate) {
  g.ppnsamp = g.pnsamp;
  g.pnsamp = data_read;
  if (g.survey) g.survey->nsamp = data_read;
  if (data_read > 0) {  // the stop call keeps the last timing (manager.cpp:656-677)
    g.timestamp = timestamp;
    g.blockRate = blockRate;
    if (g.survey) {
      g.survey->timestamp = timestamp;
      g.survey->blockRate = blockRate;
    }
  }
  if (!g.ok || g.loop_counter < 1) {
    samples = (unsigned)-1;
    return nullptr;
  }
  // the series of the chunk submitted two calls earlier (ring slot of chunk
  // loop_counter - 2); samples is the data_read of the previous call, as
  // the reference returns it (manager.cpp:652-653,670,684)
  const int j = g.loop_counter - 2;
  const int slot = ((j % NRING) + NRING) % NRING;
  if (j >= 1 && !hip_ok(hipEventSynchronize(g.ev_ring[slot]), "output download")) {
    samples = (unsigned)-1;
    return nullptr;
  }
  samples = g.ppnsamp;
  return g.out_ring[slot];  // NULL when the series stay on the device
}

int start_processing(unsigned int data_read) {
  if (!g.ok) return 0;
  const int k = g.loop_counter;
  if (data_read == 0) {
    if (!g.out_switch) return 0;
    g.out_switch = false;
    // the last chunk of the stream: finish it now
    if (g.pending && !finish_chunk(g.pending, true)) {
      g.ok = false;
      return 0;
    }
    g.pending = 0;
    return ++g.loop_counter;
  }
  const float *host = g.input[k % MDSM_STAGES];
  const size_t S = g.maxshift, N = g.nsamp;
  if (k == 0) {
    // priming read: the first maxshift samples of every channel, layout
    // [beam][chan][maxshift]; parked at the row end so the next chunk's
    // overlap copy moves them to the front
    const unsigned n = std::min<unsigned>(data_read, g.maxshift);
    for (unsigned b = 0; b < g.nbeams; ++b)
      if (!hip_ok(hipMemcpy2DAsync(raw_ptr(0, b) + N + (S - n), g.ld * 4,
                                   host + (size_t)b * g.nchans * S, S * 4, n * 4, g.nchans,
                                   hipMemcpyHostToDevice, g.s_in), "priming upload") ||
          !clean_input(raw_ptr(0, b), b, N + (S - n), n))
        return 0;
    if (!hip_ok(hipEventRecord(g.ev_h2d[0], g.s_in), "event")) return 0;
    return ++g.loop_counter;
  }
  if (data_read > N) {
    fail("data_read %u exceeds nsamp %zu", data_read, N);
    return 0;
  }
  const int p = k & 1;
  // ---- s_in: overlap-save staging of chunk k, once chunk k - 2's
  // dedispersion has read this parity's buffers
  if (!hip_ok(hipStreamWaitEvent(g.s_in, g.ev_dd[p], 0), "wait")) return 0;
  for (unsigned b = 0; b < g.nbeams; ++b) {
    float *row = raw_ptr(p, b);
    // previous tail [N, N + S) -> [0, S), then the new samples
    if (!hip_ok(hipMemcpy2DAsync(row, g.ld * 4, raw_ptr(p ^ 1, b) + N, g.ld * 4, S * 4, g.nchans,
                                 hipMemcpyDeviceToDevice, g.s_in), "overlap copy") ||
        !hip_ok(hipMemcpy2DAsync(row + S, g.ld * 4, host + (size_t)b * g.nchans * N, N * 4,
                                 (size_t)data_read * 4, g.nchans, hipMemcpyHostToDevice, g.s_in),
                "chunk upload"))
      return 0;
    if (data_read < N &&  // short last buffer: zero the stale samples
        !hip_ok(hipMemset2DAsync(row + S + data_read, g.ld * 4, 0, (N - data_read) * 4, g.nchans,
                                 g.s_in), "pad"))
      return 0;
  }
  if (!hip_ok(hipEventRecord(g.ev_h2d[k % MDSM_STAGES], g.s_in), "event")) return 0;
  for (unsigned b = 0; b < g.nbeams; ++b) {
    if (!clean_input(raw_ptr(p, b), b, S, data_read)) return 0;
    // FP16 / U8 storage: per-chunk bandpass normalisation of the whole row
    // (overlap tail and new samples), statistics over its valid samples
    if (g.storage != MDSM_F32 &&
        (!abi_ok(mdsm_channel_stats(raw_ptr(p, b), g.ld, g.nchans, S + data_read, g.d_cmean,
                                    g.d_cstd, g.s_in), "channel statistics", b) ||
         !abi_ok(mdsm_normalise(raw_ptr(p, b), g.ld, g.nchans, g.ld, g.d_cmean, g.d_cstd,
                                store_ptr(p, b), g.storage, g.ld, g.s_in), "normalise", b)))
      return 0;
  }
  if (!hip_ok(hipEventRecord(g.ev_in[p], g.s_in), "event")) return 0;
  // ---- s_dd: one batched launch over every beam, then the filters, once
  // the staging is done and chunk k - 2's detection and download have read
  // this parity's outputs
  if (!hip_ok(hipStreamWaitEvent(g.s_dd, g.ev_in[p], 0), "wait") ||
      !hip_ok(hipStreamWaitEvent(g.s_dd, g.ev_out[p], 0), "wait") ||
      !abi_ok(mdsm_batch_launch(g.batch[p], g.s_dd), "dedispersion", 0))
    return 0;
  for (unsigned b = 0; b < g.nbeams; ++b)
    if (!filter_output(p, b)) return 0;
  if (!hip_ok(hipEventRecord(g.ev_dd[p], g.s_dd), "event")) return 0;
  // ---- s_out: detection, then the download into the host ring
  if (!hip_ok(hipStreamWaitEvent(g.s_out, g.ev_dd[p], 0), "wait")) return 0;
  for (unsigned b = 0; b < g.nbeams; ++b)
    if (!enqueue_detect(p, b, data_read, data_read, g.s_out)) return 0;
  if (!hip_ok(hipEventRecord(g.ev_det[p], g.s_out), "event")) return 0;
  const int slot = k % NRING;
  for (unsigned b = 0; b < g.nbeams && g.out_ring[slot]; ++b)
    if (!hip_ok(hipMemcpyAsync(g.out_ring[slot][b], out_ptr(p, b), (size_t)g.tdms * N * 4,
                               hipMemcpyDeviceToHost, g.s_out), "output download"))
      return 0;
  if (!hip_ok(hipEventRecord(g.ev_out[p], g.s_out), "event") ||
      !hip_ok(hipEventRecord(g.ev_ring[slot], g.s_out), "event"))
    return 0;
  g.chunk[k % NRING] = ChunkInfo{data_read, g.timestamp, g.blockRate};
  // ---- host: finish chunk k - 1 while the GPU runs chunk k
  if (g.pending && !finish_chunk(g.pending, false)) {
    g.ok = false;
    return 0;
  }
  g.pending = k;
  // the caller fills input stage (k + 1) % 3 next: its previous upload
  // (chunk k - 2) must be done with it
  if (!hip_ok(hipEventSynchronize(g.ev_h2d[(k + 1) % MDSM_STAGES]), "upload")) return 0;
  return ++g.loop_counter;
}

void tearDownMDSM() {
  for (hipStream_t s : {g.s_in, g.s_dd, g.s_out, g.s_aux})
    if (s) (void)hipStreamSynchronize(s);
  close_files();
  for (auto &b : g.beams)
    for (int p = 0; p < 2; ++p)
      if (b.d_cands[p]) (void)hipFree(b.d_cands[p]);
  for (int p = 0; p < 2; ++p) {
    if (g.batch[p]) mdsm_batch_destroy(g.batch[p]);
    if (g.d_store[p] && g.d_store[p] != g.d_raw[p]) (void)hipFree(g.d_store[p]);
    if (g.d_raw[p]) (void)hipFree(g.d_raw[p]);
    if (g.d_out[p]) (void)hipFree(g.d_out[p]);
    for (hipEvent_t e : {g.ev_in[p], g.ev_dd[p], g.ev_det[p], g.ev_out[p]})
      if (e) (void)hipEventDestroy(e);
  }
  for (int k = 0; k < NRING; ++k)
    for (hipEvent_t e : {g.ev_h2d[k], g.ev_ring[k]})
      if (e) (void)hipEventDestroy(e);
  for (int k = 0; k < MDSM_STAGES; ++k)
    if (g.input[k]) (void)hipHostFree(g.input[k]);
  for (int k = 0; k < NRING; ++k)
    if (g.out_ring[k]) {
      for (unsigned i = 0; i < g.beams.size(); ++i)
        if (g.out_ring[k][i]) (void)hipHostFree(g.out_ring[k][i]);
      free(g.out_ring[k]);
    }
  if (g.h_count) (void)hipHostFree(g.h_count);
  for (double *p : {g.d_fitted, g.d_cmean, g.d_cstd})
    if (p) (void)hipFree(p);
  if (g.d_tmp) (void)hipFree(g.d_tmp);
  for (hipStream_t s : {g.s_in, g.s_dd, g.s_out, g.s_aux})
    if (s) (void)hipStreamDestroy(s);
  printf("Finished Process\n");
  g = State();
}

// C entry points for non-C++ callers (and tests): the same calls with plain
// types; mdsm_compat_open returns 0 on success.
extern "C" int mdsm_compat_open(const char *xml_path, unsigned nsamp_override) {
  SURVEY *s = processSurveyParameters(xml_path ? std::string(xml_path) : std::string());
  if (!s) {
    g_err = mdsm_compat_xml_error();
    return 1;
  }
  if (nsamp_override) s->nsamp = nsamp_override;
  initialiseMDSM(s);
  return g.ok ? 0 : 2;
}

extern "C" int mdsm_compat_survey(unsigned *nbeams, unsigned *nchans, unsigned *nsamp,
                                  unsigned *tdms, unsigned *maxshift) {
  if (!g.survey) return 1;
  if (nbeams) *nbeams = g.survey->nbeams;
  if (nchans) *nchans = g.nchans;
  if (nsamp) *nsamp = g.nsamp;
  if (tdms) *tdms = g.tdms;
  if (maxshift) *maxshift = g.maxshift;
  return 0;
}

extern "C" float *mdsm_compat_buffer(void) { return get_buffer_pointer(); }

extern "C" float **mdsm_compat_next_chunk(unsigned data_read, unsigned *samples, double timestamp,
                                          double blockRate) {
  unsigned s = 0;
  float **r = next_chunk_multibeam(data_read, s, timestamp, blockRate);
  if (samples) *samples = s;
  return r;
}

extern "C" int mdsm_compat_start(unsigned data_read) { return start_processing(data_read); }

extern "C" void mdsm_compat_close(void) { tearDownMDSM(); }

// SURVEY::last_nsamp for callers without the struct (src/MDSM.cpp:101 sets it
// when the file ends inside a chunk).
extern "C" void mdsm_compat_set_last_nsamp(unsigned last_nsamp) {
  if (g.survey) g.survey->last_nsamp = last_nsamp;
}
