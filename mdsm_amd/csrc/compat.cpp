// compat.cpp -- the MDSM multibeam manager API (initialiseMDSM,
// get_buffer_pointer, next_chunk_multibeam, start_processing, tearDownMDSM)
// on top of the C ABI, for drivers written against
// src/mdsm-multibeam/multibeam_dedispersion_manager.h:9-28.
//
// The reference runs one pthread per beam plus an output thread behind two
// barriers (multibeam_dedispersion_manager.cpp:467-590,
// multibeam_dedispersion_thread.cu:464-839, multibeam_dedispersion_output.cpp:325-452).
// On one GPU the same data flow is kept on the caller's thread:
// start_processing(k) uploads chunk k of every beam into its overlap-save
// row, dedisperses it, copies the series back and runs the reference
// detector (global FP64 mean/rms, x - mean >= rms*threshold, optional
// reference DBSCAN) writing the reference CSV lines; next_chunk_multibeam
// returns the per-beam output buffers of the chunk submitted two calls
// earlier, as the reference does (:683-690).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <ctime>
#include <string>
#include <vector>

#include "internal.h"
#include "mdsm_compat.h"

extern "C" const char *mdsm_compat_xml_error(void);

namespace {

struct BeamState {
  mdsm_survey_t s{};
  mdsm_plan plan = nullptr;
  float *d_in = nullptr;   // [nchans][ld]: [0, maxshift) tail, then nsamp new samples
  float *d_out = nullptr;  // [tdms][nsamp]
  FILE *fp = nullptr;
  unsigned clusters_written = 0;
  std::vector<double> ctime, cdm, csnr;
  std::vector<int> ccl;
};

struct State {
  SURVEY *survey = nullptr;
  unsigned nsamp = 0, maxshift = 0, nchans = 0, tdms = 0;
  uint64_t ld = 0;
  std::vector<BeamState> beams;
  float *input[MDSM_STAGES] = {nullptr, nullptr, nullptr};  // pinned [beam][chan][nsamp]
  float **out_ring[2] = {nullptr, nullptr};                  // per beam [tdms][nsamp] (pinned)
  int loop_counter = 0;
  bool out_switch = true;
  unsigned pnsamp = 0, ppnsamp = 0;
  double timestamp = 0, blockRate = 0;
  mdsm_cand_t *d_cands = nullptr;
  uint32_t *d_count = nullptr;
  uint32_t cap = 1u << 22;
  double *d_fitted = nullptr;  // fitted bandpass of the beam being cleaned
  float *d_tmp = nullptr;      // median filter output [tdms][nsamp]
  std::vector<int32_t> mask_from, mask_to;
  hipStream_t stream = nullptr;
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

// reference dmdelay (multibeam_dedispersion_manager.cpp:325-326)
inline float dmdelay(float F1, float F2) {
  return (float)(4148.741601 * ((1.0 / F1 / F1) - (1.0 / F2 / F2)));
}

void open_file(unsigned b) {
  SURVEY *s = g.survey;
  char path[512];
  if (s->single_file_mode) {
    if (s->nbeams > 1) snprintf(path, sizeof(path), "%s/%s_beam_%u.dat", s->basedir, s->fileprefix, b);
    else snprintf(path, sizeof(path), "%s/%s.dat", s->basedir, s->fileprefix);
  } else {
    time_t now = s->use_pc_time ? time(nullptr) : (time_t)g.timestamp;
    char ts[32];
    strftime(ts, sizeof(ts), "%F_%T", localtime(&now));
    snprintf(path, sizeof(path), "%s/%s_beam_%u_%s.dat", s->basedir, s->fileprefix, b, ts);
  }
  g.beams[b].fp = fopen(path, "w");
  if (!g.beams[b].fp) fail("Invalid output file path: %s", path);
}

// Reference detector on one beam's dedispersed chunk (nsamp_valid columns).
bool detect_beam(unsigned b, const float *d_x, unsigned nsamp_valid, double timestamp,
                 double blockRate) {
  SURVEY *s = g.survey;
  BeamState &bs = g.beams[b];
  double mean = 0, rms = 0;
  if (mdsm_detect_global(d_x, g.tdms, nsamp_valid, g.nsamp, nsamp_valid, s->detection_threshold,
                         g.d_cands, g.cap, g.d_count, &mean, &rms, g.stream) != MDSM_OK) {
    fail("detection failed: %s", mdsm_last_error());
    return false;
  }
  uint32_t n = 0;
  if (!hip_ok(hipMemcpyAsync(&n, g.d_count, 4, hipMemcpyDeviceToHost, g.stream), "count") ||
      !hip_ok(hipStreamSynchronize(g.stream), "sync"))
    return false;
  if (n > g.cap) {
    fprintf(stderr, "mdsm: beam %u: %u candidates, keeping the first %u\n", b, n, g.cap);
    n = g.cap;
  }
  std::vector<mdsm_cand_t> c(n);
  if (n && !hip_ok(hipMemcpy(c.data(), g.d_cands, n * sizeof(mdsm_cand_t), hipMemcpyDeviceToHost), "cands"))
    return false;
  mdsm_sort_candidates(c.data(), n);
  bs.ctime.resize(n);
  bs.cdm.resize(n);
  bs.csnr.resize(n);
  bs.ccl.assign(n, 0);
  for (uint32_t i = 0; i < n; ++i) {
    bs.ctime[i] = timestamp + c[i].sample * blockRate;
    bs.cdm[i] = (double)(s->lowdm + c[i].dm_idx * s->dmstep);
    bs.csnr[i] = c[i].snr;
  }
  if (!bs.fp) return true;
  if (!s->apply_clustering) {
    for (uint32_t i = 0; i < n; ++i)
      fprintf(bs.fp, "%lf, %f, %f\n", bs.ctime[i], (float)bs.cdm[i], (float)bs.csnr[i]);
  } else {
    uint32_t nc = 0;
    mdsm_dbscan_ref(bs.ctime.data(), bs.cdm.data(), bs.csnr.data(), n,
                    (float)(s->dbscan_time_range * blockRate), s->dbscan_dm_range,
                    s->dbscan_snr_range, s->dbscan_min_points, bs.ccl.data(), &nc);
    // clusters in id order, then noise (multibeam_dedispersion_output.cpp:262-312);
    // with classification a cluster scoring >= 0.1 is written with id -2
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
                                      s->dmstep, s->tdms, s->min_pulse_width, s->beams[0].fch1,
                                      s->beams[0].foff, s->nchans, &score, nullptr) != MDSM_OK)
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
  return true;
}

// Input conditioning of one beam's rows over [col0, col0 + n): complex
// voltages to power, then bandpass fit and RFI clipping
// (multibeam_dedispersion_thread.cu:700-752).  The reference clears the
// channel flags before channel_clipper runs (:364), so only spectrum
// clipping changes the data; the flags are not computed here.
bool clean_input(unsigned b, uint64_t col0, uint64_t n) {
  SURVEY *s = g.survey;
  BeamState &bs = g.beams[b];
  if (n == 0) return true;
  if (s->voltage && mdsm_voltage_to_power(bs.d_in, g.ld, g.nchans, col0, n, g.stream) != MDSM_OK) {
    fail("beam %u: %s", b, mdsm_last_error());
    return false;
  }
  if (!s->apply_rfi_clipper) return true;
  mdsm_rfi_params_t p{};
  p.ncoeffs = s->ncoeffs;
  p.channel_block = s->channel_block;
  p.channel_thresh = s->channel_thresh;
  p.spectrum_thresh = s->spectrum_thresh;
  p.nmasks = (uint32_t)g.mask_from.size();
  p.mask_from = g.mask_from.data();
  p.mask_to = g.mask_to.data();
  p.apply_channel_clip = 0;
  mdsm_bandpass_stats_t st{};
  if (mdsm_rfi_clean(bs.d_in, g.ld, g.nchans, col0, n, &p, g.d_fitted, nullptr, &st, g.stream) !=
      MDSM_OK) {
    fail("beam %u: RFI clipping: %s", b, mdsm_last_error());
    return false;
  }
  s->bandpass_mean = st.mean;
  s->bandpass_std = st.std;
  s->bandpass_rmse = st.rmse;
  return true;
}

// Median filter and detrending of the dedispersed series (thread.cu:766-772).
bool filter_output(unsigned b) {
  SURVEY *s = g.survey;
  BeamState &bs = g.beams[b];
  if (s->apply_median_filter) {
    if (mdsm_median_filter5(bs.d_out, g.tdms, g.nsamp, g.d_tmp, g.stream) != MDSM_OK ||
        !hip_ok(hipMemcpyAsync(bs.d_out, g.d_tmp, (size_t)g.tdms * g.nsamp * 4,
                               hipMemcpyDeviceToDevice, g.stream), "median copy")) {
      fail("beam %u: median filter: %s", b, mdsm_last_error());
      return false;
    }
  }
  if (s->apply_detrending &&
      mdsm_detrend_normalise(bs.d_out, g.tdms, g.nsamp, g.stream) != MDSM_OK) {
    fail("beam %u: detrending: %s", b, mdsm_last_error());
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

void initialiseMDSM(SURVEY *survey) {
  g = State();
  g.survey = survey;
  if (!survey || survey->nbeams == 0 || survey->nchans == 0 || survey->tdms == 0) {
    fail("invalid survey (no beams, channels or DMs)");
    return;
  }
  g.nchans = survey->nchans;
  g.tdms = survey->tdms;
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
  // nsamp: as given, else sized from device memory (capped at 2^20 samples)
  if (survey->nsamp == 0) {
    size_t fr = 0, tot = 0;
    if (hipMemGetInfo(&fr, &tot) != hipSuccess) tot = (size_t)8 << 30;
    const double per = (double)survey->nbeams * (survey->nchans + survey->tdms) * 4.0;
    double n = ((double)tot * 0.9 - (double)greatest * survey->nchans * survey->nbeams * 4.0) / per;
    survey->nsamp = (unsigned)std::min(n, (double)(1u << 20));
    survey->nsamp -= survey->nsamp % DEDISP_THREADS;
  }
  if (greatest > survey->nsamp) survey->nsamp = 1u << (unsigned)std::ceil(std::log2((double)greatest));
  g.nsamp = survey->nsamp;
  g.ld = (uint64_t)g.nsamp + g.maxshift;
  survey->num_threads = survey->nbeams;
  if (!hip_ok(hipStreamCreateWithFlags(&g.stream, hipStreamNonBlocking), "stream")) return;
  g.beams.resize(survey->nbeams);
  for (unsigned i = 0; i < survey->nbeams; ++i) {
    BeamState &b = g.beams[i];
    b.s.fch1 = survey->beams[i].fch1;
    b.s.foff = survey->beams[i].foff;
    b.s.tsamp = survey->tsamp;
    b.s.lowdm = survey->lowdm;
    b.s.dmstep = survey->dmstep;
    b.s.nchans = survey->nchans;
    b.s.ndms = survey->tdms;
    b.s.nsamp = g.nsamp;
    b.s.delay_mode = MDSM_DELAY_REFERENCE;  // the reference's float dm_shifts and truncation
    if (mdsm_plan_create(&b.s, MDSM_F32, &b.plan) != MDSM_OK) {
      fail("beam %u: %s", i, mdsm_last_error());
      return;
    }
    if (!hip_ok(hipMalloc(&b.d_in, g.ld * g.nchans * sizeof(float)), "input buffer") ||
        !hip_ok(hipMemset(b.d_in, 0, g.ld * g.nchans * sizeof(float)), "input clear") ||
        !hip_ok(hipMalloc(&b.d_out, (size_t)g.tdms * g.nsamp * sizeof(float)), "output buffer"))
      return;
  }
  for (int k = 0; k < MDSM_STAGES; ++k)
    if (!hip_ok(hipHostMalloc((void **)&g.input[k],
                              (size_t)survey->nbeams * g.nsamp * g.nchans * sizeof(float)),
                "pinned input"))
      return;
  for (int k = 0; k < 2; ++k) {
    g.out_ring[k] = (float **)calloc(survey->nbeams, sizeof(float *));
    for (unsigned i = 0; i < survey->nbeams; ++i)
      if (!hip_ok(hipHostMalloc((void **)&g.out_ring[k][i], (size_t)g.tdms * g.nsamp * sizeof(float)),
                  "pinned output"))
        return;
      else
        memset(g.out_ring[k][i], 0, (size_t)g.tdms * g.nsamp * sizeof(float));
  }
  if (!hip_ok(hipMalloc(&g.d_cands, sizeof(mdsm_cand_t) * g.cap), "candidates") ||
      !hip_ok(hipMalloc(&g.d_count, sizeof(uint32_t)), "count") ||
      !hip_ok(hipMalloc(&g.d_fitted, sizeof(double) * g.nchans), "bandpass"))
    return;
  if (survey->apply_median_filter &&
      !hip_ok(hipMalloc(&g.d_tmp, (size_t)g.tdms * g.nsamp * sizeof(float)), "median buffer"))
    return;
  for (unsigned m = 0; m < survey->num_masks; ++m) {
    g.mask_from.push_back((int32_t)survey->channel_mask[m].from);
    g.mask_to.push_back((int32_t)survey->channel_mask[m].to);
  }
  for (unsigned i = 0; i < survey->nbeams; ++i) open_file(i);
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
  // loop_counter - 2); samples is the data_read of the previous call, as the
  // reference returns it (manager.cpp:652-653,670,684)
  samples = g.ppnsamp;
  return g.out_ring[g.loop_counter & 1];
}

int start_processing(unsigned int data_read) {
  if (!g.ok) return 0;
  if (data_read == 0 && g.out_switch) g.out_switch = false;
  else if (data_read == 0 && !g.out_switch) return 0;
  const int k = g.loop_counter;
  const float *host = g.input[k % MDSM_STAGES];
  const size_t S = g.maxshift, N = g.nsamp;
  if (k == 0) {
    // priming read: the first maxshift samples of every channel, layout
    // [beam][chan][maxshift]; parked at the row end so the next chunk's
    // overlap copy moves them to the front
    const unsigned n = std::min<unsigned>(data_read, g.maxshift);
    for (unsigned b = 0; b < g.beams.size(); ++b)
      if (!hip_ok(hipMemcpy2DAsync(g.beams[b].d_in + N + (S - n), g.ld * 4,
                                   host + (size_t)b * g.nchans * S, S * 4, n * 4, g.nchans,
                                   hipMemcpyHostToDevice, g.stream), "priming upload") ||
          !clean_input(b, N + (S - n), n))
        return 0;
  } else if (data_read > 0) {
    const int slot = k & 1;
    for (unsigned b = 0; b < g.beams.size(); ++b) {
      BeamState &bs = g.beams[b];
      // overlap-save: previous tail [N, N+S) -> [0, S), then the new samples
      if (!hip_ok(hipMemcpy2DAsync(bs.d_in, g.ld * 4, bs.d_in + N, g.ld * 4, S * 4, g.nchans,
                                   hipMemcpyDeviceToDevice, g.stream), "overlap copy") ||
          !hip_ok(hipMemcpy2DAsync(bs.d_in + S, g.ld * 4, host + (size_t)b * g.nchans * N, N * 4,
                                   (size_t)data_read * 4, g.nchans, hipMemcpyHostToDevice, g.stream),
                  "chunk upload"))
        return 0;
      if (data_read < N) {  // short last buffer: zero the stale samples
        if (!hip_ok(hipMemset2DAsync(bs.d_in + S + data_read, g.ld * 4, 0, (N - data_read) * 4,
                                     g.nchans, g.stream), "pad"))
          return 0;
        g.survey->last_nsamp = data_read;
      }
      if (!clean_input(b, S, data_read)) return 0;
      if (mdsm_dedisperse(bs.plan, bs.d_in, g.ld, bs.d_out, g.stream) != MDSM_OK) {
        fail("beam %u: %s", b, mdsm_last_error());
        return 0;
      }
      if (!filter_output(b)) return 0;
      if (!hip_ok(hipMemcpyAsync(g.out_ring[slot][b], bs.d_out, (size_t)g.tdms * N * 4,
                                 hipMemcpyDeviceToHost, g.stream), "output download"))
        return 0;
      // detection over the valid columns only (row stride nsamp)
      if (!detect_beam(b, bs.d_out, data_read, g.timestamp, g.blockRate)) return 0;
    }
    if (!hip_ok(hipStreamSynchronize(g.stream), "sync")) return 0;
  }
  return ++g.loop_counter;
}

void tearDownMDSM() {
  if (g.stream) (void)hipStreamSynchronize(g.stream);
  for (auto &b : g.beams) {
    if (b.plan) mdsm_plan_destroy(b.plan);
    if (b.d_in) (void)hipFree(b.d_in);
    if (b.d_out) (void)hipFree(b.d_out);
    if (b.fp) fclose(b.fp);
  }
  for (int k = 0; k < MDSM_STAGES; ++k)
    if (g.input[k]) (void)hipHostFree(g.input[k]);
  for (int k = 0; k < 2; ++k)
    if (g.out_ring[k]) {
      for (unsigned i = 0; i < g.beams.size(); ++i)
        if (g.out_ring[k][i]) (void)hipHostFree(g.out_ring[k][i]);
      free(g.out_ring[k]);
    }
  if (g.d_cands) (void)hipFree(g.d_cands);
  if (g.d_count) (void)hipFree(g.d_count);
  if (g.d_fitted) (void)hipFree(g.d_fitted);
  if (g.d_tmp) (void)hipFree(g.d_tmp);
  if (g.stream) (void)hipStreamDestroy(g.stream);
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
