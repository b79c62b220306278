// compat_legacy.cpp -- the single-beam MDSM manager API (initialiseMDSM
// returning the input buffer, next_chunk with noise statistics,
// start_processing, tearDownMDSM) on top of the C ABI, for drivers written
// against src/mdsm/dedispersion_manager.h:7-12 (the Pelican MdsmModule).
//
// The reference runs one pthread per GPU and an output thread behind two
// barriers (src/mdsm/dedispersion_manager.cpp:269-499,
// src/mdsm/dedispersion_thread.cu:246-380, src/mdsm/dedispersion_output.cpp:215-330).
// On one GPU the data flow is kept on the caller's thread: start_processing
// uploads the chunk (carrying the last maxshift samples of every row in
// front), dedisperses it with the legacy delay table and runs the legacy
// detector, writing the reference's lines; next_chunk downloads the output
// of the chunk submitted at the previous call, as the reference's threads
// deliver it.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <ctime>
#include <string>
#include <vector>

#include "mdsm_abi.h"
#include "mdsm_legacy.h"

extern "C" const char *mdsm_compat_xml_error(void);

namespace {

struct LegacyState {
  SURVEY *survey = nullptr;
  unsigned nsamp = 0, nchans = 0, tdms = 0, maxshift = 0;  // host rows: nsamp + maxshift
  uint64_t ld = 0;                                          // device rows
  mdsm_plan plan = nullptr;
  float *h_in = nullptr;   // pinned [nchans][nsamp + maxshift]
  float *h_out = nullptr;  // pinned [tdms][nsamp]
  float *d_in = nullptr;   // [nchans][ld]
  float *d_out = nullptr;  // [tdms][nsamp]
  mdsm_cand_t *d_cands = nullptr;
  uint32_t *d_count = nullptr;
  uint32_t cap = 1u << 22;
  float *d_stats = nullptr;  // per-chunk mean then rms
  size_t nstats = 0;
  int loop_counter = 0;
  bool out_switch = true;
  unsigned pnsamp = 0, ppnsamp = 0;
  double timestamp = 0, block_rate = 0;  // of the chunk the next start_processing takes
  float *mean = nullptr, *rms = nullptr;
  FILE *fp = nullptr;
  long written = 0;
  std::vector<int32_t> cdm, csample;
  std::vector<float> csnr;
  hipStream_t stream = nullptr;
  bool ok = false;
} g;

thread_local std::string g_err;
unsigned g_spc_override = 0;

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

// dmdelay of src/mdsm/dedispersion_manager.cpp:262-265
inline float dmdelay(float f1, float f2) {
  return (float)(4148.741601 * ((1.0 / f1 / f1) - (1.0 / f2 / f2)));
}

// output file of the legacy output thread (dedispersion_output.cpp:232-281)
bool open_file() {
  SURVEY *s = g.survey;
  std::string path = std::string(s->basedir) + "/" + s->fileprefix;
  if (s->single_file_mode) {
    path += ".dat";
  } else {
    time_t now = s->use_pc_time ? time(nullptr) : (time_t)g.timestamp;
    char ts[32];
    strftime(ts, sizeof(ts), "%F_%T", localtime(&now));
    path += std::string("_") + ts + "_" + std::to_string(s->secs_per_file) + ".dat";
  }
  g.fp = fopen(path.c_str(), "w");
  if (!g.fp) fail("Invalid output file path: %s", path.c_str());
  return g.fp != nullptr;
}

// process_brute (dedispersion_output.cpp:135-213) on the chunk in d_out,
// read_nsamp = data_read valid samples
bool search_chunk(unsigned data_read) {
  SURVEY *s = g.survey;
  const bool ext = g.mean != nullptr && g.rms != nullptr;
  const unsigned spc = s->samplesPerChunk;
  size_t nstats = 0;
  if (ext) {
    if (spc == 0) {
      fail("noise statistics given but samplesPerChunk is 0");
      return false;
    }
    nstats = (size_t)(g.nsamp > 4 ? (g.nsamp - 5) / spc : 0) + 1;
    if (nstats > g.nstats) {
      if (g.d_stats) (void)hipFree(g.d_stats);
      g.d_stats = nullptr;
      if (!hip_ok(hipMalloc((void **)&g.d_stats, 2 * nstats * sizeof(float)), "statistics")) return false;
      g.nstats = nstats;
    }
    if (!hip_ok(hipMemcpyAsync(g.d_stats, g.mean, nstats * sizeof(float), hipMemcpyHostToDevice, g.stream),
                "mean upload") ||
        !hip_ok(hipMemcpyAsync(g.d_stats + g.nstats, g.rms, nstats * sizeof(float), hipMemcpyHostToDevice,
                               g.stream), "rms upload"))
      return false;
  }
  float mean = 0.f, rms = 0.f;
  if (mdsm_detect_legacy(g.d_out, g.tdms, g.nsamp, s->lowdm, s->dmstep, s->detection_threshold,
                         (uint64_t)data_read * g.tdms, ext ? g.d_stats : nullptr,
                         ext ? g.d_stats + g.nstats : nullptr, spc, (uint32_t)nstats, g.d_cands,
                         g.cap, g.d_count, &mean, &rms, g.stream) != MDSM_OK) {
    fail("detection: %s", mdsm_last_error());
    return false;
  }
  uint32_t n = 0;
  if (!hip_ok(hipMemcpyAsync(&n, g.d_count, 4, hipMemcpyDeviceToHost, g.stream), "count") ||
      !hip_ok(hipStreamSynchronize(g.stream), "sync"))
    return false;
  if (n > g.cap) {
    fprintf(stderr, "mdsm: %u candidates, keeping the first %u\n", n, g.cap);
    n = g.cap;
  }
  std::vector<mdsm_cand_t> c(n);
  if (n && !hip_ok(hipMemcpy(c.data(), g.d_cands, n * sizeof(mdsm_cand_t), hipMemcpyDeviceToHost), "cands"))
    return false;
  mdsm_sort_candidates(c.data(), n);
  g.cdm.resize(n);
  g.csample.resize(n);
  g.csnr.resize(n);
  for (uint32_t i = 0; i < n; ++i) {
    g.cdm[i] = c[i].dm_idx;
    g.csample[i] = c[i].sample;
    g.csnr[i] = c[i].snr;
  }
  if (!ext)
    printf("Mean: %f, Stddev: %f\n", mean, rms);
  if (g.written == 0 && !s->single_file_mode && !open_file()) return false;
  if (!g.fp) return true;
  for (uint32_t i = 0; i < n; ++i) {
    const float thisdm = s->lowdm + (float)c[i].dm_idx * s->dmstep;
    const double t = g.timestamp + c[i].sample * g.block_rate;
    if (!ext) {
      fprintf(g.fp, "%lf, %f, %f\n", t, thisdm, c[i].snr);
    } else {
      const unsigned k = (unsigned)c[i].sample / spc;
      fprintf(g.fp, "%lf, %f, %f, %f, %f  \n", t, thisdm, c[i].snr, g.mean[k], g.rms[k]);
    }
  }
  fflush(g.fp);
  if (!s->single_file_mode) {
    g.written += data_read;
    if (g.written * g.block_rate > s->secs_per_file) {
      g.written = 0;
      fclose(g.fp);
      g.fp = nullptr;
    }
  }
  return true;
}

void free_all() {
  if (g.stream) (void)hipStreamSynchronize(g.stream);
  if (g.plan) mdsm_plan_destroy(g.plan);
  if (g.h_in) (void)hipHostFree(g.h_in);
  if (g.h_out) (void)hipHostFree(g.h_out);
  if (g.d_in) (void)hipFree(g.d_in);
  if (g.d_out) (void)hipFree(g.d_out);
  if (g.d_cands) (void)hipFree(g.d_cands);
  if (g.d_count) (void)hipFree(g.d_count);
  if (g.d_stats) (void)hipFree(g.d_stats);
  if (g.fp) fclose(g.fp);
  if (g.stream) (void)hipStreamDestroy(g.stream);
  g = LegacyState();
}

}  // namespace

// processSurveyParameters (src/mdsm/dedispersion_manager.cpp:40-200) from a
// path string, through the library's XML reader (both dialects).
SURVEY *processSurveyParameters(const std::string &filepath) {
  mdsm_xml_params_t x;
  if (mdsm_parse_observation_xml(filepath.c_str(), &x) != MDSM_OK) {
    fail("%s", mdsm_compat_xml_error());
    return nullptr;
  }
  auto *s = (SURVEY *)calloc(1, sizeof(SURVEY));
  s->nsamp = x.nsamp;
  s->nchans = x.nchans;
  s->tdms = x.tdms;
  s->nbits = x.nbits;
  s->tsamp = x.tsamp;
  s->foff = x.foff;
  s->fch1 = x.fch1;
  s->useBruteForce = x.use_brute_force != 0;
  s->lowdm = x.lowdm;
  s->dmstep = x.dmstep;
  s->detection_threshold = x.detection_threshold;
  strncpy(s->fileprefix, x.fileprefix, sizeof(s->fileprefix) - 1);
  strncpy(s->basedir, x.basedir, sizeof(s->basedir) - 1);
  s->secs_per_file = x.secs_per_file;
  s->use_pc_time = x.use_pc_time;
  s->single_file_mode = x.single_file_mode;
  s->num_gpus = 1;
  s->gpu_ids = (unsigned *)calloc(1, sizeof(unsigned));
  return s;
}

// initialiseMDSM (src/mdsm/dedispersion_manager.cpp:269-383): dm_shifts and
// maxshift as the reference computes them, buffers, the output file.
float *initialiseMDSM(SURVEY *survey) {
  free_all();
  g.survey = survey;
  if (!survey || survey->nchans == 0 || survey->tdms == 0) {
    fail("invalid survey (no channels or DMs)");
    return nullptr;
  }
  if (!survey->useBruteForce) {
    fail("subband dedispersion is not part of this library (set useBruteForce)");
    return nullptr;
  }
  g.nchans = survey->nchans;
  g.tdms = survey->tdms;
  const float bottom = dmdelay(survey->fch1 + survey->foff * (float)(survey->nchans - 1), survey->fch1);
  const float hi_dm = survey->lowdm + survey->dmstep * (float)(survey->tdms - 1);
  g.maxshift = (unsigned)(int)(bottom * hi_dm / survey->tsamp);
  survey->maxshift = g.maxshift;
  if (survey->nsamp == 0) {  // sized from device memory (capped at 2^20 samples)
    size_t fr = 0, tot = 0;
    if (hipMemGetInfo(&fr, &tot) != hipSuccess) tot = (size_t)8 << 30;
    double n = ((double)tot * 0.9 / 4.0 - (double)g.maxshift * survey->nchans) /
               (double)(survey->nchans + survey->tdms);
    survey->nsamp = (unsigned)std::min(n, (double)(1u << 20));
    survey->nsamp -= survey->nsamp % 128;
  }
  if (g.maxshift > survey->nsamp)
    survey->nsamp = 1u << (unsigned)std::ceil(std::log2((double)g.maxshift));
  g.nsamp = survey->nsamp;
  survey->num_threads = 1;
  if (g_spc_override) survey->samplesPerChunk = g_spc_override;

  mdsm_survey_t ps{};
  ps.fch1 = survey->fch1;
  ps.foff = survey->foff;
  ps.tsamp = survey->tsamp;
  ps.lowdm = survey->lowdm;
  ps.dmstep = survey->dmstep;
  ps.nchans = survey->nchans;
  ps.ndms = survey->tdms;
  ps.nsamp = g.nsamp;
  ps.delay_mode = MDSM_DELAY_LEGACY;
  if (mdsm_plan_create(&ps, MDSM_F32, &g.plan) != MDSM_OK) {
    fail("%s", mdsm_last_error());
    return nullptr;
  }
  uint32_t plan_shift = 0;
  uint64_t min_ld = 0;
  mdsm_plan_info(g.plan, &plan_shift, &min_ld);
  g.ld = std::max<uint64_t>(min_ld, (uint64_t)g.nsamp + g.maxshift);
  const size_t row = (size_t)g.nsamp + g.maxshift;
  if (!hip_ok(hipStreamCreateWithFlags(&g.stream, hipStreamNonBlocking), "stream") ||
      !hip_ok(hipHostMalloc((void **)&g.h_in, row * g.nchans * sizeof(float)), "input buffer") ||
      !hip_ok(hipHostMalloc((void **)&g.h_out, (size_t)g.tdms * g.nsamp * sizeof(float)), "output buffer") ||
      !hip_ok(hipMalloc((void **)&g.d_in, g.ld * g.nchans * sizeof(float)), "device input") ||
      !hip_ok(hipMemset(g.d_in, 0, g.ld * g.nchans * sizeof(float)), "device input clear") ||
      !hip_ok(hipMalloc((void **)&g.d_out, (size_t)g.tdms * g.nsamp * sizeof(float)), "device output") ||
      !hip_ok(hipMalloc((void **)&g.d_cands, sizeof(mdsm_cand_t) * g.cap), "candidates") ||
      !hip_ok(hipMalloc((void **)&g.d_count, sizeof(uint32_t)), "count"))
    return nullptr;
  memset(g.h_in, 0, row * g.nchans * sizeof(float));
  memset(g.h_out, 0, (size_t)g.tdms * g.nsamp * sizeof(float));
  if (survey->single_file_mode && !open_file()) return nullptr;
  printf("nchans: %u, nsamp: %u, tsamp: %f, foff: %f, fch1: %f\n", survey->nchans, survey->nsamp,
         survey->tsamp, survey->foff, survey->fch1);
  printf("ndms: %u, max dm: %f, maxshift: %u\n", survey->tdms, hi_dm, g.maxshift);
  g.ok = true;
  return g.h_in;
}

// next_chunk (src/mdsm/dedispersion_manager.cpp:416-484)
float *next_chunk(unsigned int data_read, unsigned &samples, double timestamp, double blockRate,
                  float *mean, float *rms) {
  g.ppnsamp = g.pnsamp;
  g.pnsamp = data_read;
  if (!g.ok) {
    samples = (unsigned)-1;
    return nullptr;
  }
  if (data_read > 0) {
    g.timestamp = timestamp;
    g.block_rate = blockRate;
    g.mean = mean;
    g.rms = rms;
    g.survey->timestamp = timestamp;
    g.survey->blockRate = blockRate;
    g.survey->noiseMean = mean;
    g.survey->noiseRMS = rms;
  }
  if (g.loop_counter < 1) {
    samples = (unsigned)-1;
    return nullptr;
  }
  // the output of the chunk the previous start_processing took
  if (!hip_ok(hipMemcpyAsync(g.h_out, g.d_out, (size_t)g.tdms * g.nsamp * sizeof(float),
                             hipMemcpyDeviceToHost, g.stream), "output download") ||
      !hip_ok(hipStreamSynchronize(g.stream), "sync")) {
    samples = (unsigned)-1;
    return nullptr;
  }
  samples = g.ppnsamp;
  return g.h_out;
}

// start_processing (src/mdsm/dedispersion_manager.cpp:487-499) with the
// thread's copy-in (dedispersion_thread.cu:279-303) and the output thread's
// process_brute
int start_processing(unsigned int data_read) {
  if (!g.ok) return 0;
  if (data_read == 0 && g.out_switch) g.out_switch = false;
  else if (data_read == 0 && !g.out_switch) return 0;
  if (data_read > 0) {
    const size_t S = g.maxshift, N = g.nsamp, row = N + S;
    if (g.loop_counter == 0) {
      if (!hip_ok(hipMemcpy2DAsync(g.d_in, g.ld * 4, g.h_in, row * 4, row * 4, g.nchans,
                                   hipMemcpyHostToDevice, g.stream), "input upload"))
        return 0;
    } else {
      // the previous row's last maxshift samples move to the front (the
      // ranges are disjoint: nsamp >= maxshift)
      if (S > 0 && !hip_ok(hipMemcpy2DAsync(g.d_in, g.ld * 4, g.d_in + N, g.ld * 4, S * 4, g.nchans,
                                            hipMemcpyDeviceToDevice, g.stream), "carry"))
        return 0;
      if (!hip_ok(hipMemcpy2DAsync(g.d_in + S, g.ld * 4, g.h_in + S, row * 4, N * 4, g.nchans,
                                   hipMemcpyHostToDevice, g.stream), "input upload"))
        return 0;
    }
    if (mdsm_dedisperse(g.plan, g.d_in, g.ld, g.d_out, g.stream) != MDSM_OK) {
      fail("%s", mdsm_last_error());
      return 0;
    }
    if (!search_chunk(data_read)) return 0;
  }
  return ++g.loop_counter;
}

void tearDownMDSM() {
  free_all();
  printf("Finished Process\n");
}

extern "C" const char *mdsm_legacy_last_error(void) { return g_err.c_str(); }

extern "C" unsigned mdsm_legacy_candidates(const int32_t **dm_idx, const int32_t **sample,
                                           const float **snr) {
  if (dm_idx) *dm_idx = g.cdm.data();
  if (sample) *sample = g.csample.data();
  if (snr) *snr = g.csnr.data();
  return (unsigned)g.cdm.size();
}

static SURVEY *g_owned = nullptr;

extern "C" int mdsm_legacy_open(const char *xml_path, unsigned nsamp_override,
                                unsigned samples_per_chunk) {
  SURVEY *s = processSurveyParameters(xml_path ? xml_path : "");
  if (!s) return MDSM_ERR_ARG;
  if (nsamp_override) s->nsamp = nsamp_override;
  g_spc_override = samples_per_chunk;
  if (!initialiseMDSM(s)) {
    free(s->gpu_ids);
    free(s);
    return MDSM_ERR_ARG;
  }
  g_owned = s;
  return MDSM_OK;
}

extern "C" int mdsm_legacy_info(unsigned *nchans, unsigned *nsamp, unsigned *maxshift,
                                unsigned *tdms) {
  if (!g.ok) return MDSM_ERR_ARG;
  if (nchans) *nchans = g.nchans;
  if (nsamp) *nsamp = g.nsamp;
  if (maxshift) *maxshift = g.maxshift;
  if (tdms) *tdms = g.tdms;
  return MDSM_OK;
}

extern "C" float *mdsm_legacy_buffer(void) { return g.ok ? g.h_in : nullptr; }

extern "C" float *mdsm_legacy_next_chunk(unsigned data_read, unsigned *samples, double timestamp,
                                         double blockRate, float *mean, float *rms) {
  unsigned s = 0;
  float *p = next_chunk(data_read, s, timestamp, blockRate, mean, rms);
  if (samples) *samples = s;
  return p;
}

extern "C" int mdsm_legacy_start(unsigned data_read) { return start_processing(data_read); }

extern "C" void mdsm_legacy_close(void) {
  tearDownMDSM();
  if (g_owned) {
    free(g_owned->gpu_ids);
    free(g_owned);
    g_owned = nullptr;
  }
  g_spc_override = 0;
}
