/*
 * mdsm_legacy.h -- the single-beam (legacy) MDSM manager API on top of
 * libmdsm_amd, in libmdsm_legacy.so.
 *
 * The names, signatures, SURVEY struct and call protocol of
 * src/mdsm/dedispersion_manager.h:7-12 and src/mdsm/survey.h, which the
 * Pelican module (src/pelican-mdsm/modules/MdsmModule.cpp:30-151) drives:
 *
 *   SURVEY *s = processSurveyParameters("obs.xml");
 *   float *in = initialiseMDSM(s);           // [nchans][nsamp + maxshift]
 *   loop: fill in;  out = next_chunk(read, samples, timestamp, blockRate,
 *                                    mean, rms);
 *         if (!start_processing(read)) break;
 *   tearDownMDSM();
 *
 * Semantics kept from the reference (src/mdsm/dedispersion_thread.cu:279-303,
 * dedispersion_manager.cpp:416-499, dedispersion_output.cpp:135-300):
 *  - the input buffer holds nchans rows of nsamp + maxshift floats; the first
 *    chunk supplies the whole row, every later chunk the nsamp samples after
 *    the first maxshift (the library carries the previous row's last
 *    maxshift samples in front, on the device);
 *  - next_chunk returns the dedispersed [tdms][nsamp] output of the chunk
 *    submitted at the previous call (NULL and samples = -1 before there is
 *    one), with samples = that call's data_read; data_read = 0 twice ends
 *    the run;
 *  - each chunk is searched by the legacy detector (median of five,
 *    threshold, DM > 1) with the statistics of its own values or, when
 *    next_chunk was given mean/rms arrays, with mean/rms[l / samplesPerChunk];
 *    candidates are written as "time, dm, snr" (or "time, dm, snr, mean,
 *    rms") lines to <basedir>/<fileprefix>.dat (singleFileMode) or to
 *    timestamped files rotated every secondsPerFile.
 * Differences, by design:
 *  - processSurveyParameters takes a path string (no Qt); both XML dialects
 *    are read.  Brute force only: a survey with useBruteForce = 0 (subband
 *    dedispersion) is refused.
 *  - one GPU, no worker threads: start_processing dedisperses and searches
 *    the chunk before it returns; errors return 0 / NULL with
 *    mdsm_legacy_last_error() instead of exit().
 *  - the delays are the legacy kernel's floor(dm_shifts[c]*(DM/tsamp)) (the
 *    MDSM_DELAY_LEGACY table); where that table's largest delay exceeds the
 *    reference's maxshift by one sample, the extra column reads as 0 (the
 *    reference reads the next channel's first sample there).
 *  - nsamp below maxshift is raised to the next power of two (the
 *    multibeam rule), so that the carried samples never overlap.
 *
 * Both this header and mdsm_compat.h define SURVEY: a program includes one.
 * libmdsm_legacy.so and libmdsm_amd.so both export initialiseMDSM(SURVEY*),
 * start_processing and tearDownMDSM under the same C++ names, so a legacy
 * program links -lmdsm_legacy before -lmdsm_amd (both libraries bind their
 * own calls internally).
 */
#ifndef MDSM_LEGACY_H
#define MDSM_LEGACY_H

#include <stdint.h>
#include <stdio.h>

#include <string>

typedef struct {
  float lowdm, highdm, dmstep, sub_dmstep;
  int binsize, ndms, ncalls, calldms, mean, stddev;
} SUBBAND_PASSES;

typedef struct {
  unsigned int nsamp, nchans, tdms, maxshift, nbits;
  float tsamp, foff, fch1;
  bool useBruteForce;
  bool useL1Cache;
  float lowdm, dmstep;
  SUBBAND_PASSES *pass_parameters;
  unsigned num_passes, nsubs;
  double timestamp;
  double blockRate;
  float *noiseMean;
  float *noiseRMS;
  unsigned samplesPerChunk;
  FILE *fp;
  char fileprefix[80], basedir[120];
  unsigned secs_per_file;
  char use_pc_time, single_file_mode;
  float detection_threshold;
  unsigned num_threads;
  unsigned *gpu_ids;
  unsigned num_gpus;
} SURVEY;

SURVEY *processSurveyParameters(const std::string &filepath);
float *initialiseMDSM(SURVEY *input_survey);
float *next_chunk(unsigned int data_read, unsigned &samples, double timestamp = 0,
                  double blockRate = 0, float *mean = NULL, float *rms = NULL);
int start_processing(unsigned int data_read);
void tearDownMDSM();

/* Extensions (no reference counterpart): errors, the candidates of the last
 * searched chunk, and plain-C forms of the calls above for bindings. */
extern "C" {
const char *mdsm_legacy_last_error(void);
unsigned mdsm_legacy_candidates(const int32_t **dm_idx, const int32_t **sample, const float **snr);
int mdsm_legacy_open(const char *xml_path, unsigned nsamp_override, unsigned samples_per_chunk);
int mdsm_legacy_info(unsigned *nchans, unsigned *nsamp, unsigned *maxshift, unsigned *tdms);
float *mdsm_legacy_buffer(void);
float *mdsm_legacy_next_chunk(unsigned data_read, unsigned *samples, double timestamp,
                              double blockRate, float *mean, float *rms);
int mdsm_legacy_start(unsigned data_read);
void mdsm_legacy_close(void);
}

#endif /* MDSM_LEGACY_H */
