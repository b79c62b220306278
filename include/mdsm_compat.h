/*
 * mdsm_compat.h -- MDSM multibeam manager API on top of libmdsm_amd.
 *
 * The names, signatures, structs and call protocol of the reference's
 * multibeam manager (src/mdsm-multibeam/multibeam_dedispersion_manager.h:9-28,
 * src/mdsm-multibeam/survey.h:8-88), so that drivers written against it
 * (src/MDSM.cpp, src/medicina-transient-pipeline/pipeline.cpp) keep their
 * call sequence:
 *
 *   SURVEY *s = processSurveyParameters("obs.xml");
 *   initialiseMDSM(s);
 *   loop: float *in = get_buffer_pointer();  // fill [beam][chan][time] floats
 *         next_chunk_multibeam(read, samples, timestamp, blockRate);
 *         if (!start_processing(read)) break;
 *   tearDownMDSM();
 *
 * Differences from the reference, by design:
 *  - processSurveyParameters takes a path string (no Qt); it reads both XML
 *    dialects (multibeam names such as nchans/nsamp/detectionThreshold and
 *    the older number/threshold/ids names) and records which one it used.
 *  - errors are reported by mdsm_compat_last_error() and a 0 / NULL return
 *    instead of exit().
 *  - one GPU: all beams go through one batched dedispersion launch per
 *    chunk; uploads, dedispersion and detection + downloads of successive
 *    chunks overlap on three HIP streams (the reference's three stages).
 *  - optional FP16 / U8 input storage (<storage type="fp16"/> or "u8"):
 *    each chunk's rows are bandpass-normalised per channel (mean and
 *    standard deviation of the row's valid samples) before quantisation,
 *    so the dedispersed series are sums of normalised samples.
 *  - stream semantics: the first call supplies samples [0, maxshift) of each
 *    channel; every later chunk supplies nsamp new samples, and its
 *    dedispersed output covers the nsamp samples starting maxshift samples
 *    earlier (overlap-save, multibeam_dedispersion_thread.cu:636-673).
 *  - candidates are written per beam to <basedir>/<fileprefix>[_beam_<b>].dat
 *    in the reference's CSV formats (multibeam_dedispersion_output.cpp:179,
 *    :275-310) and are also available through mdsm_compat_candidates().
 */
#ifndef MDSM_COMPAT_H
#define MDSM_COMPAT_H

#include <stdint.h>
#include <stdio.h>

#include <string>

#define NULLVALUE -999
#define DEDISP_THREADS 128 /* input chunks are trimmed to multiples of this (src/MDSM.cpp:134) */
#define DEDISP_DMS 32
#define MDSM_STAGES 3

typedef struct {
  unsigned beam_id, gpu_id;
  float foff, fch1;
  float *dm_shifts;
  unsigned maxshift;
} BEAM;

typedef struct {
  unsigned from, to;
} RANGE;

typedef struct {
  bool voltage;
  unsigned nbeams, npols, nchans, nsamp, nbits, nantennas;
  float tsamp;
  unsigned last_nsamp;
  BEAM *beams;
  bool apply_beamforming;
  float lowdm, dmstep;
  unsigned tdms;
  double timestamp;
  double blockRate;
  FILE *fp;
  char fileprefix[120], basedir[120];
  unsigned secs_per_file;
  char use_pc_time, single_file_mode;
  float bandpass_mean, bandpass_std, bandpass_rmse;
  unsigned ncoeffs;
  bool apply_rfi_clipper;
  float spectrum_thresh, channel_thresh;
  unsigned channel_block, num_masks;
  RANGE *channel_mask;
  float detection_threshold;
  bool apply_median_filter;
  bool apply_detrending;
  bool apply_clustering, apply_classification;
  unsigned dbscan_min_points;
  float dbscan_time_range, dbscan_dm_range, dbscan_snr_range;
  float min_pulse_width;
  bool tbb_enabled;
  bool dump_to_disk;
  unsigned output_bits;
  unsigned output_compression;
  unsigned num_threads;
  unsigned *gpu_ids;
  unsigned num_gpus;
  /* Extension, after the reference's fields so that their offsets are
   * unchanged: input storage of the dedisperser, MDSM_F32 (0, the
   * reference's), MDSM_F16 (1), MDSM_U8 (2) or MDSM_F16_PAIRS (3), from
   * <storage type="fp32|fp16|u8|fp16pairs"/>. */
  unsigned input_storage;
  /* Extension: 1 keeps the dedispersed series on the device, so that
   * next_chunk_multibeam returns NULL and the chunk's series cross no bus
   * (callers such as src/MDSM.cpp ignore them; detection and the candidate
   * files are unchanged); 0, the reference's behaviour, downloads them into
   * the returned buffers.  From <series download="0|1"/>. */
  unsigned series_on_device;
} SURVEY;

/* XML dialect recorded by processSurveyParameters. */
enum { MDSM_XML_MULTIBEAM = 1, MDSM_XML_LEGACY = 2, MDSM_XML_MIXED = 3 };

SURVEY *processSurveyParameters(const std::string &filepath);
void initialiseMDSM(SURVEY *input_survey);
float **next_chunk_multibeam(unsigned int data_read, unsigned &samples, double timestamp = 0,
                             double blockRate = 0);
float *get_buffer_pointer();
unsigned char *get_antenna_pointer();
int start_processing(unsigned int data_read);
void tearDownMDSM();

/* Extensions (no reference counterpart). */
extern "C" {
const char *mdsm_compat_last_error(void);
int mdsm_compat_xml_dialect(void);
/* Candidates of the last chunk whose detection finished (start_processing(k)
 * finishes chunk k-1; the first stop call finishes the last chunk) of a
 * beam: (time, dm, snr, cluster) arrays owned by the library, valid until
 * the next start_processing. */
unsigned mdsm_compat_candidates(unsigned beam, const double **time, const double **dm,
                                const double **snr, const int **cluster);
/* Input storage in use (MDSM_F32, MDSM_F16 or MDSM_U8). */
int mdsm_compat_storage(void);
}

#endif /* MDSM_COMPAT_H */
