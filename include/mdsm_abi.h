/*
 * mdsm_abi.h -- C ABI of the MI355X (gfx950) brute-force dedispersion and
 * single-pulse detection library (libmdsm_amd.so).
 *
 * Plain C: pointers, sizes and status codes only.  No entry point calls
 * exit(); every one returns an MDSM_* status and leaves a message for
 * mdsm_last_error().  Device pointers are HIP device allocations; "stream"
 * arguments are hipStream_t passed as void* (NULL = the null stream).
 *
 * Reference interfaces replaced (paths relative to the MDSM source tree):
 *   src/mdsm-multibeam/multibeam_dedispersion_wrapper.cu:4-23   the C seam
 *       (call_dedisperse, call_initialise_devices, call_allocate*Buffer)
 *   src/mdsm-multibeam/multibeam_dedispersion_thread.cu:204-221 shared_brute_force
 *   src/mdsm-multibeam/multibeam_dedispersion_kernel.cu:44-95   shared_dedispersion
 *   src/mdsm-multibeam/multibeam_dedispersion_output.cpp:188-321 process_brute_clustering
 *   src/mdsm-multibeam/dbscan.cpp:194-276                        DBScan::performClustering
 * Individual citations are given per declaration below.
 */
#ifndef MDSM_ABI_H
#define MDSM_ABI_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ------------------------------------------------------------------ status */
enum {
  MDSM_OK = 0,
  MDSM_ERR_ARG = 1,      /* invalid argument or survey parameters */
  MDSM_ERR_HIP = 2,      /* a HIP runtime call failed */
  MDSM_ERR_NOMEM = 3,    /* device or host allocation failed */
  MDSM_ERR_OVERFLOW = 4, /* a bounded output (candidate list) overflowed */
  MDSM_ERR_STATE = 5,    /* call out of sequence */
  MDSM_ERR_IO = 6,       /* a file could not be opened or read */
  MDSM_ERR_FORMAT = 7    /* malformed SIGPROC header */
};

/* Input sample storage types.  MDSM_F16_PAIRS stores halves exactly as
 * MDSM_F16 (the same rows, produced by mdsm_normalise with either code) and
 * changes only the dedisperser's arithmetic: channels 2k and 2k+1 of a trial
 * are added in FP16 (one IEEE round-to-nearest-even) and that pair sum is
 * added to the FP32 channel-order accumulator (an odd last channel alone),
 * which halves the FP32 adds; its rounding error against FP64 is about 1.7x
 * that of MDSM_F16. */
enum { MDSM_F32 = 0, MDSM_F16 = 1, MDSM_U8 = 2, MDSM_F16_PAIRS = 3 };

/* Delay-table conventions.
 * FP64_ROUND: D[c][d] = round(4148.808*DM_d*(f_c^-2 - f_top^-2)/tsamp), FP64.
 * REFERENCE: (int)(dm_shifts[c]*(lowdm + d*dmstep)/tsamp) in float with
 *   dm_shifts[c] = (float)(4148.741601*(F1^-2 - F2^-2)), F1 = fch1 + foff*c in
 *   float (multibeam_dedispersion_manager.cpp:325-326,366-369;
 *   multibeam_dedispersion_thread.cu:551-553).
 * LEGACY: floor(dm_shifts[c]*((lowdm + d*dmstep)/tsamp)) in float, the same
 *   dm_shifts (the single-beam manager's brute force:
 *   src/mdsm/dedispersion_manager.cpp:262-265,284-286 and
 *   src/mdsm/dedispersion_kernel.cu:55,68-69). */
enum { MDSM_DELAY_FP64_ROUND = 0, MDSM_DELAY_REFERENCE = 1, MDSM_DELAY_LEGACY = 2 };

/* Survey parameters of one beam: the <observation> XML fields of the
 * reference (SURVEY/BEAM, src/mdsm-multibeam/survey.h:8-88). */
typedef struct {
  double fch1;      /* MHz, centre of channel 0 (the top of the band) */
  double foff;      /* MHz per channel, negative (descending) */
  double tsamp;     /* seconds */
  double lowdm;     /* pc cm^-3 */
  double dmstep;    /* pc cm^-3 */
  uint32_t nchans;
  uint32_t ndms;    /* "tdms" / numDMs */
  uint32_t nsamp;   /* output samples per chunk */
  int32_t delay_mode;
} mdsm_survey_t;

/* One detection: DM trial index, first sample of the boxcar, log2 of the
 * boxcar width, signal-to-noise ratio. */
typedef struct {
  int32_t dm_idx;
  int32_t sample;
  int32_t width_log2;
  float snr;
} mdsm_cand_t;

const char *mdsm_last_error(void);
int mdsm_version(void);

/* --------------------------------------------------------- device helpers
 * Replace call_initialise_devices / call_allocateInputBuffer /
 * call_allocateOutputBuffer (multibeam_dedispersion_wrapper.cu:13-23). */
int mdsm_device_count(int *count);
int mdsm_set_device(int device);
int mdsm_malloc(void **ptr, size_t bytes);
int mdsm_free(void *ptr);
int mdsm_host_malloc(void **ptr, size_t bytes); /* pinned host memory */
int mdsm_host_free(void *ptr);
int mdsm_memcpy_htod(void *dst, const void *src, size_t bytes, void *stream);
int mdsm_memcpy_dtoh(void *dst, const void *src, size_t bytes, void *stream);
int mdsm_memcpy_dtod(void *dst, const void *src, size_t bytes, void *stream);
int mdsm_memset(void *ptr, int value, size_t bytes, void *stream);
int mdsm_stream_create(void **stream);
int mdsm_stream_destroy(void *stream);
int mdsm_stream_synchronize(void *stream);
int mdsm_device_synchronize(void);
/* Event timing on a stream: elapsed milliseconds between two records. */
int mdsm_event_create(void **event);
int mdsm_event_destroy(void *event);
int mdsm_event_record(void *event, void *stream);
int mdsm_event_elapsed(void *start, void *stop, float *ms);

/* ------------------------------------------------------------ delay table
 * dmdelay + per-beam dm_shifts + all_shifts + maxshift
 * (multibeam_dedispersion_manager.cpp:325-326,366-387;
 *  multibeam_dedispersion_thread.cu:546-559).  Host only.
 * table: int32[nchans][ndms] (may be NULL); maxshift: largest delay (may be NULL). */
int mdsm_delay_table(const mdsm_survey_t *s, int32_t *table, uint32_t *maxshift);

/* ------------------------------------------------------------------- plan
 * A plan owns one beam's device delay table and launch configuration.
 * Input rows are [nchans][ld] with ld >= nsamp + maxshift (the overlap-save
 * layout of multibeam_dedispersion_thread.cu:636-673); output rows are
 * [ndms][nsamp] float32. */
typedef struct mdsm_plan_s *mdsm_plan;
int mdsm_plan_create(const mdsm_survey_t *s, int in_dtype, mdsm_plan *plan);
int mdsm_plan_destroy(mdsm_plan plan);
/* maxshift of the plan's table and the minimum input row length (nsamp+maxshift). */
int mdsm_plan_info(mdsm_plan plan, uint32_t *maxshift, uint64_t *min_ld);
/* Device pointer to the plan's int32[nchans][ndms] delay table. */
int mdsm_plan_device_table(mdsm_plan plan, const int32_t **table);
/* Kernel tile variant (0 = the storage type's default); see
 * mdsm_tile_variants(): 15 shapes per storage type, and for MDSM_F16_PAIRS
 * and MDSM_U8 seven more (16-22) that build each channel pair's sums (FP16,
 * or exact u16) once per delay difference of its two channels; those take
 * a survey only when that
 * difference is >= 0 and spans at most 3 values (19: 4) in every DM tile of
 * the shape, and a shape that does not take the survey makes
 * mdsm_dedisperse fail with MDSM_ERR_ARG (the autotuners skip it).  Every
 * shape of a storage type gives the same results.  Setting a variant builds
 * its trial-offset table (one allocation and one kernel, synchronous); a
 * variant first used by mdsm_dedisperse without being set builds it there,
 * once. */
int mdsm_plan_set_variant(mdsm_plan plan, int variant);
int mdsm_plan_get_variant(mdsm_plan plan, int *variant);
int mdsm_tile_variants(int in_dtype, int *count);
/* The untuned default tile shape of a storage type.  A plan or batch uses
 * it when it takes the survey; otherwise MDSM_F16_PAIRS and MDSM_U8 fall
 * back to shapes 15 and 2 (which take every survey whose delay spread they
 * can stage), and any type to the first shape that can. */
int mdsm_tile_preferred(int in_dtype, int *variant);
/* Channels a tile shape stages per LDS block for this plan's delay spread;
 * 0 when the shape cannot stage it. */
int mdsm_plan_variant_channels(mdsm_plan plan, int variant, uint32_t *channels);
/* Time each tile variant on the given buffers (best of three launches after
 * a warm-up) and keep the fastest; the storage type's default variant is
 * kept when it is within 2 % of the fastest. */
int mdsm_plan_autotune(mdsm_plan plan, const void *d_in, uint64_t ld, float *d_out, void *stream,
                       int *best_variant, float *best_ms);

/* ---------------------------------------------------------- dedispersion
 * out[d][t] = sum_{c=0}^{nchans-1} in[c][t + D[c][d]], accumulated in FP32 in
 * channel order.  Replaces shared_brute_force / shared_dedispersion
 * (multibeam_dedispersion_thread.cu:204-221, multibeam_dedispersion_kernel.cu:44-95),
 * cached_brute_force / cache_dedispersion (:174-201, :8-41) and the legacy
 * brute_force_dedispersion (src/mdsm/dedispersion_thread.cu:220-243).
 * d_in dtype is the plan's: F32 floats, F16 halves, or U8 codes whose value
 * is u*(8/255) - 4 in normalised units (then the sum is exact in integers and
 * out = (float)(S*(8/255) - 4*nchans)); F16_PAIRS halves are summed as
 * out = sum_k fl32(fl16(in[2k] + in[2k+1])) in channel-pair order.  Rows must
 * start on 4-byte boundaries: ld a multiple of 2 (F16) or 4 (U8), rows
 * shorter than 2 GiB;
 * mdsm_plan_info reports the smallest valid ld.  A plan (like the
 * reference's per-beam thread state) is used by one host thread at a time:
 * the first call with a tile variant builds that variant's offset table. */
int mdsm_dedisperse(mdsm_plan plan, const void *d_in, uint64_t ld, float *d_out, void *stream);

/* --------------------------------------------------- bandpass normalisation
 * Per-channel FP64 mean/std over n samples of d_raw[nchans][ld] (fixed
 * reduction order), then v = (x - mean) * (1/std) stored as F32, F16 (or
 * F16_PAIRS, the same halves) or U8
 * (U8: clip to [-4,4], u = floor((v+4)*255/8 + 0.5)).  The per-channel
 * statistics of the reference bandpass stage (bandpass_power_sum,
 * multibeam_dedispersion_kernel.cu:317-350) feed a normaliser here. */
int mdsm_channel_stats(const float *d_raw, uint64_t ld, uint32_t nchans, uint64_t n,
                       double *d_mean, double *d_std, void *stream);
int mdsm_normalise(const float *d_raw, uint64_t ld, uint32_t nchans, uint64_t n,
                   const double *d_mean, const double *d_std, void *d_out, int out_dtype,
                   uint64_t ldo, void *stream);

/* ------------------------------------------------------------- detection
 * Per-block FP64 mean/std of d_x[nrows][nsamp] over blocks of L samples
 * (block b = samples [bL, min((b+1)L, nsamp))).  Output [nrows][ceil(nsamp/L)]. */
int mdsm_block_stats(const float *d_x, uint32_t nrows, uint32_t nsamp, uint32_t L,
                     double *d_mean, double *d_std, void *stream);
/* Dyadic boxcars of widths 1..2^(nwidths-1) (nwidths <= 7) in FP64, best
 * width per sample, threshold snr >= thresh; compacted to d_cands (unordered)
 * with the total count in *d_count (which the call zeroes first on the
 * stream).  Entries beyond cap are dropped but counted.  Replaces the
 * threshold scan of process_brute / process_brute_clustering
 * (multibeam_dedispersion_output.cpp:148-185,234-248). */
int mdsm_detect(const float *d_x, uint32_t nrows, uint32_t nsamp, uint32_t L,
                const double *d_mean, const double *d_std, int nwidths, double thresh,
                mdsm_cand_t *d_cands, uint32_t cap, uint32_t *d_count, void *stream);

/* Reference-semantics detection (process_brute and process_brute_clustering,
 * multibeam_dedispersion_output.cpp:148-185,188-248) on rows [0, nrows) x
 * samples [0, nsamp) of d_x with row stride ld (>= nsamp floats): FP64 mean
 * and rms over the first stats_nsamp samples of every row (stats_nsamp =
 * nsamp normally; the reference's last buffer takes its statistics over
 * last_nsamp samples, :216-229), rms = sqrt(S2/n - mean^2); candidate iff
 * x - mean >= rms*thresh over all nsamp samples, snr = (x - mean)/rms.
 * The sums run in one fixed order: one partial per 65536-sample chunk of
 * a row (lane i of 256 sums elements i, i+256, ... of the chunk, then a
 * binary tree), the partials in row-major chunk order summed the same way;
 * the reference adds all values in one sequential loop, which rounds
 * differently in the last bits.  Mean and rms stay on the device for the
 * threshold; when h_mean or h_rms is given they are copied back and the
 * stream is synchronised, otherwise the call only enqueues.  Workspace:
 * nrows*ceil(stats_nsamp/65536) partials, kept by the library per device
 * and stream (allocated on a stream's first call, grown when needed).
 * *d_count receives the number of candidates (may exceed cap). */
int mdsm_detect_global(const float *d_x, uint32_t nrows, uint32_t nsamp, uint64_t ld,
                       uint32_t stats_nsamp, float thresh, mdsm_cand_t *d_cands, uint32_t cap,
                       uint32_t *d_count, double *h_mean, double *h_rms, void *stream);
/* mdsm_detect_global without any allocation.  No counterpart in the
 * reference: for callers that cannot allocate device memory (code under HIP
 * graph capture, serving loops, harnesses that forbid hipMalloc).  The
 * workspace is a fixed 2 MiB device array per GPU instead of the library's
 * per-stream buffers, so the region is limited to
 * nrows * ceil(stats_nsamp / 65536) <= 131072 chunk partials (e.g. 4096 rows
 * of up to 2^21 samples); larger regions are rejected with MDSM_ERR_ARG.
 * Every call on a GPU uses that GPU's one array: calls that overlap on one
 * GPU (other streams or threads) race, so callers serialise them.  Same
 * arguments, order of sums and results as mdsm_detect_global. */
int mdsm_detect_global_noalloc(const float *d_x, uint32_t nrows, uint32_t nsamp, uint64_t ld,
                               uint32_t stats_nsamp, float thresh, mdsm_cand_t *d_cands,
                               uint32_t cap, uint32_t *d_count, double *h_mean, double *h_rms,
                               void *stream);
/* The threshold of mdsm_detect_global with the caller's mean and rms: the
 * reference's process_brute thresholds detrended series at mean 0, rms 1
 * (multibeam_dedispersion_output.cpp:165-166). */
int mdsm_threshold_global(const float *d_x, uint32_t nrows, uint32_t nsamp, uint64_t ld,
                          double mean, double rms, float thresh, mdsm_cand_t *d_cands,
                          uint32_t cap, uint32_t *d_count, void *stream);

/* Legacy single-beam detector (process_brute,
 * src/mdsm/dedispersion_output.cpp:135-213, one thread's share).  For each
 * row k (DM lowdm + k*dmstep, FP32) and sample l < nsamp - 4 the median of
 * samples l..l+4 (the selection network of dedispersion_output.cpp:14-53) is
 * a candidate iff DM > 1 and
 *   d_mean == NULL: median - mean >= rms*thresh, snr = median / rms, where
 *     mean and rms come from FP64 sums over 32768-value blocks of the first
 *     nvals values of the row-major array (the reference passes
 *     read_nsamp*ndms; 0 means ndms*nsamp) in the reference's order and
 *     types (its last block is left out when there is more than one);
 *     h_mean/h_rms receive them;
 *   d_mean != NULL: median - d_mean[l/spc] >= d_rms[l/spc]*thresh, snr =
 *     (median - d_mean[l/spc]) / d_rms[l/spc], with nstats entries in each
 *     device array (at least (nsamp-5)/spc + 1); h_mean/h_rms untouched.
 * All comparisons in FP32 as the reference rounds them.  ndms*nsamp < 2^32.
 * *d_count receives the number of candidates (may exceed cap; order not
 * defined). */
int mdsm_detect_legacy(const float *d_x, uint32_t ndms, uint32_t nsamp, float lowdm,
                       float dmstep, float thresh, uint64_t nvals, const float *d_mean,
                       const float *d_rms,
                       uint32_t samples_per_chunk, uint32_t nstats, mdsm_cand_t *d_cands,
                       uint32_t cap, uint32_t *d_count, float *h_mean, float *h_rms,
                       void *stream);

/* Observation parameters of an <observation> XML file in either dialect
 * (processSurveyParameters of src/mdsm/dedispersion_manager.cpp:40-200 and
 * src/mdsm-multibeam/multibeam_dedispersion_manager.cpp:69-321), first
 * beam only, as plain C: what the single-beam (legacy) manager reads. */
typedef struct {
  float fch1, foff, tsamp, lowdm, dmstep, detection_threshold;
  uint32_t nchans, tdms, nsamp, nbits;
  int use_brute_force;            /* dm@useBruteForce; 1 when absent */
  char fileprefix[120], basedir[120];
  uint32_t secs_per_file;
  char use_pc_time, single_file_mode;
  int dialect;                    /* MDSM_XML_* of mdsm_compat.h: 1 multibeam, 2 legacy, 3 mixed */
} mdsm_xml_params_t;
int mdsm_parse_observation_xml(const char *path, mdsm_xml_params_t *out);

/* 5-tap running median per row, edges replicated (median_filter,
 * multibeam_dedispersion_kernel.cu:98-157), out of place. */
int mdsm_median_filter5(const float *d_x, uint32_t nrows, uint32_t nsamp, float *d_out,
                        void *stream);
/* Linear detrend and RMS normalisation of consecutive segments of length L,
 * in place, FP64 accumulation (detrend_normalise,
 * multibeam_dedispersion_kernel.cu:207-295). */
int mdsm_detrend_normalise(float *d_x, uint64_t nseg, uint32_t L, void *stream);

/* -------------------------------------------------------------- clustering
 * Host side.  mdsm_sort_candidates orders by (dm_idx, sample).
 * mdsm_dbscan2d: points (dm_idx, sample), Chebyshev distance, neighbour iff
 * |dd| <= eps and |dt| <= eps; core iff >= minpts points in the neighbourhood
 * (self included); clusters 0.. in order of their first core point in the
 * array order; border points join the lowest-numbered reachable cluster;
 * noise = -1.  mdsm_cluster_peaks: per cluster the highest-snr member (first
 * in array order on ties) and the member count. */
int mdsm_sort_candidates(mdsm_cand_t *cands, uint32_t n);
int mdsm_dbscan2d(const mdsm_cand_t *cands, uint32_t n, int eps, int minpts, int32_t *labels,
                  uint32_t *nclusters);
int mdsm_cluster_peaks(const mdsm_cand_t *cands, const int32_t *labels, uint32_t n,
                       uint32_t nclusters, mdsm_cand_t *peaks, uint32_t *sizes);
/* Reference 3-D box DBSCAN (DBScan::performClustering + getNeighbours,
 * dbscan.cpp:194-276,390-415): neighbours strictly within (min_time, min_dm,
 * min_snr), excluding self and already-assigned or visited points; a point
 * is a seed iff it has >= min_points such neighbours; ids from 1, noise -1. */
int mdsm_dbscan_ref(const double *time, const double *dm, const double *snr, uint32_t n,
                    float min_time, float min_dm, float min_snr, uint32_t min_points,
                    int32_t *cluster, uint32_t *nclusters);
/* Reference cluster classification (Cluster::computeTransientProbability,
 * dbscan.cpp:25-171) over the n member points of one cluster, with the
 * survey's DM grid, minimum pulse width (ms) and band (beam 0's fch1, foff,
 * nchans).  *score: 1 when a rule rejects the cluster (peak DM <= 2, pulse
 * width not above min_pulse_width), else the mean squared distance of the
 * smoothed DM-S/N curve to the smearing model; the reference keeps clusters
 * with score < 0.1 (multibeam_dedispersion_output.cpp:262-287).  info
 * (may be NULL) receives {peak DM, peak S/N, width ms, centre time}. */
int mdsm_classify_cluster_ref(const double *time, const double *dm, const double *snr, uint32_t n,
                              float min_dm, float dm_step, uint32_t num_dms, float min_pulse_width,
                              float fch1, float foff, uint32_t nchans, float *score, double *info);

/* ---------------------------------------------------------- multi-beam batch
 * Many beams of uneven nchans/nsamp/ndms dedispersed by one launch over
 * all (beam, tile) items, largest beams first (launch shapes below).
 * Replaces the per-beam host threads of
 * multibeam_dedispersion_thread.cu:464-839 (one launch per beam per chunk). */
typedef struct mdsm_batch_s *mdsm_batch;
int mdsm_batch_create(const mdsm_survey_t *beams, uint32_t nbeams, int in_dtype,
                      mdsm_batch *batch);
int mdsm_batch_destroy(mdsm_batch batch);
int mdsm_batch_info(mdsm_batch batch, uint32_t beam, uint32_t *maxshift, uint64_t *min_ld);
/* d_in[b]: beam b input [nchans_b][ld[b]]; d_out[b]: [ndms_b][nsamp_b].
 * bind uploads the beam descriptors and builds missing trial-offset tables
 * (synchronous); launch enqueues only the dedispersion kernel (and, in the
 * persistent shape, a counter clear) with no allocation and no host copy,
 * so it may be captured into a hipGraph and replayed while the bound
 * buffers stay valid.  dedisperse = bind + launch. */
int mdsm_batch_bind(mdsm_batch batch, const void *const *d_in, const uint64_t *ld,
                    float *const *d_out);
int mdsm_batch_launch(mdsm_batch batch, void *stream);
int mdsm_batch_dedisperse(mdsm_batch batch, const void *const *d_in, const uint64_t *ld,
                          float *const *d_out, void *stream);
int mdsm_batch_set_variant(mdsm_batch batch, int variant);
/* Tile autotuner for a batch (the reference has none: its tile is fixed at
 * 128 x 32, multibeam_dedispersion_thread.cu:204-221), in two stages: every
 * tile shape on every beam truncated to 2^16 samples (reading d_in[i], ld[i]
 * into a scratch output it allocates), each shape's batch time estimated as
 * the sum of its beams' times scaled to their full length; then the three
 * least estimates among the shapes that fit every beam are timed as one
 * whole launch each into d_out, and the fastest is kept (the next bind uses
 * it; the buffers are left unbound).  best_variant receives the batch's
 * variant afterwards, best_ms its launch time. */
int mdsm_batch_autotune(mdsm_batch batch, const void *const *d_in, const uint64_t *ld,
                        float *const *d_out, void *stream, int *best_variant, float *best_ms);
/* Launch shape (takes effect at the next bind): FLAT (default) launches one
 * workgroup per tile, numbered so that the DM tiles re-reading one input
 * window share an XCD (and its L2); PERSISTENT runs resident workgroups that
 * drain eight device queues of tiles (one per XCD) in the same order.  Both
 * are one launch per step. */
enum { MDSM_BATCH_PERSISTENT = 0, MDSM_BATCH_FLAT = 1 };
int mdsm_batch_set_mode(mdsm_batch batch, int mode);

/* ------------------------------------------------------- search pipeline
 * The per-chunk sequence of a multi-beam search: for every beam bandpass
 * normalisation of the raw FP32 input into the storage type, one
 * batched dedispersion, then per-block statistics and boxcar detection into
 * a per-beam candidate list.  The reference runs these stages from separate
 * host threads (multibeam_dedispersion_thread.cu:464-839,
 * multibeam_dedispersion_output.cpp:325-452).  mdsm_pipeline_run records the
 * sequence into a hipGraph on its first call with use_graph set and replays
 * the graph afterwards (counters are cleared by a kernel inside the graph).
 * The pipeline owns the normalised inputs, outputs and candidate buffers;
 * raw inputs are bound by the caller and must stay valid while the graph is
 * replayed. */
typedef struct {
  uint32_t stats_block; /* samples per statistics block (e.g. 16384) */
  int32_t nwidths;      /* boxcar widths 1..2^(nwidths-1), <= 7 */
  double thresh;        /* S/N threshold */
  uint32_t cand_cap;    /* candidates kept per beam */
} mdsm_detect_params_t;
typedef struct mdsm_pipeline_s *mdsm_pipeline;
int mdsm_pipeline_create(const mdsm_survey_t *beams, uint32_t nbeams, int storage_dtype,
                         const mdsm_detect_params_t *params, mdsm_pipeline *pipeline);
int mdsm_pipeline_destroy(mdsm_pipeline pipeline);
/* d_raw[b]: beam b raw input [nchans_b][ld_raw[b]] FP32 with
 * ld_raw[b] >= nsamp_b + maxshift_b (see mdsm_batch_info). */
int mdsm_pipeline_bind(mdsm_pipeline pipeline, const float *const *d_raw, const uint64_t *ld_raw);
int mdsm_pipeline_set_variant(mdsm_pipeline pipeline, int variant);
int mdsm_pipeline_run(mdsm_pipeline pipeline, void *stream, int use_graph);
/* Sorted candidates of beam b from the last run (host copy); *n receives
 * the total found (may exceed cap, then MDSM_ERR_OVERFLOW is returned). */
int mdsm_pipeline_candidates(mdsm_pipeline pipeline, uint32_t beam, mdsm_cand_t *host,
                             uint32_t cap, uint32_t *n);
/* Device pointer to beam b's dedispersed series [ndms][nsamp]. */
int mdsm_pipeline_output(mdsm_pipeline pipeline, uint32_t beam, float **d_out);

/* --------------------------------------------------- bandpass and RFI
 * The reference's bandpass_fitting + rfi_clipping stage
 * (multibeam_dedispersion_thread.cu:224-390; kernels
 * multibeam_dedispersion_kernel.cu:302-450) on raw [nchans][ld] FP32 rows,
 * over the n samples from column col0 (the reference's `shift`).  Device
 * kernels keep the reference's reduction orders (bit-identical results);
 * the polynomial fit runs on the host in FP64. */
typedef struct {
  uint32_t ncoeffs;         /* polynomial terms (reference default 8) */
  uint32_t channel_block;   /* samples per channel block (default 1024) */
  float channel_thresh;     /* block flagged above rz(fit) + this * rmse (default 10) */
  float spectrum_thresh;    /* spectrum clipped outside mean +- this * std (default 6) */
  uint32_t nmasks;          /* channel ranges replaced by their neighbours */
  const int32_t *mask_from, *mask_to;
  int32_t apply_channel_clip; /* 0: flags computed, not applied (the reference
                                 clears them before channel_clipper,
                                 multibeam_dedispersion_thread.cu:364) */
} mdsm_rfi_params_t;
typedef struct {
  float mean, std, rmse;    /* of the masked bandpass; rmse against the fit */
  float channel_thresh;     /* channel_thresh * rmse */
  float spectrum_hi, spectrum_lo;
} mdsm_bandpass_stats_t;
/* d_bandpass[c] = mean of row c (bandpass_power_sum, kernel.cu:317-350). */
int mdsm_bandpass_sum(const float *d_x, uint64_t ld, uint32_t nchans, uint64_t col0, uint64_t n,
                      double *d_bandpass, void *stream);
/* Host: masks, least-squares fit at x_i = i/nchans, statistics and
 * thresholds (thread.cu:224-296, 344-346).  fitted: host double[nchans]. */
int mdsm_bandpass_fit(const double *bandpass, uint32_t nchans, const mdsm_rfi_params_t *params,
                      double *fitted, mdsm_bandpass_stats_t *stats);
/* d_flags: uint8 [nchans][ceil(n/channel_block)], cleared by the call
 * (channel_flagger, kernel.cu:353-398). */
int mdsm_rfi_channel_flag(const float *d_x, uint64_t ld, uint32_t nchans, uint64_t col0, uint64_t n,
                          const double *d_fitted, uint32_t channel_block, float thresh,
                          uint8_t *d_flags, void *stream);
/* channel_clipper, kernel.cu:402-423. */
int mdsm_rfi_channel_clip(float *d_x, uint64_t ld, uint32_t nchans, uint64_t col0, uint64_t n,
                          const double *d_fitted, uint32_t channel_block, const uint8_t *d_flags,
                          void *stream);
/* spectrum_clipper, kernel.cu:427-450; d_count (may be NULL) is incremented
 * per clipped spectrum. */
int mdsm_rfi_spectrum_clip(float *d_x, uint64_t ld, uint32_t nchans, uint64_t col0, uint64_t n,
                           const double *d_fitted, float hi, float lo, uint32_t *d_count,
                           void *stream);
/* voltage_to_power, kernel.cu:302-312: (re, im) int16 pairs -> power. */
int mdsm_voltage_to_power(float *d_x, uint64_t ld, uint32_t nchans, uint64_t col0, uint64_t n,
                          void *stream);
/* The whole stage: sum, host fit, flag (when d_flags is given), clip.
 * d_fitted: device double[nchans] (receives the fitted bandpass).
 * Synchronises the stream (the fit needs the sums on the host). */
int mdsm_rfi_clean(float *d_x, uint64_t ld, uint32_t nchans, uint64_t col0, uint64_t n,
                   const mdsm_rfi_params_t *params, double *d_fitted, uint8_t *d_flags,
                   mdsm_bandpass_stats_t *stats, void *stream);

/* ------------------------------------------------------- triage network
 * A BF16 CNN that classifies candidate events from paired snippets (DM-time
 * and dedispersed subband-time, each H x W) as pulse (class 1) or not
 * (class 0).  It replaces the reference's heuristic cluster triage
 * (Cluster::computeTransientProbability, dbscan.cpp:25-171), which stays
 * available as mdsm_dbscan_ref + classification on the host.
 *   conv 3x3 (pad 1) 2->16->32->64, each + bias, ReLU, 2x2 max pool;
 *   fc 64*(H/8)*(W/8) -> 64 (ReLU) -> 2 logits; softmax cross-entropy.
 * Products run on MFMA with BF16 operands and FP32 accumulation; master
 * weights, biases and Adam moments are FP32.
 * Parameter arrays (host, float) use the canonical layout w1 b1 w2 b2 w3 b3
 * fw1 fb1 fw2 fb2: conv weights [co][ci][ky][kx], fw1[o][c*(H/8)*(W/8) +
 * y*(W/8) + x], fw2[o][i]. */
typedef struct mdsm_cnn_s *mdsm_cnn;
int mdsm_cnn_create(int height, int width, int max_batch, mdsm_cnn *net);
int mdsm_cnn_destroy(mdsm_cnn net);
int mdsm_cnn_param_count(mdsm_cnn net, uint64_t *count);
int mdsm_cnn_shape(mdsm_cnn net, int *height, int *width, int *max_batch);
int mdsm_cnn_set_params(mdsm_cnn net, const float *host_params); /* also resets Adam */
int mdsm_cnn_get_params(mdsm_cnn net, float *host_params);
int mdsm_cnn_get_grads(mdsm_cnn net, float *host_grads);    /* of the last backward */
int mdsm_cnn_init(mdsm_cnn net, uint64_t seed);              /* He-normal weights, zero biases */
/* d_x: BF16 bits [batch][2][H][W]; d_logits (may be NULL): float [batch][2]. */
int mdsm_cnn_forward(mdsm_cnn net, const uint16_t *d_x, uint32_t batch, float *d_logits,
                     void *stream);
/* Gradients of the mean cross-entropy of the last forward; d_labels int32
 * [batch] in {0,1}; d_loss (may be NULL) receives the loss. */
int mdsm_cnn_backward(mdsm_cnn net, const int32_t *d_labels, uint32_t batch, float *d_loss,
                      void *stream);
int mdsm_cnn_adam_step(mdsm_cnn net, float lr, float weight_decay, void *stream);
int mdsm_cnn_train_step(mdsm_cnn net, const uint16_t *d_x, const int32_t *d_labels,
                        uint32_t batch, float lr, float *d_loss, void *stream);

/* ------------------------------------------------------- triage data
 * Synthetic training and evaluation streams for the triage network, made by
 * the plain search (normalise, dedisperse, per-block statistics, boxcar
 * detection, 2-D DBSCAN, cluster peaks) over seeded chunks with injected
 * dispersed pulses and RFI: zero-DM broadband spikes, narrowband bursts,
 * 50 Hz-modulated channels and on/off ("telegraph") channels. */
enum { MDSM_RFI_BURST = 0, MDSM_RFI_MAINS = 1, MDSM_RFI_TELEGRAPH = 2 };
typedef struct {
  uint32_t chan;
  int32_t kind;     /* MDSM_RFI_* */
  int64_t start;    /* first sample (row index) */
  int64_t len;      /* samples */
  float amp;        /* in units of the channel's noise sigma (scaled by its gain) */
  float period;     /* MAINS: samples per cycle; TELEGRAPH: samples per on/off block */
} mdsm_rfi_item_t;
/* Adds the items to raw rows [nchans][ld] (columns [0, n)) on the device:
 * BURST amp; MAINS amp*(1+sin(2 pi t/period))/2; TELEGRAPH amp in blocks
 * chosen with probability 1/4 by Philox4x32-10 from (seed, block, chan). */
int mdsm_inject_rfi(const mdsm_rfi_item_t *host_items, uint32_t nitems, const float *d_gain,
                    uint64_t seed, float *d_raw, uint64_t ld, uint64_t n, void *stream);

typedef struct {
  mdsm_survey_t survey;        /* nsamp = output samples per chunk */
  uint32_t npulses;            /* injected pulses per chunk */
  double dm_lo, dm_hi;         /* pulse DM, uniform */
  double width_lo, width_hi;   /* pulse width in samples, log-uniform */
  double snr_lo, snr_hi;       /* matched-boxcar S/N, uniform */
  uint32_t n_spikes, n_bursts, n_mains, n_telegraph; /* RFI items per chunk */
  double rfi_lo, rfi_hi;       /* RFI strength (spike S/N; narrowband amplitude in sigma) */
  double thresh;               /* detection S/N (6) */
  uint32_t stats_block;        /* samples per statistics block (16384) */
  int32_t nwidths;             /* boxcar widths (7) */
  int32_t eps, minpts;         /* DBSCAN (3, 4) */
  int32_t snip_dms, snip_samples, snip_subbands; /* snippet shape (64, 256, 64) */
} mdsm_triage_config_t;
typedef struct {
  int32_t dm_idx, sample, width_log2;  /* cluster peak */
  float snr;
  int32_t cluster_size;
  int32_t label;                       /* 1: matches an injected pulse */
  int32_t pulse;                       /* index of that pulse in its chunk, else -1 */
  int32_t chunk;
} mdsm_event_t;
typedef struct {
  uint64_t chunks, injected, injected_found; /* pulses injected / matched by >= 1 cluster */
  uint64_t clusters, false_clusters;          /* all clusters / clusters matching no pulse */
  uint64_t candidates;                        /* threshold crossings before clustering */
} mdsm_triage_stats_t;
/* Snippets of nev events (host arrays: DM index, centre sample) from a
 * beam's dedispersed series d_out [ndms][nsamp] and its normalised FP32
 * input d_store [nchans][ld] with the plan's delay table d_table
 * [nchans][ndms]: plane 0 = out rows dm-SD/2 .. dm+SD/2-1 times scale0,
 * plane 1 = SB subband sums (channels in order) of the input dedispersed at
 * the event DM times scale1, both over samples tc-ST/2 .. tc+ST/2-1, zero
 * outside the arrays, clipped to +-clip, stored as BF16 [nev][2][SD][ST]
 * (SB == SD). */
int mdsm_triage_snippets(const float *d_out, uint32_t ndms, uint32_t nsamp, const float *d_store,
                         uint64_t ld, uint32_t nchans, const int32_t *d_table, const int32_t *ev_dm,
                         const int32_t *ev_t, uint32_t nev, int SD, int ST, int SB, float scale0,
                         float scale1, float clip, uint16_t *d_snippets, void *stream);
typedef struct mdsm_triage_sim_s *mdsm_triage_sim;
int mdsm_triage_sim_create(const mdsm_triage_config_t *config, uint64_t seed, mdsm_triage_sim *sim);
int mdsm_triage_sim_destroy(mdsm_triage_sim sim);
/* Runs chunks until nchunks are done or max_events events are stored.
 * Snippets go to d_snippets (BF16 [max_events][2][snip_dms or
 * snip_subbands][snip_samples]; both planes have snip_dms rows, so
 * snip_subbands must equal snip_dms) from index *n_events on; events to
 * host_events; *n_events is advanced.  Stats accumulate. */
int mdsm_triage_sim_run(mdsm_triage_sim sim, uint32_t nchunks, uint16_t *d_snippets,
                        mdsm_event_t *host_events, uint32_t max_events, uint32_t *n_events,
                        mdsm_triage_stats_t *stats);
/* One pass over n snippets in a seeded random order, batches of `batch`;
 * *mean_loss receives the mean of the batch losses. */
int mdsm_cnn_train_epoch(mdsm_cnn net, const uint16_t *d_snippets, const int32_t *d_labels,
                         uint32_t n, uint32_t batch, float lr, uint64_t seed, float *mean_loss,
                         void *stream);
/* Pulse probability (softmax class 1) of n snippets, device float[n]. */
int mdsm_cnn_predict(mdsm_cnn net, const uint16_t *d_snippets, uint32_t n, float *d_prob,
                     void *stream);

/* ------------------------------------------------------- SIGPROC ingest
 * Filterbank files: a keyword header (read_header,
 * src/mdsm-multibeam/file_handler.cpp:56-190) followed by time-major
 * spectra of nbits-wide samples (1/2/4 bits packed least-significant first,
 * 8/16 unsigned, 32 float; read_block, file_handler.cpp:193-287).  The
 * reference unpacks on the host and corner-turns into [beam][chan][t]
 * float rows (readBinaryData, src/MDSM.cpp:71-104); here packed bytes cross
 * PCIe and one kernel unpacks and transposes them into the device rows. */
typedef struct {
  char source_name[81], rawdatafile[81];
  int32_t machine_id, telescope_id, data_type, nchans, nbits, nifs;
  int32_t nbeams, ibeam, barycentric, pulsarcentric, nbins, nsamples;
  double tstart, tsamp, fch1, foff, refdm, az_start, za_start, src_raj, src_dej, period;
  int64_t npuls;
  int32_t nfchannel;           /* entries of frequency_table (fchannel keywords) */
  int32_t has_header;          /* 0: no HEADER_START, raw 8-bit data from byte 0 */
  int32_t total_bytes;         /* the reference's FILE_HEADER.total_bytes count */
  int64_t data_offset;         /* byte offset of the first sample */
  double frequency_table[4096];
} mdsm_sigproc_header_t;
/* Parses the header of `path` (MDSM_ERR_FORMAT on a malformed or unknown
 * keyword, where the reference exits the process).  A file without
 * HEADER_START is raw 8-bit data (src/MDSM.cpp:53-56). */
int mdsm_sigproc_read_header(const char *path, mdsm_sigproc_header_t *header);
/* Unpack + corner turn on the device: d_packed holds nspectra time-major
 * spectra of nchans_total = nbeams*nchans samples (nchans_total*nbits a
 * multiple of 8).  Sample (t, b*nchans + c) is written as float to
 * d_out[b*beam_stride + c*ld + col0 + t]. */
int mdsm_unpack_corner_turn(const void *d_packed, int nbits, uint32_t nspectra, uint32_t nchans,
                            uint32_t nbeams, float *d_out, uint64_t ld, uint64_t beam_stride,
                            uint64_t col0, void *stream);
/* Streaming reader: opens the file at its first sample and owns two pinned
 * + device staging slots of max_spectra spectra.  nchans is the channel
 * count per beam (0: the header's nchans / nbeams); the header, when given,
 * receives the parsed header. */
typedef struct mdsm_sigproc_reader_s *mdsm_sigproc_reader;
int mdsm_sigproc_open(const char *path, uint32_t nchans, uint32_t nbeams, uint32_t max_spectra,
                      mdsm_sigproc_reader *reader, mdsm_sigproc_header_t *header);
int mdsm_sigproc_close(mdsm_sigproc_reader reader);
/* Reads up to nspectra whole spectra (fewer at the end of the file; *got
 * receives the count), copies the packed bytes to the device and unpacks
 * them as mdsm_unpack_corner_turn does, on `stream`.  Returns once the file
 * read is done and the copy and unpack are queued; the slot is reused two
 * calls later, after its event. */
int mdsm_sigproc_read(mdsm_sigproc_reader reader, uint32_t nspectra, float *d_out, uint64_t ld,
                      uint64_t beam_stride, uint64_t col0, uint32_t *got, void *stream);

/* Per-DM SIGPROC output (DedispersedDataWriter,
 * src/pelican-mdsm/modules/DedispersedDataWriter.cpp:17-135): one file
 * "<prefix>_<dm %g>.fil" per requested DM with the reference's header
 * (machine_id 0, telescope_id 0, data_type 1, fch1, foff, nchans 1, tsamp,
 * nbits 32, tstart, nifs 1); each write appends, for every requested DM,
 * the FP32 host row of `series` [nrows][nsamp] whose trial DM
 * lowdm + j*dmstep is within 1e-5 of it. */
typedef struct mdsm_dm_writer_s *mdsm_dm_writer;
int mdsm_dm_writer_open(const char *prefix, const float *dms, uint32_t ndms, double fch1, double foff,
                        double tsamp, double tstart, mdsm_dm_writer *writer);
int mdsm_dm_writer_write(mdsm_dm_writer writer, const float *series, uint32_t nrows, uint32_t nsamp,
                         float lowdm, float dmstep);
int mdsm_dm_writer_close(mdsm_dm_writer writer);

/* Logarithmic ("mu-law") quantisation of raw beam data for the disk writer
 * (quantise_power_data / quantise_complex_data,
 * multibeam_dedispersion_writer.cpp:21-118).  Power: byte =
 * ((int)(d/Q)*Q + Q/2) * 2^bits with d = log10(1 + k*mu/32768)/log10(1+mu),
 * k = (int)(x / max * 32768), max = twice the maximum of the first buffer
 * (state->max_value 0 on the first call; kept afterwards).  Complex (int16
 * re, im pairs): 3-bit magnitudes + sign bits, [rs r r r is i i i].  The
 * reference overruns its table for values above max and for -32768; these
 * are clamped to the last entry. */
typedef struct {
  uint32_t bits;         /* output bits (outputBits), 1..8 */
  uint32_t compression;  /* mu (compression) */
  float max_value;       /* 0 before the first buffer */
} mdsm_mulaw_state_t;
int mdsm_mulaw_encode_power(const float *d_in, uint64_t n, mdsm_mulaw_state_t *state, uint8_t *d_out,
                            void *stream);
int mdsm_mulaw_encode_complex(const int16_t *d_in, uint64_t n, uint32_t bits, uint32_t compression,
                              uint8_t *d_out, void *stream);

/* ------------------------------------------------------- synthetic input
 * Seeded synthetic filterbank data generated on the device (Philox4x32-10
 * counters, Box-Muller normals): raw[c][t0+i] = gain[c]*(z + pulses) +
 * offset[c] for i < n, written to d_raw[c][i] (row stride ld).  Pulses:
 * p_start[p] (absolute sample), p_width[p], p_amp[p] (normalised units), and
 * p_delay[p*nchans + c] (per-channel delay in samples).  Host arrays. */
int mdsm_generate(uint64_t seed, uint32_t nchans, int64_t t0, uint64_t n, const float *gain,
                  const float *offset, uint32_t npulses, const int64_t *p_start,
                  const int32_t *p_width, const float *p_amp, const int64_t *p_delay,
                  float *d_raw, uint64_t ld, void *stream);

#ifdef __cplusplus
}
#endif
#endif /* MDSM_ABI_H */
