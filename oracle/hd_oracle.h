/*
 * hd_oracle.h -- CPU restatement of Heimdall's per-gulp search (hd_execute).
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker that the HIP
 * implementation is compared against.  Only tests/, __graft_entry__.smoke()
 * and bench.py's cpu_baseline leg may load it; the product library never
 * links, loads or calls it.
 *
 * Every function follows the sequential semantics of the reference
 * (ajameson/heimdall_multibeam, Pipeline/ *.cu built with a sequential
 * Thrust backend), with each float/double conversion written out in the
 * order the cited reference line performs it.  Citations are file:line in
 * the reference tree.
 */
#ifndef HD_ORACLE_H
#define HD_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- medians (Pipeline/median_filter.cu:21-75): order statistics ---- */
float orc_median3(float a, float b, float c);
float orc_median4(float a, float b, float c, float d);
float orc_median5(float a, float b, float c, float d, float e);

/* ---- dedisp plan: DM list, delay table, delays ---- */
/* DM trial list (dedisp_generate_dm_list, pipeline.cu:178-182).  Returns the
 * number of trials written (at most cap). */
size_t orc_generate_dm_list(float* out, size_t cap,
                            float dm_min, float dm_max, float dt_s,
                            float pulse_width_us, float f0, float df,
                            size_t nchans, float tol);
/* Per-channel dispersion delay slope in samples per unit DM. */
void orc_delay_table(float* out, size_t nchans, float dt_s, float f0, float df);
/* Integer delay of channel c at DM dm: (int)(dm*k_c + 0.5f). */
int orc_delay(float dm, float kc);

/* ---- direct dedispersion (dedisp_execute_adv, pipeline.cu:452-462) ----
 * in: [nsamps][nchans*nbits/8] bytes, channels packed low-bits-first.
 * killmask: nchans ints (0 = excluded) or NULL.
 * out: [n_dm][out_stride] floats, out[d][t] = (float)sum for t < nsamps-max_delay.
 * nthreads > 1 parallelises over DMs with OpenMP (result independent of it). */
int orc_dedisperse(const uint8_t* in, size_t nsamps, size_t nbits, size_t nchans,
                   const int* killmask, const float* dm_list, size_t n_dm,
                   const float* delay_table, size_t max_delay,
                   float* out, size_t out_stride, int nthreads);

/* ---- stage restatements operating on host arrays ---- */
/* linear_stretch / linear_stretch_beam (median_filter.cu:451-519).  The
 * caller keeps in[in_count] (one past the end) readable: the reference reads
 * it when float rounding lands the last output at the end of the input. */
void orc_linear_stretch(const float* in, size_t in_count, float* out,
                        size_t out_count);
void orc_linear_stretch_beam(const float* in, size_t in_count_beam, float* out,
                             size_t out_count_beam, size_t nbeam);
/* Baseline removal over one DM series (remove_baseline.cu:21-160).
 * work: persistent scratch object (see orc_work_*) or NULL for a fresh one. */
struct orc_work;
struct orc_work* orc_work_create(void);
void orc_work_destroy(struct orc_work* w);
int orc_remove_baseline(struct orc_work* w, float* d, size_t beam_stride,
                        size_t beam_count, size_t smooth_radius, size_t nbeams);
/* Per-beam RMS (get_rms.cu:66-100) */
int orc_get_rms_multibeam(float* d, float* rms, size_t beam_stride,
                          size_t beam_count, size_t nbeams);
/* Per-beam normalise (get_rms.cu:174-212) */
void orc_normalise_multibeam(float* d, const float* rms, size_t beam_stride,
                             size_t cur_nsamps);
/* Pooled block RMS (get_rms.cu:104-131) */
float orc_block_rms(const float* d, size_t beam_stride, size_t beam_count,
                    size_t nbeams);
/* Block normalise (get_rms.cu:189-200) */
void orc_normalise_block(float* d, float block_rms, size_t count);
/* Matched filter (matched_filter.cu:21-67): scan has count+1 entries. */
void orc_scan(const float* x, size_t count, float* scanned);
size_t orc_boxcar(const float* scanned, size_t scanned_size, size_t max_width,
                  size_t width, size_t tscrunch, float* out);

/* ---- giants / clustering / merging ---- */
typedef struct {
  float    peak;
  uint64_t ind, begin, end, filter, dm, members;
} orc_cand;

/* GiantFinder::exec (find_giants.cu:101-316), appending to out (capacity cap).
 * Returns the number of giants found (may exceed cap: then not all written). */
size_t orc_find_giants(const float* d, size_t count, float thresh,
                       size_t merge_dist, size_t beam_stride, size_t beam_count,
                       orc_cand* out, size_t cap);
/* label_candidate_clusters (label_candidate_clusters.cu:203-394) with the
 * sequential (single-thread) semantics.  Returns label_count. */
size_t orc_label_clusters(size_t n, const orc_cand* c, size_t time_tol,
                          size_t filter_tol, size_t dm_tol, size_t nsamps_beam,
                          uint64_t* labels);
/* merge_candidates (merge_candidates.cu:74-134). groups: label_count entries. */
void orc_merge_candidates(size_t n, const uint64_t* labels, const orc_cand* c,
                          orc_cand* groups);

/* ---- zero-DM RFI cleaning (clean_filterbank_rfi.cu:256-528) ---- */
/* hash() (clean_filterbank_rfi.cu:33-43): Robert Jenkins' 32-bit integer hash. */
uint32_t orc_hash32(uint32_t a);
/* thrust::default_random_engine (minstd_rand: x <- 48271 x mod 2^31-1) seeded
 * with s, as in linear_congruential_engine::seed. */
uint32_t orc_minstd_seed(uint32_t s);
/* Advances the engine state and returns the new output. */
uint32_t orc_minstd_next(uint32_t* x);
/* One draw of thrust::uniform_int_distribution<unsigned>(lo, hi) from the
 * engine state *x (double-precision uniform-real path of
 * thrust/random/detail/uniform_{int,real}_distribution.inl). */
uint32_t orc_uniform_draw(uint32_t* x, uint32_t lo, uint32_t hi);
/* zap_filterbank_rfi (clean_filterbank_rfi.cu:51-108,197-234): 32-bit words of
 * the first nsamps time samples; masked samples are replaced channel by
 * channel from random unmasked nearby samples. */
int orc_zap_filterbank_rfi(const int* mask, const uint8_t* in, size_t nsamps, size_t nbits,
                           size_t nchans, size_t max_resample_dist, uint8_t* out);
/* clean_filterbank_rfi: series at one DM (dm 0 in the pipeline), baseline,
 * per-beam RMS, sigma cut and boxcars 1..boxcar_max, then the zap.  Writes
 * the first nsamps - max_delay(dm) samples of out; mask_out (may be NULL)
 * receives that many mask entries.  Returns 0, or -1 for shapes the
 * reference cannot process (word stride, too few samples). */
int orc_clean_filterbank_rfi(const uint8_t* in, size_t nsamps, size_t nbits, size_t nchans,
                             float dm, float dt, float f0, float df, float baseline_length,
                             float rfi_tol, size_t boxcar_max, size_t nbeams, uint8_t* out,
                             int* mask_out);

/* ---- adaptive time resolution (dedisp's, as pipeline.cu:187-196 enables it) ----
 * Parity unpinned: dedisp is absent from the reference tree.  The rule and the
 * halving follow dedisp's published method; oracle/adaptive_dt.py states the
 * same rules in numpy, and the tests pin the two against each other and
 * against the invariants. */
/* Scrunch factor of each trial: 1 for the first; a trial doubles its
 * predecessor's factor when doubling the sampling time raises the effective
 * pulse width (float32) by less than a factor tol. */
void orc_scrunch_factors(const float* dms, size_t n, float dt, float pulse_width_us, float f0,
                         float df, size_t nchans, float tol, size_t* out);
/* Dedispersion at the trials' own resolutions: a trial of factor s sums the
 * unpacked input halved log2(s) times (pairwise means rounded down, four
 * zero samples appended) with delays orc_delay(dm, k_c / s), into the first
 * (nsamps - max_delay) / s floats of its row; the rest of the row is zero.
 * factors must not decrease along the list.  Returns 0, -1 (nbits) or -2
 * (max_delay >= nsamps). */
int orc_dedisperse_adaptive(const uint8_t* in, size_t nsamps, size_t nbits, size_t nchans,
                            const int* killmask, const float* dm_list, size_t n_dm,
                            const float* delay_table, size_t max_delay, const size_t* factors,
                            float* out, size_t out_stride, int nthreads);

/* ---- whole-gulp restatement of hd_execute (pipeline.cu:233-1087) ---- */
typedef struct {
  size_t nchans;
  float  dt, f0, df;
  float  dm_min, dm_max, dm_tol, dm_pulse_width;
  float  baseline_length;
  size_t boxcar_max;
  float  detect_thresh;
  size_t cand_sep_time, cand_sep_filter, cand_sep_dm;
  float  max_giant_rate;
  size_t min_tscrunch_width;
  size_t beam;                 /* first beam index (0-based) for output */
  const int* killmask;         /* nchans ints or NULL (all live) */
  int    nthreads;             /* OpenMP threads (dedispersion and the DM loop) */
  size_t dm_first, dm_count;   /* DM subset processed (dm_count 0 = all) */
  size_t dm_stride;            /* process every dm_stride-th trial of the subset (0 = 1) */
  int    cleaning;             /* run clean_filterbank_rfi first (the CLEANING build) */
  float  rfi_tol;              /* sigma cut of the cleaning (default 5) */
  int    use_scrunching;       /* adaptive time resolution (pipeline.cu:187-196; default 0 here) */
  float  scrunch_tol;          /* its smearing tolerance (default_params.C: 1.15) */
} orc_params;

typedef struct {
  int      error;              /* 0 ok, 12 = too many events, other = invalid */
  size_t   n_dm;
  size_t   max_delay;
  size_t   nsamps_processed;
  size_t   n_giants;
  orc_cand* giants;            /* malloc'd, n_giants entries */
  size_t   n_groups;
  orc_cand* groups;            /* malloc'd, n_groups entries (unfiltered) */
  float*   group_dms;          /* malloc'd, n_groups entries */
  double   seconds;            /* wall clock of the call */
} orc_result;

int  orc_execute(const orc_params* p, const uint8_t* fb, size_t nsamps,
                 size_t nbits, size_t first_idx, size_t nbeams, orc_result* r);
void orc_result_free(orc_result* r);
/* Formats the .cand lines (pipeline.cu:1012-1040) into buf (NUL-terminated).
 * Returns the number of bytes needed (excluding NUL). */
size_t orc_format_cands(const orc_result* r, const orc_params* p, size_t nsamps,
                        size_t nbeams, size_t first_idx, char* buf, size_t cap);

/* default params (default_params.C:10-66) in orc_params form */
void orc_default_params(orc_params* p);

#ifdef __cplusplus
}
#endif
#endif
