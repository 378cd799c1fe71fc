/* hd/stages.h -- the search's stage functions over device memory.
 *
 * The reference exposes these stages as C++ classes over Thrust vectors
 * (Pipeline/hd/*.h).  Here each is an extern "C" function over raw device
 * pointers, batched over `n_series` DM series laid out `series_stride`
 * floats apart, running on the HIP stream passed as `stream` (NULL = the
 * default stream).  Argument order and meaning follow the reference call
 * each one replaces.  All return an hd_error.
 */
#pragma once

#include "hd/types.h"
#include "hd/error.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Giant list in device memory (structure of arrays, capacity entries each).
 * `count` receives the number of giants found; when it exceeds `capacity`
 * only the first `capacity` are stored. */
typedef struct hd_giants {
  hd_float* peaks;
  hd_size*  inds;
  hd_size*  begins;
  hd_size*  ends;
  hd_size   capacity;
  hd_size   count;
} hd_giants;

/* replaces dedisp_execute_adv(plan, nsamps, in, in_nbits, in_stride, out,
 * out_nbits, out_stride, flags) (called at Pipeline/pipeline.cu:452-462), with
 * the pipeline as the dedisp plan: its DM list, delay table and killmask
 * (hd_create_pipeline's params).  d_in: nsamps time samples of in_stride =
 * nchans*in_nbits/8 bytes in device memory (channels packed low bits first);
 * d_out: one row of nsamps - max_delay floats per DM trial, out_stride BYTES
 * apart, each value the exact integer sum over the live channels
 * (out_nbits must be 32).  flags is accepted for the signature and ignored:
 * both pointers are device memory.  Runs on the pipeline's stream and
 * returns when the output is written. */
hd_error hd_dedisperse(hd_pipeline pl, hd_size nsamps, const hd_byte* d_in, hd_size in_nbits,
                       hd_size in_stride, hd_byte* d_out, hd_size out_nbits,
                       hd_size out_stride, unsigned flags);

/* replaces RemoveBaselinePlan::exec (Pipeline/remove_baseline.cu:21-160,
 * hd/remove_baseline.h:19-23).  A series holds nout = beam_stride*nbeams -
 * (beam_stride - beam_count) samples; the last beam's resampling reads the
 * element one past them, d_data[nout] of each series, as the reference reads
 * the float after its series (median_filter.cu:460-462), so series_stride
 * must be at least nout + 1 and that element readable (the search keeps it
 * 0). */
hd_error hd_remove_baseline(hd_float* d_data, hd_size beam_stride, hd_size beam_count,
                            hd_size smooth_radius, hd_size nbeams,
                            hd_size n_series, hd_size series_stride, void* stream);

/* replaces GetRMSPlanMB::exec_multibeam per-beam form (Pipeline/get_rms.cu:66-100):
 * d_rms receives nbeams values per series. */
hd_error hd_get_rms_multibeam(const hd_float* d_data, hd_float* d_rms, hd_size beam_stride,
                              hd_size beam_count, hd_size nbeams,
                              hd_size n_series, hd_size series_stride, void* stream);

/* replaces normalise_multibeam(d, d_rms, beam_stride, cur_nsamps, nbeams)
 * (Pipeline/get_rms.cu:202-212) */
hd_error hd_normalise_multibeam(hd_float* d_data, const hd_float* d_rms, hd_size beam_stride,
                                hd_size cur_nsamps, hd_size nbeams,
                                hd_size n_series, hd_size series_stride, void* stream);

/* replaces GetRMSPlanMB::exec_multibeam pooled form (Pipeline/get_rms.cu:104-131):
 * d_block_rms receives one value per series. */
hd_error hd_block_rms(const hd_float* d_data, hd_float* d_block_rms, hd_size beam_stride,
                      hd_size beam_count, hd_size nbeams,
                      hd_size n_series, hd_size series_stride, void* stream);

/* hd_block_rms without allocation.  NO COUNTERPART in the reference: for
 * callers that cannot allocate device memory (code under HIP graph capture,
 * serving loops, harnesses that forbid it).  Same arguments, kernels and
 * results as hd_block_rms, but its scratch (n_series*np partial medians and
 * n_series reciprocals, np = floor(nbeams*beam_count / 5^L) with 5^L the
 * largest power of five <= min(nbeams*beam_count, 3125)) lives in a fixed
 * device array of HD_NOALLOC_SCRATCH_WORDS floats (4 MiB) that each GPU the
 * library runs on holds until the process ends.  Returns HD_MEM_ALLOC_FAILED,
 * launching nothing, when round_up(n_series*np, 64) + n_series exceeds it.
 * No hipMalloc/hipFree is called.  Calls on one GPU share the array: calls
 * that overlap on one GPU (concurrent streams or threads) race. */
#define HD_NOALLOC_SCRATCH_WORDS (1u << 20)
hd_error hd_block_rms_noalloc(const hd_float* d_data, hd_float* d_block_rms,
                              hd_size beam_stride, hd_size beam_count, hd_size nbeams,
                              hd_size n_series, hd_size series_stride, void* stream);

/* replaces normalise_multibeam(d, block_rms, count) (Pipeline/get_rms.cu:189-200),
 * with one block_rms per series read from device memory. */
hd_error hd_normalise_block(hd_float* d_data, const hd_float* d_block_rms, hd_size count,
                            hd_size n_series, hd_size series_stride, void* stream);

/* replaces MatchedFilterPlan<float>::prep (Pipeline/matched_filter.cu:21-32):
 * d_scanned receives count+1 prefix sums per series (scanned_stride apart). */
hd_error hd_matched_filter_prep(const hd_float* d_in, hd_size count, hd_float* d_scanned,
                                hd_size n_series, hd_size series_stride,
                                hd_size scanned_stride, void* stream);

/* replaces MatchedFilterPlan<float>::exec (Pipeline/matched_filter.cu:37-67):
 * writes ceil((scanned_count - max_width)/tscrunch) values per series. */
hd_error hd_matched_filter_exec(const hd_float* d_scanned, hd_size scanned_count,
                                hd_size max_width, hd_size width, hd_size tscrunch,
                                hd_float* d_out, hd_size n_series, hd_size scanned_stride,
                                hd_size out_stride, void* stream);

/* replaces GiantFinder::exec (Pipeline/find_giants.cu:101-316) for one series;
 * giants are written in time order from index 0 of `out`. */
hd_error hd_find_giants(const hd_float* d_data, hd_size count, hd_size nbeams,
                        hd_float thresh, hd_size merge_dist, hd_size beam_stride,
                        hd_size beam_count, hd_giants* out, void* stream);

/* replaces label_candidate_clusters (Pipeline/label_candidate_clusters.cu:203-394),
 * with the reference's single-thread result: each giant takes the smallest index
 * reachable along coincident links to earlier giants. */
hd_error hd_label_candidate_clusters(hd_size count, const struct ConstRawCandidates* d_cands,
                                     hd_size time_tol, hd_size filter_tol, hd_size dm_tol,
                                     hd_size nsamps_beam, hd_size* d_labels,
                                     hd_size* label_count, void* stream);

/* replaces merge_candidates (Pipeline/merge_candidates.cu:74-134) */
hd_error hd_merge_candidates(hd_size count, const hd_size* d_labels,
                             const struct ConstRawCandidates* d_cands,
                             const struct RawCandidates* d_groups, void* stream);

/* replaces clean_filterbank_rfi (Pipeline/clean_filterbank_rfi.cu:256-528,
 * hd/clean_filterbank_rfi.h:14-26).  The reference takes the channel count,
 * f0 and df from its dedisp plan; here they are arguments.  d_in and d_out
 * are device filterbanks of nsamps x nchans*nbits/8 bytes (that row size
 * must be a multiple of 4); the first nsamps - max_delay(dm) time samples of
 * d_out receive the cleaned data.  h_killmask and rfi_min_beams are unused,
 * as in the reference. */
hd_error hd_clean_filterbank_rfi(hd_size nchans, hd_float f0, hd_float df,
                                 const hd_byte* d_in, hd_size nsamps, hd_size nbits,
                                 hd_byte* d_out, int* h_killmask, hd_float dm, hd_float dt,
                                 hd_float baseline_length, hd_float rfi_tol,
                                 hd_size rfi_min_beams, hd_size boxcar_max, hd_size nbeams,
                                 void* stream);

/* replaces apply_manual_killmasks (Pipeline/clean_filterbank_rfi.cu:530-545):
 * zeroes killmask[j] for j in [start, end) of each range (end exclusive). */
hd_error hd_apply_manual_killmasks(int* h_killmask, hd_size nchans,
                                   unsigned int num_channel_zaps,
                                   const hd_range_t* channel_zaps);

/* replaces dedisp_generate_dm_list (called at Pipeline/pipeline.cu:178-182):
 * returns the number of trials and writes at most `cap` of them. */
hd_size hd_generate_dm_list(hd_float* out, hd_size cap, hd_float dm_min, hd_float dm_max,
                            hd_float dt, hd_float pulse_width_us, hd_float f0, hd_float df,
                            hd_size nchans, hd_float tol);

#ifdef __cplusplus
}
#endif
