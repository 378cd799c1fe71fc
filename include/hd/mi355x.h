/* hd/mi355x.h -- additive entry points of this implementation.
 *
 * None of these has a counterpart in the reference interface; they expose
 * what the reference keeps inside hd_execute (candidates, giants, stage
 * timings) and let callers hand over device-resident filterbanks or several
 * beams of different lengths in one call.
 */
#pragma once

#include "hd/types.h"
#include "hd/error.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Candidate groups of the last hd_execute* call, before the per-beam valid
 * window filter applied when writing the .cand file.  Copies at most `cap`
 * entries into the host arrays (any may be NULL) and returns the count. */
hd_size hd_get_last_candidates(hd_pipeline pl, hd_size cap, hd_float* peaks, hd_size* inds,
                               hd_size* begins, hd_size* ends, hd_size* filter_inds,
                               hd_size* dm_inds, hd_size* members, hd_float* dms);

/* Giants of the last call in the reference's order (DM, then width, then
 * time).  Same conventions as hd_get_last_candidates. */
hd_size hd_get_last_giants(hd_pipeline pl, hd_size cap, hd_float* peaks, hd_size* inds,
                           hd_size* begins, hd_size* ends, hd_size* filter_inds,
                           hd_size* dm_inds);

/* The .cand text the last call wrote (or would have written), NUL-terminated;
 * returns the length needed. */
hd_size hd_get_last_cand_text(hd_pipeline pl, char* buf, hd_size cap);

/* hd_execute on a filterbank already in device memory (same layout as the
 * host buffer of hd_execute).  The stream is the pipeline's own. */
hd_error hd_execute_device(hd_pipeline pl, const hd_byte* d_filterbank, hd_size nsamps,
                           hd_size nbits, hd_size first_idx, hd_size nbeams,
                           hd_size* nsamps_processed);

/* One gulp of a batched call: what a separate hd_execute call would get. */
typedef struct hd_gulp_desc {
  const hd_byte* filterbank;  /* host or device bytes, [nsamps][nchans*nbits/8] */
  hd_size nsamps;             /* time samples (all STF beams of this gulp) */
  hd_size nbeams;             /* STF beams in this gulp (1 for a ragged beam batch) */
  hd_size first_idx;          /* index of the gulp's first sample in its stream */
  hd_size beam;               /* params.beam of the separate call (output beam number) */
} hd_gulp_desc;

/* Searches `count` gulps of different lengths in one call: one launch per
 * stage over all of them.  Each gulp's candidates, .cand file and status are
 * those of a separate hd_execute(pl, filterbank, nsamps, nbits, first_idx,
 * nbeams) with params.beam = beam.  The reference interface takes one gulp
 * per call (Pipeline/hd/pipeline.h:21-23); this entry point has no
 * counterpart there.  on_device: the filterbank pointers are device memory.
 * nsamps_processed (count entries) receives each gulp's processed count and
 * errors (count entries, may be NULL) each gulp's status; the return value is
 * a call error, else the first gulp error in order, else HD_NO_ERROR. */
hd_error hd_execute_batch(hd_pipeline pl, const hd_gulp_desc* gulps, hd_size count,
                          hd_size nbits, int on_device, hd_size* nsamps_processed,
                          hd_error* errors);

/* Gulp index (position in the last batch) of each candidate returned by
 * hd_get_last_candidates; candidates are in gulp order.  Returns the count. */
hd_size hd_get_last_candidate_gulps(hd_pipeline pl, hd_size* gulp_index, hd_size cap);

/* Number of DM trials and the maximum delay (samples) of the plan. */
hd_size hd_get_dm_count(hd_pipeline pl);
hd_size hd_get_max_delay(hd_pipeline pl);
/* Copies the DM trial list (at most cap values); returns the count. */
hd_size hd_get_dm_list(hd_pipeline pl, hd_float* out, hd_size cap);

/* Stage timings (milliseconds, HIP events) of the last call, in the order
 * unpack, dedisperse, baseline, rms, scan, block_rms, giants, cluster,
 * total, clean (total includes clean).  Returns the number of values
 * written (at most cap). */
hd_size hd_get_last_timings(hd_pipeline pl, double* ms, hd_size cap);

/* The dedispersion tile of the last call: trials per block (8, 16 or 32),
 * whether blocks were ordered for the 8 XCDs' L2 caches, whether samples were
 * staged as bytes (4 per LDS word: summed in 8-bit lanes for 1-, 2- and 4-bit
 * input, as even/odd 16-bit-lane sums for 8-bit input) rather than as 16-bit
 * pairs, the channels staged in LDS per step (4, 8 or 16), and the threads
 * per block (512, or 1024 for byte staging at 32 trials).  The first call of
 * a shape times every candidate on the real input and keeps the fastest (all
 * give the same sums); HD_DEDISP_TILE="db,remap,bytes,cc,wide" in the
 * environment fixes it for a process.  Any output pointer may be NULL. */
void hd_get_dedisp_tile(hd_pipeline pl, int* trials_per_block, int* xcd_remap, int* byte_lanes,
                        int* channels_per_step, int* block_threads);
/* Fixes the dedispersion tile (trials_per_block 8, 16 or 32; channels_per_step
 * 4, 8 or 16; block_threads 512 or 1024, the latter for byte staging at 32
 * trials), or returns to the autotuner (trials_per_block 0). */
hd_error hd_set_dedisp_tile(hd_pipeline pl, int trials_per_block, int xcd_remap, int byte_lanes,
                            int channels_per_step, int block_threads);

/* Runs the per-gulp stage chain from dedispersion to thresholding as a HIP
 * graph: captured on the first call of a set of launch arguments (shapes,
 * dedispersion tile, buffers) and replayed while they repeat (on by default;
 * HD_GRAPH=0 in the environment or this call with 0 launches the kernels one
 * by one).  Gated calls always launch one by one.  In graph mode the stage
 * timings report the whole chain as the dedispersion stage. */
void hd_set_graph(hd_pipeline pl, int enabled);
/* Graphs captured so far; launches (may be NULL) receives the replays. */
hd_size hd_get_graph_stats(hd_pipeline pl, hd_size* launches);

/* When enabled, hd_execute* skips the .cand file / socket output (used by
 * benchmarks and tests that read candidates through the accessors). */
void hd_set_output_enabled(hd_pipeline pl, int enabled);

/* Runs the zero-DM RFI cleaning before every search (off by default).  The
 * reference selects it at compile time (-DCLEANING, Pipeline/pipeline.cu:
 * 255-299, 447-451): clean_filterbank_rfi at DM 0 with boxcars up to 64 and
 * params.rfi_tol, then dedispersion of the cleaned copy. */
void hd_set_cleaning(hd_pipeline pl, int enabled);

#ifdef __cplusplus
}
#endif
