/* hd/dedisp_plan.h -- a dedispersion plan with the calls the reference's tools
 * make on the dedisp library (Applications/candidate_profiler.C:150-164,
 * 261-327; Pipeline/clean_filterbank_rfi.cu:361-397; Pipeline/pipeline.cu:
 * 171-196, 341-363, 452-462).  dedisp is an external library absent from the
 * reference tree; each function below names the dedisp call it stands for
 * and keeps its arguments' meaning.  Sums are exact integers over the live
 * channels, computed on the GPU by the search's dedispersion kernel.
 *
 * Adaptive time resolution (hd_dedisp_enable_adaptive_dt): trial d is
 * computed at scrunch factor s_d (hd_dedisp_get_dt_factors) from the input
 * averaged over s_d consecutive samples (pairwise means rounded down, halving
 * the resolution once per doubling), with delays in those coarser samples;
 * its output row then holds (nsamps - max_delay) / s_d samples.  These rules
 * restate dedisp's published method and are not pinned by any reference
 * vector.
 */
#pragma once

#include "hd/types.h"
#include "hd/error.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct hd_dedisp_plan_t* hd_dedisp_plan;

/* hd_dedisp_execute flags (dedisp's DEDISP_HOST_POINTERS / DEVICE_POINTERS) */
#define HD_DEDISP_HOST_POINTERS 0u
#define HD_DEDISP_DEVICE_POINTERS 1u

/* dedisp_create_plan(&plan, nchans, dt, f0, df): dt in seconds, f0 the first
 * channel's frequency and df the channel step, in MHz.  The plan runs on the
 * current HIP device; it starts with no DM list, every channel live and
 * adaptive time resolution off. */
hd_error hd_dedisp_create_plan(hd_dedisp_plan* plan, hd_size nchans, hd_float dt, hd_float f0,
                               hd_float df);
/* dedisp_destroy_plan */
void hd_dedisp_destroy_plan(hd_dedisp_plan plan);
/* dedisp_set_dm_list(plan, dm_list, count) */
hd_error hd_dedisp_set_dm_list(hd_dedisp_plan plan, const hd_float* dm_list, hd_size count);
/* dedisp_generate_dm_list(plan, dm_start, dm_end, pulse_width_us, tol): the
 * list hd_generate_dm_list gives */
hd_error hd_dedisp_generate_dm_list(hd_dedisp_plan plan, hd_float dm_start, hd_float dm_end,
                                    hd_float pulse_width_us, hd_float tol);
/* dedisp_set_killmask(plan, killmask): nchans flags, 0 = channel excluded */
hd_error hd_dedisp_set_killmask(hd_dedisp_plan plan, const int* killmask);
/* dedisp_enable_adaptive_dt(plan, pulse_width_us, tol) */
hd_error hd_dedisp_enable_adaptive_dt(hd_dedisp_plan plan, hd_float pulse_width_us,
                                      hd_float tol);
/* dedisp_disable_adaptive_dt(plan) */
hd_error hd_dedisp_disable_adaptive_dt(hd_dedisp_plan plan);
/* dedisp_get_dt_factors(plan): one factor per trial (all 1 when adaptive
 * time resolution is off); valid until the plan's DM list or adaptive
 * setting changes */
const hd_size* hd_dedisp_get_dt_factors(hd_dedisp_plan plan);
/* dedisp_get_max_delay(plan): the largest delay in samples over trials and
 * channels (full resolution) */
hd_size hd_dedisp_get_max_delay(hd_dedisp_plan plan);
/* dedisp_get_dm_delay(plan, dm_idx): trial dm_idx's largest delay */
hd_size hd_dedisp_get_dm_delay(hd_dedisp_plan plan, hd_size dm_idx);
/* dedisp_get_dm_count, dedisp_get_dm_list, dedisp_get_channel_count,
 * dedisp_get_dt, dedisp_get_f0, dedisp_get_df */
hd_size hd_dedisp_get_dm_count(hd_dedisp_plan plan);
const hd_float* hd_dedisp_get_dm_list(hd_dedisp_plan plan);
hd_size hd_dedisp_get_channel_count(hd_dedisp_plan plan);
hd_float hd_dedisp_get_dt(hd_dedisp_plan plan);
hd_float hd_dedisp_get_f0(hd_dedisp_plan plan);
hd_float hd_dedisp_get_df(hd_dedisp_plan plan);
/* dedisp_execute(plan, nsamps, in, in_nbits, out, out_nbits, flags): in is
 * nsamps samples of nchans * in_nbits / 8 bytes (in_nbits 1, 2, 4 or 8,
 * channels packed low bits first); out is one row of nsamps - max_delay
 * floats per trial (out_nbits must be 32: dedisp's 32-bit output is float),
 * each the exact integer sum.  With adaptive time resolution, trial d's row
 * holds its first (nsamps - max_delay) / s_d values and zeros after them.
 * Pointers are host memory unless flags has HD_DEDISP_DEVICE_POINTERS.
 * Returns when the output is written. */
hd_error hd_dedisp_execute(hd_dedisp_plan plan, hd_size nsamps, const hd_byte* in,
                           hd_size in_nbits, hd_byte* out, hd_size out_nbits, unsigned flags);

#ifdef __cplusplus
}
#endif
