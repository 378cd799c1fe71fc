/* hd/pipeline.h -- the per-gulp search entry points.
 * Drop-in for the reference's Pipeline/hd/pipeline.h:18-24 (same names,
 * argument order, meaning and error codes). */
#pragma once

#ifdef __cplusplus
extern "C" {
#endif

#include "hd/types.h"
#include "hd/params.h"
#include "hd/error.h"

/* replaces hd_create_pipeline (Pipeline/pipeline.cu:138-212) */
hd_error hd_create_pipeline(hd_pipeline* pipeline, hd_params params);
/* replaces hd_preallocate (Pipeline/pipeline.cu:221-231) */
hd_error hd_preallocate(hd_pipeline pipeline, hd_size nsamps, hd_size nbeams);
/* replaces hd_get_max_overlap (Pipeline/pipeline.cu:214-219) */
hd_size  hd_get_max_overlap(hd_pipeline pl, hd_size nbeams);
/* replaces hd_execute (Pipeline/pipeline.cu:233-1087) */
hd_error hd_execute(hd_pipeline pipeline,
                    const hd_byte* filterbank, hd_size nsamps, hd_size nbits,
                    hd_size first_idx, hd_size nbeams, hd_size* nsamps_processed);
/* replaces hd_destroy_pipeline (Pipeline/pipeline.cu:1089-1100) */
void     hd_destroy_pipeline(hd_pipeline pipeline);

#ifdef __cplusplus
}
#endif
