/* hd/error.h -- error codes of the Heimdall C interface.
 * Same values and order as the reference (Pipeline/hd/error.h:19-41). */
#pragma once

#ifdef __cplusplus
extern "C" {
#endif

typedef int hd_error;

enum {
  HD_NO_ERROR = 0,
  HD_MEM_ALLOC_FAILED,
  HD_MEM_COPY_FAILED,
  HD_INVALID_DEVICE_INDEX,
  HD_DEVICE_ALREADY_SET,
  HD_INVALID_PIPELINE,
  HD_INVALID_POINTER,
  HD_INVALID_STRIDE,
  HD_INVALID_NBITS,

  HD_PRIOR_GPU_ERROR,
  HD_INTERNAL_GPU_ERROR,

  HD_TOO_FEW_NSAMPS,
  HD_TOO_MANY_EVENTS,
  HD_UNKNOWN_ERROR
};

/* replaces hd_get_error_string (Pipeline/error.C:15-59) */
const char* hd_get_error_string(hd_error error);
/* replaces throw_error (Pipeline/error.C:61-63) */
hd_error throw_error(hd_error error);

#ifdef __cplusplus
}
#endif
