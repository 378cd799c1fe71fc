/* hd/types.h -- scalar and candidate types of the Heimdall C interface.
 * Same names, widths and layouts as the reference (Pipeline/hd/types.h:12-48). */
#pragma once

#include <stddef.h>

typedef unsigned char         hd_byte;
typedef size_t                hd_size;
typedef float                 hd_float;
typedef struct hd_pipeline_t* hd_pipeline;

/* Fundamental candidate quantities (structure of arrays). */
struct RawCandidates {
  hd_float* peaks;
  hd_size*  inds;
  hd_size*  begins;
  hd_size*  ends;
  hd_size*  filter_inds;
  hd_size*  dm_inds;
  hd_size*  members;
};
struct ConstRawCandidates {
  const hd_float* peaks;
  const hd_size*  inds;
  const hd_size*  begins;
  const hd_size*  ends;
  const hd_size*  filter_inds;
  const hd_size*  dm_inds;
  const hd_size*  members;
};
#ifndef __cplusplus
typedef struct RawCandidates RawCandidates;
typedef struct ConstRawCandidates ConstRawCandidates;
#endif

typedef struct hd_range {
  hd_size start;
  hd_size end;
} hd_range_t;
