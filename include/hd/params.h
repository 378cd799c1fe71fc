/* hd/params.h -- search parameters passed by value to hd_create_pipeline.
 * Same field order and types as the reference (Pipeline/hd/params.h:19-78).
 * `dada_id` exists only when HAVE_PSRDADA is defined, as in the reference;
 * on LP64 both layouts put `sigproc_file` at offset 8, and the library reads
 * the struct through the HAVE_PSRDADA layout, so callers built either way
 * are compatible. */
#pragma once

#include <sys/types.h>
#include <time.h>
#include <stdbool.h>

#include "hd/types.h"

struct hd_params {
  /* Application parameters */
  int      verbosity;
#ifdef HAVE_PSRDADA
  key_t    dada_id;          /* PSRDADA shared-memory key */
#endif
  const char* sigproc_file;  /* sigproc filterbank file name */
  bool     yield_cpu;        /* block instead of spin while waiting on the GPU */
  hd_size  nsamps_gulp;      /* samples per gulp */
  hd_size  dm_gulp_size;     /* unused (kept for layout) */
  /* Normalisation parameters */
  hd_float baseline_length;  /* seconds over which to smooth the baseline */
  /* Observational parameters */
  hd_size  beam;             /* beam index (0-based) */
  hd_size  nbeams;           /* number of beams */
  bool     override_beam;
  hd_size  nchans;           /* frequency channels */
  hd_float dt;               /* sampling time (s) */
  hd_float f0;               /* frequency of channel 0 (MHz) */
  hd_float df;               /* channel spacing (MHz) */
  /* Dedispersion parameters */
  hd_float dm_min;
  hd_float dm_max;
  hd_float dm_tol;           /* smearing tolerance between DM trials */
  hd_float dm_pulse_width;   /* intrinsic pulse width (us) */
  hd_size  dm_nbits;         /* bits per dedispersed sample (32 = float) */
  bool     use_scrunching;   /* adaptive DM-dependent time scrunching */
  hd_float scrunch_tol;
  /* RFI mitigation parameters */
  hd_float rfi_tol;
  hd_size  rfi_min_beams;
  /* Single pulse search parameters */
  hd_size  boxcar_max;       /* widest boxcar */
  hd_float detect_thresh;    /* detection threshold (sigma) */
  hd_size  cand_sep_time;    /* clustering tolerance in time (samples x width) */
  hd_size  cand_sep_filter;  /* clustering tolerance in filters */
  hd_size  cand_sep_dm;      /* clustering tolerance in DM trials */
  hd_size  cand_rfi_dm_cut;  /* unused */
  hd_float max_giant_rate;   /* giants per minute, 0 = unlimited */
  hd_size  min_tscrunch_width; /* boxcar width at which decimation starts */
  /* coincidencer socket mode */
  const char* coincidencer_host;
  int      coincidencer_port;
  /* channel zapping (end exclusive) */
  unsigned int num_channel_zaps;
  hd_range_t*  channel_zaps;
  hd_size  beam_count;
  int      gpu_id;
  time_t   utc_start;        /* UTC of the first sample */
  hd_size  spectra_per_second;
  const char* output_dir;
};
#ifndef __cplusplus
typedef struct hd_params hd_params;
#endif
