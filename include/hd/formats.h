/* hd/formats.h -- sigproc filterbank input for the gulp driver.
 *
 * Replaces the reference's SigprocFile data source (Formats/SigprocFile.C:
 * 16-68, Formats/hd/header.h:125-198, Formats/hd/DataSource.h:49-80) with a
 * C interface: read the header once, then read gulps of time samples.
 */
#pragma once

#include <time.h>

#include "hd/types.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct hd_sigproc_header {
  double tstart;      /* MJD */
  double tsamp;       /* seconds */
  double fch1, foff;  /* MHz */
  int    nchans, nbits, nifs;
  int    ibeam, nbeams;
  int    nsamples;    /* from the header, or from the file size when absent */
  int    header_size; /* bytes before the first sample */
  char   source_name[80];
} hd_sigproc_header;

/* Parses the header of a sigproc filterbank file (the keyword set and
 * length-prefixed strings of header.h:125-198; unknown keywords are reported
 * and skipped as there).  Returns 0, or -1 (cannot open), -2 (no
 * HEADER_START), -3 (truncated or malformed). */
int hd_sigproc_read_header(const char* path, hd_sigproc_header* h);

/* The reference's MJD to time_t conversion (DataSource.h:49-80), kept with
 * its rounding (the second is always incremented) and its local-time mktime. */
time_t hd_mjd_to_utc(double mjd);

typedef struct hd_sigproc_file hd_sigproc_file;
/* Opens the file positioned at its first sample; NULL on failure. */
hd_sigproc_file* hd_sigproc_open(const char* path, hd_sigproc_header* h);
/* Reads up to nsamps time samples into data; returns the number read
 * (SigprocFile::get_data, SigprocFile.C:60-68). */
hd_size hd_sigproc_read(hd_sigproc_file* f, hd_size nsamps, hd_byte* data);
void hd_sigproc_close(hd_sigproc_file* f);

#ifdef __cplusplus
}
#endif
