// heimdall_main.cpp -- the gulp driver: a sigproc filterbank searched gulp by
// gulp with overlap-save, as Applications/heimdall.C:32-308 drives the search
// over its data source.
//
// Each call processes nsamps_gulp new samples plus the overlap kept from the
// previous call (hd_get_max_overlap bounds it); after the call the samples
// the search did not finish (nsamps_to_process - nsamps_processed) are moved
// to the front of the buffer and the next gulp is read behind them
// (heimdall.C:176-292, here for one beam).  A gulp shorter than nsamps_gulp
// ends the loop without being searched, as in the reference.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <iostream>
#include <vector>

#include "hd/default_params.h"
#include "hd/error.h"
#include "hd/formats.h"
#include "hd/mi355x.h"
#include "hd/parse_command_line.h"
#include "hd/pipeline.h"

int main(int argc, char* argv[]) {
  hd_params params;
  hd_set_default_params(&params);
  if (hd_parse_command_line(argc, argv, &params) < 0) return -1;
  if (!params.sigproc_file) {
    std::cerr << "ERROR: only sigproc file input is supported by this driver (-f file)"
              << std::endl;
    return -1;
  }
  hd_sigproc_header hdr;
  hd_sigproc_file* src = hd_sigproc_open(params.sigproc_file, &hdr);
  if (!src) {
    std::cerr << "ERROR: Failed to open file '" << params.sigproc_file << "'" << std::endl;
    return -1;
  }
  if (hdr.nifs > 1) {
    std::cerr << "ERROR: only one IF is supported" << std::endl;
    hd_sigproc_close(src);
    return -1;
  }
  // data source metadata (heimdall.C:96-117 over SigprocFile.C:16-50)
  if (!params.override_beam) params.beam = hdr.ibeam > 0 ? (size_t)(hdr.ibeam - 1) : 0;
  params.nbeams = 1;
  params.f0 = (float)hdr.fch1;
  params.df = (float)hdr.foff;
  params.dt = (float)hdr.tsamp;
  params.nchans = (size_t)hdr.nchans;
  params.utc_start = (long)hd_mjd_to_utc(hdr.tstart);
  params.spectra_per_second = (size_t)(1.0 / (double)(float)hdr.tsamp);
  const size_t nbits = (size_t)hdr.nbits;
  const size_t stride = params.nchans * nbits / 8;
  if (params.verbosity > 0)
    std::cout << "processing beam " << (params.beam + 1) << std::endl;

  hd_pipeline pipeline;
  hd_error error = hd_create_pipeline(&pipeline, params);
  if (error != HD_NO_ERROR) {
    std::cerr << "ERROR: Pipeline creation failed" << std::endl;
    std::cerr << "       " << hd_get_error_string(error) << std::endl;
    hd_sigproc_close(src);
    return -1;
  }
  const size_t nsamps_gulp = params.nsamps_gulp;
  error = hd_preallocate(pipeline, nsamps_gulp, params.nbeams);
  if (error != HD_NO_ERROR) {
    std::cerr << "ERROR: Pipeline pre-allocation failed" << std::endl;
    hd_destroy_pipeline(pipeline);
    hd_sigproc_close(src);
    return -1;
  }
  const size_t max_overlap = hd_get_max_overlap(pipeline, params.nbeams);
  std::vector<hd_byte> buf((nsamps_gulp + max_overlap) * stride);
  if (params.verbosity >= 1)
    std::cout << "Beginning data processing, requesting " << nsamps_gulp << " samples"
              << std::endl;

  size_t total_nsamps = 0;
  size_t overlap = 0;
  size_t nsamps_read = hd_sigproc_read(src, nsamps_gulp, buf.data());
  while (nsamps_read == nsamps_gulp) {
    const size_t nsamps_to_process = nsamps_gulp + overlap;
    hd_size nsamps_processed = 0;
    if (params.verbosity >= 1)
      std::cout << "Executing pipeline on new gulp of " << nsamps_read
                << " samples, nsamps_to_process=" << nsamps_to_process << std::endl;
    error = hd_execute(pipeline, buf.data(), nsamps_to_process, nbits, total_nsamps,
                       params.nbeams, &nsamps_processed);
    if (error == HD_NO_ERROR) {
      if (params.verbosity >= 1)
        std::cout << "Processed " << nsamps_processed << " samples." << std::endl;
    } else if (error == HD_TOO_MANY_EVENTS) {
      if (params.verbosity >= 1)
        std::cerr << "WARNING: hd_execute produces too many events, some data skipped"
                  << std::endl;
    } else {
      std::cerr << "ERROR: Pipeline execution failed" << std::endl;
      std::cerr << "       " << hd_get_error_string(error) << std::endl;
      hd_destroy_pipeline(pipeline);
      hd_sigproc_close(src);
      return -1;
    }
    overlap = nsamps_to_process - nsamps_processed;   // heimdall.C:210
    total_nsamps += nsamps_processed;                 // :212
    if (overlap > max_overlap) {
      std::cerr << "ERROR: overlap " << overlap << " exceeds the maximum " << max_overlap
                << std::endl;
      hd_destroy_pipeline(pipeline);
      hd_sigproc_close(src);
      return -1;
    }
    // keep the unprocessed tail, then read the next gulp behind it
    std::copy(buf.begin() + (ptrdiff_t)(nsamps_processed * stride),
              buf.begin() + (ptrdiff_t)(nsamps_to_process * stride), buf.begin());
    nsamps_read = hd_sigproc_read(src, nsamps_gulp, buf.data() + overlap * stride);
  }
  if (params.verbosity >= 1)
    std::cout << "Successfully processed a total of " << total_nsamps << " samples."
              << std::endl;
  hd_destroy_pipeline(pipeline);
  hd_sigproc_close(src);
  if (params.verbosity >= 1) std::cout << "All done." << std::endl;
  return 0;
}
