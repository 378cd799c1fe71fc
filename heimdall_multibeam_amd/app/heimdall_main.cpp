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
//
// The gulps live in device memory, in two buffers used in turn: while gulp k
// is searched (hd_execute_device, the same search as hd_execute), a second
// thread reads the next nsamps_gulp samples from the file into page-locked
// memory and copies them to the device on a stream of its own.  After the
// call the next gulp is put together on the device from the unfinished tail
// and those samples, so neither the file read nor the host-to-device copy
// waits for the search or delays it.
#include <hip/hip_runtime_api.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <iostream>
#include <thread>
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
  // device buffers of (nsamps_gulp + max_overlap) samples for the gulps, one
  // of nsamps_gulp for the next gulp's new samples, page-locked host memory
  // for the file reads, and the copy stream (default priority)
  const size_t gulp_bytes = (nsamps_gulp + max_overlap) * stride;
  const size_t new_bytes = nsamps_gulp * stride;
  hd_byte* dev[2] = {nullptr, nullptr};
  hd_byte* stage = nullptr;
  hd_byte* pin = nullptr;
  hipStream_t cs = nullptr;
  auto release = [&]() {
    if (cs) (void)hipStreamDestroy(cs);
    for (hd_byte* d : dev)
      if (d) (void)hipFree(d);
    if (stage) (void)hipFree(stage);
    if (pin) (void)hipHostFree(pin);
    hd_destroy_pipeline(pipeline);
    hd_sigproc_close(src);
  };
  if (hipSetDevice((int)params.gpu_id) != hipSuccess ||
      hipMalloc((void**)&dev[0], gulp_bytes) != hipSuccess ||
      hipMalloc((void**)&dev[1], gulp_bytes) != hipSuccess ||
      hipMalloc((void**)&stage, new_bytes) != hipSuccess ||
      hipHostMalloc((void**)&pin, new_bytes, 0) != hipSuccess ||
      hipStreamCreate(&cs) != hipSuccess) {
    std::cerr << "ERROR: device buffers for the gulps could not be allocated" << std::endl;
    release();
    return -1;
  }
  if (params.verbosity >= 1)
    std::cout << "Beginning data processing, requesting " << nsamps_gulp << " samples"
              << std::endl;

  size_t total_nsamps = 0;
  size_t overlap = 0;
  int cur = 0;
  size_t nsamps_read = hd_sigproc_read(src, nsamps_gulp, pin);
  if (nsamps_read == nsamps_gulp &&
      (hipMemcpyAsync(dev[cur], pin, new_bytes, hipMemcpyHostToDevice, cs) != hipSuccess ||
       hipStreamSynchronize(cs) != hipSuccess)) {
    std::cerr << "ERROR: copy of the first gulp to the device failed" << std::endl;
    release();
    return -1;
  }
  while (nsamps_read == nsamps_gulp) {
    const size_t nsamps_to_process = nsamps_gulp + overlap;
    hd_size nsamps_processed = 0;
    if (params.verbosity >= 1)
      std::cout << "Executing pipeline on new gulp of " << nsamps_read
                << " samples, nsamps_to_process=" << nsamps_to_process << std::endl;
    // the next gulp's new samples, read and copied beside the search
    size_t nsamps_next = 0;
    hipError_t copy_err = hipSuccess;
    std::thread reader([&]() {
      copy_err = hipSetDevice((int)params.gpu_id);
      nsamps_next = hd_sigproc_read(src, nsamps_gulp, pin);
      if (copy_err == hipSuccess && nsamps_next == nsamps_gulp)
        copy_err = hipMemcpyAsync(stage, pin, new_bytes, hipMemcpyHostToDevice, cs);
    });
    error = hd_execute_device(pipeline, dev[cur], nsamps_to_process, nbits, total_nsamps,
                              params.nbeams, &nsamps_processed);
    reader.join();
    if (copy_err == hipSuccess) copy_err = hipStreamSynchronize(cs);
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
      release();
      return -1;
    }
    if (copy_err != hipSuccess) {
      std::cerr << "ERROR: copy of the next gulp to the device failed" << std::endl;
      release();
      return -1;
    }
    overlap = nsamps_to_process - nsamps_processed;   // heimdall.C:210
    total_nsamps += nsamps_processed;                 // :212
    if (overlap > max_overlap) {
      std::cerr << "ERROR: overlap " << overlap << " exceeds the maximum " << max_overlap
                << std::endl;
      release();
      return -1;
    }
    nsamps_read = nsamps_next;
    if (nsamps_read != nsamps_gulp) break;  // a short last gulp is not searched
    // the next gulp: the unprocessed tail, then the new samples behind it
    const int nxt = 1 - cur;
    if (hipMemcpyAsync(dev[nxt], dev[cur] + nsamps_processed * stride, overlap * stride,
                       hipMemcpyDeviceToDevice, cs) != hipSuccess ||
        hipMemcpyAsync(dev[nxt] + overlap * stride, stage, new_bytes, hipMemcpyDeviceToDevice,
                       cs) != hipSuccess ||
        hipStreamSynchronize(cs) != hipSuccess) {
      std::cerr << "ERROR: assembling the next gulp on the device failed" << std::endl;
      release();
      return -1;
    }
    cur = nxt;
  }
  if (params.verbosity >= 1)
    std::cout << "Successfully processed a total of " << total_nsamps << " samples."
              << std::endl;
  release();
  if (params.verbosity >= 1) std::cout << "All done." << std::endl;
  return 0;
}
