// candidate_profiler.cpp -- plots around one candidate of a sigproc file, as
// Applications/candidate_profiler.C does: a DM-time image around the
// candidate's trial (adaptive time resolution, as requested there from dedisp
// at :261-327), its DM-0 series, a frequency-time image at the candidate's
// delay (:136-230), and the S/N against DM and time along the brightest
// pixel (:591-622).  The dedispersion runs on the GPU through the library's
// dedisp plan (include/hd/dedisp_plan.h); the rest is host arithmetic in the
// reference's order and precision.
//
// Outputs in the working directory: freq_time.dat, snr_dm.dat, snr_time.dat
// and dm_time.pgm (skipped beyond 16384 samples), with the reference's
// formats.  Where the reference reads outside its buffers (a DM trial past
// the list, a series row outside the image, a time bin past the image), this
// tool writes 0 or skips the entry instead.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <iomanip>
#include <iostream>
#include <iterator>
#include <string>
#include <vector>

#include "hd/dedisp_plan.h"
#include "hd/error.h"
#include "hd/formats.h"

namespace {

struct Input {
  std::string path;
  hd_sigproc_header h{};
};

// Samples [first, first + n) of the file, each of nchans * nbits / 8 bytes;
// what lies past the end of the file (or past any representable offset)
// reads as zero.
std::vector<uint8_t> read_samples(const Input& in, uint64_t first, uint64_t n) {
  const uint64_t bytes = (uint64_t)in.h.nchans * in.h.nbits / 8;
  std::vector<uint8_t> buf(n * bytes, 0);
  FILE* f = std::fopen(in.path.c_str(), "rb");
  if (!f) return buf;
  const uint64_t off = first * bytes;
  if (bytes == 0 || off / bytes == first) {
    if (fseeko(f, (off_t)(in.h.header_size + off), SEEK_SET) == 0) {
      size_t got = std::fread(buf.data(), 1, buf.size(), f);
      (void)got;
    }
  }
  std::fclose(f);
  return buf;
}

// Per row: subtract the line through the medians of the two halves (each
// the quarter-point order statistic, :96-112), then divide by 1.4826 times
// the median absolute value (:121-128).
void normalise_rows(float* out, size_t n, size_t rows) {
  std::vector<float> buf(n);
  for (size_t r = 0; r < rows; ++r) {
    float* x = out + r * n;
    std::copy(x, x + n, buf.begin());
    std::nth_element(buf.begin(), buf.begin() + n / 4, buf.begin() + n / 2);
    const float m1 = buf[n / 4];
    std::nth_element(buf.begin() + n / 2, buf.begin() + 3 * n / 4, buf.end());
    const float m2 = buf[3 * n / 4];
    const double slope = (m2 - m1) / (n / 2);
    const float m0 = m1 - slope * (n / 4);
    for (size_t j = 0; j < n; ++j) {
      const float line = m0 + j * slope;
      x[j] -= line;
    }
    for (size_t j = 0; j < n; ++j) buf[j] = std::fabs(x[j] - 0.f);
    std::nth_element(buf.begin(), buf.begin() + n / 2, buf.end());
    const float sigma = buf[n / 2] * 1.4826;
    const float scale = 1. / sigma;
    for (size_t j = 0; j < n; ++j) x[j] = x[j] * scale;
  }
}

bool plan_for(const Input& in, hd_dedisp_plan* p) {
  if (hd_dedisp_create_plan(p, (hd_size)in.h.nchans, (float)in.h.tsamp, (float)in.h.fch1,
                            (float)in.h.foff) != HD_NO_ERROR) {
    std::cerr << "dedisp_create_plan failed" << std::endl;
    return false;
  }
  return true;
}

// DM-time image (:232-409): out_dm_count trials centred on dm_idx, out_nsamps
// bins around samp.  Without filtering each bin sums `width` samples (in the
// trial's coarser samples); with it, a boxcar of `width` full-resolution
// samples over the trial's series held back at full resolution.
int dm_time(const Input& in, const std::vector<float>& dms, size_t samp, size_t filter,
            size_t dm_idx, size_t out_nsamps, size_t out_dm_count, float* out, bool do_filter,
            bool verbose) {
  const float pulse_width = 40, tol = 1.15f;
  const size_t dm_count = dms.size();
  if (dm_idx >= dm_count) {
    std::cerr << "Invalid DM index " << dm_idx << std::endl;
    return -1;
  }
  if (verbose) {
    std::cerr << "dt = " << in.h.tsamp << std::endl;
    std::cerr << "f0 = " << in.h.fch1 << std::endl;
    std::cerr << "df = " << in.h.foff << std::endl;
  }
  hd_dedisp_plan plan;
  if (!plan_for(in, &plan)) return -1;
  const size_t lo = dm_idx >= out_dm_count / 2 ? dm_idx - out_dm_count / 2 : 0;
  const size_t hi = std::min(dm_idx < dm_count - ((out_dm_count - 1) / 2 + 1)
                                 ? dm_idx + (out_dm_count - 1) / 2 + 1
                                 : dm_count,
                             dm_count);
  if (verbose) {
    std::cerr << "Target DM      = " << dms[dm_idx] << std::endl;
    std::cerr << "DM trial range = " << lo << " : " << hi << std::endl;
  }
  const size_t ndm = hi - lo;
  if (verbose) std::cerr << "Out DM count = " << ndm << std::endl;
  if (hd_dedisp_set_dm_list(plan, &dms[lo], ndm) != HD_NO_ERROR ||
      hd_dedisp_enable_adaptive_dt(plan, pulse_width, tol) != HD_NO_ERROR) {
    hd_dedisp_destroy_plan(plan);
    return -1;
  }
  const std::vector<hd_size> f(hd_dedisp_get_dt_factors(plan),
                               hd_dedisp_get_dt_factors(plan) + ndm);
  const size_t width = (size_t)1 << filter;
  const size_t max_delay = hd_dedisp_get_max_delay(plan);
  size_t nsamps, first;
  if (!do_filter) {
    nsamps = out_nsamps * width;
    first = samp >= out_nsamps / 2 * width ? samp - out_nsamps / 2 * width : 0;
  } else {
    nsamps = out_nsamps + width - 1;
    first = samp > out_nsamps / 2 + width / 2 ? samp - out_nsamps / 2 - width / 2 : 0;
  }
  const size_t in_nsamps = nsamps + max_delay;
  if (verbose) {
    std::cerr << "Max dispersion delay = " << max_delay << std::endl;
    std::cerr << "Input sample range = " << first << " : " << first + in_nsamps << std::endl;
    std::cerr << "(" << in_nsamps << " samples)" << std::endl;
  }
  const std::vector<uint8_t> data = read_samples(in, first, in_nsamps);
  std::vector<float> dd(nsamps * ndm);
  const hd_error e = hd_dedisp_execute(plan, in_nsamps, data.data(), (hd_size)in.h.nbits,
                                       reinterpret_cast<hd_byte*>(dd.data()), 32,
                                       HD_DEDISP_HOST_POINTERS);
  hd_dedisp_destroy_plan(plan);
  if (e != HD_NO_ERROR) {
    std::cerr << "dedisp_execute failed: " << hd_get_error_string(e) << std::endl;
    return -1;
  }
  if (!do_filter) {
    for (size_t d = 0; d < ndm; ++d) {
      const size_t per = width / f[d];  // coarse samples per bin (0 when width < factor)
      for (size_t t = 0; t < out_nsamps; ++t) {
        float acc = 0.f;
        const float* row = dd.data() + d * out_nsamps * width + t * per;
        for (size_t j = 0; j < per; ++j) acc += row[j];
        out[d * out_nsamps + t] = acc;
      }
    }
  } else {
    std::vector<float> full(nsamps), scan(nsamps + 1);
    for (size_t d = 0; d < ndm; ++d) {
      for (size_t t = 0; t < nsamps; ++t) full[t] = dd[d * nsamps + t / f[d]];
      scan[0] = 0.f;
      for (size_t t = 1; t <= nsamps; ++t) scan[t] = scan[t - 1] + full[t - 1];
      const size_t hi_off = width / 2 + (width - 1) / 2 + 1;
      for (size_t j = 0; j < out_nsamps; ++j)
        out[d * out_nsamps + j] = scan[hi_off + j] - scan[width / 2 - width / 2 + j];
    }
  }
  normalise_rows(out, out_nsamps, ndm);
  return 0;
}

// Frequency-time image (:136-230) at the candidate's trial: bins of
// tscrunch samples x fscrunch channels, widened to a power of two samples
// when the window is shorter than the trial's delay.
int freq_time(const Input& in, size_t samp, size_t filter, float dm, size_t out_nsamps,
              size_t fscrunch, float* out, bool verbose) {
  hd_dedisp_plan plan;
  if (!plan_for(in, &plan)) return -1;
  hd_dedisp_set_dm_list(plan, &dm, 1);
  const size_t max_delay = hd_dedisp_get_max_delay(plan);
  hd_dedisp_destroy_plan(plan);
  size_t delay = 2;
  size_t tscrunch = (size_t)1 << filter;
  size_t in_nsamps = out_nsamps * tscrunch;
  if (in_nsamps < max_delay) {
    while (delay < max_delay) delay *= 2;
    in_nsamps = delay;
    tscrunch = in_nsamps / out_nsamps;
  }
  const size_t centre = samp + max_delay / 2;
  const size_t first = centre - in_nsamps / 2;  // wraps below sample 0, as there
  if (verbose) {
    std::cerr << "max delay= " << max_delay << std::endl;
    std::cerr << "delay= " << delay << std::endl;
    std::cerr << "tscrunch = " << tscrunch << std::endl;
    std::cerr << "Input sample range: " << first << " : " << first + in_nsamps << " = "
              << in_nsamps << " samples" << std::endl;
    std::cerr << "DM = " << dm << std::endl;
  }
  std::cout << "in_nsamps=" << in_nsamps << std::endl;
  const size_t nbits = (size_t)in.h.nbits, nchans = (size_t)in.h.nchans;
  const std::vector<uint8_t> raw = read_samples(in, first, in_nsamps);
  // 32-bit words of nchans * nbits / 32 per sample, channels low bits first
  const size_t per_word = 32 / nbits;
  const size_t words = nchans / per_word;
  std::vector<uint32_t> w(in_nsamps * words, 0);
  std::copy(raw.begin(), raw.begin() + std::min(raw.size(), w.size() * 4),
            reinterpret_cast<uint8_t*>(w.data()));
  const size_t mask = ((size_t)1 << nbits) - 1;
  const float peak = 255;
  for (size_t t = 0; t < in_nsamps; t += tscrunch) {
    for (size_t c = 0; c < nchans; c += fscrunch) {
      float sum = 0.f;
      for (size_t s = 0; s < tscrunch; ++s)
        for (size_t k = 0; k < fscrunch; ++k) {
          const size_t idx = (c + k) / per_word + (t + s) * words;
          const uint32_t word = idx < w.size() ? w[idx] : 0u;
          const uint32_t x = (word >> (((c + k) % per_word) * nbits)) & mask;
          sum += (float)x / mask * peak;
        }
      if (t / tscrunch < out_nsamps && c / fscrunch < nchans / fscrunch)
        out[c / fscrunch * out_nsamps + t / tscrunch] = sum / (tscrunch * fscrunch);
    }
  }
  normalise_rows(out, out_nsamps, nchans / fscrunch);
  return 0;
}

void usage(const char* binary) {
  std::cerr << "Usage: " << binary
            << " filename.fil dmlist samp filter dm_idx out_nsamps out_dm_count fscrunch "
               "[do_filter=0]"
            << std::endl;
  std::cerr << "    sigproc_file        2bit sigproc filterbank file" << std::endl;
  std::cerr << "    dmlist              file containing list of DMs to plot" << std::endl;
  std::cerr << "    samp                sample index of the candidate" << std::endl;
  std::cerr << "    filter              filter (log2) in which candidate was found" << std::endl;
  std::cerr << "    dm_idx              line in dmlist which corresponds to nominal DM for "
               "candidate"
            << std::endl;
  std::cerr << "    out_nsamps          " << std::endl;
  std::cerr << "    out_dm_count        should match the number of lines in dmlist" << std::endl;
  std::cerr << "    fscrunch            fscrunching factor for Freq vs Time data" << std::endl;
  std::cerr << "    do_filter           boxcar-filter the DM-time image instead of binning it"
            << std::endl;
}

}  // namespace

int main(int argc, char* argv[]) {
  if (argc <= 8) {
    std::cerr << "Error: expected 8 arguments found " << argc << std::endl;
    usage(argv[0]);
    return 0;
  }
  const bool verbose = true;
  Input in;
  in.path = argv[1];
  const std::string dm_path = argv[2];
  const size_t samp = (size_t)std::atoi(argv[3]);
  const size_t filter = (size_t)std::atoi(argv[4]);
  const size_t dm_idx = (size_t)std::atoi(argv[5]);
  const size_t out_nsamps = (size_t)std::atoi(argv[6]);
  const size_t out_dm_count = (size_t)std::atoi(argv[7]);
  const size_t fscrunch = (size_t)std::atoi(argv[8]);
  const bool do_filter = argc >= 10 && std::atoi(argv[9]) != 0;
  std::cerr << (do_filter ? "Filtering enabled" : "Filtering disabled") << std::endl;
  std::cerr << "out_nsamps   = " << out_nsamps << std::endl;
  std::cerr << "out_dm_count = " << out_dm_count << std::endl;
  std::cerr << "Reading header info..." << std::endl;
  if (hd_sigproc_read_header(in.path.c_str(), &in.h) != 0) {
    std::cerr << "ERROR: Could not open " << in.path << std::endl;
    return -1;
  }
  std::cerr << "Loading DM list '" << dm_path << "'..." << std::endl;
  std::ifstream dm_file(dm_path.c_str());
  if (!dm_file) {
    std::cerr << "ERROR: Could not open " << dm_path << std::endl;
    return -1;
  }
  const std::vector<float> dms{std::istream_iterator<float>(dm_file),
                               std::istream_iterator<float>()};
  std::cerr << "Read " << dms.size() << " dms from list" << std::endl;

  std::vector<float> dm_img(out_nsamps * out_dm_count, 0.f);
  std::cerr << "Computing DM-time plot..." << std::endl;
  if (dm_time(in, dms, samp, filter, dm_idx, out_nsamps, out_dm_count, dm_img.data(), do_filter,
              verbose)) {
    std::cerr << "gen_dm_time_plot failed" << std::endl;
    return -1;
  }
  std::vector<float> dm0(out_nsamps, 0.f);
  if (dm_time(in, dms, samp, filter, 0, out_nsamps, 1, dm0.data(), do_filter, verbose)) {
    std::cerr << "gen_dm0_time_plot failed" << std::endl;
    return -1;
  }
  const size_t out_nchans = (size_t)in.h.nchans / fscrunch;
  std::vector<float> ft(out_nsamps * out_nchans, 0.f);
  std::cerr << "Computing freq-time plot..." << std::endl;
  freq_time(in, samp, filter > 2 ? filter - 2 : 0, dms[dm_idx], out_nsamps, fscrunch, ft.data(),
            verbose);

  std::cerr << "Writing freq-time data to freq_time.dat..." << std::endl;
  {
    std::ofstream o("freq_time.dat");
    for (size_t c = 0; c < out_nchans; ++c) {
      for (size_t t = 0; t < out_nsamps; ++t) {
        float v = ft[c * out_nsamps + t];
        if (std::isnan(v)) v = 0;
        if (c * fscrunch < 150) v = 0;  // the lowest 150 channels are blanked, as there
        o << v << "\t";
      }
      o << "\n";
    }
  }
  std::cerr << "Computing result bounds..." << std::endl;
  const float lo = *std::min_element(dm_img.begin(), dm_img.end());
  const float hi = *std::max_element(dm_img.begin(), dm_img.end());
  const size_t peak = std::max_element(dm_img.begin(), dm_img.end()) - dm_img.begin();
  std::cerr << "Max value        = " << hi << std::endl;
  std::cerr << "Min value        = " << lo << std::endl;
  std::cerr << "Max data index   = " << peak << std::endl;
  std::cerr << "Max sample index = " << peak % out_nsamps << std::endl;
  // rows of the image are trials from here (the requested count, not clipped)
  const size_t row0 = dm_idx > out_dm_count / 2 ? dm_idx - out_dm_count / 2 : 0;
  std::cerr << "Max DM index     = " << row0 + peak / out_nsamps << std::endl;
  std::cerr << "Extracting SNR vs. DM..." << std::endl;
  {
    std::ofstream o("snr_dm.dat");
    o << "#dm_trial\tDM\tSNR" << std::endl;
    const size_t t = peak % out_nsamps;
    for (size_t d = 0; d < std::min(out_dm_count, dms.size()); ++d) {
      const size_t trial = row0 + d;
      if (trial >= dms.size()) break;
      o << trial << "\t" << std::setprecision(10) << dms[trial] << "\t"
        << dm_img[d * out_nsamps + t] << std::endl;
    }
  }
  std::cerr << "Extracting SNR vs. time..." << std::endl;
  {
    const float width = (float)((size_t)1 << filter);
    const size_t d = peak / out_nsamps;
    auto at = [&](size_t row, size_t t) {
      return row < out_dm_count ? dm_img[row * out_nsamps + t] : 0.f;
    };
    std::ofstream o("snr_time.dat");
    o << "#sample\ttime\tSNR\tSNR@DM-1\tSNR@DM+1\tSNR@DM0" << std::endl;
    for (size_t t = 0; t < out_nsamps; ++t) {
      const float tt = (samp + ((float)t - out_nsamps / 2) * width) * in.h.tsamp;
      o << tt << "\t" << std::setprecision(10) << tt << "\t" << at(d, t) << "\t"
        << at(d - 1, t) << "\t" << at(d + 1, t) << "\t" << dm0[t] << std::endl;
    }
  }
  if (out_nsamps > 16384) {
    std::cerr << "Skipping image output for large data" << std::endl;
    std::cerr << "Done." << std::endl;
    return 0;
  }
  std::cerr << "Writing output to 'dm_time.pgm'..." << std::endl;
  {
    const size_t levels = 256;
    std::ofstream o("dm_time.pgm");
    o << "P2 " << out_nsamps << " " << out_dm_count << " " << levels - 1 << std::endl;
    for (size_t d = 0; d < out_dm_count; ++d) {
      for (size_t t = 0; t < out_nsamps; ++t) {
        const float v = dm_img[d * out_nsamps + t];
        const float lvl = std::min(std::max((v - 0.f) / (8.f - 0.f) * (levels - 1), 0.f),
                                   (float)levels - 1);
        o << (std::isnan(lvl) ? (size_t)0 : (size_t)lvl) << "\t";
      }
      o << "\n";
    }
  }
  std::cerr << "Done" << std::endl;
  return 0;
}
