// hd_api.cpp -- the Heimdall C interface on MI355X.
//
// Host orchestration of one gulp (Pipeline/pipeline.cu:233-1087 in the
// reference): the filterbank goes to HBM once, every stage runs over all DM
// trials at once on the pipeline's own HIP stream, the DM-time plane and the
// boxcar outputs never leave the device, and only the candidate groups come
// back to the host, where the .cand file is written exactly as the reference
// writes it.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <arpa/inet.h>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <ctime>
#include <fstream>
#include <iomanip>
#include <iostream>
#include <map>
#include <tuple>
#include <netdb.h>
#include <sstream>
#include <string>
#include <sys/socket.h>
#include <unistd.h>
#include <vector>

#define HAVE_PSRDADA 1  // read hd_params through the layout that includes dada_id
#include "hd/default_params.h"
#include "hd/mi355x.h"
#include "hd/parse_command_line.h"
#include "hd/pipeline.h"
#include "hd/proposal.h"
#include "hd/stages.h"
#include "hd_internal.hpp"
#include "hd_plan.hpp"

using hd::SeriesShape;
using hd::WidthTable;
using hd::BlockRmsShape;

namespace {

// ---------------------------------------------------------------------------
// device allocations that only grow
struct DevBuf {
  void* p = nullptr;
  size_t bytes = 0;
  hipError_t ensure(size_t n) {
    if (n <= bytes && p) return hipSuccess;
    if (p) (void)hipFree(p);
    p = nullptr;
    bytes = 0;
    size_t want = n ? n : 16;
    hipError_t e = hipMalloc(&p, want);
    if (e == hipSuccess) bytes = want;
    return e;
  }
  void release() {
    if (p) (void)hipFree(p);
    p = nullptr;
    bytes = 0;
  }
  template <typename T> T* as() const { return static_cast<T*>(p); }
};

size_t round_up(size_t x, size_t a) { return (x + a - 1) / a * a; }

unsigned log2_pow2(unsigned v) {  // get_filter_index (pipeline.cu:126-136)
  unsigned r = 0;
  while ((1u << (r + 1)) <= v && r < 31) ++r;
  return r;
}

enum Stage { ST_UNPACK, ST_DEDISP, ST_BASELINE, ST_RMS, ST_SCAN, ST_BLOCKRMS, ST_GIANTS,
             ST_CLUSTER, ST_TOTAL, ST_CLEAN, ST_PROPOSAL, ST_COUNT };

// One gulp's output.  Candidates [lo, hi) of the result arrays belong to it;
// `beam` is the params.beam of the call that searched it.
struct GulpOut {
  size_t nsamps = 0, nbeams = 1, first_idx = 0, beam = 0;
  size_t processed = 0;
  bool too_many = false;
  hd_error err = HD_NO_ERROR;
  size_t lo = 0, hi = 0;
};


}  // namespace

struct hd_pipeline_t {
  hd_params params;
  std::string output_dir, sigproc_file, coinc_host;
  std::vector<hd_range_t> zaps;
  int device = 0;
  hipStream_t stream = nullptr;
  bool output_enabled = true;
  bool cleaning = false;  // zero-DM RFI cleaning first (the reference's CLEANING build)
  // plan
  std::vector<float> dm_list;
  std::vector<float> kc;
  size_t max_delay = 0;
  // adaptive time resolution (use_scrunching): each trial's scrunch factor,
  // the runs of equal factor (first trial, count, factor), and, when any
  // factor exceeds 1, the delay table with each trial's delays in samples of
  // its own resolution (the slope divided by the factor)
  std::vector<size_t> factors;
  struct ScrunchClass {
    size_t first, count, factor;
  };
  std::vector<ScrunchClass> classes;
  bool adaptive = false;
  DevBuf adelays;
  std::vector<int> adelays_h;
  std::vector<int> delays_h;  // host copy of the delay table (rows of live channels)
  // device buffers
  DevBuf fb, x, live, delays, plane, pts, rms, brms, S, partial, wrms, winv;
  DevBuf skey, sval, counter, ccount, coff, cpos, okey, oval, head, segid, scan_tmp;
  DevBuf g_peak, g_ind, g_begin, g_end, g_filter, g_dm, g_best;
  // the giants of adaptive-resolution runs, gathered on the device run by run
  DevBuf a_peak, a_ind, a_begin, a_end, a_filter, a_dm, a_group;
  DevBuf c_keys, c_lo, c_hi, c_labels, c_flag, c_best, c_members, c_rflag, c_rpos, c_tmp;
  DevBuf c_sind, c_sfd;
  DevBuf c_ecount, c_eoff, c_adj;  // coincidence lists
  DevBuf c_gb;                     // per clustering group: first giant, first list entry
  DevBuf o_peak, o_ind, o_begin, o_end, o_filter, o_dm, o_members;
  DevBuf clean_out, clean_work;
  DevBuf gdev, g_group, o_group;        // group table, group of each giant / candidate
  std::vector<hd::GroupDev> hgd;        // host copy of the group table
  std::vector<GulpOut> gulps;           // per-gulp status and candidate ranges
  std::vector<uint32_t> r_group;        // active-group index of each candidate
  std::vector<hd::GroupDev> hrg;        // row groups of the search (gulps, or gated units)
  DevBuf rdev;
  // learned proposal: the network, the coarse plan and plane, the last gating
  hd_cnn net = nullptr;
  bool gating = false;
  float prop_thresh = 0.f;
  bool coarse_only = false;             // run_gulps stops after the coarse plane
  std::vector<float> co_dms, co_kc;
  size_t co_max_delay = 0, co_wlen = 0;
  std::vector<int> co_live_cached;
  DevBuf co_delays, co_pairs, co_xc, co_plane, co_pts, co_rms, co_brms, co_bf, co_bw, co_gdev;
  DevBuf co_tiles, co_scores, co_toff, co_ntpos, co_ncols, co_flags, co_pos;
  std::vector<hd::GroupDev> co_hgd;
  int co_G = 0, co_rows = 0, co_units = 0;
  size_t co_bpitch = 0, co_wpitch = 0;
  std::vector<uint32_t> co_toff_h, co_ntpos_h, co_ncols_h;
  std::vector<uint8_t> prop_flags;      // per gulp x unit of the last gated call
  int prop_units = 0;
  unsigned* h_pinned = nullptr;  // small pinned staging word(s)
  size_t sample_cap = 1u << 22;
  // delay table state for the current live-channel set
  std::vector<int> live_list;
  std::vector<int> killmask_cached;
  size_t wlen = 0;  // staged dwords per copy of the widest DM block (any tile height)
  // dedispersion tile heights the autotuner chooses from, with their staging
  // windows; the choice per launch shape (rows, groups, live channels, pitch)
  // staging plans per tile shape, index log2(db / 8) + 3 * log2(cc / 4)
  DevBuf pairs_db[9];
  int wlen_db[9] = {};
  int wlen8_db[9] = {};  // 8-bit-lane kernel (sub-byte samples)
  std::map<std::tuple<int, int, int, int>, std::array<int, 5>> tune;
  hipEvent_t tev[2] = {};
  int last_db = 16, last_remap = 1, last_lanes8 = 0, last_cc = 4, last_wide = 0;
  // double-buffered dedispersion staging where the windows allow
  // (HD_DEDISP_DBUF=0 turns it off)
  bool use_dbuf = true;
  // the stage chain as a HIP graph (hd_set_graph)
  bool use_graph = false;
  hipGraphExec_t gexec = nullptr;
  std::string gkey;
  size_t graph_captures = 0, graph_launches = 0;
  // hd_set_dedisp_tile (fixed_db 0: autotune)
  int fixed_db = 0, fixed_remap = 1, fixed_lanes8 = 0, fixed_cc = 4, fixed_wide = 0;
  // results of the last call
  size_t n_giants = 0;
  std::vector<float> r_peak, r_dms;
  std::vector<hd_size> r_ind, r_begin, r_end, r_filter, r_dm, r_members;
  std::string cand_text;
  double timings[ST_COUNT] = {0};
  hipEvent_t ev[ST_COUNT + 1] = {};
};

// ---------------------------------------------------------------------------
// params, errors, command line (default_params.C, error.C, parse_command_line.C)

extern "C" void hd_set_default_params(hd_params* p) {
  std::memset(p, 0, sizeof(*p));
  p->verbosity = 0;
  p->dada_id = 0;
  p->sigproc_file = nullptr;
  p->yield_cpu = false;
  p->nsamps_gulp = 262144;
  p->dm_gulp_size = 2048;
  p->baseline_length = 2.0f;
  p->beam = 0;
  p->nbeams = 1;
  p->override_beam = false;
  p->nchans = 1024;
  p->dt = 64e-6f;
  p->f0 = 1581.804688f;
  p->df = -.390625f;
  p->dm_min = 0.0f;
  p->dm_max = 1000.0f;
  p->dm_tol = 1.25f;
  p->dm_pulse_width = 40;
  p->dm_nbits = 32;
  p->use_scrunching = true;
  p->scrunch_tol = 1.15f;
  p->rfi_tol = 5.0f;
  p->rfi_min_beams = 8;
  p->boxcar_max = 4096;
  p->detect_thresh = 6.0f;
  p->cand_sep_time = 3;
  p->cand_sep_filter = 3;
  p->cand_sep_dm = 200;
  p->cand_rfi_dm_cut = (hd_size)1.5;
  p->max_giant_rate = 0;
  p->min_tscrunch_width = 4096;
  p->num_channel_zaps = 0;
  p->channel_zaps = nullptr;
  p->coincidencer_host = nullptr;
  p->coincidencer_port = -1;
  p->beam_count = 13;
  p->gpu_id = 0;
  p->utc_start = 0;
  p->spectra_per_second = 0;
  p->output_dir = ".";
}

extern "C" const char* hd_get_error_string(hd_error error) {
  switch (error) {
    case HD_NO_ERROR: return "No error";
    case HD_MEM_ALLOC_FAILED: return "Memory allocation failed";
    case HD_MEM_COPY_FAILED: return "Memory copy failed";
    case HD_INVALID_DEVICE_INDEX: return "Invalid device index";
    case HD_DEVICE_ALREADY_SET: return "Device is already set and cannot be changed";
    case HD_INVALID_PIPELINE: return "Invalid pipeline";
    case HD_INVALID_POINTER: return "Invalid pointer";
    case HD_INVALID_STRIDE: return "Invalid stride";
    case HD_TOO_FEW_NSAMPS: return "No. samples < maximum delay";
    case HD_PRIOR_GPU_ERROR: return "Prior GPU error.";
    case HD_INTERNAL_GPU_ERROR: return "Internal GPU error. Please contact the author(s).";
    case HD_TOO_MANY_EVENTS: return "Too many events";
    case HD_UNKNOWN_ERROR: return "Unknown error. Please contact the author(s).";
    default: return "Invalid error code";
  }
}

extern "C" hd_error throw_error(hd_error error) { return error; }

extern "C" void hd_print_usage(void) {
  hd_params p;
  hd_set_default_params(&p);
  std::cout << "Usage: heimdall [options]\n"
            << "    -k  key                  use PSRDADA hexidecimal key\n"
            << "    -f  filename             process specified SIGPROC filterbank file\n"
            << "    -vVgG                    increase verbosity level\n"
            << "    -yield_cpu               yield CPU during GPU operations\n"
            << "    -gpu_id ID               run on specified GPU\n"
            << "    -nsamps_gulp num         number of samples to be read at a time ["
            << p.nsamps_gulp << "]\n"
            << "    -baseline_length num     number of seconds over which to smooth the baseline ["
            << p.baseline_length << "]\n"
            << "    -beam ##                 over-ride beam number\n"
            << "    -nbeams ##               give number of beams [def 1]\n"
            << "    -output_dir path         create all output files in specified path\n"
            << "    -dm min max              min and max DM\n"
            << "    -dm_tol num              SNR loss tolerance between each DM trial ["
            << p.dm_tol << "]\n"
            << "    -coincidencer host:port  connect to the coincidencer on the specified host and port\n"
            << "    -zap_chans start end     zap all channels between start and end channels (end exclusive)\n"
            << "    -max_giant_rate nevents  limit the maximum number of individual detections per minute to nevents\n"
            << "    -dm_pulse_width num      expected intrinsic width of the pulse signal in microseconds ["
            << p.dm_pulse_width << "]\n"
            << "    -dm_nbits num            number of bits per sample in dedispersed time series ["
            << p.dm_nbits << "]\n"
            << "    -scrunching num          1 = use adaptive dm-dependent time scrunching, 0 = off\n"
            << "    -scrunching_tol num      smearing tolerance factor for time scrunching\n"
            << "    -rfi_tol num             RFI exicision threshold limits [" << p.rfi_tol << "]\n"
            << "    -rfi_min_beams num       minimum beams for coincident RFI [" << p.rfi_min_beams
            << "]\n"
            << "    -boxcar_max num          maximum boxcar width in samples [" << p.boxcar_max
            << "]\n"
            << "    -min_tscrunch_width num  vary between high quality (large value) and high performance (low value)\n";
}

extern "C" int hd_parse_command_line(int argc, char* argv[], hd_params* params) {
  int i = 0;
  auto need = [&](int k) { return i + k < argc; };
  while (++i < argc) {
    const std::string a = argv[i];
    if (a == "-h") { hd_print_usage(); return -1; }
    if (a == "-v") params->verbosity = std::max(params->verbosity, 1);
    else if (a == "-V") params->verbosity = std::max(params->verbosity, 2);
    else if (a == "-g") params->verbosity = std::max(params->verbosity, 3);
    else if (a == "-G") params->verbosity = std::max(params->verbosity, 4);
    else if (a == "-k" && need(1)) { unsigned k = 0; std::sscanf(argv[++i], "%x", &k); params->dada_id = (key_t)k; }
    else if (a == "-f" && need(1)) params->sigproc_file = strdup(argv[++i]);
    else if (a == "-yield_cpu" && need(1)) params->yield_cpu = std::atoi(argv[++i]) != 0;
    else if (a == "-nsamps_gulp" && need(1)) params->nsamps_gulp = std::atoi(argv[++i]);
    else if (a == "-baseline_length" && need(1)) params->baseline_length = std::atof(argv[++i]);
    else if (a == "-dm" && need(2)) { params->dm_min = std::atof(argv[++i]); params->dm_max = std::atof(argv[++i]); }
    else if (a == "-dm_tol" && need(1)) params->dm_tol = std::atof(argv[++i]);
    else if (a == "-dm_pulse_width" && need(1)) params->dm_pulse_width = std::atof(argv[++i]);
    else if (a == "-dm_nbits" && need(1)) params->dm_nbits = std::atoi(argv[++i]);
    else if (a == "-gpu_id" && need(1)) params->gpu_id = std::atoi(argv[++i]);
    else if (a == "-scrunching" && need(1)) params->use_scrunching = std::atoi(argv[++i]) != 0;
    else if (a == "-scrunch_tol" && need(1)) params->scrunch_tol = std::atof(argv[++i]);
    else if (a == "-rfi_tol" && need(1)) params->rfi_tol = std::atof(argv[++i]);
    else if (a == "-rfi_min_beams" && need(1)) params->rfi_min_beams = std::atoi(argv[++i]);
    else if (a == "-boxcar_max" && need(1)) params->boxcar_max = std::atoi(argv[++i]);
    else if (a == "-detect_thresh" && need(1)) params->detect_thresh = std::atof(argv[++i]);
    else if (a == "-beam" && need(1)) { params->beam = std::atoi(argv[++i]) - 1; params->override_beam = true; }
    else if (a == "-nbeams" && need(1)) params->nbeams = std::atoi(argv[++i]);
    else if (a == "-beam_count" && need(1)) params->beam_count = std::atoi(argv[++i]);
    else if (a == "-cand_sep_time" && need(1)) params->cand_sep_time = std::atoi(argv[++i]);
    else if (a == "-cand_sep_filter" && need(1)) params->cand_sep_filter = std::atoi(argv[++i]);
    else if (a == "-cand_sep_dm_trial" && need(1)) params->cand_sep_dm = std::atoi(argv[++i]);
    else if (a == "-cand_rfi_dm_cut" && need(1)) params->cand_rfi_dm_cut = (hd_size)std::atof(argv[++i]);
    else if (a == "-max_giant_rate" && need(1)) params->max_giant_rate = std::atof(argv[++i]);
    else if (a == "-output_dir" && need(1)) params->output_dir = strdup(argv[++i]);
    else if (a == "-min_tscrunch_width" && need(1)) params->min_tscrunch_width = std::atoi(argv[++i]);
    else if (a == "-coincidencer" && need(1)) {
      std::istringstream iss(argv[++i]);
      std::string host, port;
      std::getline(iss, host, ':');
      std::getline(iss, port);
      params->coincidencer_host = strdup(host.c_str());
      params->coincidencer_port = std::atoi(port.c_str());
    } else if (a == "-zap_chans" && need(2)) {
      unsigned iz = params->num_channel_zaps++;
      params->channel_zaps = (hd_range_t*)realloc(params->channel_zaps,
                                                  sizeof(hd_range_t) * params->num_channel_zaps);
      params->channel_zaps[iz].start = std::atoi(argv[++i]);
      params->channel_zaps[iz].end = std::atoi(argv[++i]);
    } else {
      std::cerr << "WARNING: Unknown parameter '" << a << "'" << std::endl;
    }
  }
  if (params->sigproc_file == nullptr && params->dada_id == 0) {
    std::cerr << "ERROR: no input mechanism specified" << std::endl;
    hd_print_usage();
    return -1;
  }
  return 0;
}

extern "C" hd_error hd_apply_manual_killmasks(int* h_killmask, hd_size nchans,
                                              unsigned int num_channel_zaps,
                                              const hd_range_t* channel_zaps) {
  if (num_channel_zaps && !channel_zaps) return HD_INVALID_POINTER;
  for (unsigned i = 0; i < num_channel_zaps; ++i)
    for (hd_size j = channel_zaps[i].start; j < channel_zaps[i].end; ++j)
      if (j < nchans) h_killmask[j] = 0;
  return HD_NO_ERROR;
}

extern "C" hd_size hd_generate_dm_list(hd_float* out, hd_size cap, hd_float dm_min,
                                       hd_float dm_max, hd_float dt, hd_float pulse_width_us,
                                       hd_float f0, hd_float df, hd_size nchans, hd_float tol) {
  const std::vector<float> l = hdplan::dm_list(dm_min, dm_max, dt, pulse_width_us, f0, df, nchans, tol);
  for (size_t i = 0; i < l.size() && i < cap && out; ++i) out[i] = l[i];
  return l.size();
}

// ---------------------------------------------------------------------------
namespace hd {

// Baseline and per-beam RMS geometry of a series whose nbeams, stride, nout
// and bc are set (pipeline.cu:572-573, remove_baseline.cu:36-46,
// get_rms.cu:66-100).
int series_shape(float cur_dt, float baseline_length, SeriesShape& sh) {
  const size_t smooth = (size_t)(baseline_length / (2 * cur_dt));
  const float oversample = 2;
  sh.sc = (size_t)((double)(oversample * (float)sh.stride / (float)(2 * smooth)) + 0.5);
  if (sh.sc) {
    sh.nscr = (int)(size_t)(std::log((double)(sh.stride / sh.sc)) / std::log(5.));
    sh.count_round = (size_t)(std::pow(5., (double)sh.nscr) * (double)sh.sc);
    if (sh.count_round < 2) return HD_INVALID_STRIDE;
    int Ld = 0;
    while (Ld < std::min(sh.nscr, 3) && sh.count_round / pow5_host(Ld) > 12288) ++Ld;
    if (sh.count_round / pow5_host(Ld) > 12288) return HD_INVALID_STRIDE;
    // at least the first level straight from the leaves (the same medians)
    // when every thread of a 256-thread block still keeps a value: five loads
    // in flight per value for short series
    if (Ld == 0 && sh.nscr >= 1 && sh.count_round / 5 >= 256) Ld = 1;
    sh.bl_direct = Ld;
    sh.step_rs_beam = (float)(sh.stride - 1) / (float)sh.count_round;
    sh.step_rs_last = (float)(sh.bc - 1) / (float)(sh.count_round - 1);
    sh.step_bl_beam = (float)(sh.sc - 1) / (float)sh.stride;
    sh.step_bl_last = (float)(sh.sc - 1) / (float)(sh.bc - 1);
    // The last beam's stretch back could read one point past the pyramid
    // (median_filter.cu:462 with j+1 == sample_count), i.e. a stale scrunch
    // buffer entry.  That case is refused rather than guessed.
    for (size_t i = sh.bc; i-- > 0;) {
      const float x = (float)(unsigned)i * sh.step_bl_last;
      const unsigned j = (unsigned)x;
      const float fr = x - (float)j;
      if ((size_t)j + 1 < sh.sc) break;
      if (fr > 1e-5f) return HD_INVALID_STRIDE;
    }
  }
  int Ld = 0;
  while (sh.bc / pow5_host(Ld) > 12288 && sh.bc >= pow5_host(Ld + 1) && Ld < 3) ++Ld;
  if (sh.bc / pow5_host(Ld) > 12288) return HD_INVALID_STRIDE;
  sh.rms_direct = Ld;
  sh.rms_delta = (float)(sh.stride - sh.bc);
  return HD_NO_ERROR;
}

}  // namespace hd

namespace {

hd_error gpu_err(hipError_t e) {
  if (e == hipSuccess) return HD_NO_ERROR;
  if (e == hipErrorOutOfMemory || e == hipErrorMemoryAllocation) return HD_MEM_ALLOC_FAILED;
  if (e == hipErrorInvalidDevice) return HD_INVALID_DEVICE_INDEX;
  std::cerr << "HIP error: " << hipGetErrorString(e) << std::endl;
  return HD_INTERNAL_GPU_ERROR;
}

#define HD_CHECK(expr)                                  \
  do {                                                  \
    hipError_t _e = (expr);                             \
    if (_e != hipSuccess) return gpu_err(_e);           \
  } while (0)

struct DeviceGuard {  // binds the pipeline's device, restores the caller's
  int prev = -1;
  explicit DeviceGuard(int dev) {
    if (hipGetDevice(&prev) != hipSuccess) prev = -1;
    if (prev != dev) (void)hipSetDevice(dev);
  }
  ~DeviceGuard() {
    int cur = -1;
    if (prev >= 0 && hipGetDevice(&cur) == hipSuccess && cur != prev) (void)hipSetDevice(prev);
  }
};

// LDS window per DM block: widest (last - first) delay span over channels,
// the first trial's delay rounded down to a dword (and per chunk of channels
// staged together).  d: n_rows x nl delays; returns the pairs and the widest
// window.
std::vector<int> staging_pairs(const std::vector<int>& d, size_t n_rows, size_t nl, int DB,
                               int CC, int* hw_out, int* spread_out = nullptr) {
  const size_t nblk = (n_rows + DB - 1) / DB;
  const size_t nchunk = (nl + CC - 1) / CC;
  std::vector<int> pairs(std::max<size_t>(nblk * nchunk, 1));
  int hw_max = 2, spread_max = 0;
  for (size_t b = 0; b < nblk; ++b) {
    const size_t m0 = b * DB;
    const size_t ml = std::min(m0 + DB, n_rows) - 1;
    for (size_t k = 0; k < nchunk; ++k) {
      int spread = 0;
      for (size_t l = k * CC; l < std::min((k + 1) * CC, nl); ++l)
        spread = std::max(spread, d[ml * nl + l] - (d[m0 * nl + l] & ~3));
      const int hw = hd::dedisp_window_dwords(spread);
      pairs[b * nchunk + k] = hw / 2;
      hw_max = std::max(hw_max, hw);
      spread_max = std::max(spread_max, spread);
    }
  }
  *hw_out = hw_max;
  if (spread_out) *spread_out = spread_max;
  return pairs;
}

// Delay table of the live channels.  Its rows run past the last trial to a
// whole number of gated units (repeating the last trial), so that a gated
// group's slice of HD_PROPOSAL_UNIT trials always has rows to read.
hd_error build_live_and_delays(hd_pipeline pl) {
  const hd_params& p = pl->params;
  std::vector<int> km(p.nchans, 1);
  hd_apply_manual_killmasks(km.data(), p.nchans, (unsigned)pl->zaps.size(), pl->zaps.data());
  if (km == pl->killmask_cached && !pl->live_list.empty()) return HD_NO_ERROR;
  pl->killmask_cached = km;
  pl->live_list.clear();
  for (size_t c = 0; c < p.nchans; ++c)
    if (km[c]) pl->live_list.push_back((int)c);
  const size_t n_dm = pl->dm_list.size();
  const size_t n_rows = round_up(n_dm, HD_PROPOSAL_UNIT);
  const size_t nl = pl->live_list.size();
  std::vector<int> d(n_rows * std::max<size_t>(nl, 1));
  for (size_t m = 0; m < n_rows; ++m)
    for (size_t l = 0; l < nl; ++l)
      d[m * nl + l] = hdplan::delay(pl->dm_list[std::min(m, n_dm - 1)], pl->kc[pl->live_list[l]]);
  HD_CHECK(pl->live.ensure(std::max<size_t>(nl, 1) * sizeof(int)));
  HD_CHECK(pl->delays.ensure(d.size() * sizeof(int)));
  HD_CHECK(hipMemcpy(pl->live.p, pl->live_list.data(), nl * sizeof(int), hipMemcpyHostToDevice));
  HD_CHECK(hipMemcpy(pl->delays.p, d.data(), d.size() * sizeof(int), hipMemcpyHostToDevice));
  pl->delays_h = d;
  if (pl->adaptive) {
    // each trial's delays in samples of its own resolution (the slope over a
    // power of two is exact in float)
    pl->adelays_h.assign(n_dm * std::max<size_t>(nl, 1), 0);
    for (size_t m = 0; m < n_dm; ++m) {
      const float inv = 1.f / (float)pl->factors[m];
      for (size_t l = 0; l < nl; ++l)
        pl->adelays_h[m * nl + l] = hdplan::delay(pl->dm_list[m], pl->kc[pl->live_list[l]] * inv);
    }
    HD_CHECK(pl->adelays.ensure(pl->adelays_h.size() * sizeof(int)));
    HD_CHECK(hipMemcpy(pl->adelays.p, pl->adelays_h.data(), pl->adelays_h.size() * sizeof(int),
                       hipMemcpyHostToDevice));
  }
  pl->wlen = 0;
  for (int k = 0; k < 9; ++k) {
    int hw = 2, spread = 0;
    const std::vector<int> pairs = staging_pairs(d, n_rows, nl, 8 << (k % 3), 4 << (k / 3), &hw,
                                                 &spread);
    pl->wlen_db[k] = hw;
    pl->wlen8_db[k] = hd::dedisp8_window_dwords(spread);
    pl->wlen = std::max(pl->wlen, (size_t)hw);  // sizes the input rows' padding
    HD_CHECK(pl->pairs_db[k].ensure(pairs.size() * sizeof(int)));
    HD_CHECK(hipMemcpy(pl->pairs_db[k].p, pairs.data(), pairs.size() * sizeof(int),
                       hipMemcpyHostToDevice));
  }
  pl->tune.clear();
  return HD_NO_ERROR;
}

// Coarse plan of the proposal: trial j of the coarse plane is trial
// min(8j + 4, n_dm - 1) of the plan, dedispersed at 16x the sampling time
// (delays from the same rule with dt*16) over the live channels.
hd_error build_coarse(hd_pipeline pl) {
  if (!pl->co_dms.empty() && pl->co_live_cached == pl->live_list) return HD_NO_ERROR;
  const hd_params& p = pl->params;
  const size_t n_dm = pl->dm_list.size();
  const size_t nc = (n_dm + HD_PROPOSAL_DM - 1) / HD_PROPOSAL_DM;
  pl->co_dms.resize(nc);
  for (size_t j = 0; j < nc; ++j)
    pl->co_dms[j] = pl->dm_list[std::min(HD_PROPOSAL_DM * j + HD_PROPOSAL_DM / 2, n_dm - 1)];
  pl->co_kc = hdplan::delay_slopes(p.nchans, p.dt * (float)HD_PROPOSAL_TIME, p.f0, p.df);
  pl->co_max_delay = 0;
  for (size_t c = 0; c < p.nchans; ++c) {
    const int d = hdplan::delay(pl->co_dms.back(), pl->co_kc[c]);
    if (d > 0 && (size_t)d > pl->co_max_delay) pl->co_max_delay = (size_t)d;
  }
  const size_t nl = pl->live_list.size();
  std::vector<int> d(nc * std::max<size_t>(nl, 1));
  for (size_t m = 0; m < nc; ++m)
    for (size_t l = 0; l < nl; ++l)
      d[m * nl + l] = hdplan::delay(pl->co_dms[m], pl->co_kc[pl->live_list[l]]);
  int hw_max = 2;
  const std::vector<int> pairs =
      staging_pairs(d, nc, nl, hd::dedisp_tile_dms(), hd::dedisp_chunk_channels(), &hw_max);
  pl->co_wlen = (size_t)hw_max;
  HD_CHECK(pl->co_delays.ensure(d.size() * sizeof(int)));
  HD_CHECK(pl->co_pairs.ensure(pairs.size() * sizeof(int)));
  HD_CHECK(hipMemcpy(pl->co_delays.p, d.data(), d.size() * sizeof(int), hipMemcpyHostToDevice));
  HD_CHECK(hipMemcpy(pl->co_pairs.p, pairs.data(), pairs.size() * sizeof(int),
                     hipMemcpyHostToDevice));
  pl->co_live_cached = pl->live_list;
  return HD_NO_ERROR;
}

// Shapes of one call (pipeline.cu:488-499, 572-573, 626-692; remove_baseline.cu:30-46)
// at scrunch factor f (the trial's series holds nsamps_computed / f samples).
struct CallShape {
  SeriesShape sh;
  WidthTable wt;
  BlockRmsShape bs;
  size_t bc692;
  bool skip;  // DM loop breaks before the first trial (no giants)
};

hd_error make_shapes(hd_pipeline pl, size_t nsamps, size_t nbeams, CallShape& cs, size_t f = 1) {
  const hd_params& p = pl->params;
  std::memset(&cs, 0, sizeof(cs));
  const size_t delay = pl->max_delay / f;  // cur_dm_delay (:493-497)
  const size_t W = p.boxcar_max / f;       // rel_boxcar_max (:626)
  SeriesShape& sh = cs.sh;
  sh.nbeams = (unsigned)nbeams;
  sh.stride = nsamps / nbeams / f;          // cur_beam_stride (:498)
  sh.nout = (nsamps - pl->max_delay) / f;   // cur_nsamps (:490)
  cs.skip = (delay >= sh.stride) || (delay + W > sh.stride);  // :501-506, 633-638
  if (cs.skip) return HD_NO_ERROR;
  sh.bc = sh.stride - delay;
  const int TB = hd::dedisp_tile_time();
  sh.ppitch = round_up(sh.nout + 1, (size_t)TB);
  sh.spitch = round_up(sh.nout + 1, 64);
  {
    const int e = hd::series_shape(p.dt * (float)f, p.baseline_length, sh);  // cur_dt (:491)
    if (e != HD_NO_ERROR) return (hd_error)e;
  }
  // widths (pipeline.cu:657-692): filter_width = f, 2f, .. <= boxcar_max, that
  // is relative widths w = 1, 2, .. <= boxcar_max / f
  WidthTable& wt = cs.wt;
  wt.off = (unsigned)(W / 2);
  const size_t max_nsamps_filtered = sh.nout + 1 - W;
  const size_t rel_min_ts = std::max<size_t>(p.min_tscrunch_width / f, 1);
  int nw = 0;
  for (size_t w = 1; w <= W; w *= 2) {
    if (nw >= hd::kMaxWidths) return HD_INVALID_STRIDE;
    hd::WidthInfo& wi = wt.wi[nw++];
    const size_t ts = std::max<size_t>(2 * w / rel_min_ts, 1);
    wi.w = (unsigned)w;
    wi.ts = (unsigned)ts;
    wi.ahead = (unsigned)((w - 1) / 2 + 1);
    wi.behind = (unsigned)(w / 2);
    wi.n_w = (unsigned)((max_nsamps_filtered - 1) / ts + 1);
    wi.filter_idx = log2_pow2((unsigned)(w * f));
    wi.merge_dist = (unsigned)(p.cand_sep_time * (w / ts));
    wi.rel_off = (unsigned)((W / 2) / ts);
    wi.cs = (unsigned)(ts * f);
  }
  wt.nw = nw;
  // The first width must produce fresh output at every block-RMS position;
  // otherwise its RMS would read the previous trial's buffer contents.
  if (nw == 0 || wt.wi[0].ts != 1) return HD_INVALID_STRIDE;
  cs.bc692 = sh.stride - (delay + W);
  BlockRmsShape& bs = cs.bs;
  bs.bc = cs.bc692;
  bs.n = nbeams * cs.bc692;
  if (bs.n == 0) return HD_TOO_FEW_NSAMPS;
  bs.delta = (float)(sh.stride - cs.bc692);
  bs.L0 = 0;
  while (bs.L0 < 5 && hd::pow5_host(bs.L0 + 1) <= bs.n) ++bs.L0;
  bs.chunk = hd::pow5_host(bs.L0);
  bs.np = bs.n / bs.chunk;
  return HD_NO_ERROR;
}

void record(hd_pipeline pl, int st) { (void)hipEventRecord(pl->ev[st], pl->stream); }

// ostream float formatting as the reference's candidate writer uses it.
void write_cand_lines(hd_pipeline pl, std::ostream& os, const GulpOut& go) {
  const hd_params& p = pl->params;
  const size_t first_idx = go.first_idx;
  const unsigned nsamps_beam = (unsigned)(go.nsamps / go.nbeams);      // pipeline.cu:903
  const unsigned nsamps_valid = (unsigned)(go.processed / go.nbeams);   // :905
  const size_t pbeam = go.beam;
  for (size_t i = go.lo; i < go.hi; ++i) {
    const hd_size ind = pl->r_ind[i];
    const hd_size beam_idx = (hd_size)std::floor((float)ind / (float)nsamps_beam);
    const hd_size beam_start = beam_idx * nsamps_beam + (p.boxcar_max / 2);
    const hd_size beam_end = beam_start + nsamps_valid;
    if (!(beam_start <= ind && ind < beam_end)) continue;
    const hd_size samp_offset = ind - (nsamps_beam * beam_idx);
    const hd_size samp_idx = first_idx + samp_offset;
    const hd_float snr = std::isinf(pl->r_peak[i]) ? 0.0f : pl->r_peak[i];
    os << snr << "\t" << samp_idx << "\t" << p.dt * samp_idx << "\t" << pl->r_filter[i] << "\t"
       << pl->r_dm[i] << "\t" << pl->r_dms[i] << "\t" << pl->r_members[i] << "\t"
       << first_idx + (pl->r_begin[i] % nsamps_beam) << "\t"
       << first_idx + (pl->r_end[i] % nsamps_beam) << "\t" << (pbeam + beam_idx + 1) << "\t"
       << "\n";
  }
}

hd_error emit_output(hd_pipeline pl, const GulpOut& go) {
  const hd_params& p = pl->params;
  const size_t nsamps = go.nsamps, nbeams = go.nbeams, first_idx = go.first_idx;
  const size_t processed = go.processed;
  const size_t pbeam = go.beam;
  std::ostringstream lines;
  write_cand_lines(pl, lines, go);
  const std::string text = lines.str();
  pl->cand_text += text;
  if (!pl->output_enabled) return HD_NO_ERROR;
  // Output name (pipeline.cu:893-901).  The reference divides by
  // spectra_per_second unconditionally; a zero rate is taken as offset 0.
  char buffer[64];
  const time_t off = p.spectra_per_second ? (time_t)(first_idx / p.spectra_per_second) : 0;
  time_t now = p.utc_start + off;
  std::strftime(buffer, 64, "%Y-%m-%d-%H:%M:%S", std::gmtime(&now));
  std::ostringstream ss;
  ss << std::setw(2) << std::setfill('0') << (pbeam + 1);
  if (!pl->coinc_host.empty() && p.coincidencer_port != -1) {
    // TCP branch (pipeline.cu:907-995): header line then one line per event
    const unsigned nsamps_beam = (unsigned)(nsamps / nbeams);
    const unsigned nsamps_valid = (unsigned)(processed / nbeams);
    unsigned n_events = 0;
    for (size_t i = go.lo; i < go.hi; ++i) {
      const hd_size ind = pl->r_ind[i];
      const hd_size beam_idx = (hd_size)std::floor((float)ind / (float)nsamps_beam);
      const hd_size beam_start = beam_idx * nsamps_beam + (p.boxcar_max / 2);
      if (beam_start <= ind && ind < beam_start + nsamps_valid) ++n_events;
    }
    std::ostringstream msg;
    char ubuf[64];
    time_t utc = p.utc_start;
    std::strftime(ubuf, 64, "%Y-%m-%d-%H:%M:%S", std::gmtime(&utc));
    msg << ubuf << " " << buffer << " " << first_idx << " " << pbeam << " " << nbeams << " "
        << n_events << std::endl;
    std::ostringstream ev;
    for (size_t i = go.lo; i < go.hi; ++i) {
      const hd_size ind = pl->r_ind[i];
      const hd_size beam_idx = (hd_size)std::floor((float)ind / (float)nsamps_beam);
      const hd_size beam_start = beam_idx * nsamps_beam + (p.boxcar_max / 2);
      if (!(beam_start <= ind && ind < beam_start + nsamps_valid)) continue;
      const hd_size samp_idx = first_idx + ind - (nsamps_beam * beam_idx);
      const hd_float snr = std::isinf(pl->r_peak[i]) ? 0.0f : pl->r_peak[i];
      ev << snr << "\t" << samp_idx << "\t" << samp_idx * p.dt << "\t" << pl->r_filter[i] << "\t"
         << pl->r_dm[i] << "\t" << pl->r_dms[i] << "\t" << pl->r_members[i] << "\t"
         << first_idx + (pl->r_begin[i] % nsamps_beam) << "\t"
         << first_idx + (pl->r_end[i] % nsamps_beam) << "\t" << (pbeam + beam_idx + 1)
         << std::endl;
    }
    addrinfo hints{}, *res = nullptr;
    hints.ai_family = AF_INET;
    hints.ai_socktype = SOCK_STREAM;
    const std::string port = std::to_string(p.coincidencer_port);
    if (getaddrinfo(pl->coinc_host.c_str(), port.c_str(), &hints, &res) == 0 && res) {
      int fd = socket(res->ai_family, res->ai_socktype, res->ai_protocol);
      if (fd >= 0 && connect(fd, res->ai_addr, res->ai_addrlen) == 0) {
        const std::string a = msg.str(), b = ev.str();
        (void)!write(fd, a.data(), a.size());
        (void)!write(fd, b.data(), b.size());
      } else {
        std::cerr << "SocketException was caught:Could not connect to port." << "\n";
      }
      if (fd >= 0) close(fd);
      freeaddrinfo(res);
    } else {
      std::cerr << "SocketException was caught:Could not resolve host." << "\n";
    }
    return HD_NO_ERROR;
  }
  const std::string filename = pl->output_dir + "/" + std::string(buffer) + "_" + ss.str() + ".cand";
  std::ofstream cand_file(filename.c_str(), std::ios::out);
  if (cand_file.good()) cand_file << text;
  else std::cout << "Skipping dump due to bad file open on " << filename << std::endl;
  cand_file.close();
  return HD_NO_ERROR;
}

// Input of one gulp: what a separate hd_execute call would get, with the
// filterbank already in device memory.
struct GulpIn {
  const hd_byte* d_fb;
  size_t nsamps, nbeams, first_idx, beam;
};

// Dedispersion autotuner.  Every tile height (8, 16 or 32 trials per block),
// channels staged per step (4, 8 or 16), block order (XCD-aware or linear) and,
// for sub-byte samples, lane width (16-bit pairs or 8-bit quads) computes the
// same exact sums, so the choice only moves time.  The first launch of a
// shape (rows per group, grouped or not, live channels, sample bits) times
// each candidate on the real input with HIP events (one warm-up launch and
// one timed launch each, over a slice of the time tiles and groups, into the
// plane buffer that the real launch then overwrites) and keeps the fastest;
// HD_DEDISP_TILE="db,remap[,lanes8[,cc]]" fixes the choice instead.  Sets
// dg's tile fields and returns the staging pairs.
// Channels the byte-staged kernel sums between widenings: 8-bit lanes for
// sub-byte samples (lanes * max sample <= 255), 16-bit lanes of the even/odd
// split for 8-bit samples.
int widen8_for(size_t nbits) {
  return nbits <= 4 ? 255 / ((1 << nbits) - 1) : 256;
}

int tile_index(int db, int cc) {
  return (db == 8 ? 0 : (db == 16 ? 1 : 2)) + (cc == 4 ? 0 : (cc == 8 ? 3 : 6));
}

void apply_tile(hd_pipeline pl, hd::DedispGeom& dg, int db, int remap, int lanes8, int cc,
                int wide, size_t nbits) {
  const int k = tile_index(db, cc);
  dg.db = db;
  dg.cc = cc;
  dg.remap = remap;
  dg.wlen = pl->wlen_db[k];
  dg.widen8 = lanes8 ? widen8_for(nbits) : 0;
  dg.wlen8 = lanes8 ? pl->wlen8_db[k] : 0;
  dg.split8 = lanes8 && nbits == 8 ? 1 : 0;
  dg.wide = lanes8 && db == 32 ? wide : 0;
}

// The leading DM blocks of a byte-staged launch of 32 trials on 1024 threads
// whose windows fit two staging buffers (rows r0 .. r0 + rows - 1 of the
// host delay table d, nl live channels per row): they run double-buffered.
void plan_dbuf(hd::DedispGeom& dg, const std::vector<int>& d, size_t r0, size_t rows,
               size_t nl) {
  dg.dbuf_blocks = 0;
  dg.wlen8_dbuf = 0;
  if (!(dg.widen8 > 0 && dg.wide && dg.db == 32) || nl == 0 || d.size() < (r0 + rows) * nl)
    return;
  int wmax = 0;
  for (size_t m0 = 0; m0 < rows; m0 += 32) {
    const size_t a = (r0 + m0) * nl, b = (r0 + std::min(m0 + 32, rows) - 1) * nl;
    int spread = 0;
    for (size_t l = 0; l < nl; ++l) spread = std::max(spread, d[b + l] - (d[a + l] & ~3));
    const int w = std::max(wmax, hd::dedisp8_window_dwords(spread));
    if (hd::dedisp_lds_bytes_dbuf(dg, w) > 160 * 1024) break;
    wmax = w;
    dg.dbuf_blocks = (int)(m0 / 32) + 1;
  }
  dg.wlen8_dbuf = wmax;
}

// The same per group, for groups that start at different trials (the gated
// search's units): a group whose DM blocks' windows all fit two staging
// buffers runs double-buffered (GroupDev::dbuf).  Returns whether any does.
bool plan_dbuf_groups(hd::DedispGeom& dg, std::vector<hd::GroupDev>& groups,
                      const std::vector<int>& d, size_t rows, size_t nl) {
  dg.dbuf_blocks = 0;
  dg.wlen8_dbuf = 0;
  dg.dbuf_groups = 0;
  for (hd::GroupDev& g : groups) g.dbuf = 0;
  if (!(dg.widen8 > 0 && dg.wide && dg.db == 32) || nl == 0) return false;
  int wlim = 0;  // the widest window two buffers hold
  while (hd::dedisp_lds_bytes_dbuf(dg, wlim + 64) <= 160 * 1024) wlim += 64;
  int wmax = 0;
  for (hd::GroupDev& g : groups) {
    if (d.size() < (g.trial0 + rows) * nl) continue;
    int w = 0;
    for (size_t m0 = 0; m0 < rows; m0 += 32) {
      const size_t a = (g.trial0 + m0) * nl, b = (g.trial0 + std::min(m0 + 32, rows) - 1) * nl;
      int spread = 0;
      for (size_t l = 0; l < nl; ++l) spread = std::max(spread, d[b + l] - (d[a + l] & ~3));
      w = std::max(w, hd::dedisp8_window_dwords(spread));
    }
    if (w <= wlim) {
      g.dbuf = 1;
      wmax = std::max(wmax, w);
    }
  }
  if (wmax == 0) return false;
  dg.dbuf_groups = 1;
  dg.wlen8_dbuf = wmax;
  return true;
}

hd_error choose_dedisp_tile(hd_pipeline pl, hd::DedispGeom& dg, const hd::GroupArgs& gr,
                            size_t nbits, hipStream_t s, const int** pairs) {
  // trials per block, remap, byte staging, channels per step, 1024 threads
  using Tile = std::array<int, 5>;
  static const Tile forced = [] {
    const char* e = getenv("HD_DEDISP_TILE");
    int db = 0, rm = 1, l8 = 0, cc = 4, wide = 0;
    if (e && std::sscanf(e, "%d,%d,%d,%d,%d", &db, &rm, &l8, &cc, &wide) >= 1 &&
        (db == 8 || db == 16 || db == 32))
      return Tile{db, rm ? 1 : 0, l8 ? 1 : 0, (cc == 8 || cc == 16) ? cc : 4, wide ? 1 : 0};
    return Tile{0, 1, 0, 4, 0};
  }();
  Tile c;
  if (pl->fixed_db) {
    c = Tile{pl->fixed_db, pl->fixed_remap, pl->fixed_lanes8, pl->fixed_cc, pl->fixed_wide};
  } else if (forced[0]) {
    c = forced;
  } else {
    // rows per group, live channels, sample bits, grouped or not: the number of
    // groups and the row pitch only scale the grid
    const auto key = std::make_tuple(dg.n_dm, gr.gd ? 1 : 0, dg.n_live, (int)nbits);
    auto it = pl->tune.find(key);
    if (it == pl->tune.end()) {
      // candidates timed as they will run: with double-buffered staging for
      // the leading DM blocks when every group starts at the same trial
      bool dbuf_ok = pl->use_dbuf && gr.host;
      for (int g = 0; dbuf_ok && g < gr.G; ++g)
        dbuf_ok = gr.host[g].trial0 == gr.host[0].trial0;
      const size_t r0 = dbuf_ok ? gr.host[0].trial0 : 0;
      const size_t nl = (size_t)std::max(dg.n_live, 1);
      if (!pl->tev[0]) {
        HD_CHECK(hipEventCreate(&pl->tev[0]));
        HD_CHECK(hipEventCreate(&pl->tev[1]));
      }
      float best = 1e30f;
      Tile choice{16, 1, 0, 4, 0};
      for (int lanes = 0; lanes < 2; ++lanes)
        for (int cc = 4; cc <= 16; cc *= 2)
          for (int k = 0; k < 3; ++k)
            for (int wide = 0; wide <= (lanes && k == 2 ? 1 : 0); ++wide)
            for (int r = 0; r < 2; ++r) {
              // timed on the first (at most) 48 time tiles of the first (at most)
              // 64 groups: the tiles are alike, and large shapes would otherwise
              // spend seconds here; fewer groups would leave a launch too small
              // to rank the tiles as the full launch does
              hd::DedispGeom t = dg;
              apply_tile(pl, t, 8 << k, r, lanes, cc, wide, nbits);
              if (dbuf_ok) plan_dbuf(t, pl->delays_h, r0, (size_t)dg.n_dm, nl);
              t.ppitch = std::min(dg.ppitch, (size_t)48 * hd::dedisp_tile_time());
              hd::GroupArgs tg = gr;
              tg.G = std::min(gr.G, 64);
              if (hd::dedisp_lds_bytes(t) > 160 * 1024) continue;
              const int* tp = pl->pairs_db[tile_index(8 << k, cc)].as<int>();
              auto launch = [&]() {
                return hd::launch_dedisperse(pl->x.as<uint8_t>(), pl->delays.as<int>(), tp, t,
                                             pl->plane.as<float>(), s, tg);
              };
              if (launch() != hipSuccess) continue;
              HD_CHECK(hipEventRecord(pl->tev[0], s));
              HD_CHECK(launch());
              HD_CHECK(hipEventRecord(pl->tev[1], s));
              HD_CHECK(hipEventSynchronize(pl->tev[1]));
              float ms = 0.f;
              HD_CHECK(hipEventElapsedTime(&ms, pl->tev[0], pl->tev[1]));
              if (ms < best) {
                best = ms;
                choice = Tile{8 << k, r, lanes, cc, wide};
              }
            }
      // The block order's benefit (L2 sharing within an XCD) grows with the
      // time tiles a launch holds, so the winner's order is settled on all of
      // them (at most 64 groups).
      {
        float bt[2] = {1e30f, 1e30f};
        for (int r = 0; r < 2; ++r) {
          hd::DedispGeom t = dg;
          apply_tile(pl, t, choice[0], r, choice[2], choice[3], choice[4], nbits);
          if (dbuf_ok) plan_dbuf(t, pl->delays_h, r0, (size_t)dg.n_dm, nl);
          hd::GroupArgs tg = gr;
          tg.G = std::min(gr.G, 64);
          const int* tp = pl->pairs_db[tile_index(choice[0], choice[3])].as<int>();
          auto launch = [&]() {
            return hd::launch_dedisperse(pl->x.as<uint8_t>(), pl->delays.as<int>(), tp, t,
                                         pl->plane.as<float>(), s, tg);
          };
          if (launch() != hipSuccess) continue;
          HD_CHECK(hipEventRecord(pl->tev[0], s));
          HD_CHECK(launch());
          HD_CHECK(hipEventRecord(pl->tev[1], s));
          HD_CHECK(hipEventSynchronize(pl->tev[1]));
          HD_CHECK(hipEventElapsedTime(&bt[r], pl->tev[0], pl->tev[1]));
        }
        if (bt[0] < 1e30f || bt[1] < 1e30f) choice[1] = bt[1] <= bt[0] ? 1 : 0;
      }
      it = pl->tune.emplace(key, choice).first;
      if (pl->params.verbosity >= 2)
        std::cout << "dedispersion tile: " << choice[0] << " trials per block, " << choice[3]
                  << " channels per step, " << (choice[1] ? "XCD-aware" : "linear")
                  << " order, " << (choice[2] ? "byte" : "16-bit pair") << " staging"
                  << (choice[4] ? ", 1024 threads" : "") << " (" << best << " ms)" << std::endl;
    }
    c = it->second;
  }
  if (c[2] && widen8_for(nbits) < c[3]) c[2] = 0;  // byte lanes would carry within a step
  apply_tile(pl, dg, c[0], c[1], c[2], c[3], c[4], nbits);
  // a fixed tile whose staged copies exceed the LDS stages fewer channels per step
  while (c[3] > 4 && hd::dedisp_lds_bytes(dg) > 160 * 1024) {
    c[3] /= 2;
    apply_tile(pl, dg, c[0], c[1], c[2], c[3], c[4], nbits);
  }
  *pairs = pl->pairs_db[tile_index(c[0], c[3])].as<int>();
  pl->last_db = c[0];
  pl->last_remap = c[1];
  pl->last_lanes8 = c[2];
  pl->last_cc = c[3];
  pl->last_wide = dg.wide;
  return HD_NO_ERROR;
}

// The coarse plane of the G gulps of pl->hgd, whose unpacked rows are slots
// 0..G-1 of pl->x (pitch xpitch): decimation, coarse dedispersion, the
// search's baseline and RMS kernels on the coarse series, bfloat16 output,
// and the tile grid (units x time positions per gulp).
hd_error coarse_pass(hd_pipeline pl, int G, size_t xpitch, size_t nbits, hipStream_t s) {
  const hd_params& p = pl->params;
  hd_error e = build_coarse(pl);
  if (e != HD_NO_ERROR) return e;
  const int nc = (int)pl->co_dms.size();
  const int n_live = (int)pl->live_list.size();
  const int TB = hd::dedisp_tile_time();
  std::vector<hd::GroupDev>& cg = pl->co_hgd;
  cg.assign(G, hd::GroupDev{});
  size_t ppc = 0, spc = 0, prow = 0, nout_max = 0;
  for (int g = 0; g < G; ++g) {
    const size_t ns = pl->hgd[g].nsamps / HD_PROPOSAL_TIME;
    if (ns < pl->co_max_delay + 64) return HD_TOO_FEW_NSAMPS;
    SeriesShape& sh = cg[g].sh;
    sh.nbeams = 1;
    sh.stride = ns;
    sh.nout = ns - pl->co_max_delay;
    sh.bc = sh.nout;
    sh.ppitch = round_up(sh.nout + 1, (size_t)TB);
    sh.spitch = round_up(sh.nout + 1, 64);
    const int se = hd::series_shape(p.dt * (float)HD_PROPOSAL_TIME, p.baseline_length, sh);
    if (se != HD_NO_ERROR) return (hd_error)se;
    cg[g].nsamps = ns;
    cg[g].nsamps_beam = ns;
    cg[g].xslot = (uint32_t)g;
    cg[g].cgroup = (uint32_t)g;
    cg[g].rows_valid = (uint32_t)nc;
    ppc = std::max(ppc, sh.ppitch);
    spc = std::max(spc, sh.spitch);
    prow = std::max(prow, sh.sc);
    nout_max = std::max(nout_max, sh.nout);
  }
  for (int g = 0; g < G; ++g) {
    cg[g].sh.ppitch = ppc;
    cg[g].sh.spitch = spc;
  }
  HD_CHECK(pl->co_gdev.ensure((size_t)G * sizeof(hd::GroupDev)));
  HD_CHECK(hipMemcpyAsync(pl->co_gdev.p, cg.data(), (size_t)G * sizeof(hd::GroupDev),
                          hipMemcpyHostToDevice, s));
  hd::GroupArgs cga;
  cga.gd = pl->co_gdev.as<hd::GroupDev>();
  cga.host = cg.data();
  cga.G = G;
  cga.nbmax = 1;
  cga.prow = std::max<size_t>(prow, 1);
  const size_t xpc = round_up(ppc + pl->co_max_delay + 2 * pl->co_wlen + 256, 256);
  HD_CHECK(pl->co_xc.ensure((size_t)G * std::max(n_live, 1) * xpc));
  HD_CHECK(hd::launch_decimate(pl->x.as<uint8_t>(), xpitch, n_live, (int)nbits,
                               pl->co_xc.as<uint8_t>(), xpc, cga.gd, G, s));
  const size_t rows = (size_t)G * nc;
  HD_CHECK(pl->co_plane.ensure(rows * ppc * sizeof(float)));
  hd::DedispGeom dg{nc, n_live, cg[0].sh.nout, xpc, ppc, (int)pl->co_wlen};
  HD_CHECK(hd::launch_dedisperse(pl->co_xc.as<uint8_t>(), pl->co_delays.as<int>(),
                                 pl->co_pairs.as<int>(), dg, pl->co_plane.as<float>(), s, cga));
  HD_CHECK(pl->co_pts.ensure((rows * cga.prow + 64) * sizeof(float)));
  HD_CHECK(pl->co_rms.ensure(rows * sizeof(float)));
  HD_CHECK(hd::launch_baseline_points(pl->co_plane.as<float>(), cg[0].sh, nc,
                                      pl->co_pts.as<float>(), s, cga));
  const size_t words = hd::beam_rms_scratch_words(cg[0].sh, nc, cga);
  HD_CHECK(pl->co_brms.ensure(std::max<size_t>(words, 1) * sizeof(float)));
  HD_CHECK(hd::launch_beam_rms(pl->co_plane.as<float>(), pl->co_pts.as<float>(), cg[0].sh, nc,
                               pl->co_rms.as<float>(), s, pl->co_brms.as<float>(), words, cga));
  const size_t bpitch = round_up(nout_max, 64);
  const size_t wpitch = round_up(nout_max / 4 + 1, 64);
  HD_CHECK(pl->co_bf.ensure(rows * bpitch * sizeof(uint16_t)));
  HD_CHECK(pl->co_bw.ensure(rows * wpitch * sizeof(uint16_t)));
  HD_CHECK(hd::launch_coarse_bf16(pl->co_plane.as<float>(), ppc, pl->co_pts.as<float>(),
                                  cga.prow, pl->co_rms.as<float>(), 1, cga.gd, G, nc,
                                  pl->co_bf.as<uint16_t>(), bpitch, pl->co_bw.as<uint16_t>(),
                                  wpitch, s));
  // tile grid
  const int units = (int)((pl->dm_list.size() + HD_PROPOSAL_UNIT - 1) / HD_PROPOSAL_UNIT);
  pl->co_toff_h.assign(G + 1, 0);
  pl->co_ntpos_h.assign(G, 0);
  pl->co_ncols_h.assign(G, 0);
  for (int g = 0; g < G; ++g) {
    const size_t nv = (cg[g].sh.nout + HD_PROPOSAL_TSTEP - 1) / HD_PROPOSAL_TSTEP;
    pl->co_ntpos_h[g] = (uint32_t)nv;
    pl->co_ncols_h[g] = (uint32_t)cg[g].sh.nout;
    const uint64_t next = (uint64_t)pl->co_toff_h[g] + (uint64_t)units * nv;
    if (next >= (1ull << 32)) return HD_INVALID_STRIDE;
    pl->co_toff_h[g + 1] = (uint32_t)next;
  }
  HD_CHECK(pl->co_toff.ensure((G + 1) * sizeof(uint32_t)));
  HD_CHECK(pl->co_ntpos.ensure(G * sizeof(uint32_t)));
  HD_CHECK(pl->co_ncols.ensure(G * sizeof(uint32_t)));
  HD_CHECK(hipMemcpyAsync(pl->co_toff.p, pl->co_toff_h.data(), (G + 1) * sizeof(uint32_t),
                          hipMemcpyHostToDevice, s));
  HD_CHECK(hipMemcpyAsync(pl->co_ntpos.p, pl->co_ntpos_h.data(), G * sizeof(uint32_t),
                          hipMemcpyHostToDevice, s));
  HD_CHECK(hipMemcpyAsync(pl->co_ncols.p, pl->co_ncols_h.data(), G * sizeof(uint32_t),
                          hipMemcpyHostToDevice, s));
  pl->co_G = G;
  pl->co_rows = nc;
  pl->co_units = units;
  pl->co_bpitch = bpitch;
  pl->co_wpitch = wpitch;
  return HD_NO_ERROR;
}

hd::TileGrid tile_grid(hd_pipeline pl) {
  hd::TileGrid tg;
  tg.bf = pl->co_bf.as<uint16_t>();
  tg.bpitch = pl->co_bpitch;
  tg.bw = pl->co_bw.as<uint16_t>();
  tg.wpitch = pl->co_wpitch;
  tg.n_rows = pl->co_rows;
  tg.ncols = pl->co_ncols.as<uint32_t>();
  tg.n_units = pl->co_units;
  tg.tile_off = pl->co_toff.as<uint32_t>();
  tg.n_tpos = pl->co_ntpos.as<uint32_t>();
  tg.G = pl->co_G;
  return tg;
}

// Logits of every tile of the coarse plane into co_scores, in batches of
// the network's capacity (tiles extracted, then scored, on stream s).
hd_error score_tiles(hd_pipeline pl, hipStream_t s) {
  if (!pl->net || pl->co_G == 0) return HD_INVALID_POINTER;
  const size_t total = pl->co_toff_h.back();
  const size_t B = hd::cnn_max_batch(pl->net);
  HD_CHECK(pl->co_tiles.ensure(B * HD_PROPOSAL_TILE * HD_PROPOSAL_TILE * 2 * sizeof(uint16_t)));
  HD_CHECK(pl->co_scores.ensure(std::max<size_t>(total, 1) * sizeof(float)));
  const hd::TileGrid tg = tile_grid(pl);
  for (size_t first = 0; first < total; first += B) {
    const size_t n = std::min(B, total - first);
    HD_CHECK(hd::launch_extract_tiles(tg, nullptr, first, n, pl->co_tiles.as<uint16_t>(), s));
    HD_CHECK(hd::cnn_forward_async(pl->net, pl->co_tiles.as<uint16_t>(), n,
                                   pl->co_scores.as<float>() + first, s));
  }
  return HD_NO_ERROR;
}

// Units of every gulp whose tiles score above the threshold (host flags).
hd_error propose(hd_pipeline pl, hipStream_t s) {
  hd_error e = score_tiles(pl, s);
  if (e != HD_NO_ERROR) return e;
  const int G = pl->co_G, U = pl->co_units;
  HD_CHECK(pl->co_flags.ensure(std::max(G * U, 1)));
  HD_CHECK(hd::launch_unit_flags(pl->co_scores.as<float>(), pl->co_toff.as<uint32_t>(),
                                 pl->co_ntpos.as<uint32_t>(), G, U, pl->prop_thresh,
                                 pl->co_flags.as<uint8_t>(), s));
  pl->prop_flags.assign((size_t)G * U, 0);
  pl->prop_units = U;
  HD_CHECK(hipMemcpyAsync(pl->prop_flags.data(), pl->co_flags.p, (size_t)G * U,
                          hipMemcpyDeviceToHost, s));
  HD_CHECK(hipStreamSynchronize(s));
  return HD_NO_ERROR;
}

// The device part of a call: every gulp of `in` is searched with one launch
// per stage (the gulps are the groups of the kernels' group dimension).
// Per-gulp results land in pl->gulps; a returned error concerns the call.
hd_error run_gulps(hd_pipeline pl, const std::vector<GulpIn>& in, size_t nbits) {
  const hd_params& p = pl->params;
  hipStream_t s = pl->stream;
  // the launchers report hipGetLastError(): start from a clean state, so that
  // an error an earlier call already returned is not reported again here
  (void)hipGetLastError();
  pl->n_giants = 0;
  pl->r_peak.clear(); pl->r_ind.clear(); pl->r_begin.clear(); pl->r_end.clear();
  pl->r_filter.clear(); pl->r_dm.clear(); pl->r_members.clear(); pl->r_dms.clear();
  pl->r_group.clear();
  pl->cand_text.clear();
  pl->prop_flags.clear();
  pl->prop_units = 0;
  for (double& t : pl->timings) t = 0;
  const size_t W = p.boxcar_max;
  const size_t max_delay = pl->max_delay;
  pl->gulps.assign(in.size(), GulpOut{});
  // shapes of every gulp (pipeline.cu:366-372, 488-499, 633-638)
  std::vector<CallShape> shapes(in.size());
  std::vector<int> active;
  for (size_t k = 0; k < in.size(); ++k) {
    GulpOut& go = pl->gulps[k];
    go.nsamps = in[k].nsamps;
    go.nbeams = in[k].nbeams;
    go.first_idx = in[k].first_idx;
    go.beam = in[k].beam;
    const size_t ns = in[k].nsamps, nb = in[k].nbeams;
    go.processed = (ns < nb * (W + max_delay)) ? 0 : ns - nb * (W + max_delay);
    if (max_delay >= ns) {
      go.err = HD_TOO_FEW_NSAMPS;
      continue;
    }
    go.err = make_shapes(pl, ns, nb, shapes[k]);
    if (go.err == HD_NO_ERROR && !shapes[k].skip) active.push_back((int)k);
  }
  hd_error err = build_live_and_delays(pl);
  if (err != HD_NO_ERROR) return err;
  if (active.empty()) return HD_NO_ERROR;
  const int G = (int)active.size();
  const WidthTable& wt = shapes[active[0]].wt;  // shared but for the fresh counts
  const int n_dm = (int)pl->dm_list.size();
  const int n_live = (int)pl->live_list.size();
  // common row pitches; every group keeps its own lengths
  size_t ppitch = 0, spitch = 0;
  unsigned nbmax = 0;
  size_t prow = 0;
  for (int k : active) {
    ppitch = std::max(ppitch, shapes[k].sh.ppitch);
    spitch = std::max(spitch, shapes[k].sh.spitch);
    nbmax = std::max(nbmax, shapes[k].sh.nbeams);
    prow = std::max(prow, (size_t)shapes[k].sh.nbeams * shapes[k].sh.sc);
  }
  const uint64_t span = (uint64_t)p.cand_sep_time << wt.wi[wt.nw - 1].filter_idx;
  pl->hgd.assign(G, hd::GroupDev{});
  for (hd::GroupDev& d : pl->hgd) std::memset(&d, 0, sizeof(d));  // bytes key the graph
  uint64_t t_off = 0;
  for (int g = 0; g < G; ++g) {
    const CallShape& cs = shapes[active[g]];
    hd::GroupDev& d = pl->hgd[g];
    d.sh = cs.sh;
    d.sh.ppitch = ppitch;
    d.sh.spitch = spitch;
    d.bs = cs.bs;
    d.bc692 = cs.bc692;
    d.fb = in[active[g]].d_fb;
    d.nsamps = in[active[g]].nsamps;
    for (int j = 0; j < cs.wt.nw; ++j) d.n_w[j] = cs.wt.wi[j].n_w;
    d.t_off = t_off;
    d.nsamps_beam = d.nsamps / cs.sh.nbeams;
    t_off += d.nsamps + 2 * span + 64;  // groups never fall within one coincidence window
  }
  if (t_off >= (1ull << 32)) return HD_INVALID_STRIDE;  // clustering coordinate is 32-bit

  record(pl, 9);
  if (pl->cleaning) {
    // pipeline.cu:263-299: clean at DM 0 with boxcars up to 64, then search
    // the cleaned copy (:447-451)
    const size_t row_bytes = p.nchans * nbits / 8;
    size_t total = 0, work = 0;
    std::vector<hd::CleanPlan> cps(G);
    for (int g = 0; g < G; ++g) {
      const int ce = hd::clean_plan(pl->hgd[g].nsamps, nbits, p.nchans, pl->hgd[g].sh.nbeams, 0.f,
                                    p.dt, p.f0, p.df, p.baseline_length, p.rfi_tol, 64, cps[g]);
      if (ce != HD_NO_ERROR) return (hd_error)ce;
      total += pl->hgd[g].nsamps * row_bytes;
      work = std::max(work, cps[g].work_bytes);
    }
    HD_CHECK(pl->clean_out.ensure(total));
    HD_CHECK(pl->clean_work.ensure(work));
    size_t off = 0;
    for (int g = 0; g < G; ++g) {
      hd_byte* dst = pl->clean_out.as<hd_byte>() + off;
      HD_CHECK(hd::run_clean(cps[g], pl->hgd[g].fb, dst, pl->clean_work.p, s));
      pl->hgd[g].fb = dst;
      off += pl->hgd[g].nsamps * row_bytes;
    }
  }
  for (int g = 0; g < G; ++g) {  // every trial of the gulp, its own input slot
    pl->hgd[g].trial0 = 0;
    pl->hgd[g].rows_valid = (uint32_t)n_dm;
    pl->hgd[g].xslot = (uint32_t)g;
    pl->hgd[g].cgroup = (uint32_t)g;
  }
  HD_CHECK(pl->gdev.ensure((size_t)G * sizeof(hd::GroupDev)));
  HD_CHECK(hipMemcpyAsync(pl->gdev.p, pl->hgd.data(), (size_t)G * sizeof(hd::GroupDev),
                          hipMemcpyHostToDevice, s));
  hd::GroupArgs ga;
  ga.gd = pl->gdev.as<hd::GroupDev>();
  ga.host = pl->hgd.data();
  ga.G = G;
  ga.nbmax = nbmax;
  ga.prow = std::max<size_t>(prow, 1);
  const SeriesShape& sh0 = pl->hgd[0].sh;

  // Adaptive time resolution searches each run of equal scrunch factor from
  // the input at that resolution; the gated search has no such path.
  if (pl->adaptive && (pl->gating || pl->coarse_only)) return HD_INVALID_PIPELINE;
  const bool adaptive = pl->adaptive;
  int n_levels = 1;  // input resolutions: level l holds the input halved l times
  if (adaptive)
    for (const auto& k : pl->classes)
      if (k.factor <= W) n_levels = std::max(n_levels, (int)log2_pow2((unsigned)k.factor) + 1);

  record(pl, 0);
  // unpack to channel-major bytes
  const size_t xpitch = round_up(ppitch + max_delay + 2 * pl->wlen + 256, 256);
  HD_CHECK(pl->x.ensure((size_t)n_levels * G * std::max(n_live, 1) * xpitch));
  // (the proposal's decimation reads whole rows, so rows are trimmed only
  // for the plain search of several gulps)
  HD_CHECK(hd::launch_unpack(nullptr, p.nchans * nbits / 8, (int)nbits, 0, pl->live.as<int>(),
                             n_live, pl->x.as<uint8_t>(), xpitch, s, ga,
                             (size_t)n_live == p.nchans,
                             G > 1 && !(pl->gating || pl->coarse_only)));
  // level l, slots l*G .. l*G + G - 1: level l - 1 halved (pairwise means
  // rounded down), then zeros as far as the dedispersion windows reach: a
  // tile starting at t0 <= nout / f reads bytes below t0 + (max delay of its
  // trials) + tile + 52 <= nsamps / f + tile + 53
  for (int l = 1; l < n_levels && n_live > 0; ++l)
    for (int g = 0; g < G; ++g) {
      const size_t slot = (size_t)n_live * xpitch;
      const size_t ns = pl->hgd[g].nsamps >> (l - 1);
      const size_t limit = round_up(ns / 2 + 2 * (size_t)hd::dedisp_tile_time(), 256);
      HD_CHECK(hd::launch_scrunch2(pl->x.as<uint8_t>() + ((size_t)(l - 1) * G + g) * slot, n_live,
                                   xpitch, ns, pl->x.as<uint8_t>() + ((size_t)l * G + g) * slot, s,
                                   limit));
    }
  record(pl, 10);
  // The search runs over row groups: a whole gulp each, or with gating one
  // per proposed unit of HD_PROPOSAL_UNIT trials (rows past the last trial
  // are padding).  Row groups keep their gulp's shapes, input and
  // clustering group, so every value they compute is the full search's.
  int R = n_dm;
  std::vector<hd::GroupDev>& hrg = pl->hrg;
  if (pl->gating || pl->coarse_only) {
    for (int k : active)
      if (in[k].nbeams != 1) return HD_INVALID_STRIDE;  // one beam per gulp
    hd_error e = coarse_pass(pl, G, xpitch, nbits, s);
    if (e != HD_NO_ERROR) return e;
    if (pl->coarse_only) {
      for (int ev = 1; ev <= 8; ++ev) record(pl, ev);
      return HD_NO_ERROR;
    }
    e = propose(pl, s);
    if (e != HD_NO_ERROR) return e;
    R = HD_PROPOSAL_UNIT;
    hrg.clear();
    for (int g = 0; g < G; ++g)
      for (int u = 0; u < pl->prop_units; ++u) {
        if (!pl->prop_flags[(size_t)g * pl->prop_units + u]) continue;
        hd::GroupDev d = pl->hgd[g];
        d.trial0 = (uint32_t)(u * HD_PROPOSAL_UNIT);
        d.rows_valid = (uint32_t)std::min<size_t>(HD_PROPOSAL_UNIT, n_dm - d.trial0);
        hrg.push_back(d);
      }
  } else {
    hrg = pl->hgd;
  }
  record(pl, 1);
  if (hrg.empty() && !adaptive) {  // nothing proposed: no giants, no candidates
    for (int ev = 2; ev <= 8; ++ev) record(pl, ev);
    return HD_NO_ERROR;
  }
  // The stage chain (dedispersion .. thresholding) and the giants of row
  // groups hrg of R rows each: width table wt (shared by the groups but for
  // their fresh counts), baseline points per row prow, unpacked input xb,
  // delay table dly, launch shape sh0; the dedispersion tile fixed (tile) or
  // the autotuner's.  Leaves the giants in g_* and their count in n_giants.
  auto search_rows = [&](std::vector<hd::GroupDev>& hrg, const int R, const WidthTable& wt,
                         const size_t prow, const uint8_t* xb, const int* dly,
                         const hd::DedispGeom* tile, const SeriesShape& sh0,
                         size_t& n_giants) -> hd_error {
  const int Gr = (int)hrg.size();
  n_giants = 0;
  HD_CHECK(pl->rdev.ensure((size_t)Gr * sizeof(hd::GroupDev)));
  HD_CHECK(hipMemcpyAsync(pl->rdev.p, hrg.data(), (size_t)Gr * sizeof(hd::GroupDev),
                          hipMemcpyHostToDevice, s));
  hd::GroupArgs gr = ga;
  gr.gd = pl->rdev.as<hd::GroupDev>();
  gr.host = hrg.data();
  gr.G = Gr;
  gr.prow = prow;
  const size_t rows = (size_t)Gr * R;
  // Every buffer of the stage chain (dedispersion .. thresholding) is sized
  // first: the chain may be captured as a graph, which must not allocate.
  HD_CHECK(pl->plane.ensure(rows * ppitch * sizeof(float)));
  HD_CHECK(pl->pts.ensure((rows * gr.prow + 64) * sizeof(float)));
  HD_CHECK(pl->rms.ensure(rows * nbmax * sizeof(float)));
  const size_t rms_words = hd::beam_rms_scratch_words(sh0, R, gr);
  HD_CHECK(pl->brms.ensure(std::max<size_t>(rms_words, 1) * sizeof(float)));
  // normalised prefix sums; rows start 3 floats into the allocation so that
  // S[1..] is 16-byte aligned for the scan's vector stores
  HD_CHECK(pl->S.ensure((rows * spitch + 8) * sizeof(float)));
  float* const Sbase = pl->S.as<float>() + 3;
  const size_t np_max = hd::block_rms_np_max(gr);
  HD_CHECK(pl->partial.ensure(rows * wt.nw * std::max<size_t>(np_max, 1) * sizeof(float)));
  HD_CHECK(pl->wrms.ensure(rows * wt.nw * sizeof(float)));
  HD_CHECK(pl->winv.ensure(rows * wt.nw * sizeof(float)));
  int n_fresh = 0;
  while (n_fresh < wt.nw && wt.wi[n_fresh].ts == 1) ++n_fresh;
  hd::GiantBuffers gb{};
  gb.chunk = hd::threshold_chunk(wt);
  size_t n_w0 = 0;
  for (const hd::GroupDev& d : hrg) n_w0 = std::max<size_t>(n_w0, d.n_w[0]);
  gb.nchunks_per_w = (n_w0 + gb.chunk - 1) / gb.chunk;
  const size_t nkeys = rows * wt.nw * gb.nchunks_per_w;
  HD_CHECK(pl->ccount.ensure(nkeys * sizeof(unsigned)));
  HD_CHECK(pl->coff.ensure(nkeys * sizeof(unsigned)));
  HD_CHECK(pl->cpos.ensure(nkeys * sizeof(unsigned)));
  HD_CHECK(pl->counter.ensure(sizeof(unsigned) * 4));
  HD_CHECK(pl->scan_tmp.ensure(hd::exclusive_scan_tmp_words(std::max<size_t>(nkeys, 1 << 20)) *
                               sizeof(unsigned)));
  HD_CHECK(pl->skey.ensure(pl->sample_cap * sizeof(uint64_t)));
  HD_CHECK(pl->sval.ensure(pl->sample_cap * sizeof(float)));
  auto bind_gb = [&]() {
    gb.skey = pl->skey.as<uint64_t>();
    gb.sval = pl->sval.as<float>();
    gb.scap = pl->sample_cap;
    gb.counter = pl->counter.as<unsigned>();
    gb.ccount = pl->ccount.as<unsigned>();
    gb.coff = pl->coff.as<unsigned>();
    gb.cpos = pl->cpos.as<unsigned>();
    gb.scan_tmp = pl->scan_tmp.as<unsigned>();
  };
  bind_gb();
  hd::DedispGeom dg;
  std::memset(&dg, 0, sizeof(dg));  // padding too: the struct's bytes key the graph
  dg.n_dm = R;
  dg.n_live = n_live;
  dg.nout = sh0.nout;
  dg.xpitch = xpitch;
  dg.ppitch = ppitch;
  dg.wlen = (int)pl->wlen;
  const int* pairs = nullptr;
  if (tile) {
    dg.db = tile->db;
    dg.cc = tile->cc;
    dg.remap = tile->remap;
    dg.wlen = tile->wlen;
    dg.widen8 = tile->widen8;
    dg.wlen8 = tile->wlen8;
    dg.split8 = tile->split8;
    dg.wide = tile->wide;
    // the groups' own series lengths set the grid's time tiles
    size_t nmax = 0;
    for (const hd::GroupDev& d : hrg) nmax = std::max(nmax, d.sh.nout);
    dg.tiles = nmax / hd::dedisp_tile_time() + 1;
  } else {
    hd_error te = choose_dedisp_tile(pl, dg, gr, nbits, s, &pairs);  // may time candidates
    if (te != HD_NO_ERROR) return te;
  }
  // double-buffered staging for the leading DM blocks, when every group
  // starts at the same trial (its blocks then have the same windows)
  {
    bool same = true;
    for (const hd::GroupDev& d : hrg) same = same && d.trial0 == hrg[0].trial0;
    const std::vector<int>& table = dly == pl->delays.as<int>() ? pl->delays_h : pl->adelays_h;
    if (same && pl->use_dbuf) {
      plan_dbuf(dg, table, hrg[0].trial0, (size_t)R, (size_t)std::max(n_live, 1));
    } else if (pl->use_dbuf &&
               plan_dbuf_groups(dg, hrg, table, (size_t)R, (size_t)std::max(n_live, 1))) {
      // the groups' flags go to the device table
      HD_CHECK(hipMemcpyAsync(pl->rdev.p, hrg.data(), (size_t)Gr * sizeof(hd::GroupDev),
                              hipMemcpyHostToDevice, s));
    }
  }
  auto threshold = [&]() -> hipError_t {
    hipError_t e = hd::launch_fill_u32(gb.counter, 1, 0u, s);
    if (e == hipSuccess)
      e = hd::launch_threshold(Sbase, spitch, wt, pl->winv.as<float>(), R, p.detect_thresh, gb,
                               s, gr);
    return e;
  };
  // The chain: kernel launches only.  With `timed`, stage events in between.
  auto chain = [&](bool timed) -> hipError_t {
    hipError_t e = hd::launch_dedisperse(xb, dly, pairs, dg, pl->plane.as<float>(), s, gr);
    if (e != hipSuccess) return e;
    if (timed) record(pl, 2);
    e = hd::launch_baseline_points(pl->plane.as<float>(), sh0, R, pl->pts.as<float>(), s, gr);
    if (e != hipSuccess) return e;
    if (timed) record(pl, 3);
    e = hd::launch_beam_rms(pl->plane.as<float>(), pl->pts.as<float>(), sh0, R,
                            pl->rms.as<float>(), s, pl->brms.as<float>(), rms_words, gr);
    if (e != hipSuccess) return e;
    if (timed) record(pl, 4);
    e = hd::launch_scan(pl->plane.as<float>(), ppitch, 0, R, Sbase, spitch, s,
                        pl->pts.as<float>(), pl->rms.as<float>(), nullptr, gr);
    if (e != hipSuccess) return e;
    if (timed) record(pl, 5);
    // block RMS: the ts == 1 widths together, then each decimated width in order
    e = hd::launch_block_rms_partials(Sbase, spitch, wt, 0, n_fresh, R, pl->partial.as<float>(),
                                      pl->winv.as<float>(), s, gr);
    if (e == hipSuccess)
      e = hd::launch_block_rms_finish(pl->partial.as<float>(), wt.nw, 0, n_fresh, R,
                                      pl->wrms.as<float>(), pl->winv.as<float>(), s, gr);
    for (int j = n_fresh; j < wt.nw && e == hipSuccess; ++j) {
      e = hd::launch_block_rms_partials(Sbase, spitch, wt, j, 1, R, pl->partial.as<float>(),
                                        pl->winv.as<float>(), s, gr);
      if (e == hipSuccess)
        e = hd::launch_block_rms_finish(pl->partial.as<float>(), wt.nw, j, 1, R,
                                        pl->wrms.as<float>(), pl->winv.as<float>(), s, gr);
    }
    if (e != hipSuccess) return e;
    if (timed) record(pl, 6);
    return threshold();
  };
  if (pl->use_graph && !tile && !(pl->gating || pl->coarse_only)) {
    // Graph of the chain, captured once per set of launch arguments (shapes,
    // tile, buffers, the row-group table without its input pointers) and
    // replayed while they repeat.  The row-group table itself is uploaded
    // above on every call.  Its span is timed as the dedispersion stage.
    // (the structs below are zero-filled when made, so their padding is too)
    std::string key;
    auto put = [&](const void* v, size_t n) { key.append(static_cast<const char*>(v), n); };
    for (const hd::GroupDev& d : hrg) {
      hd::GroupDev k;
      std::memcpy(&k, &d, sizeof(k));
      k.fb = nullptr;
      put(&k, sizeof(k));
    }
    put(&sh0, sizeof(sh0));
    put(&wt, sizeof(wt));
    put(&dg, sizeof(dg));
    put(&gb, sizeof(gb));
    const void* ptrs[] = {pl->x.p, pl->delays.p, pairs, pl->plane.p, pl->pts.p, pl->rms.p,
                          pl->brms.p, pl->S.p, pl->partial.p, pl->wrms.p, pl->winv.p,
                          pl->rdev.p};
    put(ptrs, sizeof(ptrs));
    const size_t sc[] = {rows, (size_t)R, (size_t)Gr, gr.prow, gr.nbmax, rms_words, np_max,
                         nkeys, spitch, (size_t)n_fresh};
    put(sc, sizeof(sc));
    put(&p.detect_thresh, sizeof(p.detect_thresh));
    if (!pl->gexec || key != pl->gkey) {
      if (pl->gexec) {
        (void)hipGraphExecDestroy(pl->gexec);
        pl->gexec = nullptr;
      }
      hipGraph_t graph = nullptr;
      HD_CHECK(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
      const hipError_t ce = chain(false);
      const hipError_t ee = hipStreamEndCapture(s, &graph);
      if (ce != hipSuccess || ee != hipSuccess) {
        if (graph) (void)hipGraphDestroy(graph);
        return gpu_err(ce != hipSuccess ? ce : ee);
      }
      const hipError_t ie = hipGraphInstantiate(&pl->gexec, graph, nullptr, nullptr, 0);
      (void)hipGraphDestroy(graph);
      if (ie != hipSuccess) {
        pl->gexec = nullptr;
        return gpu_err(ie);
      }
      pl->gkey = key;
      ++pl->graph_captures;
    }
    HD_CHECK(hipGraphLaunch(pl->gexec, s));
    ++pl->graph_launches;
    for (int ev = 2; ev <= 6; ++ev) record(pl, ev);
  } else {
    HD_CHECK(chain(true));
  }
  // samples above threshold; with too many for the buffer, grow it and redo
  // the threshold (eagerly: the graph is recaptured for the new buffer)
  unsigned nsamp = 0;
  for (int attempt = 0; attempt < 3; ++attempt) {
    if (attempt > 0) {
      HD_CHECK(pl->skey.ensure(pl->sample_cap * sizeof(uint64_t)));
      HD_CHECK(pl->sval.ensure(pl->sample_cap * sizeof(float)));
      bind_gb();
      HD_CHECK(threshold());
    }
    HD_CHECK(hipMemcpyAsync(pl->h_pinned, gb.counter, sizeof(unsigned), hipMemcpyDeviceToHost, s));
    HD_CHECK(hipStreamSynchronize(s));
    nsamp = pl->h_pinned[0];
    if (nsamp <= pl->sample_cap) break;
    pl->sample_cap = round_up((size_t)nsamp + nsamp / 4, 1 << 20);
  }
  if (nsamp > pl->sample_cap) return HD_MEM_ALLOC_FAILED;
  if (nsamp > 0) {
    HD_CHECK(hd::launch_exclusive_scan_u32(gb.ccount, gb.cpos, nkeys, gb.scan_tmp, s));
    HD_CHECK(pl->okey.ensure((size_t)nsamp * sizeof(uint64_t)));
    HD_CHECK(pl->oval.ensure((size_t)nsamp * sizeof(float)));
    HD_CHECK(pl->head.ensure((size_t)nsamp * sizeof(unsigned)));
    HD_CHECK(pl->segid.ensure((size_t)nsamp * sizeof(unsigned)));
    HD_CHECK(pl->scan_tmp.ensure(hd::exclusive_scan_tmp_words(std::max<size_t>(nkeys, nsamp)) *
                                 sizeof(unsigned)));
    gb.scan_tmp = pl->scan_tmp.as<unsigned>();
    gb.okey = pl->okey.as<uint64_t>();
    gb.oval = pl->oval.as<float>();
    gb.head = pl->head.as<unsigned>();
    gb.segid = pl->segid.as<unsigned>();
    HD_CHECK(hd::launch_order_samples(gb, wt, (unsigned)nsamp, s));
    HD_CHECK(hd::launch_segment_heads(gb, wt, nsamp, s));
    HD_CHECK(hd::launch_exclusive_scan_u32(gb.head, gb.segid, nsamp, gb.scan_tmp, s));
    HD_CHECK(hipMemcpyAsync(pl->h_pinned, gb.segid + (nsamp - 1), sizeof(unsigned),
                            hipMemcpyDeviceToHost, s));
    HD_CHECK(hipMemcpyAsync(pl->h_pinned + 1, gb.head + (nsamp - 1), sizeof(unsigned),
                            hipMemcpyDeviceToHost, s));
    HD_CHECK(hipStreamSynchronize(s));
    n_giants = (size_t)pl->h_pinned[0] + pl->h_pinned[1];
    HD_CHECK(pl->g_peak.ensure(n_giants * sizeof(float)));
    HD_CHECK(pl->g_ind.ensure(n_giants * sizeof(uint64_t)));
    HD_CHECK(pl->g_begin.ensure(n_giants * sizeof(uint64_t)));
    HD_CHECK(pl->g_end.ensure(n_giants * sizeof(uint64_t)));
    HD_CHECK(pl->g_filter.ensure(n_giants * sizeof(uint32_t)));
    HD_CHECK(pl->g_dm.ensure(n_giants * sizeof(uint32_t)));
    HD_CHECK(pl->g_group.ensure(n_giants * sizeof(uint32_t)));
    hd::GiantOut go{pl->g_peak.as<float>(), pl->g_ind.as<uint64_t>(), pl->g_begin.as<uint64_t>(),
                    pl->g_end.as<uint64_t>(), pl->g_filter.as<uint32_t>(),
                    pl->g_dm.as<uint32_t>(), pl->g_group.as<uint32_t>()};
    HD_CHECK(pl->g_best.ensure(std::max<size_t>(n_giants, 1) * sizeof(uint64_t)));
    HD_CHECK(hd::launch_segment_reduce(gb, wt, nsamp, (unsigned)n_giants, R, go,
                                       pl->g_best.as<uint64_t>(), s, gr.gd));
  }
  return HD_NO_ERROR;
  };
  size_t n_giants = 0;
  if (!adaptive) {
    const hd_error e = search_rows(hrg, R, wt, ga.prow, pl->x.as<uint8_t>(),
                                   pl->delays.as<int>(), nullptr, sh0, n_giants);
    if (e != HD_NO_ERROR) return e;
  } else {
    // Adaptive time resolution (pipeline.cu:488-790): the runs of trials of
    // equal scrunch factor in turn, each over every gulp with the shapes of
    // its factor, from the input at its resolution.  The giants are gathered
    // per gulp in trial order, as the reference's DM loop appends them: each
    // run's giants (in gulp, then trial order) are appended on the device,
    // and with several gulps put in gulp order at the end.
    size_t acc = 0;  // giants gathered so far
    // room for `need` elements of `elem` bytes, keeping the first acc
    auto room = [&](DevBuf& b, size_t elem, size_t need) -> hipError_t {
      if (b.p && b.bytes >= need * elem) return hipSuccess;
      DevBuf nb;
      hipError_t e = nb.ensure(std::max(need * elem, 2 * b.bytes));
      if (e == hipSuccess && acc)
        e = hipMemcpyAsync(nb.p, b.p, acc * elem, hipMemcpyDeviceToDevice, s);
      if (e != hipSuccess) {
        nb.release();
        return e;
      }
      b.release();  // (hipFree waits for the copy above)
      b = nb;
      return hipSuccess;
    };
    auto append = [&](DevBuf& dst, const DevBuf& src, size_t elem, size_t n) -> hipError_t {
      hipError_t e = room(dst, elem, acc + n);
      if (e == hipSuccess)
        e = hipMemcpyAsync(dst.as<char>() + acc * elem, src.p, n * elem,
                           hipMemcpyDeviceToDevice, s);
      return e;
    };
    const size_t nl = std::max(n_live, 1);
    for (const auto& k : pl->classes) {
      if (k.factor > W) break;  // no width to search (:657-659), here or later
      std::vector<hd::GroupDev> crg;
      WidthTable cwt{};
      size_t cprow = 1;
      for (int g = 0; g < G; ++g) {
        CallShape cs;
        const hd_error e =
            make_shapes(pl, pl->hgd[g].nsamps, pl->hgd[g].sh.nbeams, cs, k.factor);
        if (e != HD_NO_ERROR) return e;
        if (cs.skip) continue;
        hd::GroupDev d = pl->hgd[g];
        d.sh = cs.sh;
        d.sh.ppitch = ppitch;
        d.sh.spitch = spitch;
        d.bs = cs.bs;
        d.bc692 = cs.bc692;
        for (int j = 0; j < hd::kMaxWidths; ++j) d.n_w[j] = j < cs.wt.nw ? cs.wt.wi[j].n_w : 0;
        d.trial0 = (uint32_t)k.first;
        d.rows_valid = (uint32_t)k.count;
        if (crg.empty()) cwt = cs.wt;
        cprow = std::max(cprow, (size_t)cs.sh.nbeams * cs.sh.sc);
        crg.push_back(d);
      }
      if (crg.empty()) continue;
      hd::DedispGeom tile;
      std::memset(&tile, 0, sizeof(tile));
      const std::vector<int> dc(pl->adelays_h.begin() + k.first * nl,
                                pl->adelays_h.begin() + (k.first + k.count) * nl);
      if (!hd::pick_byte_tile(dc, k.count, nl, nbits, tile)) return HD_INVALID_STRIDE;
      tile.remap = 1;
      const uint8_t* xb =
          pl->x.as<uint8_t>() + (size_t)log2_pow2((unsigned)k.factor) * G * nl * xpitch;
      size_t ng = 0;
      const hd_error e = search_rows(crg, (int)k.count, cwt, cprow, xb, pl->adelays.as<int>(),
                                     &tile, crg[0].sh, ng);
      if (e != HD_NO_ERROR) return e;
      if (ng == 0) continue;
      HD_CHECK(append(pl->a_peak, pl->g_peak, 4, ng));
      HD_CHECK(append(pl->a_ind, pl->g_ind, 8, ng));
      HD_CHECK(append(pl->a_begin, pl->g_begin, 8, ng));
      HD_CHECK(append(pl->a_end, pl->g_end, 8, ng));
      HD_CHECK(append(pl->a_filter, pl->g_filter, 4, ng));
      HD_CHECK(append(pl->a_dm, pl->g_dm, 4, ng));
      HD_CHECK(append(pl->a_group, pl->g_group, 4, ng));
      acc += ng;
    }
    // the stage events of the runs are not split: the whole search counts as
    // dedispersion
    for (int ev = 2; ev <= 6; ++ev) record(pl, ev);
    n_giants = acc;
    if (n_giants > 0 && G == 1) {
      // one gulp: the runs appended in trial order are the giants
      std::swap(pl->g_peak, pl->a_peak);
      std::swap(pl->g_ind, pl->a_ind);
      std::swap(pl->g_begin, pl->a_begin);
      std::swap(pl->g_end, pl->a_end);
      std::swap(pl->g_filter, pl->a_filter);
      std::swap(pl->g_dm, pl->a_dm);
      std::swap(pl->g_group, pl->a_group);
    } else if (n_giants > 0) {
      std::vector<float> a_peak(n_giants);
      std::vector<uint64_t> a_ind(n_giants), a_begin(n_giants), a_end(n_giants);
      std::vector<uint32_t> a_filter(n_giants), a_dm(n_giants), a_group(n_giants);
      HD_CHECK(hipMemcpyAsync(a_peak.data(), pl->a_peak.p, n_giants * 4, hipMemcpyDeviceToHost, s));
      HD_CHECK(hipMemcpyAsync(a_ind.data(), pl->a_ind.p, n_giants * 8, hipMemcpyDeviceToHost, s));
      HD_CHECK(hipMemcpyAsync(a_begin.data(), pl->a_begin.p, n_giants * 8, hipMemcpyDeviceToHost, s));
      HD_CHECK(hipMemcpyAsync(a_end.data(), pl->a_end.p, n_giants * 8, hipMemcpyDeviceToHost, s));
      HD_CHECK(hipMemcpyAsync(a_filter.data(), pl->a_filter.p, n_giants * 4, hipMemcpyDeviceToHost, s));
      HD_CHECK(hipMemcpyAsync(a_dm.data(), pl->a_dm.p, n_giants * 4, hipMemcpyDeviceToHost, s));
      HD_CHECK(hipMemcpyAsync(a_group.data(), pl->a_group.p, n_giants * 4, hipMemcpyDeviceToHost, s));
      HD_CHECK(hipStreamSynchronize(s));
      // gulp by gulp (stable: each gulp's giants stay in trial order)
      std::vector<size_t> ord(n_giants);
      for (size_t i = 0; i < n_giants; ++i) ord[i] = i;
      std::stable_sort(ord.begin(), ord.end(),
                       [&](size_t a, size_t b) { return a_group[a] < a_group[b]; });
      auto permute = [&](auto& v) {
        auto c = v;
        for (size_t i = 0; i < n_giants; ++i) v[i] = c[ord[i]];
      };
      permute(a_peak); permute(a_ind); permute(a_begin); permute(a_end);
      permute(a_filter); permute(a_dm); permute(a_group);
      HD_CHECK(pl->g_peak.ensure(n_giants * sizeof(float)));
      HD_CHECK(pl->g_ind.ensure(n_giants * sizeof(uint64_t)));
      HD_CHECK(pl->g_begin.ensure(n_giants * sizeof(uint64_t)));
      HD_CHECK(pl->g_end.ensure(n_giants * sizeof(uint64_t)));
      HD_CHECK(pl->g_filter.ensure(n_giants * sizeof(uint32_t)));
      HD_CHECK(pl->g_dm.ensure(n_giants * sizeof(uint32_t)));
      HD_CHECK(pl->g_group.ensure(n_giants * sizeof(uint32_t)));
      HD_CHECK(hipMemcpyAsync(pl->g_peak.p, a_peak.data(), n_giants * 4, hipMemcpyHostToDevice, s));
      HD_CHECK(hipMemcpyAsync(pl->g_ind.p, a_ind.data(), n_giants * 8, hipMemcpyHostToDevice, s));
      HD_CHECK(hipMemcpyAsync(pl->g_begin.p, a_begin.data(), n_giants * 8, hipMemcpyHostToDevice, s));
      HD_CHECK(hipMemcpyAsync(pl->g_end.p, a_end.data(), n_giants * 8, hipMemcpyHostToDevice, s));
      HD_CHECK(hipMemcpyAsync(pl->g_filter.p, a_filter.data(), n_giants * 4, hipMemcpyHostToDevice, s));
      HD_CHECK(hipMemcpyAsync(pl->g_dm.p, a_dm.data(), n_giants * 4, hipMemcpyHostToDevice, s));
      HD_CHECK(hipMemcpyAsync(pl->g_group.p, a_group.data(), n_giants * 4, hipMemcpyHostToDevice, s));
      HD_CHECK(hipStreamSynchronize(s));
    }
  }
  pl->n_giants = n_giants;
  record(pl, 7);
  // giant-rate limit (pipeline.cu:780-787), per gulp: the running total only
  // grows, so the limit trips during the loop exactly when the final total
  // exceeds it.  A gulp that trips it has no clustering and no candidates.
  std::vector<char> excluded(G, 0);
  bool any_excluded = false;
  if (p.max_giant_rate != 0 && n_giants > 0) {
    std::vector<uint32_t> grp(n_giants);
    HD_CHECK(hipMemcpyAsync(grp.data(), pl->g_group.p, n_giants * 4, hipMemcpyDeviceToHost, s));
    HD_CHECK(hipStreamSynchronize(s));
    std::vector<size_t> cnt(G, 0);
    for (uint32_t g : grp) ++cnt[g];
    for (int g = 0; g < G; ++g) {
      const float mins = (float)((double)((float)pl->hgd[g].nsamps * p.dt) / 60.0);
      if ((float)cnt[g] / mins > p.max_giant_rate) {
        excluded[g] = 1;
        any_excluded = true;
        pl->gulps[active[g]].too_many = true;
      }
    }
  }
  size_t n_cluster = n_giants;
  if (any_excluded) {
    // keep the giants of the other gulps, in order (a host round trip: the
    // limiter is a rare, off-by-default path)
    std::vector<uint32_t> grp(n_giants), filt(n_giants), dm(n_giants);
    std::vector<float> peak(n_giants);
    std::vector<uint64_t> ind(n_giants), beg(n_giants), end(n_giants);
    HD_CHECK(hipMemcpyAsync(grp.data(), pl->g_group.p, n_giants * 4, hipMemcpyDeviceToHost, s));
    HD_CHECK(hipMemcpyAsync(filt.data(), pl->g_filter.p, n_giants * 4, hipMemcpyDeviceToHost, s));
    HD_CHECK(hipMemcpyAsync(dm.data(), pl->g_dm.p, n_giants * 4, hipMemcpyDeviceToHost, s));
    HD_CHECK(hipMemcpyAsync(peak.data(), pl->g_peak.p, n_giants * 4, hipMemcpyDeviceToHost, s));
    HD_CHECK(hipMemcpyAsync(ind.data(), pl->g_ind.p, n_giants * 8, hipMemcpyDeviceToHost, s));
    HD_CHECK(hipMemcpyAsync(beg.data(), pl->g_begin.p, n_giants * 8, hipMemcpyDeviceToHost, s));
    HD_CHECK(hipMemcpyAsync(end.data(), pl->g_end.p, n_giants * 8, hipMemcpyDeviceToHost, s));
    HD_CHECK(hipStreamSynchronize(s));
    size_t w = 0;
    for (size_t k = 0; k < n_giants; ++k) {
      if (excluded[grp[k]]) continue;
      grp[w] = grp[k]; filt[w] = filt[k]; dm[w] = dm[k]; peak[w] = peak[k];
      ind[w] = ind[k]; beg[w] = beg[k]; end[w] = end[k];
      ++w;
    }
    n_cluster = w;
    if (w) {
      pl->n_giants = w;  // the giants of the gulps that go on to clustering
      HD_CHECK(hipMemcpyAsync(pl->g_group.p, grp.data(), w * 4, hipMemcpyHostToDevice, s));
      HD_CHECK(hipMemcpyAsync(pl->g_filter.p, filt.data(), w * 4, hipMemcpyHostToDevice, s));
      HD_CHECK(hipMemcpyAsync(pl->g_dm.p, dm.data(), w * 4, hipMemcpyHostToDevice, s));
      HD_CHECK(hipMemcpyAsync(pl->g_peak.p, peak.data(), w * 4, hipMemcpyHostToDevice, s));
      HD_CHECK(hipMemcpyAsync(pl->g_ind.p, ind.data(), w * 8, hipMemcpyHostToDevice, s));
      HD_CHECK(hipMemcpyAsync(pl->g_begin.p, beg.data(), w * 8, hipMemcpyHostToDevice, s));
      HD_CHECK(hipMemcpyAsync(pl->g_end.p, end.data(), w * 8, hipMemcpyHostToDevice, s));
      HD_CHECK(hipStreamSynchronize(s));
    }
  }
  if (n_cluster == 0) {
    record(pl, 8);
    return HD_NO_ERROR;
  }
  // clustering and merging (pipeline.cu:808-887)
  const unsigned ng = (unsigned)n_cluster;
  size_t npad = 1;
  while (npad < ng) npad <<= 1;
  HD_CHECK(pl->c_keys.ensure(npad * sizeof(uint64_t)));
  // window bounds: one pair per giant, or one per filter band
  const unsigned nbands = (unsigned)std::min<size_t>(2 * p.cand_sep_filter + 1, 16);
  HD_CHECK(pl->c_lo.ensure((size_t)ng * nbands * sizeof(uint32_t)));
  HD_CHECK(pl->c_hi.ensure((size_t)ng * nbands * sizeof(uint32_t)));
  HD_CHECK(pl->c_labels.ensure(ng * sizeof(int32_t)));
  HD_CHECK(pl->c_flag.ensure(sizeof(unsigned) * 4));
  HD_CHECK(pl->c_best.ensure(ng * sizeof(uint64_t)));
  HD_CHECK(pl->c_members.ensure(ng * sizeof(unsigned)));
  HD_CHECK(pl->c_rflag.ensure(ng * sizeof(unsigned)));
  HD_CHECK(pl->c_rpos.ensure(ng * sizeof(unsigned)));
  HD_CHECK(pl->c_tmp.ensure(hd::exclusive_scan_tmp_words(ng + 1) * sizeof(unsigned)));
  HD_CHECK(pl->c_sind.ensure(ng * sizeof(uint64_t)));
  HD_CHECK(pl->c_sfd.ensure(ng * sizeof(uint32_t)));
  HD_CHECK(pl->c_ecount.ensure((ng + 1) * sizeof(uint32_t)));
  HD_CHECK(pl->c_eoff.ensure((ng + 1) * sizeof(uint32_t)));
  hd::ClusterBuffers cb{pl->c_keys.as<uint64_t>(), pl->c_lo.as<uint32_t>(), pl->c_hi.as<uint32_t>(),
                        pl->c_labels.as<int32_t>(), pl->c_flag.as<unsigned>(),
                        pl->c_best.as<uint64_t>(), pl->c_members.as<unsigned>(),
                        pl->c_rflag.as<unsigned>(), pl->c_rpos.as<unsigned>(),
                        pl->c_tmp.as<unsigned>(), npad, pl->c_sind.as<uint64_t>(),
                        pl->c_sfd.as<uint32_t>(), nbands};
  cb.ecount = pl->c_ecount.as<uint32_t>();
  cb.eoff = pl->c_eoff.as<uint32_t>();
  cb.adj_alloc = [pl](size_t entries) -> uint32_t* {
    if (pl->c_adj.ensure(std::max<size_t>(entries, 1) * sizeof(uint32_t)) != hipSuccess)
      return nullptr;
    return pl->c_adj.as<uint32_t>();
  };
  // the giants' groups are the gulps, in index order (giants are kept in
  // (gulp, trial, width, time) order)
  const unsigned n_cg = (unsigned)pl->hgd.size();
  HD_CHECK(pl->c_gb.ensure((2 * (size_t)n_cg + 3) * sizeof(uint32_t)));
  cb.gbounds = pl->c_gb.as<uint32_t>();
  cb.n_cgroups = n_cg;
  const unsigned max_filter = wt.wi[wt.nw - 1].filter_idx;
  int iters = 0;
  HD_CHECK(hd::launch_cluster(pl->g_ind.as<uint64_t>(), pl->g_filter.as<uint32_t>(),
                              pl->g_dm.as<uint32_t>(), nullptr, ng, p.cand_sep_time,
                              p.cand_sep_filter, p.cand_sep_dm, pl->hgd[0].nsamps_beam,
                              max_filter, cb, s, pl->h_pinned, &iters,
                              pl->g_group.as<uint32_t>(), ga.gd, t_off));
  HD_CHECK(pl->o_peak.ensure(ng * sizeof(float)));
  HD_CHECK(pl->o_ind.ensure(ng * sizeof(uint64_t)));
  HD_CHECK(pl->o_begin.ensure(ng * sizeof(uint64_t)));
  HD_CHECK(pl->o_end.ensure(ng * sizeof(uint64_t)));
  HD_CHECK(pl->o_filter.ensure(ng * sizeof(uint64_t)));
  HD_CHECK(pl->o_dm.ensure(ng * sizeof(uint64_t)));
  HD_CHECK(pl->o_members.ensure(ng * sizeof(uint64_t)));
  HD_CHECK(pl->o_group.ensure(ng * sizeof(uint32_t)));
  HD_CHECK(hd::launch_merge(pl->g_peak.as<float>(), pl->g_ind.as<uint64_t>(),
                            pl->g_begin.as<uint64_t>(), pl->g_end.as<uint64_t>(),
                            pl->g_filter.as<uint32_t>(), pl->g_dm.as<uint32_t>(), nullptr, ng, cb,
                            pl->o_peak.as<float>(), pl->o_ind.as<uint64_t>(),
                            pl->o_begin.as<uint64_t>(), pl->o_end.as<uint64_t>(),
                            pl->o_filter.as<uint64_t>(), pl->o_dm.as<uint64_t>(),
                            pl->o_members.as<uint64_t>(), s, pl->g_group.as<uint32_t>(),
                            pl->o_group.as<uint32_t>()));
  HD_CHECK(hipMemcpyAsync(pl->h_pinned, cb.root_pos + (ng - 1), sizeof(unsigned),
                          hipMemcpyDeviceToHost, s));
  HD_CHECK(hipMemcpyAsync(pl->h_pinned + 1, cb.root_flag + (ng - 1), sizeof(unsigned),
                          hipMemcpyDeviceToHost, s));
  HD_CHECK(hipStreamSynchronize(s));
  const size_t n_groups = (size_t)pl->h_pinned[0] + pl->h_pinned[1];
  pl->r_peak.resize(n_groups);
  pl->r_ind.resize(n_groups);
  pl->r_begin.resize(n_groups);
  pl->r_end.resize(n_groups);
  pl->r_filter.resize(n_groups);
  pl->r_dm.resize(n_groups);
  pl->r_members.resize(n_groups);
  pl->r_dms.resize(n_groups);
  pl->r_group.resize(n_groups);
  static_assert(sizeof(hd_size) == sizeof(uint64_t), "hd_size must be 64-bit");
  HD_CHECK(hipMemcpyAsync(pl->r_peak.data(), pl->o_peak.p, n_groups * 4, hipMemcpyDeviceToHost, s));
  HD_CHECK(hipMemcpyAsync(pl->r_ind.data(), pl->o_ind.p, n_groups * 8, hipMemcpyDeviceToHost, s));
  HD_CHECK(hipMemcpyAsync(pl->r_begin.data(), pl->o_begin.p, n_groups * 8, hipMemcpyDeviceToHost, s));
  HD_CHECK(hipMemcpyAsync(pl->r_end.data(), pl->o_end.p, n_groups * 8, hipMemcpyDeviceToHost, s));
  HD_CHECK(hipMemcpyAsync(pl->r_filter.data(), pl->o_filter.p, n_groups * 8, hipMemcpyDeviceToHost, s));
  HD_CHECK(hipMemcpyAsync(pl->r_dm.data(), pl->o_dm.p, n_groups * 8, hipMemcpyDeviceToHost, s));
  HD_CHECK(hipMemcpyAsync(pl->r_members.data(), pl->o_members.p, n_groups * 8, hipMemcpyDeviceToHost, s));
  HD_CHECK(hipMemcpyAsync(pl->r_group.data(), pl->o_group.p, n_groups * 4, hipMemcpyDeviceToHost, s));
  record(pl, 8);
  HD_CHECK(hipStreamSynchronize(s));
  for (size_t g2 = 0; g2 < n_groups; ++g2) pl->r_dms[g2] = pl->dm_list[pl->r_dm[g2]];  // :872-875
  // candidates come out in root order, so each gulp's are one run
  for (size_t k = 0; k < n_groups; ++k) {
    GulpOut& go = pl->gulps[active[pl->r_group[k]]];
    if (go.hi == 0) go.lo = k;
    go.hi = k + 1;
  }
  return HD_NO_ERROR;
}

// Timings, then each gulp's output in order; returns the first gulp's status
// (the call's status for a single gulp).
hd_error finish_call(hd_pipeline pl, hd_error err) {
  if (err != HD_NO_ERROR) return err;
  (void)hipEventSynchronize(pl->ev[8]);
  static const int stage_ev[ST_COUNT][2] = {{0, 10}, {1, 2}, {2, 3}, {3, 4}, {4, 5}, {5, 6},
                                            {6, 7}, {7, 8}, {9, 8}, {9, 0}, {10, 1}};
  for (int k = 0; k < ST_COUNT; ++k) {
    float ms = 0.f;
    if (hipEventElapsedTime(&ms, pl->ev[stage_ev[k][0]], pl->ev[stage_ev[k][1]]) == hipSuccess)
      pl->timings[k] = ms;
  }
  hd_error first = HD_NO_ERROR;
  for (size_t k = 0; k < pl->gulps.size(); ++k) {
    GulpOut& go = pl->gulps[k];
    if (go.err != HD_NO_ERROR) {
      if (k == 0) first = go.err;
      continue;
    }
    if (go.too_many) go.lo = go.hi = 0;  // no clustering, empty output (pipeline.cu:808)
    hd_error e = emit_output(pl, go);
    if (e != HD_NO_ERROR) return e;
    if (go.too_many) go.err = HD_TOO_MANY_EVENTS;
    if (k == 0) first = go.err;
    if (pl->output_enabled) {
      char buffer[64];
      const hd_params& p = pl->params;
      const time_t off = p.spectra_per_second ? (time_t)(go.first_idx / p.spectra_per_second) : 0;
      time_t now = p.utc_start + off;
      std::strftime(buffer, 64, "%Y-%m-%d-%H:%M:%S", std::gmtime(&now));
      std::cout << "Dump: " << buffer << " Total Time: " << pl->timings[ST_TOTAL] * 1e-3
                << std::endl;  // pipeline.cu:1081
    }
  }
  if (pl->params.verbosity >= 1 && pl->output_enabled) {
    const double* t = pl->timings;
    std::cout << "0-DM cleaning time:      " << t[ST_CLEAN] * 1e-3 << "\n"
              << "Unpack time:             " << t[ST_UNPACK] * 1e-3 << "\n"
              << "Proposal time:           " << t[ST_PROPOSAL] * 1e-3 << "\n"
              << "Dedispersion time:       " << t[ST_DEDISP] * 1e-3 << "\n"
              << "Baselining time:         " << t[ST_BASELINE] * 1e-3 << "\n"
              << "Normalisation time:      " << t[ST_RMS] * 1e-3 << "\n"
              << "Filtering time:          " << (t[ST_SCAN] + t[ST_BLOCKRMS]) * 1e-3 << "\n"
              << "Find giants time:        " << t[ST_GIANTS] * 1e-3 << "\n"
              << "Process candidates time: " << t[ST_CLUSTER] * 1e-3 << "\n"
              << "Total time:              " << t[ST_TOTAL] * 1e-3 << std::endl;
  }
  return first;
}

}  // namespace

// ---------------------------------------------------------------------------
extern "C" hd_error hd_create_pipeline(hd_pipeline* pipeline_, hd_params params) {
  if (!pipeline_) return HD_INVALID_POINTER;
  *pipeline_ = nullptr;
  if (params.nchans == 0 || !(params.dt > 0)) return HD_INVALID_STRIDE;
  if (!(params.df < 0)) return HD_INVALID_STRIDE;  // delays assume channel 0 is the top
  if (params.dm_nbits != 32) return HD_INVALID_NBITS;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return HD_INVALID_DEVICE_INDEX;
  if (params.gpu_id < 0 || params.gpu_id >= ndev) return HD_INVALID_DEVICE_INDEX;
  hd_pipeline_t* pl = new (std::nothrow) hd_pipeline_t();
  if (!pl) return HD_MEM_ALLOC_FAILED;
  pl->params = params;
  if (params.output_dir) pl->output_dir = params.output_dir;
  else pl->output_dir = ".";
  if (params.coincidencer_host) pl->coinc_host = params.coincidencer_host;
  if (params.num_channel_zaps && params.channel_zaps)
    pl->zaps.assign(params.channel_zaps, params.channel_zaps + params.num_channel_zaps);
  pl->params.output_dir = nullptr;
  pl->params.coincidencer_host = nullptr;
  pl->params.channel_zaps = nullptr;
  pl->params.sigproc_file = nullptr;
  pl->device = params.gpu_id;
  {
    const char* g = getenv("HD_GRAPH");
    pl->use_graph = g ? atoi(g) != 0 : true;
    const char* db = getenv("HD_DEDISP_DBUF");
    pl->use_dbuf = db ? atoi(db) != 0 : true;
  }
  DeviceGuard dg(pl->device);
  if (params.verbosity >= 1)
    std::cout << "Process " << params.beam << " using GPU " << params.gpu_id << std::endl;
  if (params.use_scrunching && params.verbosity >= 2)
    std::cout << "Enabling adaptive dt" << std::endl;  // pipeline.cu:187-196
  hipError_t e = hipStreamCreateWithFlags(&pl->stream, hipStreamNonBlocking);
  if (e == hipSuccess) e = hipHostMalloc((void**)&pl->h_pinned, 64, hipHostMallocDefault);
  for (int k = 0; k <= ST_COUNT && e == hipSuccess; ++k) e = hipEventCreate(&pl->ev[k]);
  if (e != hipSuccess) {
    hd_destroy_pipeline(pl);
    return gpu_err(e);
  }
  pl->dm_list = hdplan::dm_list(params.dm_min, params.dm_max, params.dt, params.dm_pulse_width,
                                params.f0, params.df, params.nchans, params.dm_tol);
  pl->kc = hdplan::delay_slopes(params.nchans, params.dt, params.f0, params.df);
  pl->max_delay = 0;
  for (size_t c = 0; c < params.nchans; ++c) {
    const int d = hdplan::delay(pl->dm_list.back(), pl->kc[c]);
    if (d > 0 && (size_t)d > pl->max_delay) pl->max_delay = (size_t)d;
  }
  // scrunch factors (dedisp_enable_adaptive_dt with dm_pulse_width and
  // scrunch_tol, pipeline.cu:187-196, 344-360)
  if (params.use_scrunching)
    pl->factors = hdplan::scrunch_list(pl->dm_list, params.dt, params.dm_pulse_width, params.f0,
                                       params.df, params.nchans, params.scrunch_tol);
  else
    pl->factors.assign(pl->dm_list.size(), 1);
  for (size_t a = 0; a < pl->factors.size();) {
    size_t b = a;
    while (b < pl->factors.size() && pl->factors[b] == pl->factors[a]) ++b;
    pl->classes.push_back({a, b - a, pl->factors[a]});
    a = b;
  }
  pl->adaptive = pl->classes.size() > 1;
  if (params.verbosity >= 2) {
    std::cout << "Scrunch factors:" << std::endl;
    for (size_t f : pl->factors) std::cout << f << " ";
    std::cout << std::endl;
  }
  *pipeline_ = pl;
  return HD_NO_ERROR;
}

extern "C" hd_size hd_get_max_overlap(hd_pipeline pl, hd_size nbeams) {
  if (!pl) return 0;
  return (pl->params.boxcar_max + pl->max_delay) * nbeams;  // pipeline.cu:214-219
}

extern "C" hd_error hd_preallocate(hd_pipeline pl, hd_size nsamps, hd_size nbeams) {
  if (!pl) return HD_INVALID_PIPELINE;
  if (nbeams == 0) return HD_INVALID_STRIDE;
  DeviceGuard dg(pl->device);
  const size_t max_alloc = nsamps + hd_get_max_overlap(pl, nbeams);
  const size_t n_dm = pl->dm_list.size();
  const size_t ppitch = round_up(max_alloc + 1, (size_t)hd::dedisp_tile_time());
  HD_CHECK(pl->fb.ensure(max_alloc * pl->params.nchans));
  HD_CHECK(pl->plane.ensure(n_dm * ppitch * sizeof(float)));
  HD_CHECK(pl->S.ensure((n_dm * round_up(max_alloc + 1, 64) + 8) * sizeof(float)));
  return HD_NO_ERROR;
}

extern "C" hd_error hd_execute(hd_pipeline pl, const hd_byte* h_filterbank, hd_size nsamps,
                               hd_size nbits, hd_size first_idx, hd_size nbeams,
                               hd_size* nsamps_processed) {
  if (!pl) return HD_INVALID_PIPELINE;
  if (!h_filterbank || !nsamps_processed) return HD_INVALID_POINTER;
  if (!(nbits == 1 || nbits == 2 || nbits == 4 || nbits == 8)) return HD_INVALID_NBITS;
  if (nbeams == 0 || nsamps < nbeams) return HD_INVALID_STRIDE;
  DeviceGuard dg(pl->device);
  const size_t bytes = nsamps * pl->params.nchans * nbits / 8;
  HD_CHECK(pl->fb.ensure(bytes));
  HD_CHECK(hipMemcpyAsync(pl->fb.p, h_filterbank, bytes, hipMemcpyHostToDevice, pl->stream));
  const std::vector<GulpIn> in{{pl->fb.as<hd_byte>(), nsamps, nbeams, first_idx, pl->params.beam}};
  hd_error err = run_gulps(pl, in, nbits);
  *nsamps_processed = pl->gulps.empty() ? 0 : pl->gulps[0].processed;
  return finish_call(pl, err);
}

extern "C" hd_error hd_execute_device(hd_pipeline pl, const hd_byte* d_filterbank,
                                      hd_size nsamps, hd_size nbits, hd_size first_idx,
                                      hd_size nbeams, hd_size* nsamps_processed) {
  if (!pl) return HD_INVALID_PIPELINE;
  if (!d_filterbank || !nsamps_processed) return HD_INVALID_POINTER;
  if (!(nbits == 1 || nbits == 2 || nbits == 4 || nbits == 8)) return HD_INVALID_NBITS;
  if (nbeams == 0 || nsamps < nbeams) return HD_INVALID_STRIDE;
  DeviceGuard dg(pl->device);
  const std::vector<GulpIn> in{{d_filterbank, nsamps, nbeams, first_idx, pl->params.beam}};
  hd_error err = run_gulps(pl, in, nbits);
  *nsamps_processed = pl->gulps.empty() ? 0 : pl->gulps[0].processed;
  return finish_call(pl, err);
}

extern "C" hd_error hd_execute_batch(hd_pipeline pl, const hd_gulp_desc* gulps, hd_size count,
                                     hd_size nbits, int on_device, hd_size* nsamps_processed,
                                     hd_error* errors) {
  if (!pl) return HD_INVALID_PIPELINE;
  if (!gulps || !nsamps_processed) return HD_INVALID_POINTER;
  if (!(nbits == 1 || nbits == 2 || nbits == 4 || nbits == 8)) return HD_INVALID_NBITS;
  for (hd_size k = 0; k < count; ++k) {
    if (!gulps[k].filterbank) return HD_INVALID_POINTER;
    if (gulps[k].nbeams == 0 || gulps[k].nsamps < gulps[k].nbeams) return HD_INVALID_STRIDE;
  }
  DeviceGuard dg(pl->device);
  const size_t row = pl->params.nchans * nbits / 8;
  std::vector<GulpIn> in(count);
  if (!on_device) {
    // one staging buffer for all gulps' host bytes
    size_t total = 0;
    for (hd_size k = 0; k < count; ++k) total += gulps[k].nsamps * row;
    HD_CHECK(pl->fb.ensure(total));
    size_t off = 0;
    for (hd_size k = 0; k < count; ++k) {
      const size_t b = gulps[k].nsamps * row;
      HD_CHECK(hipMemcpyAsync(pl->fb.as<hd_byte>() + off, gulps[k].filterbank, b,
                              hipMemcpyHostToDevice, pl->stream));
      in[k] = {pl->fb.as<hd_byte>() + off, gulps[k].nsamps, gulps[k].nbeams, gulps[k].first_idx,
               gulps[k].beam};
      off += b;
    }
  } else {
    for (hd_size k = 0; k < count; ++k)
      in[k] = {gulps[k].filterbank, gulps[k].nsamps, gulps[k].nbeams, gulps[k].first_idx,
               gulps[k].beam};
  }
  const hd_error err = run_gulps(pl, in, nbits);
  // finish_call writes every gulp's output and sets its status in pl->gulps
  if (err == HD_NO_ERROR) (void)finish_call(pl, HD_NO_ERROR);
  hd_error first = err;
  for (hd_size k = 0; k < count; ++k) {
    const GulpOut go = pl->gulps.size() == count ? pl->gulps[k] : GulpOut{};
    nsamps_processed[k] = go.processed;
    if (errors) errors[k] = err != HD_NO_ERROR ? err : go.err;
    if (first == HD_NO_ERROR && go.err != HD_NO_ERROR) first = go.err;
  }
  return first;
}

// dedisp_execute_adv(plan, nsamps, in, in_nbits, in_stride, out, out_nbits,
// out_stride, flags) as called at Pipeline/pipeline.cu:452-462, with the
// pipeline as the plan: its trial list, delay table and killmask.  The
// unpack and dedispersion kernels are the search's own (same tile choice).
extern "C" hd_error hd_dedisperse(hd_pipeline pl, hd_size nsamps, const hd_byte* d_in,
                                  hd_size in_nbits, hd_size in_stride, hd_byte* d_out,
                                  hd_size out_nbits, hd_size out_stride, unsigned flags) {
  (void)flags;  // pointers are always device memory here
  if (!pl) return HD_INVALID_PIPELINE;
  if (!d_in || !d_out) return HD_INVALID_POINTER;
  if (!(in_nbits == 1 || in_nbits == 2 || in_nbits == 4 || in_nbits == 8)) return HD_INVALID_NBITS;
  if (out_nbits != 32) return HD_INVALID_NBITS;  // float output only (pipeline.cu:544-550)
  const hd_params& p = pl->params;
  if (in_stride != p.nchans * in_nbits / 8 || (p.nchans * in_nbits) % 8) return HD_INVALID_STRIDE;
  if (nsamps <= pl->max_delay) return HD_TOO_FEW_NSAMPS;
  const size_t nout = nsamps - pl->max_delay;
  if (out_stride < nout * sizeof(float) || out_stride % sizeof(float)) return HD_INVALID_STRIDE;
  DeviceGuard dgd(pl->device);
  (void)hipGetLastError();
  hd_error err = build_live_and_delays(pl);
  if (err != HD_NO_ERROR) return err;
  hipStream_t s = pl->stream;
  const int n_dm = (int)pl->dm_list.size();
  const int n_live = (int)pl->live_list.size();
  const size_t ppitch = round_up(nout + 1, (size_t)hd::dedisp_tile_time());
  const size_t xpitch = round_up(ppitch + pl->max_delay + 2 * pl->wlen + 256, 256);
  HD_CHECK(pl->x.ensure(std::max(n_live, 1) * xpitch));
  HD_CHECK(pl->plane.ensure((size_t)n_dm * ppitch * sizeof(float)));
  if (n_live > 0)
    HD_CHECK(hd::launch_unpack(d_in, in_stride, (int)in_nbits, nsamps, pl->live.as<int>(), n_live,
                               pl->x.as<uint8_t>(), xpitch, s, hd::GroupArgs{},
                               (size_t)n_live == p.nchans));
  else
    HD_CHECK(hipMemsetAsync(pl->x.p, 0, xpitch, s));
  hd::DedispGeom dg;
  std::memset(&dg, 0, sizeof(dg));
  dg.n_dm = n_dm;
  dg.n_live = std::max(n_live, 1);  // no live channel: one zero row sums to 0
  dg.nout = nout;
  dg.xpitch = xpitch;
  dg.ppitch = ppitch;
  dg.wlen = (int)pl->wlen;
  const int* pairs = nullptr;
  if (n_live == 0) {
    HD_CHECK(hipMemsetAsync(pl->plane.p, 0, (size_t)n_dm * ppitch * sizeof(float), s));
  } else {
    err = choose_dedisp_tile(pl, dg, hd::GroupArgs{}, in_nbits, s, &pairs);
    if (err != HD_NO_ERROR) return err;
    HD_CHECK(hd::launch_dedisperse(pl->x.as<uint8_t>(), pl->delays.as<int>(), pairs, dg,
                                   pl->plane.as<float>(), s));
  }
  HD_CHECK(hipMemcpy2DAsync(d_out, out_stride, pl->plane.p, ppitch * sizeof(float),
                            nout * sizeof(float), (size_t)n_dm, hipMemcpyDeviceToDevice, s));
  HD_CHECK(hipStreamSynchronize(s));
  return HD_NO_ERROR;
}

extern "C" hd_size hd_get_last_candidate_gulps(hd_pipeline pl, hd_size* gulp_index, hd_size cap) {
  if (!pl) return 0;
  const size_t n = pl->r_peak.size();
  if (gulp_index) {
    // candidates are in gulp order; each gulp holds the run [lo, hi)
    for (size_t g = 0; g < pl->gulps.size(); ++g)
      for (size_t k = pl->gulps[g].lo; k < pl->gulps[g].hi && k < cap; ++k) gulp_index[k] = g;
  }
  return n;
}

extern "C" void hd_destroy_pipeline(hd_pipeline pl) {
  if (!pl) return;
  {
    DeviceGuard dg(pl->device);
    if (pl->stream) (void)hipStreamSynchronize(pl->stream);
    DevBuf* bufs[] = {&pl->clean_out, &pl->clean_work, &pl->fb, &pl->x, &pl->live, &pl->delays, &pl->plane, &pl->pts, &pl->rms, &pl->brms,
                      &pl->S, &pl->partial, &pl->wrms, &pl->winv, &pl->skey, &pl->sval,
                      &pl->counter, &pl->ccount, &pl->coff, &pl->cpos, &pl->c_gb,
                      &pl->a_peak, &pl->a_ind, &pl->a_begin, &pl->a_end, &pl->a_filter,
                      &pl->a_dm, &pl->a_group,
                      &pl->okey, &pl->oval,
                      &pl->head, &pl->segid, &pl->scan_tmp, &pl->g_peak, &pl->g_ind,
                      &pl->g_begin, &pl->g_end, &pl->g_filter, &pl->g_dm, &pl->g_best, &pl->c_keys,
                      &pl->c_lo, &pl->c_hi, &pl->c_labels, &pl->c_flag, &pl->c_best,
                      &pl->c_members, &pl->c_rflag, &pl->c_rpos, &pl->c_tmp, &pl->o_peak,
                      &pl->o_ind, &pl->o_begin, &pl->o_end, &pl->o_filter, &pl->o_dm,
                      &pl->o_members, &pl->gdev, &pl->g_group, &pl->o_group, &pl->c_sind,
                      &pl->c_sfd, &pl->c_ecount, &pl->c_eoff, &pl->c_adj, &pl->rdev, &pl->co_delays, &pl->co_pairs, &pl->co_xc,
                      &pl->co_plane, &pl->co_pts, &pl->co_rms, &pl->co_brms, &pl->co_bf, &pl->co_bw,
                      &pl->co_gdev, &pl->co_tiles, &pl->co_scores, &pl->co_toff, &pl->co_ntpos,
                      &pl->co_ncols, &pl->co_flags, &pl->co_pos, &pl->adelays};
    for (DevBuf* b : bufs) b->release();
    for (DevBuf& b : pl->pairs_db) b.release();
    if (pl->gexec) (void)hipGraphExecDestroy(pl->gexec);
    for (hipEvent_t& e : pl->tev)
      if (e) (void)hipEventDestroy(e);
    if (pl->net) hd_cnn_destroy(pl->net);
    for (hipEvent_t& ev : pl->ev)
      if (ev) (void)hipEventDestroy(ev);
    if (pl->h_pinned) (void)hipHostFree(pl->h_pinned);
    if (pl->stream) (void)hipStreamDestroy(pl->stream);
  }
  delete pl;
}

// ---------------------------------------------------------------------------
// accessors (hd/mi355x.h)
namespace {
template <typename T>
void copy_out(T* dst, const std::vector<T>& src, size_t n) {
  if (dst) std::memcpy(dst, src.data(), n * sizeof(T));
}
}  // namespace

extern "C" hd_size hd_get_last_candidates(hd_pipeline pl, hd_size cap, hd_float* peaks,
                                          hd_size* inds, hd_size* begins, hd_size* ends,
                                          hd_size* filter_inds, hd_size* dm_inds,
                                          hd_size* members, hd_float* dms) {
  if (!pl) return 0;
  const size_t n = pl->r_peak.size();
  const size_t k = std::min<size_t>(n, cap);
  copy_out(peaks, pl->r_peak, k);
  copy_out(inds, pl->r_ind, k);
  copy_out(begins, pl->r_begin, k);
  copy_out(ends, pl->r_end, k);
  copy_out(filter_inds, pl->r_filter, k);
  copy_out(dm_inds, pl->r_dm, k);
  copy_out(members, pl->r_members, k);
  copy_out(dms, pl->r_dms, k);
  return n;
}

extern "C" hd_size hd_get_last_giants(hd_pipeline pl, hd_size cap, hd_float* peaks,
                                      hd_size* inds, hd_size* begins, hd_size* ends,
                                      hd_size* filter_inds, hd_size* dm_inds) {
  if (!pl) return 0;
  const size_t n = pl->n_giants;
  const size_t k = std::min<size_t>(n, cap);
  if (k == 0) return n;
  DeviceGuard dg(pl->device);
  std::vector<uint32_t> tmp(k);
  if (peaks) (void)hipMemcpy(peaks, pl->g_peak.p, k * sizeof(float), hipMemcpyDeviceToHost);
  if (inds) (void)hipMemcpy(inds, pl->g_ind.p, k * 8, hipMemcpyDeviceToHost);
  if (begins) (void)hipMemcpy(begins, pl->g_begin.p, k * 8, hipMemcpyDeviceToHost);
  if (ends) (void)hipMemcpy(ends, pl->g_end.p, k * 8, hipMemcpyDeviceToHost);
  if (filter_inds) {
    (void)hipMemcpy(tmp.data(), pl->g_filter.p, k * 4, hipMemcpyDeviceToHost);
    for (size_t i = 0; i < k; ++i) filter_inds[i] = tmp[i];
  }
  if (dm_inds) {
    (void)hipMemcpy(tmp.data(), pl->g_dm.p, k * 4, hipMemcpyDeviceToHost);
    for (size_t i = 0; i < k; ++i) dm_inds[i] = tmp[i];
  }
  return n;
}

extern "C" hd_size hd_get_last_cand_text(hd_pipeline pl, char* buf, hd_size cap) {
  if (!pl) return 0;
  const size_t n = pl->cand_text.size();
  if (buf && cap) {
    const size_t k = std::min<size_t>(n, cap - 1);
    std::memcpy(buf, pl->cand_text.data(), k);
    buf[k] = 0;
  }
  return n;
}

extern "C" hd_size hd_get_dm_count(hd_pipeline pl) { return pl ? pl->dm_list.size() : 0; }
extern "C" hd_size hd_get_max_delay(hd_pipeline pl) { return pl ? pl->max_delay : 0; }
extern "C" hd_size hd_get_dm_list(hd_pipeline pl, hd_float* out, hd_size cap) {
  if (!pl) return 0;
  const size_t n = pl->dm_list.size();
  if (out) std::memcpy(out, pl->dm_list.data(), std::min<size_t>(n, cap) * sizeof(float));
  return n;
}
extern "C" hd_size hd_get_last_timings(hd_pipeline pl, double* ms, hd_size cap) {
  if (!pl || !ms) return 0;
  const size_t k = std::min<size_t>(ST_COUNT, cap);
  for (size_t i = 0; i < k; ++i) ms[i] = pl->timings[i];
  return k;
}
extern "C" void hd_set_output_enabled(hd_pipeline pl, int enabled) {
  if (pl) pl->output_enabled = enabled != 0;
}
extern "C" void hd_set_cleaning(hd_pipeline pl, int enabled) {
  if (pl) pl->cleaning = enabled != 0;
}

// ---------------------------------------------------------------------------
// learned proposal (hd/proposal.h)

extern "C" hd_error hd_set_proposal(hd_pipeline pl, const hd_float* params, hd_float threshold) {
  if (!pl) return HD_INVALID_PIPELINE;
  if (!params) {
    pl->gating = false;
    return HD_NO_ERROR;
  }
  DeviceGuard dg(pl->device);
  if (!pl->net) {
    const hd_error e = hd_cnn_create(&pl->net, pl->device, 4096);
    if (e != HD_NO_ERROR) return e;
  }
  const hd_error e = hd_cnn_set_params(pl->net, params);
  if (e != HD_NO_ERROR) return e;
  pl->prop_thresh = threshold;
  pl->gating = true;
  return HD_NO_ERROR;
}

extern "C" hd_error hd_proposal_coarse(hd_pipeline pl, const hd_byte* d_filterbank,
                                       hd_size nsamps, hd_size nbits, hd_size* n_units,
                                       hd_size* n_tpos) {
  if (!pl) return HD_INVALID_PIPELINE;
  if (!d_filterbank) return HD_INVALID_POINTER;
  if (!(nbits == 1 || nbits == 2 || nbits == 4 || nbits == 8)) return HD_INVALID_NBITS;
  DeviceGuard dg(pl->device);
  const std::vector<GulpIn> in{{d_filterbank, nsamps, 1, 0, pl->params.beam}};
  pl->coarse_only = true;
  hd_error err = run_gulps(pl, in, nbits);
  pl->coarse_only = false;
  if (err == HD_NO_ERROR && !pl->gulps.empty()) err = pl->gulps[0].err;
  if (err == HD_NO_ERROR && pl->co_G != 1) err = HD_TOO_FEW_NSAMPS;
  if (err != HD_NO_ERROR) {
    pl->co_G = 0;
    return err;
  }
  HD_CHECK(hipStreamSynchronize(pl->stream));
  if (n_units) *n_units = (hd_size)pl->co_units;
  if (n_tpos) *n_tpos = (hd_size)pl->co_ntpos_h[0];
  return HD_NO_ERROR;
}

extern "C" hd_error hd_proposal_extract(hd_pipeline pl, const unsigned* pos, hd_size n,
                                        unsigned short* d_tiles) {
  if (!pl) return HD_INVALID_PIPELINE;
  if ((!pos || !d_tiles) && n) return HD_INVALID_POINTER;
  if (pl->co_G == 0) return HD_INVALID_POINTER;
  if (n == 0) return HD_NO_ERROR;
  for (hd_size i = 0; i < n; ++i)
    if (pos[2 * i] >= (unsigned)pl->co_units || pos[2 * i + 1] >= pl->co_ntpos_h[0])
      return HD_INVALID_STRIDE;
  DeviceGuard dg(pl->device);
  hipStream_t s = pl->stream;
  HD_CHECK(pl->co_pos.ensure(2 * n * sizeof(uint32_t)));
  HD_CHECK(hipMemcpyAsync(pl->co_pos.p, pos, 2 * n * sizeof(uint32_t), hipMemcpyHostToDevice, s));
  HD_CHECK(hd::launch_extract_tiles(tile_grid(pl), pl->co_pos.as<uint32_t>(), 0, n, d_tiles, s));
  HD_CHECK(hipStreamSynchronize(s));
  return HD_NO_ERROR;
}

extern "C" hd_size hd_proposal_get_plane(hd_pipeline pl, int channel, unsigned short* host,
                                         hd_size cap, hd_size* rows, hd_size* cols) {
  if (!pl || pl->co_G == 0 || channel < 0 || channel > 1) return 0;
  const size_t r = (size_t)pl->co_rows;
  const size_t c = channel ? pl->co_ncols_h[0] / 4 : pl->co_ncols_h[0];
  const size_t pitch = channel ? pl->co_wpitch : pl->co_bpitch;
  const void* src = channel ? pl->co_bw.p : pl->co_bf.p;
  if (rows) *rows = r;
  if (cols) *cols = c;
  if (host && cap >= r * c) {
    DeviceGuard dg(pl->device);
    if (hipMemcpy2D(host, c * sizeof(uint16_t), src, pitch * sizeof(uint16_t),
                    c * sizeof(uint16_t), r, hipMemcpyDeviceToHost) != hipSuccess)
      return 0;
  }
  return r * c;
}

extern "C" hd_size hd_proposal_scores(hd_pipeline pl, hd_float* host, hd_size cap) {
  if (!pl || !pl->net || pl->co_G == 0) return 0;
  DeviceGuard dg(pl->device);
  if (score_tiles(pl, pl->stream) != HD_NO_ERROR) return 0;
  const size_t n = pl->co_toff_h.back();
  if (host && cap >= n) {
    if (hipMemcpyAsync(host, pl->co_scores.p, n * sizeof(float), hipMemcpyDeviceToHost,
                       pl->stream) != hipSuccess)
      return 0;
  }
  if (hipStreamSynchronize(pl->stream) != hipSuccess) return 0;
  return n;
}

extern "C" hd_size hd_get_last_proposal(hd_pipeline pl, unsigned char* flags, hd_size cap,
                                        hd_size* units_per_gulp) {
  if (!pl) return 0;
  if (units_per_gulp) *units_per_gulp = (hd_size)pl->prop_units;
  const size_t n = pl->prop_flags.size();
  if (flags) std::memcpy(flags, pl->prop_flags.data(), std::min<size_t>(n, cap));
  return n;
}

extern "C" void hd_get_dedisp_tile(hd_pipeline pl, int* trials_per_block, int* xcd_remap,
                                   int* byte_lanes, int* channels_per_step, int* block_threads) {
  if (!pl) return;
  if (trials_per_block) *trials_per_block = pl->last_db;
  if (xcd_remap) *xcd_remap = pl->last_remap;
  if (byte_lanes) *byte_lanes = pl->last_lanes8;
  if (channels_per_step) *channels_per_step = pl->last_cc;
  if (block_threads) *block_threads = pl->last_wide ? 1024 : 512;
}

extern "C" hd_error hd_set_dedisp_tile(hd_pipeline pl, int trials_per_block, int xcd_remap,
                                      int byte_lanes, int channels_per_step, int block_threads) {
  if (!pl) return HD_INVALID_PIPELINE;
  if (trials_per_block != 0 && trials_per_block != 8 && trials_per_block != 16 &&
      trials_per_block != 32)
    return HD_INVALID_STRIDE;
  if (channels_per_step != 4 && channels_per_step != 8 && channels_per_step != 16)
    return HD_INVALID_STRIDE;
  if (block_threads != 0 && block_threads != 512 && block_threads != 1024) return HD_INVALID_STRIDE;
  pl->fixed_db = trials_per_block;
  pl->fixed_remap = xcd_remap ? 1 : 0;
  pl->fixed_lanes8 = byte_lanes ? 1 : 0;
  pl->fixed_cc = channels_per_step;
  pl->fixed_wide = block_threads == 1024 ? 1 : 0;
  return HD_NO_ERROR;
}

extern "C" void hd_set_graph(hd_pipeline pl, int enabled) {
  if (!pl) return;
  pl->use_graph = enabled != 0;
  if (!pl->use_graph && pl->gexec) {
    DeviceGuard dg(pl->device);
    (void)hipStreamSynchronize(pl->stream);
    (void)hipGraphExecDestroy(pl->gexec);
    pl->gexec = nullptr;
    pl->gkey.clear();
  }
}

extern "C" hd_size hd_get_graph_stats(hd_pipeline pl, hd_size* launches) {
  if (!pl) return 0;
  if (launches) *launches = pl->graph_launches;
  return pl->graph_captures;
}
