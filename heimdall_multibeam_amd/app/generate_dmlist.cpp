// generate_dmlist.cpp -- prints the DM trial list of a search setup, one
// value per line with the default ostream precision, as
// Applications/generate_dmlist.C:41-110 does (same flags and defaults).
#include <algorithm>
#include <cstdlib>
#include <iostream>
#include <string>
#include <vector>

#include "hd/default_params.h"
#include "hd/stages.h"

static void usage(const char* binary, const hd_params& p) {
  std::cerr << "Usage: " << binary << " [options]\n"
            << "  generate the list of DM trials\n"
            << "    -f0 f               centre frequency of highest channel [default " << p.f0
            << "] MHz\n"
            << "    -df f               width of frequecy channel [default " << p.df << "] MHz\n"
            << "    -nchan n            number of frequency channels [default " << p.nchans
            << "]\n"
            << "    -dt s               sampling time [default " << p.dt << "] s\n"
            << "    -dm min max         DM range [default " << p.dm_min << " " << p.dm_max
            << "]\n"
            << "    -dm_pulse_width w   [default " << p.dm_pulse_width << "] \n"
            << "    -dm_tol t           tolerance between DM trials [default " << p.dm_tol
            << "]" << std::endl;
}

int main(int argc, char* argv[]) {
  hd_params params, defaults;
  hd_set_default_params(&defaults);
  hd_set_default_params(&params);
  for (int i = 1; i < argc; ++i) {
    const std::string a = argv[i];
    auto need = [&](int k) { return i + k < argc; };
    if (a == "-h") {
      usage(argv[0], defaults);
      return -1;
    } else if (a == "-v") {
      params.verbosity = std::max(params.verbosity, 1);
    } else if (a == "-f0" && need(1)) {
      params.f0 = (float)std::atof(argv[++i]);
    } else if (a == "-df" && need(1)) {
      params.df = (float)std::atof(argv[++i]);
    } else if (a == "-nchan" && need(1)) {
      params.nchans = (hd_size)std::atoi(argv[++i]);
    } else if (a == "-dt" && need(1)) {
      params.dt = (float)std::atof(argv[++i]);
    } else if (a == "-dm" && need(2)) {
      params.dm_min = (float)std::atof(argv[++i]);
      params.dm_max = (float)std::atof(argv[++i]);
    } else if (a == "-dm_pulse_width" && need(1)) {
      params.dm_pulse_width = (float)std::atof(argv[++i]);
    } else if (a == "-dm_tol" && need(1)) {
      params.dm_tol = (float)std::atof(argv[++i]);
    } else {
      std::cerr << "WARNING: Unknown parameter '" << a << "'" << std::endl;
    }
  }
  const hd_size n = hd_generate_dm_list(nullptr, 0, params.dm_min, params.dm_max, params.dt,
                                        params.dm_pulse_width, params.f0, params.df,
                                        params.nchans, params.dm_tol);
  std::vector<hd_float> dms(n);
  hd_generate_dm_list(dms.data(), n, params.dm_min, params.dm_max, params.dt,
                      params.dm_pulse_width, params.f0, params.df, params.nchans,
                      params.dm_tol);
  for (hd_size i = 0; i < n; ++i) std::cout << dms[i] << std::endl;
  return 0;
}
