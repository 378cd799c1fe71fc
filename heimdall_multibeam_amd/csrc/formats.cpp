// formats.cpp -- sigproc filterbank reader (include/hd/formats.h).
//
// The header is a sequence of length-prefixed keyword strings, each followed
// by a value of the keyword's type (Formats/hd/header.h:125-198): int32 for
// counts, float64 for times and frequencies, one byte for "signed", and a
// further length-prefixed string after source_name / rawdatafile.
#include <cstdio>
#include <cstring>
#include <ctime>
#include <iostream>
#include <string>

#include "hd/formats.h"

namespace {

bool read_str(FILE* f, std::string& s) {
  int len = 0;
  if (std::fread(&len, sizeof(int), 1, f) != 1) return false;
  if (len <= 0 || len >= 80) return false;  // header.h:113
  char buf[80];
  if (std::fread(buf, 1, (size_t)len, f) != (size_t)len) return false;
  s.assign(buf, (size_t)len);
  return true;
}

template <typename T>
bool read_val(FILE* f, T* v) {
  return std::fread(v, sizeof(T), 1, f) == 1;
}

int parse(FILE* f, hd_sigproc_header* h) {
  std::memset(h, 0, sizeof(*h));
  std::string s;
  if (!read_str(f, s) || s != "HEADER_START") return -2;
  bool want_source = false, want_raw = false;
  for (;;) {
    if (!read_str(f, s)) return -3;
    bool ok = true;
    if (s == "HEADER_END") break;
    else if (s == "source_name") want_source = true;
    else if (s == "rawdatafile") want_raw = true;
    else if (s == "tstart") ok = read_val(f, &h->tstart);
    else if (s == "tsamp") ok = read_val(f, &h->tsamp);
    else if (s == "fch1") ok = read_val(f, &h->fch1);
    else if (s == "foff") ok = read_val(f, &h->foff);
    else if (s == "nchans") ok = read_val(f, &h->nchans);
    else if (s == "nbits") ok = read_val(f, &h->nbits);
    else if (s == "nifs") ok = read_val(f, &h->nifs);
    else if (s == "ibeam") ok = read_val(f, &h->ibeam);
    else if (s == "nbeams") ok = read_val(f, &h->nbeams);
    else if (s == "nsamples") ok = read_val(f, &h->nsamples);
    else if (s == "az_start" || s == "za_start" || s == "src_raj" || s == "src_dej" ||
             s == "period" || s == "refdm") {
      double d;
      ok = read_val(f, &d);
    } else if (s == "telescope_id" || s == "machine_id" || s == "data_type" ||
               s == "barycentric" || s == "pulsarcentric" || s == "nbins" || s == "npuls") {
      int i;
      ok = read_val(f, &i);
    } else if (s == "signed") {
      unsigned char c;
      ok = read_val(f, &c);
    } else if (want_source) {
      std::strncpy(h->source_name, s.c_str(), sizeof(h->source_name) - 1);
      want_source = false;
    } else if (want_raw) {
      want_raw = false;
    } else {
      std::cerr << "Warning: read_header: unknown parameter " << s << std::endl;
    }
    if (!ok) return -3;
  }
  h->header_size = (int)std::ftell(f);
  if (h->nsamples == 0 && h->nchans > 0 && h->nbits > 0) {
    // header.h:189-195: from the file size (the header bytes included)
    std::fseek(f, 0, SEEK_END);
    const long end = std::ftell(f);
    h->nsamples = (int)(end / h->nchans * 8 / h->nbits);
    std::fseek(f, h->header_size, SEEK_SET);
  }
  return 0;
}

}  // namespace

struct hd_sigproc_file {
  FILE* f;
  size_t stride;
};

extern "C" int hd_sigproc_read_header(const char* path, hd_sigproc_header* h) {
  if (!path || !h) return -1;
  FILE* f = std::fopen(path, "rb");
  if (!f) return -1;
  const int rc = parse(f, h);
  std::fclose(f);
  return rc;
}

extern "C" time_t hd_mjd_to_utc(double mjd) {
  const int seconds_in_day = 86400;
  int days = (int)mjd;
  const double fdays = mjd - (double)days;
  const double seconds = fdays * (double)seconds_in_day;
  int secs = (int)seconds;
  const double fracsec = seconds - (double)secs;
  if (fracsec - 1 < 0.0000001) secs++;  // DataSource.h:57-58, true for any fraction
  const int julian_day = days + 2400001;
  const int n_four = 4 * (julian_day + ((6 * ((4 * julian_day - 17918) / 146097)) / 4 + 1) / 2 - 37);
  const int n_dten = 10 * (((n_four - 237) % 1461) / 4) + 5;
  struct tm g;
  std::memset(&g, 0, sizeof(g));
  g.tm_year = n_four / 1461 - 4712 - 1900;
  g.tm_mon = (n_dten / 306 + 2) % 12;
  g.tm_mday = (n_dten % 306) / 10 + 1;
  g.tm_hour = secs / 3600;
  secs -= 3600 * g.tm_hour;
  g.tm_min = secs / 60;
  secs -= 60 * g.tm_min;
  g.tm_sec = secs;
  g.tm_isdst = -1;
  return std::mktime(&g);
}

extern "C" hd_sigproc_file* hd_sigproc_open(const char* path, hd_sigproc_header* h) {
  if (!path || !h) return nullptr;
  FILE* f = std::fopen(path, "rb");
  if (!f) return nullptr;
  if (parse(f, h) != 0 || h->nchans <= 0 || h->nbits <= 0) {
    std::fclose(f);
    return nullptr;
  }
  std::fseek(f, h->header_size, SEEK_SET);
  hd_sigproc_file* sf = new hd_sigproc_file;
  sf->f = f;
  sf->stride = (size_t)h->nchans * (size_t)h->nbits / 8;
  return sf;
}

extern "C" hd_size hd_sigproc_read(hd_sigproc_file* sf, hd_size nsamps, hd_byte* data) {
  if (!sf || !data || sf->stride == 0) return 0;
  const size_t got = std::fread(data, 1, nsamps * sf->stride, sf->f);
  return got / sf->stride;
}

extern "C" void hd_sigproc_close(hd_sigproc_file* sf) {
  if (!sf) return;
  std::fclose(sf->f);
  delete sf;
}
