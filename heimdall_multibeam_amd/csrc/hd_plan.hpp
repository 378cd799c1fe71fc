// hd_plan.hpp -- host-side dedispersion plan: DM trial list, per-channel
// delay slopes and integer delays.  Shared by the pipeline library and the
// synthetic-data generator so that injected pulses land on the trial grid.
#pragma once

#include <cmath>
#include <cstddef>
#include <vector>

namespace hdplan {

// DM trial list as the reference requests it from dedisp
// (Pipeline/pipeline.cu:178-182): trials are spaced so that the extra
// smearing between neighbours stays within `tol`.  With f the band centre in
// GHz, a = 8.3*df/f^3, b2 = a^2*nchans^2/16 and c = (dt_us^2+ti^2)*(tol^2-1),
//   DM' = (b2*DM + sqrt(-a2*b2*DM^2 + (a2+b2)*(c + tol^2*a2*DM^2))) / (a2+b2).
// Trials are stored as float and each step starts from the stored value
// (this reproduces the reference fixture Share/dm2000.list exactly).
inline std::vector<float> dm_list(float dm_min, float dm_max, float dt_s,
                                  float pulse_width_us, float f0, float df,
                                  std::size_t nchans, float tol) {
  const double dt_us = double(dt_s) * 1e6;
  const double fc = (double(f0) + (double(nchans / 2) - 0.5) * double(df)) * 1e-3;
  const double t2 = double(tol) * double(tol);
  const double a = 8.3 * double(df) / (fc * fc * fc);
  const double a2 = a * a;
  const double b2 = a2 * double(nchans * nchans) / 16.0;
  const double c = (dt_us * dt_us + double(pulse_width_us) * double(pulse_width_us)) * (t2 - 1.0);
  std::vector<float> out;
  float dm = dm_min;
  out.push_back(dm);
  while (dm < dm_max) {
    const double x = double(dm);
    const double x2 = x * x;
    const double disc = -a2 * b2 * x2 + (a2 + b2) * (c + t2 * a2 * x2);
    dm = float((b2 * x + std::sqrt(disc)) / (a2 + b2));
    out.push_back(dm);
  }
  return out;
}

// Dispersion delay slope of channel c (samples per unit DM) for a channel
// frequency f0 + c*df MHz: 4.15e3/dt * (f_c^-2 - f0^-2), in double, stored
// as float.
inline std::vector<float> delay_slopes(std::size_t nchans, float dt_s, float f0, float df) {
  std::vector<float> k(nchans);
  const double i0 = 1.0 / double(f0);
  for (std::size_t c = 0; c < nchans; ++c) {
    const double inv = 1.0 / (double(f0) + double(c) * double(df));
    k[c] = float(4.15e3 / double(dt_s) * (inv * inv - i0 * i0));
  }
  return k;
}

// Integer delay in samples: the float product plus one half, truncated.
// Two statements, so the product is rounded before the add (no contraction).
inline int delay(float dm, float kc) {
  const float x = dm * kc;
  return int(x + 0.5f);
}

// Adaptive time resolution (dedisp_enable_adaptive_dt, requested at
// Pipeline/pipeline.cu:187-196 with dm_pulse_width and scrunch_tol).  The
// dedisp library is not part of the reference tree; this restates its
// published rule.  The effective width of a pulse at DM with trial step dDM,
// in float throughout (f0, df in MHz, sampling time and width in seconds):
//   BW = nchans |df|, fc = f0 - BW / 2,
//   t_DM = 8.3 BW DM / fc^3, t_dDM = 8.3 / 4 BW nchans dDM / fc^3,
//   W_eff = sqrt(dt^2 + W^2 + t_DM^2 + t_dDM^2).
inline float smearing(float dt, float width, float f0, std::size_t nchans, float df, float dm,
                      float ddm) {
  const float bw = float(nchans) * std::fabs(df);
  const float fc = f0 - bw / 2.f;
  const float inv_fc3 = 1.f / (fc * fc * fc);
  const float t_dm = 8.3f * bw * dm * inv_fc3;
  const float t_ddm = 8.3f / 4.f * bw * float(nchans) * ddm * inv_fc3;
  return std::sqrt(dt * dt + width * width + t_dm * t_dm + t_ddm * t_ddm);
}

// Time-scrunch factor of each trial: 1 for the first; each next trial doubles
// its predecessor's factor when doubling the sampling time would raise the
// effective width by less than a factor tol, else keeps it.  Factors are
// powers of two and never decrease along the list.
inline std::vector<std::size_t> scrunch_list(const std::vector<float>& dms, float dt_s,
                                             float pulse_width_us, float f0, float df,
                                             std::size_t nchans, float tol) {
  std::vector<std::size_t> s(dms.size(), 1);
  const float w = pulse_width_us * 1e-6f;
  for (std::size_t d = 1; d < dms.size(); ++d) {
    const float ddm = dms[d] - dms[d - 1];
    const float cur = float(s[d - 1]) * dt_s;
    const float a = smearing(cur, w, f0, nchans, df, dms[d], ddm);
    const float b = smearing(2.f * cur, w, f0, nchans, df, dms[d], ddm);
    s[d] = (b / a < tol) ? 2 * s[d - 1] : s[d - 1];
  }
  return s;
}

}  // namespace hdplan
