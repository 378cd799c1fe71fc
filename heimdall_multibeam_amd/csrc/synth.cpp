// synth.cpp -- seeded synthetic filterbank generator (host C++, OpenMP).
//
// Produces [nsamps][nchans*nbits/8] bytes in the sigproc/STF layout that the
// reference readers deliver (channel 0 = f0, low bits first for nbits < 8):
//   noise    N(mean, sigma^2) per sample, rounded half-to-even and clipped;
//   pulses   Gaussian profiles (FWHM = width samples) centred at
//            t0 + delay(dm, c) on the plan's integer delay table, with a
//            per-channel amplitude that gives the requested optimal
//            matched-filter S/N: A = snr*sigma / (sqrt(N_live)*sqrt(sum g^2));
//   tones    sinusoids on chosen channels (narrow-band RFI);
//   bursts   zero-DM Gaussian pulses in every live channel (broad-band RFI);
//   dead     channels held at a constant value.
// The noise is a counter-based hash of (seed, sample, channel), so the output
// is independent of the thread count.
#include <algorithm>
#include <cmath>
#include <functional>
#include <cstdint>
#include <cstring>
#include <vector>

#include "hd/synth.h"
#include "hd_plan.hpp"


namespace {

inline std::uint64_t mix64(std::uint64_t z) {
  z += 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

inline double gauss(std::uint64_t seed, std::uint64_t idx) {
  const std::uint64_t h1 = mix64(seed * 0xD1342543DE82EF95ull ^ mix64(idx));
  const std::uint64_t h2 = mix64(h1 ^ 0xA0761D6478BD642Full);
  const double u1 = (double((h1 >> 11) + 1)) * (1.0 / 9007199254740993.0);  // (0,1]
  const double u2 = double(h2 >> 11) * (1.0 / 9007199254740992.0);          // [0,1)
  return std::sqrt(-2.0 * std::log(u1)) * std::cos(6.283185307179586 * u2);
}

struct Profile {
  double centre_ch0;  // arrival at channel 0
  double sigma_t;     // Gaussian sigma in samples
  double amp;         // per-channel peak amplitude
  std::vector<double> delay;
  long half;          // support half-width in samples
  bool ascending, descending;
};

Profile make_profile(const hd_synth_pulse& p, const std::vector<float>& k, std::size_t nlive,
                     double noise_sigma, bool zero_dm) {
  Profile pr;
  const double w = p.width > 0.05f ? double(p.width) : 0.05;
  pr.sigma_t = w / (2.0 * std::sqrt(2.0 * std::log(2.0)));
  pr.centre_ch0 = p.t0;
  pr.half = long(std::ceil(6.0 * pr.sigma_t)) + 1;
  // sum of g^2 over integer samples around an integer-aligned centre
  double s2 = 0.0;
  for (long t = -pr.half; t <= pr.half; ++t) {
    const double g = std::exp(-0.5 * (t * t) / (pr.sigma_t * pr.sigma_t));
    s2 += g * g;
  }
  pr.amp = double(p.snr) * noise_sigma / (std::sqrt(double(nlive ? nlive : 1)) * std::sqrt(s2));
  pr.delay.resize(k.size());
  for (std::size_t c = 0; c < k.size(); ++c) pr.delay[c] = zero_dm ? 0 : hdplan::delay(p.dm, k[c]);
  pr.ascending = std::is_sorted(pr.delay.begin(), pr.delay.end());
  pr.descending = std::is_sorted(pr.delay.begin(), pr.delay.end(), std::greater<double>());
  return pr;
}

}  // namespace

extern "C" int hd_synth_filterbank(const hd_synth_spec* s, std::uint8_t* out) {
  if (!(s->nbits == 1 || s->nbits == 2 || s->nbits == 4 || s->nbits == 8)) return 1;
  const std::size_t nch = s->nchans;
  const std::vector<float> k = hdplan::delay_slopes(nch, s->dt, s->f0, s->df);
  std::size_t nlive = 0;
  for (std::size_t c = 0; c < nch; ++c) nlive += !(s->dead && s->dead[c]);
  std::vector<Profile> prof;
  for (std::size_t i = 0; i < s->npulses; ++i)
    prof.push_back(make_profile(s->pulses[i], k, nlive, s->sigma, false));
  for (std::size_t i = 0; i < s->nbursts; ++i)
    prof.push_back(make_profile(s->bursts[i], k, nlive, s->sigma, true));
  std::vector<float> tone_amp(nch, 0.f), tone_per(nch, 1.f), tone_ph(nch, 0.f);
  for (std::size_t i = 0; i < s->ntones; ++i) {
    const int c = s->tones[i].chan;
    if (c < 0 || std::size_t(c) >= nch) continue;
    tone_amp[c] = s->tones[i].amp_sigma * s->sigma;
    tone_per[c] = s->tones[i].period;
    tone_ph[c] = s->tones[i].phase;
  }
  const double vmax = double((1u << s->nbits) - 1u);
  const std::size_t per = 8 / s->nbits;
  const std::size_t stride = nch * s->nbits / 8;
  const long ns = long(s->nsamps);
#pragma omp parallel
  {
    std::vector<double> row(nch);
#pragma omp for schedule(static)
    for (long t = 0; t < ns; ++t) {
      for (std::size_t c = 0; c < nch; ++c) {
        if (s->dead && s->dead[c]) { row[c] = s->dead_value; continue; }
        double v = s->mean + s->sigma * gauss(s->seed, std::uint64_t(t) * nch + c);
        if (tone_amp[c] != 0.f)
          v += tone_amp[c] * std::sin(6.283185307179586 * double(t) / tone_per[c] + tone_ph[c]);
        row[c] = v;
      }
      for (const Profile& pr : prof) {
        // delays are monotone in the channel index, so the channels whose
        // profile covers sample t form one contiguous range
        const double lo = double(t) - pr.half - pr.centre_ch0;  // need delay >= lo
        const double hi = double(t) + pr.half - pr.centre_ch0;  // need delay <= hi
        if (!pr.ascending && !pr.descending) continue;
        std::size_t c0 = 0, c1 = nch;
        if (pr.ascending) {
          c0 = std::size_t(std::lower_bound(pr.delay.begin(), pr.delay.end(), lo) - pr.delay.begin());
          c1 = std::size_t(std::upper_bound(pr.delay.begin(), pr.delay.end(), hi) - pr.delay.begin());
        } else {
          c0 = std::size_t(std::lower_bound(pr.delay.begin(), pr.delay.end(), hi, std::greater<double>()) - pr.delay.begin());
          c1 = std::size_t(std::upper_bound(pr.delay.begin(), pr.delay.end(), lo, std::greater<double>()) - pr.delay.begin());
        }
        for (std::size_t c = c0; c < c1; ++c) {
          if (s->dead && s->dead[c]) continue;
          const double dtc = double(t) - (pr.centre_ch0 + pr.delay[c]);
          if (dtc < -pr.half || dtc > pr.half) continue;
          row[c] += pr.amp * std::exp(-0.5 * dtc * dtc / (pr.sigma_t * pr.sigma_t));
        }
      }
      std::uint8_t* o = out + std::size_t(t) * stride;
      if (s->nbits == 8) {
        for (std::size_t c = 0; c < nch; ++c) {
          double q = std::nearbyint(row[c]);
          q = q < 0 ? 0 : (q > vmax ? vmax : q);
          o[c] = std::uint8_t(q);
        }
      } else {
        std::memset(o, 0, stride);
        for (std::size_t c = 0; c < nch; ++c) {
          double q = std::nearbyint(row[c]);
          q = q < 0 ? 0 : (q > vmax ? vmax : q);
          o[c / per] |= std::uint8_t(unsigned(q) << ((c % per) * s->nbits));
        }
      }
    }
  }
  return 0;
}

// The plan's delays, exported so callers can place pulses on the trial grid.
extern "C" int hd_synth_delay(float dm, std::size_t chan, float dt, float f0, float df) {
  const double i0 = 1.0 / double(f0);
  const double inv = 1.0 / (double(f0) + double(chan) * double(df));
  const float kc = float(4.15e3 / double(dt) * (inv * inv - i0 * i0));
  return hdplan::delay(dm, kc);
}

// The plan's DM trials (same recurrence as the pipeline); returns the count
// and writes at most cap values.
extern "C" std::size_t hd_synth_dm_list(float* out, std::size_t cap, float dm_min, float dm_max,
                                        float dt, float pulse_width_us, float f0, float df,
                                        std::size_t nchans, float tol) {
  const std::vector<float> l = hdplan::dm_list(dm_min, dm_max, dt, pulse_width_us, f0, df, nchans, tol);
  for (std::size_t i = 0; i < l.size() && i < cap; ++i) out[i] = l[i];
  return l.size();
}
