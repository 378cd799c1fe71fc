// k_synth.hip -- the synthetic filterbank generator on the GPU.
//
// The same noise, tones and injected profiles as the host generator
// (synth.cpp), evaluated per output byte in double precision in the host's
// order: noise, tone, then every profile that covers the sample in list
// order.  A block owns 16 time samples x 256 output bytes and first selects
// (in order) the profiles whose dispersed track crosses its tile.
#include <hip/hip_runtime.h>

#include <cmath>
#include <vector>

#include "hd/synth.h"
#include "hd_plan.hpp"

namespace hd {
namespace {

struct DevProfile {
  double centre, sigma_t, amp;
  long half;
  float dm;
  int zero_dm;
};

struct SynthArgs {
  size_t nsamps, nch, stride;
  int nbits, per;
  uint64_t seed;
  float mean, sigma, dead_value;
  const float* kc;
  const int* dead;
  const float* tone_amp;
  const float* tone_per;
  const float* tone_ph;
  const DevProfile* prof;
  int nprof;
  uint8_t* out;
  unsigned* overflow;
};

constexpr int ST = 16;        // time samples per block
constexpr int SB = 256;       // output bytes per block (one per thread)
constexpr int MAXL = 512;     // profiles crossing one block

__device__ __forceinline__ uint64_t mix64(uint64_t z) {
  z += 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

__device__ __forceinline__ double gauss(uint64_t seed, uint64_t idx) {
  const uint64_t h1 = mix64(seed * 0xD1342543DE82EF95ull ^ mix64(idx));
  const uint64_t h2 = mix64(h1 ^ 0xA0761D6478BD642Full);
  const double u1 = (double((h1 >> 11) + 1)) * (1.0 / 9007199254740993.0);
  const double u2 = double(h2 >> 11) * (1.0 / 9007199254740992.0);
  return sqrt(-2.0 * log(u1)) * cos(6.283185307179586 * u2);
}

__device__ __forceinline__ int delay_of(float dm, float kc) {
  const float x = dm * kc;
  return int(x + 0.5f);
}

__global__ __launch_bounds__(SB) void k_synth(SynthArgs a) {
  __shared__ int list[MAXL];
  __shared__ unsigned char flag[SB];
  __shared__ int nlist;
  const long t0b = (long)blockIdx.x * ST;
  const long t1b = t0b + ST - 1;
  const size_t j = (size_t)blockIdx.y * SB + threadIdx.x;
  const size_t c0b = (size_t)blockIdx.y * SB * a.per;
  size_t c1b = c0b + (size_t)SB * a.per;
  if (c1b > a.nch) c1b = a.nch;
  if (threadIdx.x == 0) nlist = 0;
  __syncthreads();
  for (int base = 0; base < a.nprof; base += SB) {
    const int p = base + (int)threadIdx.x;
    unsigned char f = 0;
    if (p < a.nprof) {
      const DevProfile& pr = a.prof[p];
      int d0 = 0, d1 = 0;
      if (!pr.zero_dm) {
        d0 = delay_of(pr.dm, a.kc[c0b]);
        d1 = delay_of(pr.dm, a.kc[c1b - 1]);
      }
      const double lo = pr.centre + (double)(d0 < d1 ? d0 : d1) - (double)pr.half;
      const double hi = pr.centre + (double)(d0 < d1 ? d1 : d0) + (double)pr.half;
      f = (lo <= (double)t1b + 1.0 && hi >= (double)t0b - 1.0) ? 1 : 0;
    }
    flag[threadIdx.x] = f;
    __syncthreads();
    if (threadIdx.x == 0) {
      for (int k = 0; k < SB && base + k < a.nprof; ++k)
        if (flag[k]) {
          if (nlist < MAXL) list[nlist] = base + k;
          ++nlist;
        }
    }
    __syncthreads();
  }
  const int nl = nlist < MAXL ? nlist : MAXL;
  if (nlist > MAXL && threadIdx.x == 0) atomicAdd(a.overflow, 1u);
  if (j >= a.stride) return;
  const double vmax = double((1u << a.nbits) - 1u);
  for (int tt = 0; tt < ST; ++tt) {
    const long t = t0b + tt;
    if ((size_t)t >= a.nsamps) break;
    unsigned byte = 0;
    for (int k = 0; k < a.per; ++k) {
      const size_t c = j * a.per + k;
      if (c >= a.nch) break;
      double v;
      const bool dead = a.dead && a.dead[c];
      if (dead) {
        v = a.dead_value;
      } else {
        v = (double)a.mean + (double)a.sigma * gauss(a.seed, (uint64_t)t * a.nch + c);
        if (a.tone_amp[c] != 0.f)
          v += (double)a.tone_amp[c] *
               sin(6.283185307179586 * double(t) / (double)a.tone_per[c] + (double)a.tone_ph[c]);
        for (int q = 0; q < nl; ++q) {
          const DevProfile& pr = a.prof[list[q]];
          const double d = pr.zero_dm ? 0.0 : (double)delay_of(pr.dm, a.kc[c]);
          const double dtc = double(t) - (pr.centre + d);
          if (dtc < -(double)pr.half || dtc > (double)pr.half) continue;
          v += pr.amp * exp(-0.5 * dtc * dtc / (pr.sigma_t * pr.sigma_t));
        }
      }
      double qv = rint(v);
      qv = qv < 0 ? 0 : (qv > vmax ? vmax : qv);
      byte |= (unsigned)qv << (k * a.nbits);
    }
    a.out[(size_t)t * a.stride + j] = (uint8_t)byte;
  }
}

DevProfile profile_of(const hd_synth_pulse& p, size_t nlive, double noise_sigma, bool zero_dm) {
  DevProfile pr;
  const double w = p.width > 0.05f ? double(p.width) : 0.05;
  pr.sigma_t = w / (2.0 * std::sqrt(2.0 * std::log(2.0)));
  pr.centre = p.t0;
  pr.half = long(std::ceil(6.0 * pr.sigma_t)) + 1;
  double s2 = 0.0;
  for (long t = -pr.half; t <= pr.half; ++t) {
    const double g = std::exp(-0.5 * (t * t) / (pr.sigma_t * pr.sigma_t));
    s2 += g * g;
  }
  pr.amp = double(p.snr) * noise_sigma / (std::sqrt(double(nlive ? nlive : 1)) * std::sqrt(s2));
  pr.dm = p.dm;
  pr.zero_dm = zero_dm ? 1 : 0;
  return pr;
}

template <typename T>
hipError_t upload(T** d, const T* h, size_t n) {
  *d = nullptr;
  hipError_t e = hipMalloc((void**)d, std::max<size_t>(n, 1) * sizeof(T));
  if (e == hipSuccess && n) e = hipMemcpy(*d, h, n * sizeof(T), hipMemcpyHostToDevice);
  return e;
}

}  // namespace
}  // namespace hd

extern "C" hd_error hd_synth_filterbank_device(const hd_synth_spec* s, hd_byte* d_out,
                                               int gpu_id) {
  using namespace hd;
  if (!s || !d_out) return HD_INVALID_POINTER;
  if (!(s->nbits == 1 || s->nbits == 2 || s->nbits == 4 || s->nbits == 8)) return HD_INVALID_NBITS;
  if (s->nchans == 0 || (s->nchans * s->nbits) % 8) return HD_INVALID_STRIDE;
  int prev = -1;
  (void)hipGetDevice(&prev);
  if (hipSetDevice(gpu_id) != hipSuccess) return HD_INVALID_DEVICE_INDEX;
  const size_t nch = s->nchans;
  const std::vector<float> kc = hdplan::delay_slopes(nch, s->dt, s->f0, s->df);
  size_t nlive = 0;
  std::vector<int> dead(nch, 0);
  for (size_t c = 0; c < nch; ++c) {
    dead[c] = (s->dead && s->dead[c]) ? 1 : 0;
    nlive += !dead[c];
  }
  std::vector<DevProfile> prof;
  for (size_t i = 0; i < s->npulses; ++i) prof.push_back(profile_of(s->pulses[i], nlive, s->sigma, false));
  for (size_t i = 0; i < s->nbursts; ++i) prof.push_back(profile_of(s->bursts[i], nlive, s->sigma, true));
  std::vector<float> ta(nch, 0.f), tp(nch, 1.f), tph(nch, 0.f);
  for (size_t i = 0; i < s->ntones; ++i) {
    const int c = s->tones[i].chan;
    if (c < 0 || (size_t)c >= nch) continue;
    ta[c] = s->tones[i].amp_sigma * s->sigma;
    tp[c] = s->tones[i].period;
    tph[c] = s->tones[i].phase;
  }
  SynthArgs a{};
  a.nsamps = s->nsamps;
  a.nch = nch;
  a.nbits = (int)s->nbits;
  a.per = 8 / (int)s->nbits;
  a.stride = nch * s->nbits / 8;
  a.seed = s->seed;
  a.mean = s->mean;
  a.sigma = s->sigma;
  a.dead_value = s->dead_value;
  a.nprof = (int)prof.size();
  a.out = d_out;
  float *dkc = nullptr, *dta = nullptr, *dtp = nullptr, *dtph = nullptr;
  int* ddead = nullptr;
  DevProfile* dprof = nullptr;
  unsigned* dovf = nullptr;
  const unsigned zero = 0;
  hipError_t e = upload(&dkc, kc.data(), nch);
  if (e == hipSuccess) e = upload(&ddead, dead.data(), nch);
  if (e == hipSuccess) e = upload(&dta, ta.data(), nch);
  if (e == hipSuccess) e = upload(&dtp, tp.data(), nch);
  if (e == hipSuccess) e = upload(&dtph, tph.data(), nch);
  if (e == hipSuccess) e = upload(&dprof, prof.data(), prof.size());
  if (e == hipSuccess) e = upload(&dovf, &zero, 1);
  unsigned ovf = 0;
  if (e == hipSuccess && s->nsamps > 0) {
    a.kc = dkc;
    a.dead = ddead;
    a.tone_amp = dta;
    a.tone_per = dtp;
    a.tone_ph = dtph;
    a.prof = dprof;
    a.overflow = dovf;
    dim3 grid((unsigned)((s->nsamps + ST - 1) / ST), (unsigned)((a.stride + SB - 1) / SB));
    hipLaunchKernelGGL(k_synth, grid, dim3(SB), 0, 0, a);
    e = hipGetLastError();
    if (e == hipSuccess) e = hipDeviceSynchronize();
    if (e == hipSuccess) e = hipMemcpy(&ovf, dovf, sizeof(unsigned), hipMemcpyDeviceToHost);
  }
  void* bufs[] = {dkc, ddead, dta, dtp, dtph, dprof, dovf};
  for (void* b : bufs)
    if (b) (void)hipFree(b);
  if (prev >= 0) (void)hipSetDevice(prev);
  if (e != hipSuccess) return HD_INTERNAL_GPU_ERROR;
  return ovf ? HD_TOO_MANY_EVENTS : HD_NO_ERROR;  // more overlapping profiles than a tile holds
}
