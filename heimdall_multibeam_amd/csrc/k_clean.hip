// k_clean.hip -- zero-DM RFI cleaning of a filterbank gulp
// (Pipeline/clean_filterbank_rfi.cu:256-528, which hd_execute runs first in
// the reference's CLEANING build, pipeline.cu:263-299).
//
//   series  the gulp dedispersed at one DM over every channel (dedisp with a
//           fresh single-DM plan, :361-397).  At DM 0 that is each time
//           sample's channel sum: a wave reads the sample's 32-bit words and
//           adds them with byte sums (8-bit) or per-bit-plane popcounts.
//   mask    the series' baseline, per-beam RMS and normalised prefix sums come
//           from the search's own kernels; one tiled pass then marks sample i
//           when |z[i]| > rfi_tol or when, for some width w = 1..boxcar_max,
//           the boxcar of width w scaled by 1/sqrt(w) exceeds rfi_tol at an
//           output whose w-sample window covers i (:448-504).
//   zap     every 32-bit word of a marked time sample is rebuilt channel by
//           channel from random nearby unmarked samples (:51-108, 197-234),
//           drawing with the reference's hash, minstd engine and
//           double-precision uniform mapping, so the cleaned bytes match.
#include <cmath>
#include <vector>

#include "hd/error.h"
#include "hd_device.hpp"
#include "hd_internal.hpp"
#include "hd_plan.hpp"
#include "hd_series.hpp"

namespace hd {

namespace {

constexpr int CL_BS = 256;
constexpr int MK_T = 1024;  // mask positions per block

// ---- zero-DM series -------------------------------------------------------
__device__ __forceinline__ uint32_t word_sample_sum(uint32_t v, int nbits, uint32_t acc) {
  if (nbits == 8) return __builtin_amdgcn_sad_u8(v, 0u, acc);  // sum of the four bytes
  if (nbits == 16) return acc + (v & 0xffffu) + (v >> 16);
  // sum of the fields = sum over bit planes b of 2^b * popcount(plane b)
  const uint32_t plane0 = nbits == 4 ? 0x11111111u : nbits == 2 ? 0x55555555u : 0xffffffffu;
  for (int b = 0; b < nbits; ++b) acc += (uint32_t)__popc(v & (plane0 << b)) << b;
  return acc;
}

// series[t] = sum of every channel of time sample t (t < n); zero for
// n <= t < pitch.  One wave per time sample.
__global__ __launch_bounds__(CL_BS) void k_zero_dm_series(const uint32_t* __restrict__ fb,
                                                          unsigned stride_w, int nbits,
                                                          size_t n, size_t pitch,
                                                          float* __restrict__ series) {
  const unsigned lane = threadIdx.x & 63;
  const size_t wave = ((size_t)blockIdx.x * CL_BS + threadIdx.x) >> 6;
  const size_t nwaves = ((size_t)gridDim.x * CL_BS) >> 6;
  for (size_t t = wave; t < pitch; t += nwaves) {
    uint32_t acc = 0;
    if (t < n) {
      const uint32_t* row = fb + t * stride_w;
      for (unsigned w = lane; w < stride_w; w += 64) acc = word_sample_sum(row[w], nbits, acc);
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) acc += __shfl_xor(acc, o);
    if (lane == 0) series[t] = (float)acc;
  }
}

__device__ __forceinline__ unsigned sample_of(const uint8_t* row, unsigned c, int nbits) {
  if (nbits == 8) return row[c];
  if (nbits == 16) return (unsigned)row[2 * c] | ((unsigned)row[2 * c + 1] << 8);
  const unsigned per = 8u / (unsigned)nbits;
  return ((unsigned)row[c / per] >> ((c % per) * (unsigned)nbits)) & ((1u << nbits) - 1u);
}

// Any other DM: series[t] = sum_c x[t + delay_c][c] for t < n (zero to pitch).
__global__ __launch_bounds__(CL_BS) void k_series_at_dm(const uint8_t* __restrict__ fb,
                                                        size_t stride_bytes, int nbits,
                                                        unsigned nchans,
                                                        const int* __restrict__ delays,
                                                        size_t n, size_t pitch,
                                                        float* __restrict__ series) {
  const size_t t = (size_t)blockIdx.x * CL_BS + threadIdx.x;
  if (t >= pitch) return;
  uint32_t acc = 0;
  if (t < n)
    for (unsigned c = 0; c < nchans; ++c)
      acc += sample_of(fb + (t + (size_t)delays[c]) * stride_bytes, c, nbits);
  series[t] = (float)acc;
}

// ---- RFI mask ----------------------------------------------------------------
struct MaskWidths {
  int nw;
  float norm[16];  // (float)(1.0 / sqrt((float)w)), clean_filterbank_rfi.cu:477-478
};

// One block per MK_T positions.  T: prefix sums S[p0 - 2h .. p0 + MK_T + 2h]
// (h = boxcar_max/2, zero outside [0, n]); F: for positions j in
// [p0 - h, p0 + MK_T + h), bit l set when the width-2^l boxcar output at j
// (filtered index j - h) exceeds the cut.
__global__ __launch_bounds__(CL_BS) void k_rfi_mask(const float* __restrict__ S,
                                                    const float* __restrict__ x,
                                                    const float* __restrict__ P,
                                                    const float* __restrict__ R, SeriesShape sh,
                                                    size_t n, unsigned B, MaskWidths mw,
                                                    float tol, uint8_t* __restrict__ mask) {
  extern __shared__ float T[];
  const unsigned h = B / 2;
  const size_t p0 = (size_t)blockIdx.x * MK_T;
  const long long base = (long long)p0 - 2 * (long long)h;  // S index of T[0]
  const unsigned tl = MK_T + 4 * h + 1;
  const unsigned fl = MK_T + 2 * h;
  uint16_t* F = reinterpret_cast<uint16_t*>(T + tl);
  for (unsigned u = threadIdx.x; u < tl; u += CL_BS) {
    const long long si = base + u;
    T[u] = (si >= 0 && si <= (long long)n) ? S[si] : 0.f;
  }
  __syncthreads();
  const size_t L = n + 1 - B;  // boxcar outputs (filtered index k <-> position k + h)
  for (unsigned u = threadIdx.x; u < fl; u += CL_BS) {
    const long long j = (long long)p0 - h + u;
    uint16_t bits = 0;
    if (j >= (long long)h && j < (long long)(h + L)) {
      unsigned w = 1;
      for (int l = 0; l < mw.nw; ++l, w *= 2) {
        const unsigned ahead = (w - 1) / 2 + 1, behind = w / 2;
        const float a = T[j + ahead - base];
        const float b = T[j - behind - base];
        const float f = (a - b) * mw.norm[l];
        if (fabsf(f) > tol) bits |= (uint16_t)(1u << l);
      }
    }
    F[u] = bits;
  }
  __syncthreads();
  for (unsigned u = threadIdx.x; u < MK_T; u += CL_BS) {
    const size_t i = p0 + u;
    if (i >= n) break;
    const float z = normalised_at(x[i], P, R, sh, i);
    bool m = fabsf(z) > tol;
    if (!m && i >= h && i < h + L) {
      // the boxcar of the cut over [i - behind, i + ahead) is non-zero
      unsigned w = 1;
      for (int l = 0; l < mw.nw && !m; ++l, w *= 2) {
        const unsigned ahead = (w - 1) / 2 + 1, behind = w / 2;
        const unsigned f0 = u + h - behind, f1 = u + h + ahead;  // F indices
        for (unsigned f = f0; f < f1; ++f)
          if (F[f] & (1u << l)) {
            m = true;
            break;
          }
      }
    }
    mask[i] = m ? 1 : 0;
  }
}

// ---- zap -------------------------------------------------------------------------
// hash() of clean_filterbank_rfi.cu:33-43 (Robert Jenkins' 32-bit integer hash)
__device__ __forceinline__ uint32_t hash32(uint32_t a) {
  a = (a + 0x7ed55d16u) + (a << 12);
  a = (a ^ 0xc761c23cu) ^ (a >> 19);
  a = (a + 0x165667b1u) + (a << 5);
  a = (a + 0xd3a2646cu) ^ (a << 9);
  a = (a + 0xfd7046c5u) + (a << 3);
  a = (a ^ 0xb55a4f09u) ^ (a >> 16);
  return a;
}

// minstd_rand: x <- 48271 x mod (2^31 - 1); seed s -> s mod m, 0 -> 1
constexpr uint32_t kMinstdM = 2147483647u;
__device__ __forceinline__ uint32_t minstd_seed(uint32_t s) {
  const uint32_t x = s % kMinstdM;
  return x == 0 ? 1u : x;
}

// uniform_int_distribution<unsigned>(lo, hi) over the engine: the double
// (u - 1) / (m - 1) scaled onto [lo, hi + 1) and truncated
__device__ __forceinline__ uint32_t uniform_draw(uint32_t& x, uint32_t lo, uint32_t hi) {
  x = (uint32_t)(((uint64_t)x * 48271u) % kMinstdM);
  double r = (double)(x - 1u);
  r = r / 2147483646.0;
  const double a = (double)lo;
  const double b = (double)hi + 1.0;
  const double v = (r * (b - a)) + a;
  return (uint32_t)v;
}

__global__ __launch_bounds__(CL_BS) void k_zap(const uint32_t* __restrict__ in,
                                               uint32_t* __restrict__ out,
                                               const uint8_t* __restrict__ mask,
                                               uint32_t stride_w, uint32_t n, uint32_t maxd,
                                               int nbits) {
  const uint32_t nwords = n * stride_w;
  const uint32_t bitmask = (1u << nbits) - 1u;
  for (uint32_t i = blockIdx.x * CL_BS + threadIdx.x; i < nwords; i += gridDim.x * CL_BS) {
    const uint32_t t = i / stride_w, c = i - t * stride_w;
    uint32_t result;
    if (mask[t]) {
      uint32_t x = minstd_seed(hash32(i));
      const uint32_t min_t = t > maxd ? t - maxd : 0u;
      const uint32_t max_t = t < n - 1u - maxd ? t + maxd : n - 1u;
      result = 0;
      for (uint32_t k = 0; k < 32; k += (uint32_t)nbits) {
        uint32_t nt = uniform_draw(x, min_t, max_t);
        int attempts = 0;
        while (mask[nt] && ++attempts < 10 + 1) nt = uniform_draw(x, min_t, max_t);
        result |= ((in[(size_t)nt * stride_w + c] >> k) & bitmask) << k;
      }
    } else {
      result = in[i];
    }
    out[i] = result;
  }
}

size_t align256(size_t b) { return (b + 255) / 256 * 256; }

}  // namespace

int clean_plan(size_t nsamps, size_t nbits, size_t nchans, size_t nbeams, float dm, float dt,
               float f0, float df, float baseline_length, float rfi_tol, size_t boxcar_max,
               CleanPlan& cp) {
  cp = CleanPlan{};
  if (!(nbits == 1 || nbits == 2 || nbits == 4 || nbits == 8 || nbits == 16)) return HD_INVALID_NBITS;
  if (nbeams == 0 || boxcar_max == 0 || boxcar_max > 4096) return HD_INVALID_STRIDE;
  const size_t stride_bytes = nchans * nbits / 8;
  if (stride_bytes == 0 || stride_bytes % 4 != 0) return HD_INVALID_STRIDE;  // 32-bit words
  cp.nsamps = nsamps;
  cp.nbits = nbits;
  cp.nchans = nchans;
  cp.stride_w = stride_bytes / 4;
  cp.dm = dm;
  cp.rfi_tol = rfi_tol;
  cp.boxcar_max = boxcar_max;
  // the single-DM plan's delays (:361-377)
  const std::vector<float> kc = hdplan::delay_slopes(nchans, dt, f0, df);
  cp.delays.assign(nchans, 0);
  size_t max_delay = 0;
  for (size_t c = 0; c < nchans; ++c) {
    const int d = hdplan::delay(dm, kc[c]);
    cp.delays[c] = d;
    if (d > 0 && (size_t)d > max_delay) max_delay = (size_t)d;
  }
  cp.max_delay = max_delay;
  if (max_delay >= nsamps) return HD_TOO_FEW_NSAMPS;
  SeriesShape& sh = cp.sh;
  sh.nbeams = (unsigned)nbeams;
  sh.stride = nsamps / nbeams;                       // :379
  sh.nout = nsamps - max_delay;                      // :380
  if (max_delay >= sh.stride || sh.nout + 1 < boxcar_max) return HD_TOO_FEW_NSAMPS;
  sh.bc = sh.stride - max_delay;                     // :381
  sh.ppitch = (sh.nout + 1 + 63) / 64 * 64 + 64;
  sh.spitch = sh.ppitch;
  const int e = series_shape(dt, baseline_length, sh);
  if (e != HD_NO_ERROR) return e;
  cp.nsmooth = (size_t)(baseline_length / (2 * dt));  // :405
  const size_t maxd = cp.nsmooth / 4;                 // :516
  if (sh.nout - 1 < maxd) return HD_TOO_FEW_NSAMPS;   // resampling window wider than the data
  if ((uint64_t)sh.nout * cp.stride_w > 0xffffffffull) return HD_INVALID_STRIDE;  // 32-bit words
  // workspace: series | S | pts | rms | beam-rms scratch | mask | delays
  const size_t pts_words = (size_t)nbeams * (sh.sc ? sh.sc : 1) + 64;
  cp.brms_words = beam_rms_scratch_words(sh, 1);
  cp.off_series = 0;
  cp.off_S = align256(cp.off_series + sh.ppitch * sizeof(float));
  cp.off_pts = align256(cp.off_S + (sh.spitch + 8) * sizeof(float));
  cp.off_rms = align256(cp.off_pts + pts_words * sizeof(float));
  cp.off_brms = align256(cp.off_rms + nbeams * sizeof(float));
  cp.off_mask = align256(cp.off_brms + (cp.brms_words + 1) * sizeof(float));
  cp.off_delays = align256(cp.off_mask + sh.nout + 1);
  cp.work_bytes = align256(cp.off_delays + nchans * sizeof(int));
  return HD_NO_ERROR;
}

hipError_t run_clean(const CleanPlan& cp, const uint8_t* d_in, uint8_t* d_out, void* d_work,
                     hipStream_t s) {
  uint8_t* wb = static_cast<uint8_t*>(d_work);
  float* series = reinterpret_cast<float*>(wb + cp.off_series);
  float* S = reinterpret_cast<float*>(wb + cp.off_S) + 3;  // S[1..] 16-byte aligned
  float* pts = reinterpret_cast<float*>(wb + cp.off_pts);
  float* rms = reinterpret_cast<float*>(wb + cp.off_rms);
  float* brms = reinterpret_cast<float*>(wb + cp.off_brms);
  uint8_t* mask = wb + cp.off_mask;
  const SeriesShape& sh = cp.sh;
  const size_t n = sh.nout;
  hipError_t e;
  // 1. the series at the cleaning DM
  if (cp.max_delay == 0) {
    const size_t waves = (sh.ppitch < 8192 ? sh.ppitch : 8192);
    hipLaunchKernelGGL(k_zero_dm_series, dim3((unsigned)((waves + 3) / 4)), dim3(CL_BS), 0, s,
                       reinterpret_cast<const uint32_t*>(d_in), (unsigned)cp.stride_w,
                       (int)cp.nbits, n, sh.ppitch, series);
  } else {
    int* delays = reinterpret_cast<int*>(wb + cp.off_delays);
    e = hipMemcpyAsync(delays, cp.delays.data(), cp.delays.size() * sizeof(int),
                       hipMemcpyHostToDevice, s);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_series_at_dm, dim3((unsigned)((sh.ppitch + CL_BS - 1) / CL_BS)),
                       dim3(CL_BS), 0, s, d_in, cp.stride_w * 4, (int)cp.nbits,
                       (unsigned)cp.nchans, delays, n, sh.ppitch, series);
  }
  e = hipGetLastError();
  if (e != hipSuccess) return e;
  // 2. baseline points, per-beam RMS and the normalised prefix sums
  e = launch_baseline_points(series, sh, 1, pts, s);
  if (e != hipSuccess) return e;
  e = launch_beam_rms(series, pts, sh, 1, rms, s, brms, cp.brms_words);
  if (e != hipSuccess) return e;
  e = launch_scan(series, sh.ppitch, n, 1, S, sh.spitch, s, pts, rms, &sh);
  if (e != hipSuccess) return e;
  // 3. the mask
  MaskWidths mw{};
  for (size_t w = 1; w <= cp.boxcar_max; w *= 2)
    mw.norm[mw.nw++] = (float)(1.0 / (double)std::sqrt((float)w));
  const unsigned h = (unsigned)(cp.boxcar_max / 2);
  const size_t lds = (MK_T + 4 * (size_t)h + 1) * sizeof(float) +
                     (MK_T + 2 * (size_t)h) * sizeof(uint16_t);
  hipLaunchKernelGGL(k_rfi_mask, dim3((unsigned)((n + MK_T - 1) / MK_T)), dim3(CL_BS), lds, s, S,
                     series, pts, rms, sh, n, (unsigned)cp.boxcar_max, mw, cp.rfi_tol, mask);
  e = hipGetLastError();
  if (e != hipSuccess) return e;
  // 4. the zap of the first n time samples
  const uint32_t nwords = (uint32_t)(n * cp.stride_w);
  const unsigned blocks = (nwords + CL_BS - 1) / CL_BS < 16384 ? (nwords + CL_BS - 1) / CL_BS
                                                                : 16384;
  hipLaunchKernelGGL(k_zap, dim3(blocks ? blocks : 1), dim3(CL_BS), 0, s,
                     reinterpret_cast<const uint32_t*>(d_in), reinterpret_cast<uint32_t*>(d_out),
                     mask, (uint32_t)cp.stride_w, (uint32_t)n, (uint32_t)(cp.nsmooth / 4),
                     (int)cp.nbits);
  return hipGetLastError();
}

}  // namespace hd
