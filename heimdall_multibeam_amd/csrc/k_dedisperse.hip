// k_dedisperse.hip -- filterbank unpack and direct integer dedispersion.
//
// Replaces the reference's external dedisp call (Pipeline/pipeline.cu:459-462):
//   out[dm][t] = sum over live channels c of x[c][t + delay(dm, c)],
// summed exactly in integers and stored as float (exact below 2^24).
//
// Layout: the filterbank [t][channel bytes] is unpacked once to one byte per
// sample in channel-major rows x[live][t] (row pitch xpitch, zero padded), so
// every channel row is a contiguous byte stream in HBM.
//
// Dedispersion tile: a 256-thread block owns 8 consecutive DM trials x 512
// output samples.  For each chunk of 16 channels it stages, per channel, the
// byte window [t0 + d_min, t0 + d_max + 512) into LDS (d_min/d_max = delays
// of the block's first/last trial; delays are monotone in DM).  Each thread
// then sums 16 consecutive outputs of one trial: 5 aligned LDS dwords,
// byte-aligned with v_alignbyte, accumulated as packed 16-bit lanes
// (even/odd bytes) and widened to 32 bits every <= 256 channels.
#include "hd_device.hpp"
#include "hd_internal.hpp"

namespace hd {

namespace {

constexpr int DD_DB = 8;                 // DM trials per block
constexpr int DD_TPD = 32;               // threads per trial
constexpr int DD_R = 16;                 // outputs per thread
constexpr int DD_TB = DD_TPD * DD_R;     // outputs per block (time)
constexpr int DD_CC = 16;                // channels staged per step
constexpr int DD_BS = DD_DB * DD_TPD;    // 256 threads

__device__ __forceinline__ unsigned unpack_sample(const uint8_t* row, int c, int nbits) {
  if (nbits == 8) return row[c];
  const int per = 8 / nbits;
  const unsigned byte = row[c / per];
  const int shift = (c % per) * nbits;
  return (byte >> shift) & ((1u << nbits) - 1u);
}

// 64 time x 64 channel tiles: reads along channels, writes along time.
__global__ __launch_bounds__(256) void k_unpack(const uint8_t* __restrict__ fb,
                                                size_t stride_bytes, int nbits, size_t nsamps,
                                                const int* __restrict__ live, int n_live,
                                                uint8_t* __restrict__ x, size_t xpitch) {
  __shared__ uint8_t tile[64][68];
  const size_t t0 = (size_t)blockIdx.x * 64;
  const int l0 = blockIdx.y * 64;
  const int tx = threadIdx.x & 63;
  const int ty = threadIdx.x >> 6;
  const int l = l0 + tx;
  const int chan = l < n_live ? live[l] : 0;
  for (int r = ty; r < 64; r += 4) {
    const size_t t = t0 + r;
    uint8_t v = 0;
    if (t < nsamps && l < n_live) v = (uint8_t)unpack_sample(fb + t * stride_bytes, chan, nbits);
    tile[r][tx] = v;
  }
  __syncthreads();
  for (int r = ty; r < 64; r += 4) {
    const int lo = l0 + r;
    const size_t t = t0 + tx;
    if (lo < n_live && t < xpitch) x[(size_t)lo * xpitch + t] = tile[tx][r];
  }
}

__global__ __launch_bounds__(DD_BS) void k_dedisperse(const uint8_t* __restrict__ x,
                                                      const int* __restrict__ delays,
                                                      DedispGeom g, float* __restrict__ out) {
  extern __shared__ uint32_t win[];          // DD_CC rows of wlen bytes
  __shared__ int sdly[DD_DB][DD_CC];
  const int tid = threadIdx.x;
  const int dml = tid / DD_TPD;
  const int tl = (tid % DD_TPD) * DD_R;
  const size_t t0 = (size_t)blockIdx.x * DD_TB;
  const int m0 = blockIdx.y * DD_DB;
  const int wd = g.wlen >> 2;
  const int* drow0 = delays + (size_t)m0 * g.n_live;

  uint32_t acc[DD_R];
#pragma unroll
  for (int i = 0; i < DD_R; ++i) acc[i] = 0;
  uint32_t ae[DD_R / 4], ao[DD_R / 4];
#pragma unroll
  for (int i = 0; i < DD_R / 4; ++i) { ae[i] = 0; ao[i] = 0; }
  int pending = 0;

  for (int c0 = 0; c0 < g.n_live; c0 += DD_CC) {
    const int ncc = min(DD_CC, g.n_live - c0);
    __syncthreads();
    if (tid < DD_DB * DD_CC) {
      const int d = tid / DD_CC, cc = tid % DD_CC;
      if (cc < ncc) {
        const int mm = min(m0 + d, g.n_dm - 1);
        const int base = drow0[c0 + cc] & ~3;
        sdly[d][cc] = delays[(size_t)mm * g.n_live + c0 + cc] - base;
      }
    }
    for (int idx = tid; idx < ncc * wd; idx += DD_BS) {
      const int cc = idx / wd;
      const int k = idx - cc * wd;
      const int base = drow0[c0 + cc] & ~3;
      const uint32_t* src =
          reinterpret_cast<const uint32_t*>(x + (size_t)(c0 + cc) * g.xpitch + t0 + base);
      win[cc * wd + k] = src[k];
    }
    __syncthreads();
    for (int cc = 0; cc < ncc; ++cc) {
      const int off = sdly[dml][cc] + tl;
      const uint32_t* p = win + cc * wd + (off >> 2);
      const uint32_t sh = (uint32_t)(off & 3);
      uint32_t w[DD_R / 4 + 1];
#pragma unroll
      for (int i = 0; i <= DD_R / 4; ++i) w[i] = p[i];
#pragma unroll
      for (int i = 0; i < DD_R / 4; ++i) {
        const uint32_t a = __builtin_amdgcn_alignbyte(w[i + 1], w[i], sh);
        ae[i] += a & 0x00FF00FFu;
        ao[i] += (a >> 8) & 0x00FF00FFu;
      }
    }
    pending += ncc;
    if (pending > 256 - DD_CC || c0 + DD_CC >= g.n_live) {
#pragma unroll
      for (int i = 0; i < DD_R / 4; ++i) {
        acc[4 * i + 0] += ae[i] & 0xFFFFu;
        acc[4 * i + 2] += ae[i] >> 16;
        acc[4 * i + 1] += ao[i] & 0xFFFFu;
        acc[4 * i + 3] += ao[i] >> 16;
        ae[i] = 0;
        ao[i] = 0;
      }
      pending = 0;
    }
  }
  const int m = m0 + dml;
  if (m >= g.n_dm) return;
  float* orow = out + (size_t)m * g.ppitch;
  const size_t tb = t0 + tl;
#pragma unroll
  for (int q = 0; q < DD_R / 4; ++q) {
    const size_t t = tb + 4 * q;
    if (t + 3 < g.ppitch) {
      float4 v;
      v.x = (t + 0 < g.nout) ? (float)acc[4 * q + 0] : 0.f;
      v.y = (t + 1 < g.nout) ? (float)acc[4 * q + 1] : 0.f;
      v.z = (t + 2 < g.nout) ? (float)acc[4 * q + 2] : 0.f;
      v.w = (t + 3 < g.nout) ? (float)acc[4 * q + 3] : 0.f;
      *reinterpret_cast<float4*>(orow + t) = v;
    }
  }
}

}  // namespace

int dedisp_tile_time() { return DD_TB; }
int dedisp_tile_dms() { return DD_DB; }

hipError_t launch_unpack(const uint8_t* fb, size_t stride_bytes, int nbits, size_t nsamps,
                         const int* d_live, int n_live, uint8_t* x, size_t xpitch,
                         hipStream_t s) {
  dim3 grid((unsigned)((xpitch + 63) / 64), (unsigned)((n_live + 63) / 64));
  hipLaunchKernelGGL(k_unpack, grid, dim3(256), 0, s, fb, stride_bytes, nbits, nsamps, d_live,
                     n_live, x, xpitch);
  return hipGetLastError();
}

hipError_t launch_dedisperse(const uint8_t* x, const int* d_delays, const DedispGeom& g,
                             float* plane, hipStream_t s) {
  dim3 grid((unsigned)(g.ppitch / DD_TB), (unsigned)((g.n_dm + DD_DB - 1) / DD_DB));
  const size_t lds = (size_t)DD_CC * (size_t)g.wlen;
  if (lds > 65536) return hipErrorInvalidValue;  // delay span of a DM block too wide
  hipLaunchKernelGGL(k_dedisperse, grid, dim3(DD_BS), lds, s, x, d_delays, g, plane);
  return hipGetLastError();
}

}  // namespace hd
