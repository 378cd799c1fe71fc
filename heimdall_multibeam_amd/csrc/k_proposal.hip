// k_proposal.hip -- the coarse DM-time plane and the tiles the proposal
// network scores.
//
// The coarse plane is the search's own front half run at 16x the sampling
// time on every 8th trial: the unpacked channel rows are summed over 16
// samples (k_decimate), dedispersed by k_dedisperse16 with the coarse delay
// table, baseline-removed and RMS-normalised by the search's kernels, and
// stored as bfloat16 (k_coarse_bf16) together with a 4x wider time scale
// (the mean of 4 coarse samples, times 2).  Tiles are 64x64 windows of both,
// centred alike (k_extract_tiles); a unit of 128 trials is searched when any
// of its tiles scores above the threshold (k_unit_flags).
#include "hd_device.hpp"
#include "hd_internal.hpp"
#include "hd_series.hpp"

namespace hd {

namespace {

constexpr int TS = 64;          // tile side
constexpr int TROW0 = 24;       // coarse rows above a tile's centre band
constexpr int TCOL0 = 16;       // coarse samples before its centre span
constexpr int UROWS = 16;       // coarse rows per unit
constexpr int VSTEP = 32;       // coarse samples per time position
constexpr int WIDE = 4;         // coarse samples per wide sample
constexpr int WCOL0 = 28;       // wide samples before a tile centre's first (8v - 28)

// One output byte per thread: 16 input bytes (one 16-byte load, aligned since
// both pitches are multiples of 256 and the group's slot starts on a row).
__global__ __launch_bounds__(256) void k_decimate(const uint8_t* __restrict__ x, size_t xpitch,
                                                  int n_live, int nbits,
                                                  uint8_t* __restrict__ xc, size_t xpitch_c,
                                                  const GroupDev* __restrict__ gd) {
  const size_t tc = (size_t)blockIdx.x * 256 + threadIdx.x;
  const int l = blockIdx.y;
  const int g = blockIdx.z;
  if (tc >= xpitch_c) return;
  const size_t nc = gd[g].nsamps;
  uint8_t out = 0;
  if (tc < nc) {
    const uint4 w = *reinterpret_cast<const uint4*>(
        x + ((size_t)gd[g].xslot * n_live + l) * xpitch + 16 * tc);
    unsigned s = 0;
    const uint32_t ws[4] = {w.x, w.y, w.z, w.w};
#pragma unroll
    for (int k = 0; k < 4; ++k)
      s += (ws[k] & 0xffu) + ((ws[k] >> 8) & 0xffu) + ((ws[k] >> 16) & 0xffu) + (ws[k] >> 24);
    out = (uint8_t)(nbits == 8 ? (s + 8) >> 4 : s);
  }
  xc[((size_t)g * n_live + l) * xpitch_c + tc] = out;
}

// blockIdx: (sample chunk, row, group).
__global__ __launch_bounds__(256) void k_coarse_bf16(const float* __restrict__ plane,
                                                     size_t ppitch, const float* __restrict__ pts,
                                                     size_t prow, const float* __restrict__ rms,
                                                     unsigned nbmax,
                                                     const GroupDev* __restrict__ gd,
                                                     int n_rows, uint16_t* __restrict__ bf,
                                                     size_t bpitch, uint16_t* __restrict__ bw,
                                                     size_t wpitch) {
  const int g = blockIdx.z;
  const size_t row = (size_t)g * n_rows + blockIdx.y;
  const SeriesShape sh = gd[g].sh;
  const float* P = pts + row * prow;
  const float* R = rms + row * nbmax;
  const float* xr = plane + row * ppitch;
  uint16_t* o = bf + row * bpitch;
  uint16_t* ow = bw + row * wpitch;
  auto z_at = [&](size_t t) {
    const float z = normalised_at(xr[t], P, R, sh, t);
    return isfinite(z) ? z : 0.f;
  };
  for (size_t t = (size_t)blockIdx.x * 256 + threadIdx.x; t < bpitch;
       t += (size_t)gridDim.x * 256)
    o[t] = t < sh.nout ? f2bf(z_at(t)) : (uint16_t)0;
  const size_t nw = sh.nout / WIDE;
  for (size_t j = (size_t)blockIdx.x * 256 + threadIdx.x; j < wpitch;
       j += (size_t)gridDim.x * 256) {
    uint16_t h = 0;
    if (j < nw) {
      float s = 0.f;
      for (int i = 0; i < WIDE; ++i) s += z_at(WIDE * j + i);
      h = f2bf(0.5f * s);
    }
    ow[j] = h;
  }
}

// One block per tile; 256 threads move its 4096 values a row at a time.
__global__ __launch_bounds__(256) void k_extract_tiles(TileGrid tg, const uint32_t* __restrict__ pos,
                                                       size_t first, uint16_t* __restrict__ tiles) {
  const size_t i = blockIdx.x;
  int g = 0, u, v;
  if (pos) {
    u = (int)pos[2 * i];
    v = (int)pos[2 * i + 1];
  } else {
    const uint32_t t = (uint32_t)(first + i);
    int lo = 0, hi = tg.G - 1;  // last g with tile_off[g] <= t
    while (lo < hi) {
      const int mid = (lo + hi + 1) / 2;
      if (tg.tile_off[mid] <= t) lo = mid;
      else hi = mid - 1;
    }
    g = lo;
    const uint32_t k = t - tg.tile_off[g];
    const uint32_t nv = tg.n_tpos[g];
    u = (int)(k / nv);
    v = (int)(k % nv);
  }
  const long ncols = (long)tg.ncols[g];
  const long nwide = ncols / WIDE;
  const uint16_t* base = tg.bf + (size_t)g * tg.n_rows * tg.bpitch;
  const uint16_t* wbase = tg.bw + (size_t)g * tg.n_rows * tg.wpitch;
  uint32_t* out = reinterpret_cast<uint32_t*>(tiles + i * (size_t)(TS * TS * 2));
  const int r0 = u * UROWS - TROW0;
  const long c0 = (long)v * VSTEP - TCOL0;
  const long w0 = (long)v * (VSTEP / WIDE) - WCOL0;
  for (int e = threadIdx.x; e < TS * TS; e += 256) {
    const int y = e / TS, xx = e % TS;
    const int row = r0 + y;
    const long col = c0 + xx, wc = w0 + xx;
    uint32_t h0 = 0, h1 = 0;
    if (row >= 0 && row < tg.n_rows) {
      if (col >= 0 && col < ncols) h0 = base[(size_t)row * tg.bpitch + col];
      if (wc >= 0 && wc < nwide) h1 = wbase[(size_t)row * tg.wpitch + wc];
    }
    out[e] = h0 | (h1 << 16);
  }
}

__global__ void k_unit_flags(const float* __restrict__ scores, const uint32_t* __restrict__ tile_off,
                             const uint32_t* __restrict__ n_tpos, int G, int n_units, float thresh,
                             uint8_t* __restrict__ flags) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= G * n_units) return;
  const int g = i / n_units, u = i % n_units;
  const uint32_t nv = n_tpos[g];
  const float* s = scores + tile_off[g] + (size_t)u * nv;
  uint8_t f = 0;
  for (uint32_t v = 0; v < nv; ++v)
    if (s[v] > thresh) {
      f = 1;
      break;
    }
  flags[i] = f;
}

}  // namespace

hipError_t launch_decimate(const uint8_t* x, size_t xpitch, int n_live, int nbits, uint8_t* xc,
                           size_t xpitch_c, const GroupDev* gd, int G, hipStream_t s) {
  if (G <= 0 || n_live <= 0) return hipSuccess;
  if (xpitch % 16 || xpitch_c % 16) return hipErrorInvalidValue;
  hipLaunchKernelGGL(k_decimate, dim3((unsigned)((xpitch_c + 255) / 256), n_live, G), dim3(256),
                     0, s, x, xpitch, n_live, nbits, xc, xpitch_c, gd);
  return hipGetLastError();
}

hipError_t launch_coarse_bf16(const float* plane, size_t ppitch, const float* pts, size_t prow,
                              const float* rms, unsigned nbmax, const GroupDev* gd, int G,
                              int n_rows, uint16_t* bf, size_t bpitch, uint16_t* bw,
                              size_t wpitch, hipStream_t s) {
  if (G <= 0 || n_rows <= 0) return hipSuccess;
  size_t gx = (bpitch + 255) / 256;
  if (gx > 64) gx = 64;
  hipLaunchKernelGGL(k_coarse_bf16, dim3((unsigned)gx, n_rows, G), dim3(256), 0, s, plane, ppitch,
                     pts, prow, rms, nbmax, gd, n_rows, bf, bpitch, bw, wpitch);
  return hipGetLastError();
}

hipError_t launch_extract_tiles(const TileGrid& tg, const uint32_t* pos, size_t first, size_t n,
                                uint16_t* tiles, hipStream_t s) {
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(k_extract_tiles, dim3((unsigned)n), dim3(256), 0, s, tg, pos, first, tiles);
  return hipGetLastError();
}

hipError_t launch_unit_flags(const float* scores, const uint32_t* tile_off,
                             const uint32_t* n_tpos, int G, int n_units, float thresh,
                             uint8_t* flags, hipStream_t s) {
  const int n = G * n_units;
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_unit_flags, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, scores,
                     tile_off, n_tpos, G, n_units, thresh, flags);
  return hipGetLastError();
}

}  // namespace hd
