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
// Dedispersion stages, per block of DM trials and channel, the byte window
// [t0 + d_min, t0 + d_max + tile) into LDS (delays are monotone in DM), so
// each input byte is read from L2 once per block and many times from LDS.
#include <cstdlib>

#include "hd_device.hpp"
#include "hd_internal.hpp"

namespace hd {

namespace {

__device__ __forceinline__ unsigned unpack_sample(const uint8_t* row, int c, int nbits) {
  if (nbits == 8) return row[c];
  const int per = 8 / nbits;
  const unsigned byte = row[c / per];
  const int shift = (c % per) * nbits;
  return (byte >> shift) & ((1u << nbits) - 1u);
}

// Bytes of a group's unpacked rows that are written: the whole padded row,
// or with trim > 0 the first nsamps + trim (rounded to 64, at most the
// row).  The dedispersion tiles of a group read its rows below
// nsamps + D2_TB + 64 (a tile starts at most at the group's last output,
// its window spans D2_TB samples past the largest delay plus the staging
// overrun), so trim = 2 * D2_TB covers every read of a short group of a
// ragged batch while skipping the rest of its padding.
__device__ __forceinline__ size_t row_limit(size_t nsamps, size_t xpitch, size_t trim) {
  if (!trim) return xpitch;
  const size_t l = (nsamps + trim + 63) / 64 * 64;
  return l < xpitch ? l : xpitch;
}

// 64 time x 64 channel tiles: reads along channels, writes along time.
// blockIdx.z: the group (its input, sample count and rows of x).
__global__ __launch_bounds__(256) void k_unpack(const uint8_t* __restrict__ fb,
                                                size_t stride_bytes, int nbits, size_t nsamps,
                                                const int* __restrict__ live, int n_live,
                                                uint8_t* __restrict__ x, size_t xpitch,
                                                const GroupDev* __restrict__ gd,
                                                size_t trim) {
  __shared__ uint8_t tile[64][68];
  if (gd) {
    fb = gd[blockIdx.z].fb;
    nsamps = gd[blockIdx.z].nsamps;
  }
  x += (size_t)blockIdx.z * n_live * xpitch;
  const size_t t0 = (size_t)blockIdx.x * 64;
  const size_t tlim = row_limit(nsamps, xpitch, trim);
  if (t0 >= tlim) return;
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
    if (lo < n_live && t < tlim) x[(size_t)lo * xpitch + t] = tile[tx][r];
  }
}

// 8-bit samples with every channel live: tiles of 64 time samples x tcw
// channels moved as 32-bit words both ways (four channels per word read, four
// time samples per word written).  tcw is a multiple of 4: the whole row when
// it has at most 512 channels, so that a tile is one stretch of the input
// and no block reads part of a cache line another block also fetches; 256
// otherwise.  Dynamic LDS: 64 rows of tcw / 4 + 1 words.
constexpr int UP_T = 64;
constexpr int UP_ROW_MAX = 512;
__global__ __launch_bounds__(256) void k_unpack8(const uint8_t* __restrict__ fb,
                                                 size_t stride_bytes, size_t nsamps, int nchans,
                                                 uint8_t* __restrict__ x, size_t xpitch,
                                                 const GroupDev* __restrict__ gd, size_t trim,
                                                 int tcw) {
  extern __shared__ uint32_t tile[];
  if (gd) {
    fb = gd[blockIdx.z].fb;
    nsamps = gd[blockIdx.z].nsamps;
  }
  x += (size_t)blockIdx.z * nchans * xpitch;
  const size_t t0 = (size_t)blockIdx.x * UP_T;
  const size_t tlim = row_limit(nsamps, xpitch, trim);
  if (t0 >= tlim) return;
  const int c0 = blockIdx.y * tcw;
  const int tid = threadIdx.x;
  const int nw = tcw / 4, rw = nw + 1;  // words per tile row, with the pad
  const bool words = (stride_bytes % 4 == 0) && (reinterpret_cast<uintptr_t>(fb) % 4 == 0);
  // read: word idx of the tile is row idx / nw, word idx % nw; U loads in
  // flight per thread before their LDS writes
  constexpr int U = 8;
  const int nread = UP_T * nw;
  for (int i0 = tid; i0 < nread; i0 += U * 256) {
    uint32_t v[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int idx = i0 + u * 256;
      const int r = idx / nw, dw = idx - r * nw;
      const size_t t = t0 + r;
      const int c = c0 + 4 * dw;
      v[u] = 0;
      if (idx < nread && t < nsamps && c < nchans) {
        const uint8_t* row = fb + t * stride_bytes;
        if (words && c + 3 < nchans) {
          v[u] = *reinterpret_cast<const uint32_t*>(row + c);
        } else {
          for (int i = 0; i < 4 && c + i < nchans; ++i) v[u] |= (uint32_t)row[c + i] << (8 * i);
        }
      }
    }
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int idx = i0 + u * 256;
      if (idx < nread) {
        const int r = idx / nw;
        tile[r * rw + (idx - r * nw)] = v[u];
      }
    }
  }
  __syncthreads();
  // write: channel idx / 16 of the tile, time word idx % 16 (16 lanes write
  // 64 consecutive bytes of one channel row)
  const uint8_t* tb = reinterpret_cast<const uint8_t*>(tile);
  const int RB = rw * 4;  // bytes per tile row
  const int cend = min(tcw, nchans - c0);
  const int nwrite = cend * (UP_T / 4);
  for (int idx = tid; idx < nwrite; idx += 256) {
    const int cc = idx / (UP_T / 4), q = idx % (UP_T / 4);
    const size_t t = t0 + 4 * q;
    if (t >= tlim) continue;
    const uint32_t v = (uint32_t)tb[(4 * q + 0) * RB + cc] |
                       ((uint32_t)tb[(4 * q + 1) * RB + cc] << 8) |
                       ((uint32_t)tb[(4 * q + 2) * RB + cc] << 16) |
                       ((uint32_t)tb[(4 * q + 3) * RB + cc] << 24);
    *reinterpret_cast<uint32_t*>(x + (size_t)(c0 + cc) * xpitch + t) = v;
  }
}

// ---------------------------------------------------------------------------
// Dedispersion, 16-bit staging.  A 512-thread block owns 16 consecutive DM
// trials x 1024 output samples; wave w sums trials w and w+8.  Per channel
// the block's byte window [t0 + base, ...) (base = first trial's delay,
// dword aligned) is staged twice as 16-bit sample pairs: E[k] = (x[2k],
// x[2k+1]) and O[k] = (x[2k+1], x[2k+2]).  A trial's delay offset o is the
// same for every lane of the wave, so it reads E (o even) or O (o odd) at
// dword (o>>1) + lane + 64q: consecutive lanes hit consecutive banks, and one
// 32-bit add accumulates two samples (16-bit lanes never carry for <= 257
// channels; they are widened to 32 bits every <= 256 channels).
constexpr int D2_DB = 16;                 // trials per block
constexpr int D2_Q = 8;                   // sample pairs per lane per trial
constexpr int D2_TB = 64 * 2 * D2_Q;      // 1024 outputs per block
constexpr int D2_CC = 4;                  // channels staged per step
constexpr int D2_BS = 512;
constexpr int D2_WAVES = D2_BS / 64;     // 8: wave w sums trials w, w+8
constexpr int D2_PF = 6;                  // staging items prefetched per thread

// Waves per SIMD each tile shape is compiled for (its register budget is
// 512 / waves): the largest at which its accumulators and the compiler's
// load hoisting fit without spilling; register spills cost more than the
// occupancy given up.
constexpr int d16_waves(int db, int cc) { return db == 32 ? 2 : (db == 16 && cc >= 8 ? 3 : 4); }

// chunk_pairs[block][chunk]: staged E/O dword pairs each channel of a
// 4-channel chunk needs for this block of trials (host-computed from the
// delay spans, so top-of-band chunks stage short windows).
template <int DB, int CC>
__global__ __launch_bounds__(D2_BS, d16_waves(DB, CC)) void k_dedisperse16(const uint8_t* __restrict__ x,
                                                       const int* __restrict__ delays,
                                                       const int* __restrict__ chunk_pairs,
                                                       DedispGeom g, int hw_max, int xcd_remap,
                                                       float* __restrict__ out,
                                                       const GroupDev* __restrict__ gd) {
  extern __shared__ uint32_t lds16[];  // [CC][2][hw_max]
  constexpr int NA = DB / D2_WAVES;  // trials per wave
  __shared__ __align__(16) int sdly[DB][CC];
  // blockIdx.z: the group, with its own unpacked rows, output rows and length
  // its trials start at trial0 of the plan's delay table (a whole number of
  // blocks in) and its unpacked input is slot xslot of x
  const size_t nout = gd ? gd[blockIdx.z].sh.nout : g.nout;
  const int trial0 = gd ? (int)gd[blockIdx.z].trial0 : 0;
  x += (size_t)(gd ? gd[blockIdx.z].xslot : blockIdx.z) * g.n_live * g.xpitch;
  out += (size_t)blockIdx.z * g.n_dm * g.ppitch;
  const int tid = threadIdx.x;
  const int wave = tid >> 6;
  const int lane = tid & 63;
  // Workgroups are dispatched round-robin over the 8 XCDs in linear order.
  // With the remap, XCD x takes the contiguous run [start_x, start_x + n_x)
  // of the (trial-block fastest) tile order, so the trial blocks it runs
  // together are neighbours whose input windows overlap in its own L2.
  const unsigned nbx = gridDim.x;
  unsigned lin = blockIdx.y * nbx + blockIdx.x;
  if (xcd_remap) {
    const unsigned nblk = nbx * gridDim.y;
    const unsigned q = nblk / 8, r = nblk % 8, x = lin % 8;
    lin = x * q + min(x, r) + lin / 8;
  }
  const int mb = (int)(lin % nbx);
  const int m0 = mb * DB;
  const size_t t0 = (size_t)(lin / nbx) * D2_TB;
  if (t0 > nout) return;  // past this group's series (and its zero at nout)
  const int* drow0 = delays + (size_t)(trial0 + m0) * g.n_live;

  uint32_t pa[NA][D2_Q];
  uint32_t ua[NA][D2_Q][2];
#pragma unroll
  for (int a = 0; a < NA; ++a)
#pragma unroll
    for (int q = 0; q < D2_Q; ++q) {
      pa[a][q] = 0;
      ua[a][q][0] = 0;
      ua[a][q][1] = 0;
    }
  int pending = 0;

  // Staging of a channel chunk: item (cc, j) turns source dwords j, j+1 of
  // channel c0+cc into E/O dwords 2j, 2j+1.  The global loads of chunk k+1
  // are issued into registers before chunk k is summed, so their latency
  // hides behind the LDS work.
  const int nchunk = (g.n_live + CC - 1) / CC;
  const int* cpairs = chunk_pairs + (size_t)(trial0 / DB + mb) * nchunk;
  // D2_BS/CC threads per staged channel; thread jj handles pairs jj + TPC*r
  constexpr int TPC = D2_BS / CC;
  const int scc = tid / TPC;
  const int sjj = tid % TPC;
  uint32_t pf_lo[D2_PF], pf_hi[D2_PF];
  auto prefetch = [&](int k) {
    if (k >= nchunk) return;
    const int c = k * CC + scc;
    if (c < g.n_live) {
      const int pairs = cpairs[k];
      const uint32_t* src = reinterpret_cast<const uint32_t*>(
          x + (size_t)c * g.xpitch + t0 + (drow0[c] & ~3));
#pragma unroll
      for (int r = 0; r < D2_PF; ++r) {
        const int j = sjj + r * TPC;
        if (j < pairs) {
          pf_lo[r] = src[j];
          pf_hi[r] = src[j + 1];
        }
      }
    }
  };
  // bytes b0..b3 of lo and b4 of hi -> E pair (b0,b1)(b2,b3), O pair (b1,b2)(b3,b4)
  auto put = [&](int j, uint32_t lo, uint32_t hi) {
    uint32_t* E = lds16 + (size_t)(scc * 2) * hw_max;
    uint32_t* O = E + hw_max;
    uint2 e, o;
    e.x = __builtin_amdgcn_perm(hi, lo, 0x0C010C00u);
    e.y = __builtin_amdgcn_perm(hi, lo, 0x0C030C02u);
    o.x = __builtin_amdgcn_perm(hi, lo, 0x0C020C01u);
    o.y = __builtin_amdgcn_perm(hi, lo, 0x0C040C03u);
    *reinterpret_cast<uint2*>(E + 2 * j) = e;
    *reinterpret_cast<uint2*>(O + 2 * j) = o;
  };
  prefetch(0);
  int pf_delay = 0;
  auto prefetch_delay = [&](int k) {
    if (k < nchunk && tid < DB * CC) {
      const int c = min(k * CC + tid % CC, g.n_live - 1);
      const int mm = min(m0 + tid / CC, g.n_dm - 1);
      // the dword of the trial's staged copy (E or O of channel slot cc)
      // where lane 0 reads: cc*2 + (o&1) selects the copy, o>>1 the pair
      const int o = delays[(size_t)(trial0 + mm) * g.n_live + c] - (drow0[c] & ~3);
      pf_delay = ((tid % CC) * 2 + (o & 1)) * hw_max + (o >> 1);
    }
  };
  prefetch_delay(0);
  for (int k = 0; k < nchunk; ++k) {
    const int c0 = k * CC;
    const int ncc = min(CC, g.n_live - c0);
    const int pairs = cpairs[k];
    __syncthreads();
    if (tid < DB * CC) sdly[tid / CC][tid % CC] = pf_delay;
    if (scc < ncc) {
#pragma unroll
      for (int r = 0; r < D2_PF; ++r) {
        const int j = sjj + r * TPC;
        if (j < pairs) put(j, pf_lo[r], pf_hi[r]);
      }
      const uint32_t* src = reinterpret_cast<const uint32_t*>(
          x + (size_t)(c0 + scc) * g.xpitch + t0 + (drow0[c0 + scc] & ~3));
      for (int j = sjj + D2_PF * TPC; j < pairs; j += TPC) put(j, src[j], src[j + 1]);
    }
    __syncthreads();
    prefetch(k + 1);
    prefetch_delay(k + 1);
    if (ncc == CC) {
      // a full chunk, four channels at a time: their offsets for the wave's
      // trials in one 16-byte read per trial
#pragma unroll
      for (int v = 0; v < CC / 4; ++v) {
        int off[NA][4];
#pragma unroll
        for (int a = 0; a < NA; ++a) {
          const int4 o4 = *reinterpret_cast<const int4*>(&sdly[wave + D2_WAVES * a][4 * v]);
          off[a][0] = o4.x;
          off[a][1] = o4.y;
          off[a][2] = o4.z;
          off[a][3] = o4.w;
        }
#pragma unroll
        for (int cc = 0; cc < 4; ++cc) {
#pragma unroll
          for (int a = 0; a < NA; ++a) {
            const uint32_t* src = lds16 + off[a][cc] + lane;
#pragma unroll
            for (int q = 0; q < D2_Q; ++q) pa[a][q] += src[64 * q];
          }
        }
      }
    } else {
      for (int cc = 0; cc < ncc; ++cc) {
#pragma unroll
        for (int a = 0; a < NA; ++a) {
          const uint32_t* src = lds16 + sdly[wave + D2_WAVES * a][cc] + lane;
#pragma unroll
          for (int q = 0; q < D2_Q; ++q) pa[a][q] += src[64 * q];
        }
      }
    }
    pending += ncc;
    if (pending > 256 - CC || c0 + CC >= g.n_live) {
#pragma unroll
      for (int a = 0; a < NA; ++a)
#pragma unroll
        for (int q = 0; q < D2_Q; ++q) {
          ua[a][q][0] += pa[a][q] & 0xFFFFu;
          ua[a][q][1] += pa[a][q] >> 16;
          pa[a][q] = 0;
        }
      pending = 0;
    }
  }
#pragma unroll
  for (int a = 0; a < NA; ++a) {
    const int m = m0 + wave + D2_WAVES * a;
    if (m >= g.n_dm) continue;
    float* orow = out + (size_t)m * g.ppitch;
#pragma unroll
    for (int q = 0; q < D2_Q; ++q) {
      const size_t t = t0 + 2 * lane + 128 * q;
      if (t + 1 < g.ppitch) {
        float2 v;
        v.x = (t < nout) ? (float)ua[a][q][0] : 0.f;
        v.y = (t + 1 < nout) ? (float)ua[a][q][1] : 0.f;
        *reinterpret_cast<float2*>(orow + t) = v;
      }
    }
  }
}

// ---------------------------------------------------------------------------
// Dedispersion from byte-staged samples.  Same block shape as k_dedisperse16,
// but each staged dword holds 4 samples: the channel's byte window is staged
// as 4 copies shifted by 0..3 bytes, C_k[j] = (x[4j+k], .., x[4j+k+3]), so a
// trial's offset o reads copy o & 3 at dword (o>>2) + lane + 64q: half the
// LDS reads of the 16-bit kernel for the same sums.
//
// SPLIT = false (sub-byte samples, values <= 15): one 32-bit add accumulates
// 4 samples in 8-bit lanes, which never carry within `widen` channels
// (widen * max sample <= 255); they are then added into 32-bit sums.
//
// SPLIT = true (8-bit samples): two sums per dword.  A = sum of the raw dwords
// (mod 2^32) and O = sum of bytes 1 and 3 moved to 16-bit lanes, (v >> 8) &
// 0x00ff00ff (one byte permute).  O's lanes hold at most 257 * 255, and the
// even bytes' 16-bit-lane sum is E = A - (O << 8) mod 2^32, exact because E
// itself fits 32 bits; so after up to 256 channels E and O give all 4
// samples' sums.  One LDS read, one permute and one add per channel per
// 4 samples (the adds pair up as 3-operand adds).
constexpr int D8_Q = 4;                   // dwords (4 samples each) per lane per trial
// (32-trial tiles at 3 waves; the split sums at 32 trials need 2)
constexpr int d8_waves(int db, int cc, bool split) {
  return db == 32 ? (split ? 2 : 3) : (split && db == 16 && cc >= 8 ? 3 : 4);
}
constexpr int D8_PF = 2;                  // staging items (4 dwords) prefetched per thread

// DBUF: two staging buffers, for DM blocks whose windows fit twice in the
// LDS.  Chunk k + 1 is staged into one buffer while chunk k is summed from
// the other, with one barrier per chunk instead of two; the blocks of a
// launch are DM blocks g.mb_off .. of the grid's.
template <int DB, int CC, bool SPLIT, int BS = D2_BS, bool DBUF = false>
__global__ __launch_bounds__(BS, BS == D2_BS ? d8_waves(DB, CC, SPLIT) : 4) void k_dedisperse8(const uint8_t* __restrict__ x,
                                                      const int* __restrict__ delays,
                                                      DedispGeom g, int hw_max, int widen,
                                                      float* __restrict__ out,
                                                      const GroupDev* __restrict__ gd) {
  // [buffer][CC][4][hw_max] (one buffer, or two with DBUF), then an int2 per
  // live channel
  extern __shared__ uint32_t lds8[];
  constexpr int WAVES = BS / 64;
  constexpr int NA = DB / WAVES;
  constexpr int NBUF = DBUF ? 2 : 1;
  __shared__ __align__(16) int sdly[NBUF][DB][CC];
  const int stage_words = CC * 4 * hw_max;
  int2* ctab = reinterpret_cast<int2*>(lds8 + (size_t)NBUF * stage_words);
  // groups chosen per group: this instance takes only its own kind
  if (g.dbuf_groups && gd && (gd[blockIdx.z].dbuf != 0) != DBUF) return;
  const size_t nout = gd ? gd[blockIdx.z].sh.nout : g.nout;
  const int trial0 = gd ? (int)gd[blockIdx.z].trial0 : 0;
  x += (size_t)(gd ? gd[blockIdx.z].xslot : blockIdx.z) * g.n_live * g.xpitch;
  out += (size_t)blockIdx.z * g.n_dm * g.ppitch;
  const int tid = threadIdx.x;
  const int wave = tid >> 6;
  const int lane = tid & 63;
  const unsigned nbx = gridDim.x;
  unsigned lin = blockIdx.y * nbx + blockIdx.x;
  if (g.remap) {
    const unsigned nblk = nbx * gridDim.y;
    const unsigned q = nblk / 8, r = nblk % 8, xx = lin % 8;
    lin = xx * q + min(xx, r) + lin / 8;
  }
  const int mb = (int)(lin % nbx) + g.mb_off;
  const int m0 = mb * DB;
  const int ml = min(m0 + DB, g.n_dm) - 1;
  const size_t t0 = (size_t)(lin / nbx) * D2_TB;
  if (t0 > nout) return;
  const int* drow0 = delays + (size_t)(trial0 + m0) * g.n_live;
  const int* drowl = delays + (size_t)(trial0 + ml) * g.n_live;
  // Per channel, once per block: the window's first sample (the first
  // trial's delay rounded down to a dword) and its staging items.  The window
  // spans the block's delays plus 1024 samples; staging item u gives dwords
  // 4u .. 4u+3 of each of the 4 copies from source dwords 4u .. 4u+4, one
  // 16-byte LDS write per copy (the copies' rows are 16-byte aligned: hw_max
  // is a multiple of 64).  The last item may read up to 7 dwords past the
  // window, inside the row's padding.  Keeping these in LDS takes the delay
  // table's global-memory latency off the per-chunk staging.
  for (int c = tid; c < g.n_live; c += BS) {
    const int b = drow0[c] & ~3;
    const int ndw = (D2_TB + drowl[c] - b + 3) / 4 + 1;
    ctab[c] = make_int2(b, (ndw + 3) / 4);
  }

  uint32_t pa[NA][D8_Q];
  uint32_t po[NA][SPLIT ? D8_Q : 1];  // odd bytes in 16-bit lanes (SPLIT)
  uint32_t ua[NA][D8_Q][4];
#pragma unroll
  for (int a = 0; a < NA; ++a)
#pragma unroll
    for (int q = 0; q < D8_Q; ++q) {
      pa[a][q] = 0;
      if (SPLIT) po[a][SPLIT ? q : 0] = 0;
#pragma unroll
      for (int e = 0; e < 4; ++e) ua[a][q][e] = 0;
    }
  const int nchunk = (g.n_live + CC - 1) / CC;
  constexpr int TPC = BS / CC;
  const int scc = tid / TPC;
  const int sjj = tid % TPC;
  __syncthreads();
  uint32_t pf[D8_PF][5];
  int pf_ng = 0;  // items of this thread's channel in the prefetched chunk; -1: past the last
  auto prefetch = [&](int k) {
    const int c = k * CC + scc;
    pf_ng = 0;
    if (k >= nchunk) return;
    if (c >= g.n_live) {
      pf_ng = -1;
      return;
    }
    const int2 t = ctab[c];
    pf_ng = t.y;
    // item sjj + r * TPC at a constant offset from item sjj
    const uint32_t* src =
        reinterpret_cast<const uint32_t*>(x + (size_t)c * g.xpitch + t0 + t.x) + 4 * sjj;
#pragma unroll
    for (int r = 0; r < D8_PF; ++r) {
      if (sjj + r * TPC < t.y) {
#pragma unroll
        for (int i = 0; i < 5; ++i) pf[r][i] = src[4 * r * TPC + i];
      }
    }
  };
  // copy k, dword j: bytes 4j+k .. 4j+k+3 of the window (in buffer `buf`)
  auto put4 = [&](int buf, int u, const uint32_t (&d)[5]) {
    uint32_t* C = lds8 + (size_t)buf * stage_words + (size_t)(scc * 4) * hw_max + 4 * u;
    *reinterpret_cast<uint4*>(C) = make_uint4(d[0], d[1], d[2], d[3]);
#pragma unroll
    for (int k = 1; k < 4; ++k)
      *reinterpret_cast<uint4*>(C + k * hw_max) = make_uint4(
          __builtin_amdgcn_alignbyte(d[1], d[0], k), __builtin_amdgcn_alignbyte(d[2], d[1], k),
          __builtin_amdgcn_alignbyte(d[3], d[2], k), __builtin_amdgcn_alignbyte(d[4], d[3], k));
  };
  // the raw delay of (trial tid / CC, channel tid % CC) of a chunk, loaded
  // one chunk ahead of its use
  int rd = 0;
  auto load_delay = [&](int k) {
    if (k < nchunk && tid < DB * CC) {
      const int c = min(k * CC + tid % CC, g.n_live - 1);
      const int mm = min(m0 + tid / CC, g.n_dm - 1);
      rd = delays[(size_t)(trial0 + mm) * g.n_live + c];
    }
  };
  // stage chunk kk into buffer buf: its trial offsets from rd, its window
  // from the prefetched items (and the rest loaded here), or zeros for
  // channels past the last live one
  auto stage = [&](int kk, int buf) {
    const int c0 = kk * CC;
    if (tid < DB * CC) {
      const int c = c0 + tid % CC;
      int p = buf * stage_words + (tid % CC) * 4 * hw_max;  // past the last channel: copy 0, zeros
      if (c < g.n_live) {
        const int o = rd - ctab[c].x;
        p += (o & 3) * hw_max + (o >> 2);
      }
      sdly[buf][tid / CC][tid % CC] = p;
    }
    if (pf_ng > 0) {
#pragma unroll
      for (int r = 0; r < D8_PF; ++r) {
        const int u = sjj + r * TPC;
        if (u < pf_ng) put4(buf, u, pf[r]);
      }
      if (pf_ng > D8_PF * TPC) {
        const uint32_t* src = reinterpret_cast<const uint32_t*>(
            x + (size_t)(c0 + scc) * g.xpitch + t0 + ctab[c0 + scc].x);
        for (int u = sjj + D8_PF * TPC; u < pf_ng; u += TPC) {
          uint32_t d[5];
#pragma unroll
          for (int i = 0; i < 5; ++i) d[i] = src[4 * u + i];
          put4(buf, u, d);
        }
      }
    } else if (pf_ng < 0) {
      // the 64 * D8_Q dwords a trial reads at offset 0
      uint32_t* C = lds8 + (size_t)buf * stage_words + (size_t)(scc * 4) * hw_max;
      for (int u = sjj; u < 64 * D8_Q / 4; u += TPC)
        *reinterpret_cast<uint4*>(C + 4 * u) = make_uint4(0u, 0u, 0u, 0u);
    }
  };
  auto add = [&](int a, const uint32_t* src) {
#pragma unroll
    for (int q = 0; q < D8_Q; ++q) {
      const uint32_t v = src[64 * q];
      pa[a][q] += v;
      if (SPLIT) po[a][SPLIT ? q : 0] += __builtin_amdgcn_perm(0u, v, 0x0C030C01u);
    }
  };
  // sums of the chunk staged in buffer buf
  auto sum_chunk = [&](int buf) {
    // four channels at a time: their offsets for the wave's trials in one
    // 16-byte read per trial
#pragma unroll
    for (int v = 0; v < CC / 4; ++v) {
      int off[NA][4];
#pragma unroll
      for (int a = 0; a < NA; ++a) {
        const int4 o4 = *reinterpret_cast<const int4*>(&sdly[buf][wave + WAVES * a][4 * v]);
        off[a][0] = o4.x;
        off[a][1] = o4.y;
        off[a][2] = o4.z;
        off[a][3] = o4.w;
      }
#pragma unroll
      for (int cc = 0; cc < 4; ++cc)
#pragma unroll
        for (int a = 0; a < NA; ++a) add(a, lds8 + off[a][cc] + lane);
    }
  };
  prefetch(0);
  load_delay(0);
  if (DBUF) {
    // chunk 0 staged ahead of the loop
    stage(0, 0);
    prefetch(1);
    load_delay(1);
    __syncthreads();
  }
  // Chunks in runs of widen / CC between widenings (widen >= CC), so that the
  // summing loop has no widening branch; a chunk's channels past the last
  // live one are staged as zeros.
  const int per_widen = widen / CC;
  int k = 0;
  while (k < nchunk) {
    const int kend = min(nchunk, k + per_widen);
    for (; k < kend; ++k) {
      if (DBUF) {
        // chunk k + 1 into the other buffer (every read of it, chunk k - 1,
        // ended at the previous barrier), then chunk k's sums
        if (k + 1 < nchunk) stage(k + 1, (k + 1) & 1);
        prefetch(k + 2);
        load_delay(k + 2);
        sum_chunk(k & 1);
        __syncthreads();
      } else {
        __syncthreads();
        stage(k, 0);
        // the next chunk's loads are issued before the barrier, as early as
        // this chunk's staging registers are free
        prefetch(k + 1);
        load_delay(k + 1);
        __syncthreads();
        sum_chunk(0);
      }
    }
#pragma unroll
    for (int a = 0; a < NA; ++a)
#pragma unroll
      for (int q = 0; q < D8_Q; ++q) {
        if (SPLIT) {
          const uint32_t od = po[a][SPLIT ? q : 0];
          const uint32_t ev = pa[a][q] - (od << 8);
          ua[a][q][0] += ev & 0xFFFFu;
          ua[a][q][1] += od & 0xFFFFu;
          ua[a][q][2] += ev >> 16;
          ua[a][q][3] += od >> 16;
          po[a][SPLIT ? q : 0] = 0;
        } else {
#pragma unroll
          for (int e = 0; e < 4; ++e) ua[a][q][e] += (pa[a][q] >> (8 * e)) & 0xFFu;
        }
        pa[a][q] = 0;
      }
  }
#pragma unroll
  for (int a = 0; a < NA; ++a) {
    const int m = m0 + wave + WAVES * a;
    if (m >= g.n_dm) continue;
    float* orow = out + (size_t)m * g.ppitch;
#pragma unroll
    for (int q = 0; q < D8_Q; ++q) {
      const size_t t = t0 + 4 * (lane + 64 * q);
      if (t + 3 < g.ppitch) {
        float4 v;
        v.x = (t + 0 < nout) ? (float)ua[a][q][0] : 0.f;
        v.y = (t + 1 < nout) ? (float)ua[a][q][1] : 0.f;
        v.z = (t + 2 < nout) ? (float)ua[a][q][2] : 0.f;
        v.w = (t + 3 < nout) ? (float)ua[a][q][3] : 0.f;
        *reinterpret_cast<float4*>(orow + t) = v;
      }
    }
  }
}

// Adaptive time resolution: one thread per 4 output samples of a row, from
// one 8-byte load (xpitch is a multiple of 256, so every load is aligned and
// inside the row).  The pairwise mean rounds down, so samples keep their bit
// width.  Bytes limit .. xpitch of the output rows are left as they are.
__global__ __launch_bounds__(256) void k_scrunch2(const uint8_t* __restrict__ x, size_t xpitch,
                                                  size_t nsamps, uint8_t* __restrict__ y,
                                                  size_t limit) {
  const size_t t4 = ((size_t)blockIdx.x * 256 + threadIdx.x) * 4;
  const size_t row = blockIdx.y;
  if (t4 >= limit) return;
  const size_t nh = nsamps / 2;
  uint32_t out = 0;
  if (t4 < nh) {
    const uint2 w = *reinterpret_cast<const uint2*>(x + row * xpitch + 2 * t4);
    const uint32_t b[8] = {w.x & 0xffu, (w.x >> 8) & 0xffu, (w.x >> 16) & 0xffu, w.x >> 24,
                           w.y & 0xffu, (w.y >> 8) & 0xffu, (w.y >> 16) & 0xffu, w.y >> 24};
#pragma unroll
    for (int k = 0; k < 4; ++k)
      if (t4 + k < nh) out |= ((b[2 * k] + b[2 * k + 1]) >> 1) << (8 * k);
  }
  *reinterpret_cast<uint32_t*>(y + row * xpitch + t4) = out;
}

}  // namespace

hipError_t launch_scrunch2(const uint8_t* x, int n_live, size_t xpitch, size_t nsamps,
                           uint8_t* y, hipStream_t s, size_t limit) {
  // (a thread's 8-byte load ends at most 5 bytes past the row's last sample)
  if (xpitch % 256 || nsamps + 8 > xpitch || n_live <= 0) return hipErrorInvalidValue;
  if (limit == 0 || limit > xpitch) limit = xpitch;
  if (limit % 4 || limit < nsamps / 2) return hipErrorInvalidValue;
  dim3 grid((unsigned)((limit / 4 + 255) / 256), (unsigned)n_live);
  hipLaunchKernelGGL(k_scrunch2, grid, dim3(256), 0, s, x, xpitch, nsamps, y, limit);
  return hipGetLastError();
}

int dedisp_tile_time() { return D2_TB; }
// Default trials per block (the pipeline's autotuner may choose 8 or 32).
int dedisp_tile_dms() { return D2_DB; }
int dedisp_chunk_channels() { return D2_CC; }
int dedisp8_window_dwords(int spread) {
  // copies of 4 samples per dword: the 1024-sample tile plus the span, one
  // extra dword for the shifted copies' last bytes, rounded to 64
  const int need = (D2_TB + spread + 3) / 4 + 1;
  return (need + 63) / 64 * 64;
}
int dedisp_window_dwords(int spread) {
  // E/O copies: dword (o>>1) + 511 with o <= spread + 3, rounded to 64
  const int need = (spread + 3) / 2 + 512 + 1;
  return (need + 63) / 64 * 64;
}

hipError_t launch_unpack(const uint8_t* fb, size_t stride_bytes, int nbits, size_t nsamps,
                         const int* d_live, int n_live, uint8_t* x, size_t xpitch,
                         hipStream_t s, const GroupArgs& ga, bool live_is_identity,
                         bool trim_rows) {
  const size_t trim = trim_rows ? (size_t)2 * D2_TB : 0;
  if (nbits == 8 && live_is_identity && xpitch % 4 == 0) {
    const int tcw = n_live <= UP_ROW_MAX ? std::max(4, (n_live + 3) / 4 * 4) : 256;
    dim3 grid((unsigned)((xpitch + UP_T - 1) / UP_T), (unsigned)((n_live + tcw - 1) / tcw),
              (unsigned)ga.G);
    const size_t lds = (size_t)UP_T * (tcw / 4 + 1) * sizeof(uint32_t);
    hipLaunchKernelGGL(k_unpack8, grid, dim3(256), lds, s, fb, stride_bytes, nsamps, n_live, x,
                       xpitch, ga.gd, trim, tcw);
    return hipGetLastError();
  }
  dim3 grid((unsigned)((xpitch + 63) / 64), (unsigned)((n_live + 63) / 64), (unsigned)ga.G);
  hipLaunchKernelGGL(k_unpack, grid, dim3(256), 0, s, fb, stride_bytes, nbits, nsamps, d_live,
                     n_live, x, xpitch, ga.gd, trim);
  return hipGetLastError();
}

// Raises a kernel's dynamic-LDS limit once (never during a later capture:
// every shape is launched eagerly before its chain is captured).
hipError_t lds_limit(const void* fn, int slot, size_t lds) {
  static size_t set[40] = {0};
  if (lds <= 64 * 1024 || lds <= set[slot]) return hipSuccess;
  const hipError_t e = hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
  if (e == hipSuccess) set[slot] = lds;
  return e;
}

// Kernel instances by tile index: (log2(db / 8)) + 3 * log2(cc / 4).
int tile_slot(int db, int cc) {
  return (db == 8 ? 0 : db == 16 ? 1 : 2) + (cc == 4 ? 0 : cc == 8 ? 3 : 6);
}
template <int DB, int CC>
const void* fn16() {
  return reinterpret_cast<const void*>(&k_dedisperse16<DB, CC>);
}
template <int DB, int CC, bool SPLIT, int BS = D2_BS, bool DBUF = false>
const void* fn8() {
  return reinterpret_cast<const void*>(&k_dedisperse8<DB, CC, SPLIT, BS, DBUF>);
}
const void* kernel16(int slot) {
  static const void* const t[9] = {fn16<8, 4>(),  fn16<16, 4>(),  fn16<32, 4>(),
                                   fn16<8, 8>(),  fn16<16, 8>(),  fn16<32, 8>(),
                                   fn16<8, 16>(), fn16<16, 16>(), fn16<32, 16>()};
  return t[slot];
}
const void* kernel8(int slot, bool split) {
  static const void* const t[9] = {
      fn8<8, 4, false>(),  fn8<16, 4, false>(),  fn8<32, 4, false>(),
      fn8<8, 8, false>(),  fn8<16, 8, false>(),  fn8<32, 8, false>(),
      fn8<8, 16, false>(), fn8<16, 16, false>(), fn8<32, 16, false>()};
  static const void* const ts[9] = {
      fn8<8, 4, true>(),  fn8<16, 4, true>(),  fn8<32, 4, true>(),
      fn8<8, 8, true>(),  fn8<16, 8, true>(),  fn8<32, 8, true>(),
      fn8<8, 16, true>(), fn8<16, 16, true>(), fn8<32, 16, true>()};
  return split ? ts[slot] : t[slot];
}
// 32 trials on 1024 threads (16 waves, 2 trials each), by channels per step
const void* kernel8_wide(int cc, bool split) {
  static const void* const t[3] = {fn8<32, 4, false, 1024>(), fn8<32, 8, false, 1024>(),
                                   fn8<32, 16, false, 1024>()};
  static const void* const ts[3] = {fn8<32, 4, true, 1024>(), fn8<32, 8, true, 1024>(),
                                    fn8<32, 16, true, 1024>()};
  const int k = cc == 4 ? 0 : (cc == 8 ? 1 : 2);
  return split ? ts[k] : t[k];
}
// the same with two staging buffers
const void* kernel8_wide_dbuf(int cc, bool split) {
  static const void* const t[3] = {fn8<32, 4, false, 1024, true>(),
                                   fn8<32, 8, false, 1024, true>(),
                                   fn8<32, 16, false, 1024, true>()};
  static const void* const ts[3] = {fn8<32, 4, true, 1024, true>(), fn8<32, 8, true, 1024, true>(),
                                    fn8<32, 16, true, 1024, true>()};
  const int k = cc == 4 ? 0 : (cc == 8 ? 1 : 2);
  return split ? ts[k] : t[k];
}

// Dynamic LDS of the byte kernel: the staged copies (nbuf buffers of wlen8
// dwords per copy), then each live channel's window (int2).
static size_t dedisp8_dynamic_lds(const DedispGeom& g, int nbuf = 1, int wlen8 = 0) {
  return (size_t)nbuf * g.cc * 4 * (size_t)(wlen8 ? wlen8 : g.wlen8) * sizeof(uint32_t) +
         (size_t)g.n_live * 8;
}

// LDS of a tile: the staged copies (dynamic) plus the trial offsets (static).
size_t dedisp_lds_bytes(const DedispGeom& g) {
  const size_t offs = (size_t)g.db * g.cc * sizeof(int);
  if (g.widen8 > 0) return dedisp8_dynamic_lds(g) + offs;
  return (size_t)g.cc * 2 * (size_t)g.wlen * sizeof(uint32_t) + offs;
}
size_t dedisp_lds_bytes_dbuf(const DedispGeom& g, int wlen8) {
  return dedisp8_dynamic_lds(g, 2, wlen8) + 2 * (size_t)g.db * g.cc * sizeof(int);
}

hipError_t launch_dedisperse(const uint8_t* x, const int* d_delays, const int* d_chunk_pairs,
                             const DedispGeom& g, float* plane, hipStream_t s,
                             const GroupArgs& ga) {
  // DM blocks fastest: blocks resident together share one time tile's input
  const int db = g.db, cc = g.cc;
  if (db != 8 && db != 16 && db != 32) return hipErrorInvalidValue;
  if (cc != 4 && cc != 8 && cc != 16) return hipErrorInvalidValue;
  // time tiles: the row pitch's, or fewer when every group's series ends
  // sooner (the later tiles would only return)
  const size_t tiles =
      (g.tiles && g.tiles < g.ppitch / D2_TB) ? g.tiles : g.ppitch / D2_TB;
  dim3 grid((unsigned)((g.n_dm + db - 1) / db), (unsigned)tiles,
            (unsigned)ga.G);
  const int slot = tile_slot(db, cc);
  // the delay span of a DM block too wide for this tile's staged copies
  if (dedisp_lds_bytes(g) > 160 * 1024) return hipErrorInvalidValue;
  if (g.widen8 > 0) {
    // (a copy row holds at least the 64 * D8_Q dwords a trial reads)
    if (g.widen8 < cc || g.wlen8 < 64 * D8_Q || g.wlen8 % 64) return hipErrorInvalidValue;
    if (g.split8 && g.widen8 > 256) return hipErrorInvalidValue;   // 16-bit lanes would carry
    const size_t lds8 = dedisp8_dynamic_lds(g);
    const bool wide = g.wide && db == 32;
    using K8 = void (*)(const uint8_t*, const int*, DedispGeom, int, int, float*, const GroupDev*);
    // the leading g.dbuf_blocks DM blocks (windows of at most g.wlen8_dbuf
    // dwords) on the double-buffered kernel, the rest on the single one
    const bool dbuf_fits = wide && g.wlen8_dbuf >= 64 * D8_Q && g.wlen8_dbuf % 64 == 0 &&
                           dedisp_lds_bytes_dbuf(g, g.wlen8_dbuf) <= 160 * 1024;
    if (g.dbuf_groups) {
      // per group: both instances over the whole grid, each taking its groups
      if (!dbuf_fits || !ga.gd) return hipErrorInvalidValue;
      const int ci = cc == 4 ? 0 : cc == 8 ? 1 : 2;
      const void* fd = kernel8_wide_dbuf(cc, g.split8 != 0);
      const size_t ldsd = dedisp8_dynamic_lds(g, 2, g.wlen8_dbuf);
      hipError_t ed = lds_limit(fd, 33 + ci + (g.split8 ? 3 : 0), ldsd);
      if (ed != hipSuccess) return ed;
      DedispGeom g1 = g;
      g1.mb_off = 0;
      hipLaunchKernelGGL(reinterpret_cast<K8>(const_cast<void*>(fd)), grid, dim3(1024), ldsd, s, x,
                         d_delays, g1, g.wlen8_dbuf, g.widen8, plane, ga.gd);
      ed = hipGetLastError();
      if (ed != hipSuccess) return ed;
      const void* f8 = kernel8_wide(cc, g.split8 != 0);
      ed = lds_limit(f8, 27 + ci + (g.split8 ? 3 : 0), lds8);
      if (ed != hipSuccess) return ed;
      hipLaunchKernelGGL(reinterpret_cast<K8>(const_cast<void*>(f8)), grid, dim3(1024), lds8, s,
                         x, d_delays, g1, g.wlen8, g.widen8, plane, ga.gd);
      return hipGetLastError();
    }
    const int nb_dbuf = dbuf_fits ? min(g.dbuf_blocks, (int)grid.x) : 0;
    if (nb_dbuf > 0) {
      const int ci = cc == 4 ? 0 : cc == 8 ? 1 : 2;
      const void* fd = kernel8_wide_dbuf(cc, g.split8 != 0);
      const size_t ldsd = dedisp8_dynamic_lds(g, 2, g.wlen8_dbuf);
      const hipError_t ed = lds_limit(fd, 33 + ci + (g.split8 ? 3 : 0), ldsd);
      if (ed != hipSuccess) return ed;
      DedispGeom gd1 = g;
      gd1.mb_off = 0;
      hipLaunchKernelGGL(reinterpret_cast<K8>(const_cast<void*>(fd)),
                         dim3((unsigned)nb_dbuf, grid.y, grid.z), dim3(1024), ldsd, s, x, d_delays,
                         gd1, g.wlen8_dbuf, g.widen8, plane, ga.gd);
      const hipError_t el = hipGetLastError();
      if (el != hipSuccess || nb_dbuf == (int)grid.x) return el;
    }
    const void* f8 = wide ? kernel8_wide(cc, g.split8 != 0) : kernel8(slot, g.split8 != 0);
    const hipError_t e =
        lds_limit(f8, wide ? 27 + (cc == 4 ? 0 : cc == 8 ? 1 : 2) + (g.split8 ? 3 : 0)
                           : (g.split8 ? 18 : 9) + slot,
                  lds8);
    if (e != hipSuccess) return e;
    DedispGeom g2 = g;
    g2.mb_off = nb_dbuf;
    hipLaunchKernelGGL(reinterpret_cast<K8>(const_cast<void*>(f8)),
                       dim3(grid.x - (unsigned)nb_dbuf, grid.y, grid.z), dim3(wide ? 1024 : D2_BS),
                       lds8, s, x, d_delays, g2, g.wlen8, g.widen8, plane, ga.gd);
    return hipGetLastError();
  }
  const size_t lds = (size_t)cc * 2 * (size_t)g.wlen * sizeof(uint32_t);
  const void* fn = kernel16(slot);
  const hipError_t e = lds_limit(fn, slot, lds);
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(reinterpret_cast<void (*)(const uint8_t*, const int*, const int*, DedispGeom,
                                               int, int, float*, const GroupDev*)>(
                         const_cast<void*>(fn)),
                     grid, dim3(D2_BS), lds, s, x, d_delays, d_chunk_pairs, g, g.wlen, g.remap,
                     plane, ga.gd);
  return hipGetLastError();
}

}  // namespace hd
