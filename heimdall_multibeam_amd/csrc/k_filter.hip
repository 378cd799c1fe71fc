// k_filter.hip -- boxcar filtering, block RMS and giant finding for all
// (DM, width) pairs of a gulp.
//
// The boxcar output of width w at index i is never materialised; every
// consumer evaluates it from the prefix sums (Pipeline/matched_filter.cu:
// 43-64):  F_w[i] = S[off + ahead + i*ts] - S[off - behind + i*ts].
//
//   block RMS  GetRMSPlanMB::exec_multibeam pooled form (get_rms.cu:104-131):
//              each block reduces a 5^L0-leaf chunk of |F_w| to one value for
//              every width of a group at once (the chunk's prefix sums are
//              staged in LDS once); a finishing wave per (DM, width) runs the
//              remaining levels and the small-count rule, then
//              rms = (float)((double)med * 1.4862) and inv = 1.0f/rms.
//              Widths with ts > 1 read what the reference's filtered buffer
//              holds past the fresh decimated values: the normalised output of
//              the latest earlier width that reached that position
//              (pipeline.cu:680-712 reuse d_filtered_series).
//   threshold  GiantFinder's selection (find_giants.cu:47-60,137-183): per
//              (DM, width, chunk of 8192 or 4096 positions) the samples above
//              threshold are compacted in time order and appended to an
//              unordered buffer; the per-chunk counts are scanned so that a
//              second pass lays the samples out in the reference's order
//              (DM, width, time).
//   segments   head flags where a new giant starts (adjacent gap > merge_dist,
//              find_giants.cu:79-87,197-210), their scan, and one thread per
//              giant reducing its run to the first maximum, begin and end
//              (:260-303), with the index rescaling of pipeline.cu:755-775.
#include "hd_device.hpp"
#include "hd_internal.hpp"

namespace hd {

namespace {

constexpr int FB_BS = 256;

// LDS words of one wave's level buffers on the block-RMS tile path.  With at
// least 25 leaves per chunk the first two levels run in registers and the
// buffers hold levels 2 and 3 (n1/5 and n1/25 values), else levels 1 and 2.
__host__ __device__ constexpr size_t block_rms_wave_words(size_t n1) {
  return n1 >= 5 ? n1 / 5 + n1 / 25 + 2 : n1 + n1 / 5 + 2;
}

__device__ __forceinline__ float raw_val(const float* S, const WidthInfo& wi, unsigned off,
                                         size_t i) {
  const size_t base = i * wi.ts;
  const float a = S[off + wi.ahead + base];
  const float b = S[off - wi.behind + base];
  return a - b;
}

// What the reference's filtered buffer holds at `pos` while width j is being
// processed: fresh output below n_w, otherwise the normalised value left by
// the latest earlier width that wrote `pos`.
// nwv: the group's fresh output counts per width.
__device__ __forceinline__ float filtered_at(const float* S, const WidthTable& wt,
                                             const unsigned* nwv, const float* winv, int j,
                                             size_t pos) {
  if (pos < nwv[j]) return raw_val(S, wt.wi[j], wt.off, pos);
  for (int jj = j - 1; jj >= 0; --jj)
    if (pos < nwv[jj]) return raw_val(S, wt.wi[jj], wt.off, pos) * winv[jj];
  return 0.f;
}

// dst[t] = src[t * stride] for t < n (LDS from global), with SU loads in
// flight per thread before their LDS writes: a plain loop waits out the
// memory latency once per element.  24 per thread stage a tile of up to 6144
// entries (a 3125-leaf chunk and a 2048-sample halo) in one round.
__device__ __forceinline__ void stage_lds(float* dst, const float* __restrict__ src, size_t n,
                                 size_t stride) {
  constexpr int SU = 24;
  for (size_t t0 = threadIdx.x; t0 < n; t0 += (size_t)SU * FB_BS) {
    float r[SU];
#pragma unroll
    for (int u = 0; u < SU; ++u) {
      const size_t t = t0 + (size_t)u * FB_BS;
      r[u] = t < n ? src[t * stride] : 0.f;
    }
#pragma unroll
    for (int u = 0; u < SU; ++u) {
      const size_t t = t0 + (size_t)u * FB_BS;
      if (t < n) dst[t] = r[u];
    }
  }
}

// One block per (chunk, DM): level-L0 medians of |F_w| for widths
// [first_w, first_w + n_group).  `tile_ok` selects the LDS-staged path for a
// group of ts == 1 widths.  A single decimated width (ts > 1) stages the
// prefix sums that the chunk's values come from: those of its own decimated
// boxcar where the chunk is fresh, or of the one earlier width whose
// normalised output the reference's buffer still holds there.  Either way
// the staged entries are every ts-th prefix sum, E[x] = S[off - behind +
// ts*(pos0 + x)], so that F[pos] = E[pos - pos0 + w/ts] - E[pos - pos0]
// (w = ahead + behind is a multiple of ts); up to tile_words of them.
// Grid: (chunks, trials, groups), blocks taken in the XCD-contiguous order;
// rows g*n_dm + m.
__global__ __launch_bounds__(FB_BS) void k_block_rms_partials(
    const float* __restrict__ Sall, size_t spitch, WidthTable wt, const GroupDev* __restrict__ gd,
    size_t np_max, int first_w, int n_group, int tile_ok, unsigned amax, unsigned bmax,
    unsigned tile_words, const float* __restrict__ winv_all, float* __restrict__ partial) {
  extern __shared__ float lds[];
  const BlockRmsShape bs = gd[blockIdx.z].bs;
  const unsigned* nwv = gd[blockIdx.z].n_w;
  const uint2 xb = xcd_contiguous_block();  // chunk halos shared within an XCD's L2
  const size_t c = xb.x;
  if (c >= bs.np) return;
  const size_t row = (size_t)blockIdx.z * gridDim.y + xb.y;
  const float* S = Sall + row * spitch;
  const float* winv = winv_all + row * wt.nw;
  const size_t k0 = c * bs.chunk;
  const size_t pos0 = reblock_pos((unsigned)k0, bs.bc, bs.delta);
  const size_t posl = reblock_pos((unsigned)(k0 + bs.chunk - 1), bs.bc, bs.delta);
  const bool contiguous = (posl - pos0 == bs.chunk - 1);
  const bool use_tile = tile_ok && contiguous;
  // single decimated width: which width's boxcar this chunk reads
  int src_j = -1;
  if (!tile_ok && n_group == 1 && contiguous && tile_words > 0) {
    const unsigned nwj = nwv[first_w];
    if (posl < nwj) {
      src_j = first_w;
    } else if (pos0 >= nwj) {
      int a = -1, b = -1;
      for (int jj = first_w - 1; jj >= 0 && a < 0; --jj)
        if (pos0 < nwv[jj]) a = jj;
      for (int jj = first_w - 1; jj >= 0 && b < 0; --jj)
        if (posl < nwv[jj]) b = jj;
      if (a == b) src_j = a;
    }
    if (src_j >= 0) {
      const WidthInfo ws = wt.wi[src_j];
      const size_t span = bs.chunk + (ws.ahead + ws.behind) / ws.ts;
      if (span > tile_words || (ws.ahead + ws.behind) % ws.ts != 0) src_j = -1;
    }
    // A stale chunk holds width src_j's normalised values |F * inv|, and the
    // chunk is all fresh for src_j, whose partial (an earlier launch) is the
    // level-L0 median of |F| here.  The pyramid only selects values and
    // rounding x * inv is monotone for a positive finite inv, so the median
    // of |F * inv| is that partial times inv, bit for bit.
    if (src_j >= 0 && src_j < first_w) {
      const float sv = partial[(row * wt.nw + src_j) * np_max + c];
      const float sc = winv[src_j];
      if (sc > 0.f && sc < INFINITY && sv < INFINITY) {  // (sv >= 0; NaN fails both)
        if (threadIdx.x == 0) partial[(row * wt.nw + first_w) * np_max + c] = sv * sc;
        return;
      }
    }
  }
  const size_t tlen = use_tile ? bs.chunk + amax + bmax : (src_j >= 0 ? tile_words : 0);
  float* T = lds;
  float* A = lds + tlen;
  const size_t n1 = bs.chunk / 5;
  float* B = A + (n1 ? n1 : 1);
  if (use_tile) {
    stage_lds(T, S + pos0 + wt.off - bmax, tlen, 1);
    __syncthreads();
  }
  unsigned wts = 0;
  float sscale = 1.f;
  if (src_j >= 0) {
    const WidthInfo ws = wt.wi[src_j];
    sscale = winv[src_j];
    wts = (ws.ahead + ws.behind) / ws.ts;
    const size_t span = bs.chunk + wts;
    stage_lds(T, S + wt.off - ws.behind + (size_t)ws.ts * pos0, span, ws.ts);
    __syncthreads();
  }
  // level-L0 median of the chunk's leaves for one width (each mode below
  // passes its own branch-free leaf function)
  // (the value is meaningful in thread 0, which writes it: the first level
  // runs block-wide, the others in wave 0 with no block-wide barrier)
  auto chunk_value = [&](auto leaf) -> float {
    if (bs.L0 == 0) return leaf(k0);
    for (size_t q = threadIdx.x; q < n1; q += FB_BS) {
      const size_t b = k0 + 5 * q;
      A[q] = med5(leaf(b), leaf(b + 1), leaf(b + 2), leaf(b + 3), leaf(b + 4));
    }
    __syncthreads();
    return threadIdx.x < 64 ? wave_pyramid(A, B, n1, bs.L0 - 1)[0] : 0.f;
  };
  if (use_tile) {
    // One wave per width (widths first_w + w, + 4, ...): the chunk's median
    // pyramid runs inside the wave, in its own LDS buffers, with no
    // block-wide barrier; the groups of 5 are those of the reference.
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const size_t n2 = n1 / 5;
    float* WA = A + (size_t)wave * block_rms_wave_words(n1);
    float* WB = WA + (n2 ? n2 : (n1 ? n1 : 1));
    for (int g = wave; g < n_group; g += FB_BS / 64) {
      const int j = first_w + g;
      const WidthInfo wi = wt.wi[j];
      const unsigned oa = bmax + wi.ahead, ob = bmax - wi.behind;
      auto leaf = [&](unsigned r) -> float { return fabsf(T[r + oa] - T[r + ob]); };
      float value;
      if (bs.L0 == 0) {
        value = leaf(0);
      } else if (n2 > 0) {
        // levels 1 and 2 in registers: a lane reduces 25 consecutive leaves
        // (its lanes read 25 words apart, so no LDS bank conflicts)
        for (size_t p = lane; p < n2; p += 64) {
          const unsigned r = 25 * (unsigned)p;
          float m[5];
#pragma unroll
          for (int i = 0; i < 5; ++i) {
            const unsigned b = r + 5 * i;
            m[i] = med5(leaf(b), leaf(b + 1), leaf(b + 2), leaf(b + 3), leaf(b + 4));
          }
          WA[p] = med5(m[0], m[1], m[2], m[3], m[4]);
        }
        value = wave_pyramid(WA, WB, n2, bs.L0 - 2)[0];
      } else {
        for (size_t q = lane; q < n1; q += 64) {
          const unsigned r = 5 * (unsigned)q;  // pos - pos0 of the group's first leaf
          WA[q] = med5(leaf(r), leaf(r + 1), leaf(r + 2), leaf(r + 3), leaf(r + 4));
        }
        value = wave_pyramid(WA, WB, n1, bs.L0 - 1)[0];
      }
      if (lane == 0) partial[(row * wt.nw + j) * np_max + c] = value;
      wave_sync();  // the buffers are rewritten for the wave's next width
    }
    return;
  }
  for (int g = 0; g < n_group; ++g) {
    const int j = first_w + g;
    const WidthInfo wi = wt.wi[j];
    float value;
    if (src_j == j) {
      value = chunk_value([&](size_t k) -> float {
        const unsigned r = (unsigned)(k - k0);
        return fabsf(T[r + wts] - T[r]);
      });
    } else if (src_j >= 0) {
      value = chunk_value([&](size_t k) -> float {
        const unsigned r = (unsigned)(k - k0);
        return fabsf((T[r + wts] - T[r]) * sscale);
      });
    } else {
      value = chunk_value([&](size_t k) -> float {
        const size_t pos = reblock_pos((unsigned)k, bs.bc, bs.delta);
        return fabsf(filtered_at(S, wt, nwv, winv, j, pos));
      });
    }
    if (threadIdx.x == 0) partial[(row * wt.nw + j) * np_max + c] = value;
    __syncthreads();
  }
}

// blockIdx: (width - first_w, trial, group)
__global__ __launch_bounds__(64) void k_block_rms_finish(const float* __restrict__ partial,
                                                         const GroupDev* __restrict__ gd,
                                                         size_t np_one, size_t np_max,
                                                         int nw_total, int first_w,
                                                         float* __restrict__ wrms,
                                                         float* __restrict__ winv) {
  extern __shared__ float lds[];
  const size_t np = gd ? gd[blockIdx.z].bs.np : np_one;
  const int j = first_w + blockIdx.x;
  const size_t row = (size_t)blockIdx.z * gridDim.y + blockIdx.y;
  const float* p = partial + (row * nw_total + j) * np_max;
  float* A = lds;
  float* B = lds + np;
  for (size_t q = threadIdx.x; q < np; q += 64) A[q] = p[q];
  __syncthreads();
  const float med = lds_finish<64>(A, B, np);
  if (threadIdx.x == 0) {
    const float r = (float)((double)med * 1.4862);  // get_rms.cu:127-128
    wrms[row * nw_total + j] = r;
    winv[row * nw_total + j] = 1.0f / r;            // get_rms.cu:195
  }
}

// The same with at most four partials per (trial, width), the short series of
// a ragged batch: no median level remains, only the tail, so one thread per
// (row, width) takes it in registers.  Items: rows x n_group, width-fastest.
__global__ __launch_bounds__(256) void k_block_rms_finish_short(
    const float* __restrict__ partial, const GroupDev* __restrict__ gd, size_t np_one,
    size_t np_max, int n_dm, size_t n_items, int nw_total, int first_w, int n_group,
    float* __restrict__ wrms, float* __restrict__ winv) {
  const size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= n_items) return;
  const size_t row = idx / (unsigned)n_group;
  const int j = first_w + (int)(idx - row * (unsigned)n_group);
  const size_t np = gd ? gd[row / (unsigned)n_dm].bs.np : np_one;  // <= 4
  const float* p = partial + (row * nw_total + j) * np_max;
  const float v0 = np > 0 ? p[0] : 0.f, v1 = np > 1 ? p[1] : 0.f;
  const float v2 = np > 2 ? p[2] : 0.f, v3 = np > 3 ? p[3] : 0.f;
  // med_tail of the np values (the pyramid's end, median_filter.cu:217-235)
  const float med = np >= 4 ? med4(v0, v1, v2, v3)
                    : np == 3 ? med3(v0, v1, v2)
                    : np == 2 ? 0.5f * (v0 + v1)
                    : np == 1 ? v0 : 0.f;
  const float r = (float)((double)med * 1.4862);  // get_rms.cu:127-128
  wrms[row * nw_total + j] = r;
  winv[row * nw_total + j] = 1.0f / r;            // get_rms.cu:195
}

// Block-wide exclusive scan of one unsigned per thread; returns the prefix,
// writes the total to *total.
template <int BS>
__device__ __forceinline__ unsigned block_exclusive_scan(unsigned v, unsigned* total) {
  // inclusive scan within each wave (shuffles), then the wave totals
  __shared__ unsigned wsum[BS / 64];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  unsigned x = v;
#pragma unroll
  for (int d = 1; d < 64; d <<= 1) {
    const unsigned y = __shfl_up(x, d);
    if (lane >= d) x += y;
  }
  if (lane == 63) wsum[w] = x;
  __syncthreads();
  unsigned pre = 0, tot = 0;
#pragma unroll
  for (int k = 0; k < BS / 64; ++k) {
    const unsigned t = wsum[k];
    if (k < w) pre += t;
    tot += t;
  }
  __syncthreads();  // wsum is reused by the next call
  *total = tot;
  return pre + x - v;
}

// Positions per block: FB_BS * TH_PER, 8192 by default; 4096 when the widest
// boxcar's halo is at most 2048 (short rows and boxcars: twice the blocks
// resident per CU hide the staging better, for a halo that stays small).
constexpr int TH_PER_LONG = 32;
constexpr int TH_PER_SHORT = 16;
constexpr unsigned TH_SHORT_HALO = 2048;

// One block per (chunk of TH_CHUNK positions i*ts, DM): the chunk's prefix
// sums (plus the widest boxcar's halo) are staged in LDS once and every width
// is thresholded from them.  Output i of a width belongs to the chunk holding
// i*ts, so each (DM, width, chunk) key covers a time-ordered run of outputs.
// blockIdx: (chunk, trial, group); rows g*n_dm + m.  Chunk keys are
// (row*nw + width)*nchunks_per_w + chunk, so the compacted samples come out
// in (group, trial, width, time) order.
//
// Thread t evaluates outputs p = q*FB_BS + t (q < TH_PER) of the chunk, so the
// lanes of a wave read consecutive LDS words (no bank conflicts; with ts == 1
// the q*FB_BS offsets are immediates).  Hits are written in time order:
// q-major, then wave, then lane (one ballot per q and wave, wave counts
// prefixed in LDS).
template <int TH_PER>
__global__ __launch_bounds__(FB_BS) void k_threshold(
    const float* __restrict__ Sall, size_t spitch, const GroupDev* __restrict__ gd,
    WidthTable wt, unsigned amax, unsigned bmax, const float* __restrict__ winv_all,
    float thresh, GiantBuffers gb) {
  constexpr int TH_CHUNK = FB_BS * TH_PER;
  extern __shared__ float4 T4[];  // TH_CHUNK + amax + bmax prefix sums (+ alignment)
  __shared__ unsigned wcnt[TH_PER][FB_BS / 64];
  __shared__ unsigned sbase;
  __shared__ unsigned wmask;  // widths with an output above the threshold in the chunk
  static_assert(kMaxWidths <= 32, "one bit per width");
  const size_t c = blockIdx.x;
  const size_t row = (size_t)blockIdx.z * gridDim.y + blockIdx.y;
  const size_t stride = gd[blockIdx.z].sh.stride;
  const size_t bc692 = gd[blockIdx.z].bc692;
  const size_t s_len = gd[blockIdx.z].sh.nout + 1;
  const unsigned* nwv = gd[blockIdx.z].n_w;
  const float* S = Sall + row * spitch;
  const size_t p0 = c * TH_CHUNK;                 // first position of the chunk
  {
    // chunks past every width's output (shorter groups of a batch) only
    // record empty counts (and padding rows of a gated group); lane j of
    // each wave tests width j
    bool mine = false;
    const int lj = (int)(threadIdx.x & 63);
    if (lj < wt.nw && blockIdx.y < gd[blockIdx.z].rows_valid) {
      const unsigned ts = wt.wi[lj].ts;
      mine = (ts == 1 ? p0 : (p0 + ts - 1) / ts) < nwv[lj];
    }
    if (!__ballot(mine)) {
      for (int j = threadIdx.x; j < wt.nw; j += FB_BS)
        gb.ccount[(row * wt.nw + j) * gb.nchunks_per_w + c] = 0;
      return;
    }
  }
  // T[t] holds S[tbase + t] (0 outside [0, s_len)).  Staged in 16-byte
  // pieces from the 16-byte boundary at or below S + tbase, several loads in
  // flight per thread before their LDS writes.
  const long long tbase = (long long)p0 + (long long)wt.off - (long long)bmax;
  const unsigned tlen = (unsigned)(TH_CHUNK + amax + bmax);
  const unsigned sh = (unsigned)((((uintptr_t)S >> 2) + (unsigned long long)tbase) & 3u);
  const long long tal = tbase - sh;               // S + tal is 16-byte aligned
  const long long slen = (long long)s_len;
  // Every output a width keeps reads prefix sums below s_len, so a row that
  // ends inside the chunk (short gulps of a batch) stages only that far; the
  // tile past it is read only for outputs that are never kept.
  const long long tavail = slen - tbase;
  const unsigned tneed = tavail <= 0 ? 1u : (tavail < (long long)tlen ? (unsigned)tavail : tlen);
  const unsigned n4 = (tneed + sh + 3) / 4;
  constexpr int SU = 12;  // (the widest tile, 8192 + 4096 floats, in one round)
  // Plain loads, no per-piece tests, when the pieces lie within three floats
  // of the row (a thread past the last piece loads the last one again and
  // stores nothing).  Rows of S start 3 floats into their allocation and are
  // padded to a multiple of 64 floats, with 5 spare floats after the last, so
  // those floats are readable; the tile entries they fill (before S[0], past
  // S[s_len - 1]) are read only for outputs that are never kept.
  const bool whole = tal >= -3 && tal + 4ll * n4 <= slen + 3;
  for (unsigned v0 = threadIdx.x; v0 < n4; v0 += SU * FB_BS) {
    float4 r[SU];
    if (whole) {
#pragma unroll
      for (int u = 0; u < SU; ++u) {
        const unsigned v = v0 + u * FB_BS;
        r[u] = *reinterpret_cast<const float4*>(S + tal + 4ll * (v < n4 ? v : n4 - 1));
      }
    } else {
#pragma unroll
      for (int u = 0; u < SU; ++u) {
        const unsigned v = v0 + u * FB_BS;
        const long long si = tal + 4ll * v;
        if (v < n4) {
          if (si >= 0 && si + 3 < slen) {
            r[u] = *reinterpret_cast<const float4*>(S + si);
          } else {
            r[u].x = (si >= 0 && si < slen) ? S[si] : 0.f;
            r[u].y = (si + 1 >= 0 && si + 1 < slen) ? S[si + 1] : 0.f;
            r[u].z = (si + 2 >= 0 && si + 2 < slen) ? S[si + 2] : 0.f;
            r[u].w = (si + 3 >= 0 && si + 3 < slen) ? S[si + 3] : 0.f;
          }
        }
      }
    }
#pragma unroll
    for (int u = 0; u < SU; ++u)
      if (v0 + u * FB_BS < n4) T4[v0 + u * FB_BS] = r[u];
  }
  const float* T = reinterpret_cast<const float*>(T4) + sh;
  if (threadIdx.x == 0) wmask = 0u;
  __syncthreads();
  const unsigned tid = threadIdx.x;
  const int lane = (int)(tid & 63), wave = (int)(tid >> 6);
  const uint64_t below = lane ? (~0ull >> (64 - lane)) : 0ull;  // lanes under this one
  const unsigned tmax = tlen - 1;
  // Per-width index arithmetic on the scalar unit: 64-bit divisions are long
  // software sequences there, so the undecimated widths (output i at
  // position i) take the chunk's own range and beam period, found once:
  // positions p0 .. p0 + TH_CHUNK - 1 in period pbeam (find_giants.cu:55-58
  // keeps i with (i mod stride) < bc692), which ends at pnext.
  const size_t pbeam = (size_t)(unsigned)p0 / stride;
  const size_t pnext = (pbeam + 1) * stride;
  const size_t plim = pbeam * stride + bc692;
  // outputs i with i*ts in the chunk, i0 .. i1 - 1 (i1 not yet clipped to
  // the width's count), whether they lie in one beam period, and its bound
  auto chunk_outputs = [&](const WidthInfo& wi, size_t& i0, size_t& i1) {
    if (wi.ts == 1) {
      i0 = p0;
      i1 = p0 + TH_CHUNK;
    } else {
      i0 = (p0 + wi.ts - 1) / wi.ts;
      i1 = (p0 + TH_CHUNK + wi.ts - 1) / wi.ts;
    }
  };
  auto beam_period = [&](const WidthInfo& wi, size_t i0, size_t i1, size_t& lim) -> bool {
    if (wi.ts == 1) {  // i0 == p0
      lim = plim;
      return i1 - 1 < pnext && i1 - 1 <= 0xffffffffull;
    }
    const size_t beam0 = (size_t)(unsigned)i0 / stride;
    lim = beam0 * stride + bc692;
    return (size_t)(unsigned)(i1 - 1) / stride == beam0 && i1 - 1 <= 0xffffffffull;
  };
  // Outputs p = q*FB_BS + tid (q < TH_PER) of width j in this chunk: their
  // values, and the bits of those kept and above the threshold.  Returns
  // false when the chunk holds no output of the width (outputs i0 .. i1-1).
  auto eval_width = [&](int j, float (&vals)[TH_PER], unsigned& mask, size_t& i0) -> bool {
    const WidthInfo wi = wt.wi[j];
    const float inv = winv_all[row * wt.nw + j];
    // outputs i with i*ts in [p0, p0 + TH_CHUNK)
    size_t i1;
    chunk_outputs(wi, i0, i1);
    const size_t n_w = nwv[j];
    if (i1 > n_w) i1 = n_w;
    mask = 0;
    if (i0 >= i1) return false;
    const unsigned nout = (unsigned)(i1 - i0);
    // find_giants.cu:55-58 keeps i with (i mod stride) < bc692.  When the
    // chunk's outputs all lie in one beam period that is a single bound.
    size_t lim;
    const bool one_period = beam_period(wi, i0, i1, lim);
    const unsigned rbase = (unsigned)(i0 * wi.ts - p0 + bmax);   // T index of output i0
    if (one_period) {
      // kept: outputs p < pk (i < min(i1, lim)); the others are evaluated
      // from in-tile reads and dropped
      const size_t kend = i1 < lim ? i1 : lim;
      const unsigned pk = kend > i0 ? (unsigned)(kend - i0) : 0u;
      if (wi.ts == 1) {
        // p*ts + ahead + rbase <= TH_CHUNK - 1 + amax + bmax: always in the tile
        const float* Ta = T + rbase + wi.ahead + tid;
        const float* Tb = T + rbase - wi.behind + tid;
#pragma unroll
        for (int q = 0; q < TH_PER; ++q) {
          const float v = (Ta[q * FB_BS] - Tb[q * FB_BS]) * inv;   // get_rms.cu:194-197
          vals[q] = v;
          if ((unsigned)q * FB_BS + tid < pk && v > thresh) mask |= 1u << q;
        }
      } else {
#pragma unroll
        for (int q = 0; q < TH_PER; ++q) {
          const unsigned pp = (unsigned)q * FB_BS + tid;
          const unsigned ra = min(rbase + wi.ahead + pp * wi.ts, tmax);
          const unsigned rb = min(rbase - wi.behind + pp * wi.ts, tmax);
          const float v = (T[ra] - T[rb]) * inv;
          vals[q] = v;
          if (pp < pk && v > thresh) mask |= 1u << q;
        }
      }
    } else {
#pragma unroll
      for (int q = 0; q < TH_PER; ++q) {
        vals[q] = 0.f;
        const unsigned pp = (unsigned)q * FB_BS + tid;
        if (pp < nout) {
          const size_t i = i0 + pp;
          const unsigned ra = rbase + wi.ahead + pp * wi.ts, rb = rbase - wi.behind + pp * wi.ts;
          const float v = (T[ra] - T[rb]) * inv;
          vals[q] = v;
          const bool in_beam = (size_t)(unsigned)i % stride < bc692;
          if (in_beam && v > thresh) mask |= 1u << q;
        }
      }
    }
    return true;
  };
  // Whether any kept output of width j in this thread's positions exceeds the
  // threshold.  For undecimated widths in one beam period it tests only the
  // largest difference: rounding is monotonic, so for inv > 0
  // max_q RN(d_q * inv) = RN(max_q d_q * inv), and the test gives exactly
  // the answer of testing every output (NaN differences never exceed it in
  // either form); other widths evaluate every output.
  auto screen_width = [&](int j) -> bool {
    const WidthInfo wi = wt.wi[j];
    size_t i0, i1;
    chunk_outputs(wi, i0, i1);
    if (i1 > nwv[j]) i1 = nwv[j];
    if (i0 >= i1) return false;
    size_t lim;
    const bool one_period = beam_period(wi, i0, i1, lim);
    if (wi.ts != 1 || !one_period) {
      float vals[TH_PER];
      unsigned mask;
      size_t i0x;
      eval_width(j, vals, mask, i0x);
      return mask != 0;
    }
    const float inv = winv_all[row * wt.nw + j];
    const size_t kend = i1 < lim ? i1 : lim;
    const unsigned pk = kend > i0 ? (unsigned)(kend - i0) : 0u;
    const unsigned rbase = (unsigned)(i0 - p0 + bmax);
    const float* Ta = T + rbase + wi.ahead + tid;
    const float* Tb = T + rbase - wi.behind + tid;
    float mx = -INFINITY;
    if (pk >= (unsigned)TH_CHUNK) {
#pragma unroll
      for (int q = 0; q < TH_PER; ++q) mx = fmaxf(mx, Ta[q * FB_BS] - Tb[q * FB_BS]);
    } else {
      // a row ending in the chunk (short gulps of a batch): steps of four
      // outputs up to its last kept one, block-uniformly; inside a step the
      // reads are all issued (every index lies in the tile) and the outputs
      // past the row are selected away
#pragma unroll
      for (int q0 = 0; q0 < TH_PER; q0 += 4) {
        if ((unsigned)q0 * FB_BS >= pk) break;
        float d[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) d[u] = Ta[(q0 + u) * FB_BS] - Tb[(q0 + u) * FB_BS];
#pragma unroll
        for (int u = 0; u < 4; ++u)
          mx = fmaxf(mx, (unsigned)(q0 + u) * FB_BS + tid < pk ? d[u] : -INFINITY);
      }
    }
    return mx * inv > thresh;
  };
  // Screening pass without barriers: most chunks hold no output above the
  // threshold at any width, and then every key of the chunk is empty.
  {
    unsigned hw = 0;  // this thread's widths with a hit (nw <= 24 bits)
    for (int j = 0; j < wt.nw; ++j)
      if (screen_width(j)) hw |= 1u << j;
    // the wave's widths with a hit, one ballot per width, booked by one lane
    unsigned wb = 0;
    if (__ballot(hw != 0))
      for (int j = 0; j < wt.nw; ++j)
        if (__ballot((hw >> j) & 1u)) wb |= 1u << j;
    if (lane == 0 && wb) atomicOr(&wmask, wb);
    if (!__syncthreads_or(wb != 0)) {
      for (int j = threadIdx.x; j < wt.nw; j += FB_BS)
        gb.ccount[(row * wt.nw + j) * gb.nchunks_per_w + c] = 0;
      return;
    }
  }
  for (int j = 0; j < wt.nw; ++j) {
    const size_t key = (row * wt.nw + j) * gb.nchunks_per_w + c;
    float vals[TH_PER];
    unsigned mask;
    size_t i0;
    // widths the screening found without a hit are empty here (wmask is the
    // same for the whole block, so every thread skips together)
    if (!(wmask & (1u << j)) || !eval_width(j, vals, mask, i0)) {
      if (tid == 0) gb.ccount[key] = 0;
      continue;
    }
    // a flagged width has a hit in some thread (the screening's test is the
    // same as mask != 0), so no block-wide test here; the previous width's
    // last barrier already guards wcnt and sbase
    uint64_t bq[TH_PER];
#pragma unroll
    for (int q = 0; q < TH_PER; ++q) bq[q] = __ballot((mask >> q) & 1u);
    if (lane == 0) {
#pragma unroll
      for (int q = 0; q < TH_PER; ++q) wcnt[q][wave] = (unsigned)__popcll(bq[q]);
    }
    __syncthreads();
    if (wave == 0) {
      // exclusive prefix of the TH_PER * 4 counts in (q, wave) order: lane l
      // takes entries EP*l .. EP*l + EP - 1, the lanes' sums are scanned
      // across the wave, and lane 0 books the total
      constexpr int EP = TH_PER * (FB_BS / 64) / 64;
      static_assert(EP >= 1 && EP * 64 == TH_PER * (FB_BS / 64), "whole counts per lane");
      unsigned* e = &wcnt[0][0];
      unsigned cv[EP];
      unsigned own = 0;
#pragma unroll
      for (int k = 0; k < EP; ++k) {
        cv[k] = e[EP * lane + k];
        own += cv[k];
      }
      unsigned incl = own;
#pragma unroll
      for (int o = 1; o < 64; o <<= 1) {
        const unsigned u = __shfl_up(incl, o);
        if (lane >= o) incl += u;
      }
      unsigned run = incl - own;
#pragma unroll
      for (int k = 0; k < EP; ++k) {
        e[EP * lane + k] = run;
        run += cv[k];
      }
      const unsigned total = __shfl(incl, 63);
      if (lane == 0) {
        sbase = atomicAdd(gb.counter, total);
        gb.ccount[key] = total;
        gb.coff[key] = sbase;
      }
    }
    __syncthreads();
    const size_t base = sbase;
    if (mask) {
      const uint64_t dmw = (uint64_t)row * wt.nw + j;
#pragma unroll
      for (int q = 0; q < TH_PER; ++q) {
        if (mask & (1u << q)) {
          const size_t pos = base + wcnt[q][wave] + (unsigned)__popcll(bq[q] & below);
          if (pos < gb.scap) {  // overflow: detected by the host from the counter
            gb.skey[pos] = (dmw << 32) | (uint64_t)(i0 + (unsigned)q * FB_BS + tid);
            gb.sval[pos] = vals[q];
          }
        }
      }
    }
    __syncthreads();
  }
}

// ---- device-wide exclusive scan of unsigned ints (1024 per block) ----
constexpr int SC_BS = 256;
constexpr int SC_PER = 4;
constexpr int SC_TILE = SC_BS * SC_PER;

__global__ __launch_bounds__(SC_BS) void k_scan_blocks(const unsigned* __restrict__ in,
                                                       unsigned* __restrict__ out, size_t n,
                                                       unsigned* __restrict__ sums) {
  const size_t b0 = (size_t)blockIdx.x * SC_TILE + (size_t)threadIdx.x * SC_PER;
  unsigned v[SC_PER];
  unsigned local = 0;
#pragma unroll
  for (int q = 0; q < SC_PER; ++q) {
    v[q] = (b0 + q < n) ? in[b0 + q] : 0u;
    local += v[q];
  }
  unsigned total;
  unsigned pre = block_exclusive_scan<SC_BS>(local, &total);
#pragma unroll
  for (int q = 0; q < SC_PER; ++q) {
    if (b0 + q < n) out[b0 + q] = pre;
    pre += v[q];
  }
  if (threadIdx.x == 0) sums[blockIdx.x] = total;
}

__global__ __launch_bounds__(SC_BS) void k_scan_add(unsigned* __restrict__ out, size_t n,
                                                    const unsigned* __restrict__ offs) {
  const size_t b0 = (size_t)blockIdx.x * SC_TILE;
  const unsigned add = offs[blockIdx.x];
  for (int q = threadIdx.x; q < SC_TILE; q += SC_BS)
    if (b0 + q < n) out[b0 + q] += add;
}

__global__ void k_order_samples(GiantBuffers gb, size_t nkeys) {
  const size_t key = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (key >= nkeys) return;
  const unsigned cnt = gb.ccount[key];
  if (!cnt) return;
  const size_t src = gb.coff[key];
  const size_t dst = gb.cpos[key];
  for (unsigned q = 0; q < cnt; ++q) {
    gb.okey[dst + q] = gb.skey[src + q];
    gb.oval[dst + q] = gb.sval[src + q];
  }
}

// The same ordering with one thread per sample: a sample's key (trial, width,
// chunk) follows from its record, output i of width j lying in the chunk of
// position i*ts; its place is the key's ordered offset plus its rank in the
// key's unordered run.
__global__ void k_order_samples_by_sample(GiantBuffers gb, WidthTable wt, unsigned nsamp) {
  const unsigned p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= nsamp) return;
  const uint64_t kv = gb.skey[p];
  const unsigned dmw = (unsigned)(kv >> 32);  // a 32-bit field: 32-bit divisions
  const unsigned j = dmw % (unsigned)wt.nw;
  const uint64_t pos = (kv & 0xffffffffull) * wt.wi[j].ts;
  const size_t kch = (pos >> 32) == 0 && gb.chunk <= 0xffffffffull
                         ? (size_t)((unsigned)pos / (unsigned)gb.chunk)
                         : (size_t)(pos / gb.chunk);
  const size_t key = (size_t)dmw * gb.nchunks_per_w + kch;
  const size_t dst = (size_t)gb.cpos[key] + (p - gb.coff[key]);
  gb.okey[dst] = kv;
  gb.oval[dst] = gb.sval[p];
}
__global__ void k_segment_heads(GiantBuffers gb, WidthTable wt, unsigned nsamp) {
  const unsigned k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= nsamp) return;
  const uint64_t key = gb.okey[k];
  unsigned h = 1;
  if (k > 0) {
    const uint64_t prev = gb.okey[k - 1];
    if ((prev >> 32) == (key >> 32)) {
      const unsigned j = (unsigned)(key >> 32) % (unsigned)wt.nw;
      const uint64_t i = key & 0xffffffffull;
      const uint64_t ip = prev & 0xffffffffull;
      h = (i > ip + wt.wi[j].merge_dist) ? 1u : 0u;
    }
  }
  gb.head[k] = h;
}

// Segment reduction in one pass over the samples: every sample offers
// (value, earliest position) to its giant with a 64-bit atomic max, which
// keeps the first maximum (maximum_first); the head sample writes the
// begin, the last sample of a run the end.
__global__ void k_segment_scatter(GiantBuffers gb, WidthTable wt, unsigned nsamp, unsigned n_dm,
                                  GiantOut go, uint64_t* best, const GroupDev* __restrict__ gd) {
  const unsigned k = blockIdx.x * blockDim.x + threadIdx.x;
  const unsigned lane = threadIdx.x & 63;
  const bool valid = k < nsamp;
  const unsigned h = valid ? gb.head[k] : 1u;
  const unsigned g = valid ? gb.segid[k] + h - 1u : 0xffffffffu;  // no segment past nsamp
  const uint64_t key = valid ? gb.okey[k] : 0ull;
  // the segment's best sample: (orderable S/N, ~index) maximised over each run
  // of equal g inside the wave (segmented scan, runs are contiguous), then one
  // atomic per run piece instead of one per sample
  unsigned long long bk =
      valid ? (((uint64_t)float_key(gb.oval[k]) << 32) | (uint64_t)(0xffffffffu - k)) : 0ull;
#pragma unroll
  for (int d = 1; d < 64; d <<= 1) {
    const unsigned long long ov = __shfl_up(bk, d);
    const unsigned og = __shfl_up(g, d);
    if (lane >= (unsigned)d && og == g && ov > bk) bk = ov;
  }
  const unsigned ng = __shfl_down(g, 1);
  if (!valid) return;
  if (lane == 63 || ng != g) atomicMax((unsigned long long*)&best[g], bk);
  const unsigned dmw = (unsigned)(key >> 32);  // a 32-bit field: 32-bit divisions
  const unsigned j = dmw % (unsigned)wt.nw;
  const WidthInfo wi = wt.wi[j];
  const uint64_t ro = wi.rel_off, sc = wi.cs;
  const uint64_t i = key & 0xffffffffull;
  if (h) {
    const unsigned row = dmw / (unsigned)wt.nw;  // group * n_dm + trial
    go.begin[g] = (i + ro) * sc;
    go.filter[g] = wi.filter_idx;
    const uint32_t grp = row / n_dm;
    go.dm[g] = row % n_dm + (gd ? gd[grp].trial0 : 0u);
    if (go.group) go.group[g] = gd ? gd[grp].cgroup : grp;
  }
  if (k + 1 == nsamp || gb.head[k + 1]) go.end[g] = (i + 1 + ro) * sc;
}

__global__ void k_segment_finish(GiantBuffers gb, WidthTable wt, unsigned ngiants, GiantOut go,
                                 const uint64_t* best) {
  const unsigned g = blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= ngiants) return;
  const unsigned k = 0xffffffffu - (unsigned)(best[g] & 0xffffffffull);
  const uint64_t key = gb.okey[k];
  const WidthInfo wi = wt.wi[(unsigned)(key >> 32) % (unsigned)wt.nw];
  go.peak[g] = gb.oval[k];
  go.ind[g] = ((key & 0xffffffffull) + wi.rel_off) * (uint64_t)wi.cs;
}

__global__ void k_zero_u64(uint64_t* p, size_t n) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0;
}

}  // namespace

size_t block_rms_np_max(const GroupArgs& ga) {
  size_t np = 0;
  for (int g = 0; g < ga.G; ++g) np = np > ga.host[g].bs.np ? np : ga.host[g].bs.np;
  return np;
}

hipError_t launch_block_rms_partials(const float* S, size_t spitch, const WidthTable& wt,
                                     int first_w, int n_group, int n_dm, float* partial,
                                     const float* winv, hipStream_t s, const GroupArgs& ga) {
  if (!ga.gd || n_group == 0) return hipSuccess;
  const size_t np_max = block_rms_np_max(ga);
  if (np_max == 0) return hipSuccess;
  size_t chunk = 0;
  for (int g = 0; g < ga.G; ++g) chunk = chunk > ga.host[g].bs.chunk ? chunk : ga.host[g].bs.chunk;
  int tile_ok = 1;
  unsigned amax = 0, bmax = 0;
  for (int g = 0; g < n_group; ++g) {
    const WidthInfo& wi = wt.wi[first_w + g];
    if (wi.ts != 1) tile_ok = 0;
    if (wi.ahead > amax) amax = wi.ahead;
    if (wi.behind > bmax) bmax = wi.behind;
  }
  const size_t tlen = chunk + amax + bmax;
  const size_t n1 = chunk / 5;
  // level buffers: one pair per wave (the LDS-tile path), else block-wide
  const size_t rest_block = (n1 ? n1 : 1) + n1 / 5 + 8;
  const size_t rest_wave = std::max(rest_block, (size_t)(FB_BS / 64) * block_rms_wave_words(n1));
  constexpr size_t kMaxLds = 65536 / sizeof(float);
  if (tlen + rest_wave > kMaxLds) tile_ok = 0;  // keep LDS <= 64 KiB
  // the block-wide form asks for less LDS, so more blocks stay resident
  const size_t rest = tile_ok ? rest_wave : rest_block;
  // single decimated width: room for the widest range it or an earlier width reads
  size_t tile_words = 0;
  if (!tile_ok && n_group == 1 && chunk > 0) {
    for (int jj = 0; jj <= first_w; ++jj) {
      const WidthInfo& w = wt.wi[jj];
      const size_t span = chunk + (w.ahead + w.behind) / w.ts;
      if (span > tile_words) tile_words = span;
    }
    if (tile_words + rest > kMaxLds) tile_words = kMaxLds > rest ? kMaxLds - rest : 0;
  }
  size_t lds = ((tile_ok ? tlen : tile_words) + rest) * sizeof(float);
  hipLaunchKernelGGL(k_block_rms_partials, dim3((unsigned)np_max, n_dm, ga.G), dim3(FB_BS), lds,
                     s, S, spitch, wt, ga.gd, np_max, first_w, n_group, tile_ok, amax, bmax,
                     (unsigned)tile_words, winv, partial);
  return hipGetLastError();
}

hipError_t launch_block_rms_finish(const float* partial, int nw_total, int first_w,
                                   int n_group, int n_dm, float* wrms, float* winv,
                                   hipStream_t s, const GroupArgs& ga,
                                   const BlockRmsShape* single) {
  if ((!ga.gd && !single) || n_group == 0) return hipSuccess;
  const size_t np_max = ga.gd ? block_rms_np_max(ga) : single->np;
  if (np_max == 0) return hipSuccess;
  if (np_max <= 4) {
    const size_t n_items = (size_t)n_dm * (ga.gd ? ga.G : 1) * n_group;
    hipLaunchKernelGGL(k_block_rms_finish_short, dim3((unsigned)((n_items + 255) / 256)),
                       dim3(256), 0, s, partial, ga.gd, single ? single->np : (size_t)0, np_max,
                       n_dm, n_items, nw_total, first_w, n_group, wrms, winv);
    return hipGetLastError();
  }
  const size_t lds = (np_max + np_max / 5 + 8) * sizeof(float);
  if (lds > 65536) return hipErrorInvalidValue;  // pooled count beyond ~40M samples
  hipLaunchKernelGGL(k_block_rms_finish, dim3(n_group, n_dm, ga.gd ? ga.G : 1), dim3(64), lds, s,
                     partial, ga.gd, single ? single->np : (size_t)0, np_max, nw_total, first_w,
                     wrms, winv);
  return hipGetLastError();
}

size_t threshold_chunk(const WidthTable& wt) {
  unsigned amax = 0, bmax = 0;
  for (int j = 0; j < wt.nw; ++j) {
    amax = wt.wi[j].ahead > amax ? wt.wi[j].ahead : amax;
    bmax = wt.wi[j].behind > bmax ? wt.wi[j].behind : bmax;
  }
  return (size_t)FB_BS * (amax + bmax <= TH_SHORT_HALO ? TH_PER_SHORT : TH_PER_LONG);
}

hipError_t launch_threshold(const float* S, size_t spitch, const WidthTable& wt,
                            const float* winv, int n_dm, float thresh, GiantBuffers& gb,
                            hipStream_t s, const GroupArgs& ga) {
  if (!ga.gd) return hipErrorInvalidValue;
  gb.chunk = threshold_chunk(wt);
  unsigned amax = 0, bmax = 0;
  for (int j = 0; j < wt.nw; ++j) {
    amax = wt.wi[j].ahead > amax ? wt.wi[j].ahead : amax;
    bmax = wt.wi[j].behind > bmax ? wt.wi[j].behind : bmax;
  }
  const size_t tlen = gb.chunk + amax + bmax;
  const size_t lds = (tlen + 8) * sizeof(float);  // + the 16-byte alignment of the staging
  if (lds > 65536) return hipErrorInvalidValue;   // boxcar_max beyond ~12000
  dim3 grid((unsigned)gb.nchunks_per_w, n_dm, ga.G);
  if (gb.chunk == (size_t)FB_BS * TH_PER_SHORT)
    hipLaunchKernelGGL(k_threshold<TH_PER_SHORT>, grid, dim3(FB_BS), lds, s, S, spitch, ga.gd, wt,
                       amax, bmax, winv, thresh, gb);
  else
    hipLaunchKernelGGL(k_threshold<TH_PER_LONG>, grid, dim3(FB_BS), lds, s, S, spitch, ga.gd, wt,
                       amax, bmax, winv, thresh, gb);
  return hipGetLastError();
}

size_t exclusive_scan_tmp_words(size_t n) {
  size_t words = 0;
  while (n > 1) {
    const size_t nb = (n + SC_TILE - 1) / SC_TILE;
    words += 2 * nb + 2;
    n = nb;
  }
  return words + 4;
}

hipError_t launch_exclusive_scan_u32(const unsigned* in, unsigned* out, size_t n, unsigned* tmp,
                                     hipStream_t s) {
  if (n == 0) return hipSuccess;
  const size_t nb = (n + SC_TILE - 1) / SC_TILE;
  unsigned* sums = tmp;
  unsigned* offs = tmp + nb + 1;
  hipLaunchKernelGGL(k_scan_blocks, dim3((unsigned)nb), dim3(SC_BS), 0, s, in, out, n, sums);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess || nb == 1) return e;
  e = launch_exclusive_scan_u32(sums, offs, nb, tmp + 2 * nb + 2, s);
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(k_scan_add, dim3((unsigned)nb), dim3(SC_BS), 0, s, out, n, offs);
  return hipGetLastError();
}


hipError_t launch_order_samples(GiantBuffers& gb, size_t nkeys, hipStream_t s) {
  hipLaunchKernelGGL(k_order_samples, dim3((unsigned)((nkeys + 255) / 256)), dim3(256), 0, s, gb,
                     nkeys);
  return hipGetLastError();
}

hipError_t launch_order_samples(GiantBuffers& gb, const WidthTable& wt, unsigned nsamp,
                                hipStream_t s) {
  if (nsamp == 0) return hipSuccess;
  hipLaunchKernelGGL(k_order_samples_by_sample, dim3((nsamp + 255) / 256), dim3(256), 0, s, gb,
                     wt, nsamp);
  return hipGetLastError();
}

hipError_t launch_segment_heads(const GiantBuffers& gb, const WidthTable& wt, unsigned nsamp,
                                hipStream_t s) {
  if (!nsamp) return hipSuccess;
  hipLaunchKernelGGL(k_segment_heads, dim3((nsamp + 255) / 256), dim3(256), 0, s, gb, wt, nsamp);
  return hipGetLastError();
}

__global__ void k_fill_u32(unsigned* __restrict__ p, size_t n, unsigned v) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

hipError_t launch_fill_u32(unsigned* p, size_t n, unsigned v, hipStream_t s) {
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(k_fill_u32, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, p, n, v);
  return hipGetLastError();
}

hipError_t launch_segment_reduce(const GiantBuffers& gb, const WidthTable& wt, unsigned nsamp,
                                 unsigned ngiants, int n_dm, GiantOut go, uint64_t* best,
                                 hipStream_t s, const GroupDev* gd) {
  if (!nsamp || !ngiants) return hipSuccess;
  hipLaunchKernelGGL(k_zero_u64, dim3((ngiants + 255) / 256), dim3(256), 0, s, best,
                     (size_t)ngiants);
  hipLaunchKernelGGL(k_segment_scatter, dim3((nsamp + 255) / 256), dim3(256), 0, s, gb, wt,
                     nsamp, (unsigned)n_dm, go, best, gd);
  hipLaunchKernelGGL(k_segment_finish, dim3((ngiants + 255) / 256), dim3(256), 0, s, gb, wt,
                     ngiants, go, best);
  return hipGetLastError();
}

}  // namespace hd
