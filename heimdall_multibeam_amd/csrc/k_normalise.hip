// k_normalise.hip -- baseline removal, per-beam RMS, normalisation and the
// matched filter's prefix sum, over all DM series of a gulp at once.
//
//   baseline   RemoveBaselinePlan::exec (Pipeline/remove_baseline.cu:21-160):
//              one block per (beam, DM) resamples the beam to count_round
//              points, median-of-5 scrunches them down to sample_count points
//              and stores those points; the stretched baseline is recomputed
//              from them wherever it is needed (never materialised).
//   beam rms   GetRMSPlanMB::exec_multibeam per-beam form (get_rms.cu:66-100):
//              median-of-5 pyramid of |x - baseline| per (beam, DM) block.
//   normalise  normalise_multibeam (get_rms.cu:202-212): z = (x-baseline)/rms,
//              in place (stage entry point), or fused into the scan's loads.
//   scan       MatchedFilterPlan::prep (matched_filter.cu:21-32): the exclusive
//              prefix sum in the reference's sequential order, 8 lanes per
//              series taking turns over row tiles staged through LDS so that
//              HBM reads and writes stay coalesced.
#include <algorithm>

#include "hd_device.hpp"
#include "hd_internal.hpp"
#include "hd_series.hpp"

namespace hd {

namespace {

constexpr int NB_BS = 256;

__device__ __forceinline__ size_t pow5(int L) {
  size_t p = 1;
  for (int i = 0; i < L; ++i) p *= 5;
  return p;
}

// Finishes a median-of-5 pyramid held in A (n values) with the reference's
// loop `for (size = n; size > 1; size /= 5)`: groups of five while size > 4,
// then the small-count rule.  Returns the final value to every thread.
template <int BS>
__device__ __forceinline__ float lds_finish(float* A, float* B, size_t n) {
  __shared__ float result;
  while (n > 4) {
    const size_t m = n / 5;
    __syncthreads();
    for (size_t j = threadIdx.x; j < m; j += BS) {
      const float* p = A + 5 * j;
      B[j] = med5(p[0], p[1], p[2], p[3], p[4]);
    }
    float* t = A;
    A = B;
    B = t;
    n = m;
  }
  __syncthreads();
  if (threadIdx.x == 0) result = (n >= 1) ? med_tail(A, (int)n) : 0.f;
  __syncthreads();
  return result;
}

// blockIdx: (beam, trial, group); the group's rows are g*n_dm + m and each
// row holds prow baseline points.
__global__ __launch_bounds__(NB_BS) void k_baseline_points(const float* __restrict__ plane,
                                                           SeriesShape sh_arg,
                                                           float* __restrict__ pts,
                                                           const GroupDev* __restrict__ gd,
                                                           size_t prow_arg) {
  extern __shared__ float lds[];
  const SeriesShape sh = gd ? gd[blockIdx.z].sh : sh_arg;
  const unsigned b = blockIdx.x;
  if (b >= sh.nbeams || sh.sc == 0) return;
  const int m = blockIdx.y;
  const size_t row = (size_t)blockIdx.z * gridDim.y + m;
  const size_t prow = gd ? prow_arg : (size_t)sh.nbeams * sh.sc;
  const float* d = plane + row * sh.ppitch;
  const int Ld = sh.bl_direct;
  const size_t per = pow5(Ld);
  const size_t nl = sh.count_round / per;
  float* A = lds;
  float* B = lds + nl;
  const bool last = (b == sh.nbeams - 1);
  const float* dl = d + (size_t)(sh.nbeams - 1) * sh.stride;
  const size_t gbase = (size_t)b * sh.count_round;
  // resampled value i of beam b (remove_baseline.cu:73-75)
  auto leaf = [&](size_t i) -> float {
    if (last) {
      Stretch f{sh.step_rs_last};
      return f(dl, (unsigned)i);
    }
    StretchBeam f{sh.step_rs_beam, (unsigned)sh.count_round};
    return f(d, (unsigned)(gbase + i));
  };
  for (size_t j = threadIdx.x; j < nl; j += NB_BS) A[j] = tree_value(leaf, j * per, Ld);
  float* res = lds_pyramid<NB_BS>(A, B, (int)nl, sh.nscr - Ld);
  float* o = pts + row * prow + (size_t)b * sh.sc;
  for (size_t j = threadIdx.x; j < sh.sc; j += NB_BS) o[j] = res[j];
}

__global__ __launch_bounds__(NB_BS) void k_beam_rms(const float* __restrict__ plane,
                                                    const float* __restrict__ pts,
                                                    SeriesShape sh, float* __restrict__ rms) {
  extern __shared__ float lds[];
  const unsigned b = blockIdx.x;
  const int m = blockIdx.y;
  const float* d = plane + (size_t)m * sh.ppitch;
  const float* P = pts + (size_t)m * sh.nbeams * sh.sc;
  const int Ld = sh.rms_direct;
  const size_t per = pow5(Ld);
  const size_t nl = sh.bc / per;
  float* A = lds;
  float* B = lds + nl;
  const size_t obase = (size_t)b * sh.bc;
  // |x - baseline| at the reblocked position (get_rms.cu:81, 214-229)
  auto leaf = [&](size_t k) -> float {
    const size_t pos = reblock_pos((unsigned)(obase + k), sh.bc, sh.rms_delta);
    const float y = d[pos] - baseline_at(P, sh, pos);
    return fabsf(y);
  };
  for (size_t j = threadIdx.x; j < nl; j += NB_BS) A[j] = tree_value(leaf, j * per, Ld);
  const float med = lds_finish<NB_BS>(A, B, nl);
  if (threadIdx.x == 0) rms[(size_t)m * sh.nbeams + b] = med * 1.4862f;  // get_rms.cu:94-97
}

// Two-phase per-beam RMS for large series: blocks of 5^L0 consecutive leaves
// (staged in LDS with coalesced reads) reduce to one median each; a finishing
// wave per (beam, DM) runs the remaining levels and the small-count rule.
// Exact because the first L0 levels of the reference pyramid only combine
// leaves within aligned 5^L0 groups (bc >= 5^L0 keeps every such level in
// the groups-of-five regime).

// Leaf k of the beam's pyramid, |x - baseline|, from x = the sample at the
// leaf's reblocked position, already loaded.  `direct`: one beam with exact
// float positions, where the reblocked position is k itself.
__device__ __forceinline__ float leaf_abs_dev_of(float x, const float* P, const SeriesShape& sh,
                                                 size_t k, bool direct) {
  if (direct) {
    const float bl = (sh.sc == 0) ? 0.f : Stretch{sh.step_bl_last}(P, (unsigned)k);
    return fabsf(x - bl);
  }
  const size_t pos = reblock_pos((unsigned)k, sh.bc, sh.rms_delta);
  return fabsf(x - baseline_at(P, sh, pos));
}

__device__ __forceinline__ int beam_rms_levels(size_t bc) {
  int L0 = 0;
  while (L0 < 5 && bc >= pow5(L0 + 1)) ++L0;
  return L0;
}

// blockIdx: (chunk, trial, group * nbm + beam).  Rows g*n_dm + m; the RMS
// table has nbm entries per row and the partials np_max per (row, beam).
__global__ __launch_bounds__(NB_BS) void k_beam_rms_partials(const float* __restrict__ plane,
                                                             const float* __restrict__ pts,
                                                             SeriesShape sh_arg, size_t np_max,
                                                             int stage_pts,
                                                             float* __restrict__ partial,
                                                             const GroupDev* __restrict__ gd,
                                                             size_t prow_arg, unsigned nbm) {
  extern __shared__ float lds[];
  const unsigned gz = blockIdx.z / nbm, b = blockIdx.z % nbm;
  const SeriesShape sh = gd ? gd[gz].sh : sh_arg;
  if (b >= sh.nbeams) return;
  const int L0 = beam_rms_levels(sh.bc);
  const size_t chunk = pow5(L0);
  const size_t np = sh.bc / chunk;
  const size_t c = blockIdx.x;
  if (c >= np) return;
  const int m = blockIdx.y;
  const size_t row = (size_t)gz * gridDim.y + m;
  const size_t prow = gd ? prow_arg : (size_t)sh.nbeams * sh.sc;
  const float* d = plane + row * sh.ppitch;
  float* A = lds;
  float* B = lds + chunk;
  // the row's baseline points (and the float after them) staged in LDS
  float* sP = B + chunk / 5 + 8;
  const float* P = pts + row * prow;
  if (stage_pts) {
    for (size_t i = threadIdx.x; i <= prow; i += NB_BS) sP[i] = P[i];
    __syncthreads();
    P = sP;
  }
  const size_t k0 = (size_t)b * sh.bc + c * chunk;
  // every load of the thread's leaves in flight before the first use (a
  // chunk is at most 5^5 = 3125 leaves, 13 per thread)
  constexpr int NLEAF = (3125 + NB_BS - 1) / NB_BS;
  const bool direct = sh.nbeams == 1 && sh.bc <= (1u << 24);
  float x[NLEAF];
#pragma unroll
  for (int u = 0; u < NLEAF; ++u) {
    const size_t i = threadIdx.x + (size_t)u * NB_BS;
    if (i < chunk) x[u] = d[direct ? k0 + i : reblock_pos((unsigned)(k0 + i), sh.bc, sh.rms_delta)];
  }
#pragma unroll
  for (int u = 0; u < NLEAF; ++u) {
    const size_t i = threadIdx.x + (size_t)u * NB_BS;
    if (i < chunk) A[i] = leaf_abs_dev_of(x[u], P, sh, k0 + i, direct);
  }
  float* res = lds_pyramid<NB_BS>(A, B, (int)chunk, L0);
  if (threadIdx.x == 0) partial[(row * nbm + b) * np_max + c] = res[0];
}

// blockIdx: (beam, trial, group)
__global__ __launch_bounds__(64) void k_beam_rms_finish(const float* __restrict__ partial,
                                                        SeriesShape sh_arg, size_t np_max,
                                                        unsigned nbm, float* __restrict__ rms,
                                                        const GroupDev* __restrict__ gd) {
  extern __shared__ float lds[];
  const SeriesShape sh = gd ? gd[blockIdx.z].sh : sh_arg;
  const unsigned b = blockIdx.x;
  if (b >= sh.nbeams) return;
  const size_t np = sh.bc / pow5(beam_rms_levels(sh.bc));
  const size_t row = (size_t)blockIdx.z * gridDim.y + blockIdx.y;
  const float* p = partial + (row * nbm + b) * np_max;
  float* A = lds;
  float* B = lds + np;
  for (size_t i = threadIdx.x; i < np; i += 64) A[i] = p[i];
  const float med = lds_finish<64>(A, B, np);
  if (threadIdx.x == 0) rms[row * nbm + b] = med * 1.4862f;  // get_rms.cu:94-97
}

// z = (x - baseline) / rms[pos / stride], in place (remove_baseline.cu:99-102,
// get_rms.cu:183-184).  Position nout (one past the series) is left as is.
__global__ __launch_bounds__(256) void k_normalise(float* __restrict__ plane,
                                                   const float* __restrict__ pts,
                                                   const float* __restrict__ rms,
                                                   SeriesShape sh) {
  const int m = blockIdx.y;
  float* d = plane + (size_t)m * sh.ppitch;
  const float* P = pts + (size_t)m * sh.nbeams * sh.sc;
  const float* R = rms + (size_t)m * sh.nbeams;
  for (size_t pos = (size_t)blockIdx.x * 256 + threadIdx.x; pos < sh.nout;
       pos += (size_t)gridDim.x * 256) {
    d[pos] = normalised_at(d[pos], P, R, sh, pos);
  }
}

// Sequential exclusive scan per series: S[0] = 0, S[k] = S[k-1] + z[k-1].
//
// A block owns SC_ROWS series.  Wave 0 runs the dependent adds over a
// SC_T-sample tile held in LDS while waves 1..8 store the previous tile's
// prefix sums and load the next tile (double buffer), so the add chain never
// waits on HBM latency.  Wave 0 gives each series 8 lanes: the tile is walked
// in steps of 8 * SC_SEG samples, lane (series, seg) holds samples
// seg*SC_SEG .. seg*SC_SEG + SC_SEG-1 of the step in registers, and the 8
// segments take turns adding theirs onto the running sum, which a DPP move
// hands from each lane to the next.  Every sample is still added in order,
// one float add at a time, and the chain pays one LDS read and one write per
// SC_SEG samples instead of per 4.
//
// When rms is given the loader normalises on the fly, z = (x - baseline)/rms
// exactly as k_normalise does, so the pipeline never writes the normalised
// plane back to HBM.
constexpr int SC_ROWS = 8;
constexpr int SC_T = 1024;
constexpr int SC_PAD = 4;
constexpr int SC_MOVERS = 512;                  // waves 1..8
constexpr int SC_THREADS = 64 + SC_MOVERS;
constexpr int SC_VEC = SC_ROWS * SC_T / 4;       // float4 per tile
constexpr int SC_PER = SC_VEC / SC_MOVERS;       // float4 per mover thread
// samples per lane and step (32 measured slower)
constexpr int SC_SEG = 16;
static_assert(SC_T % (16 * SC_SEG) == 0, "two steps per loop trip");

__device__ __forceinline__ void scan4(float4& v, float& run) {
  run = run + v.x;
  v.x = run;
  run = run + v.y;
  v.y = run;
  run = run + v.z;
  v.z = run;
  run = run + v.w;
  v.w = run;
}

// DPP row moves (within each 16-lane row): lane i takes lane i-1's value
// (row_shr:1), or lane (i-9) mod 16's (row_ror:9), which for the two 8-lane
// groups of a row is each group's lane 7 into its lane 0.
template <int CTRL>
__device__ __forceinline__ float dpp_move(float v) {
  const int i = __float_as_int(v);
  return __int_as_float(__builtin_amdgcn_update_dpp(i, i, CTRL, 0xF, 0xF, false));
}

// One step of the segmented chain: segment k (lane seg == k of each 8-lane
// group) adds its SEG samples onto run, then run moves one lane on; after
// the last segment it wraps to the group's lane 0, ready for the next step.
template <int SEG>
__device__ __forceinline__ void scan_segments(float4 (&v)[SEG / 4], float& run, int seg) {
#pragma unroll
  for (int k = 0; k < 8; ++k) {
    if (seg == k) {
#pragma unroll
      for (int i = 0; i < SEG / 4; ++i) scan4(v[i], run);
    }
    run = k < 7 ? dpp_move<0x111>(run) : dpp_move<0x129>(run);
  }
}

constexpr int SC_PMAX = 1024;                    // staged baseline points per block
constexpr int SC_RMAX = 256;                     // staged beam RMS values per block

// STAGED: the block's baseline points and beam RMS values live in LDS (the
// normalising loader reads them for every sample); otherwise they are read
// from global memory.
template <bool STAGED>
__global__ __launch_bounds__(SC_THREADS) void k_scan(const float* __restrict__ z, size_t zpitch,
                                                     size_t count_arg, int n_series,
                                                     float* __restrict__ S, size_t spitch,
                                                     int s_vec_ok, const float* __restrict__ pts,
                                                     const float* __restrict__ rms,
                                                     SeriesShape sh_arg,
                                                     const GroupDev* __restrict__ gd,
                                                     size_t prow_arg, unsigned nbm_arg) {
  __shared__ float buf[2][SC_ROWS][SC_T + SC_PAD];
  __shared__ float sP[STAGED ? SC_PMAX + 1 : 1];
  __shared__ float sR[STAGED ? SC_RMAX : 1];
  const int tid = threadIdx.x;
  // the add chain is the critical path: its wave goes first when it shares
  // a SIMD with a mover wave
  if (tid < 64) __builtin_amdgcn_s_setprio(3);
  // blocks: (group, SC_ROWS-row slice of its n_series rows); rows are
  // g*n_series + r in z, S, pts and rms
  const int bpg = (n_series + SC_ROWS - 1) / SC_ROWS;
  const int gz = blockIdx.x / bpg;
  const int r0 = (blockIdx.x % bpg) * SC_ROWS;
  const SeriesShape sh = gd ? gd[gz].sh : sh_arg;
  const size_t count = gd ? sh.nout : count_arg;
  const size_t rowbase = (size_t)gz * n_series;
  z += rowbase * zpitch;
  S += rowbase * spitch;
  const size_t prow = gd ? prow_arg : (size_t)sh.nbeams * sh.sc;  // baseline points per row
  const unsigned nbm = gd ? nbm_arg : sh.nbeams;                   // RMS values per row
  if (rms) {
    pts += rowbase * prow;
    rms += rowbase * nbm;
  }
  const size_t ntiles = (count + SC_T - 1) / SC_T;
  const bool z_vec = (zpitch % 4 == 0) && ((reinterpret_cast<uintptr_t>(z) & 15) == 0);
  if (STAGED) {
    // rows r0.. of pts (plus the float after them, which the stretch may
    // touch at the very end, as it would in global memory) and of rms
    const int nrow = min(SC_ROWS, n_series - r0);
    const size_t np = (size_t)nrow * prow + 1;
    for (size_t i = tid; i < np; i += SC_THREADS) sP[i] = pts[(size_t)r0 * prow + i];
    for (size_t i = tid; i < (size_t)nrow * nbm; i += SC_THREADS) sR[i] = rms[(size_t)r0 * nbm + i];
  }

  auto norm = [&](int row, size_t t, float x) -> float {
    const float* P = STAGED ? sP + (size_t)row * prow : pts + (size_t)(r0 + row) * prow;
    const float* R = STAGED ? sR + (size_t)row * nbm : rms + (size_t)(r0 + row) * nbm;
    return normalised_at(x, P, R, sh, t);
  };
  // Tile loads go into registers (nxt) one iteration before they are
  // normalised and written to LDS, so the barrier between iterations (which
  // waits for LDS traffic only) does not wait for HBM: the loads of tile p+2
  // are in flight while the scan lane works through tile p.
  float4 nxt[SC_PER];
  auto load_raw = [&](size_t tile) {
    const int mt = tid - 64;
    const bool full = z_vec && (tile + 1) * SC_T <= count;
    // rows past the last series load the last series (zeroed in commit)
    if (full) {
#pragma unroll
      for (int q = 0; q < SC_PER; ++q) {
        const int idx = mt + SC_MOVERS * q;
        const int row = idx / (SC_T / 4);
        const int c4 = (idx % (SC_T / 4)) * 4;
        const int rr = min(r0 + row, n_series - 1);
        nxt[q] = *reinterpret_cast<const float4*>(z + (size_t)rr * zpitch + tile * SC_T + c4);
      }
    } else {
#pragma unroll
      for (int q = 0; q < SC_PER; ++q) {
        const int idx = mt + SC_MOVERS * q;
        const int row = idx / (SC_T / 4);
        const int c4 = (idx % (SC_T / 4)) * 4;
        const size_t t = tile * SC_T + c4;
        nxt[q] = make_float4(0.f, 0.f, 0.f, 0.f);
        if (r0 + row < n_series) {
          const float* src = z + (size_t)(r0 + row) * zpitch + t;
          if (t + 0 < count) nxt[q].x = src[0];
          if (t + 1 < count) nxt[q].y = src[1];
          if (t + 2 < count) nxt[q].z = src[2];
          if (t + 3 < count) nxt[q].w = src[3];
        }
      }
    }
  };
  // nxt (tile `tile`) normalised into buf[b]
  auto commit = [&](size_t tile, int b) {
    const int mt = tid - 64;
#pragma unroll
    for (int q = 0; q < SC_PER; ++q) {
      const int idx = mt + SC_MOVERS * q;
      const int row = idx / (SC_T / 4);
      const int c4 = (idx % (SC_T / 4)) * 4;
      const size_t t = tile * SC_T + c4;
      float4 v = nxt[q];
      if (r0 + row >= n_series) {
        v = make_float4(0.f, 0.f, 0.f, 0.f);
      } else if (rms) {
        if (t + 0 < count) v.x = norm(row, t + 0, v.x);
        if (t + 1 < count) v.y = norm(row, t + 1, v.y);
        if (t + 2 < count) v.z = norm(row, t + 2, v.z);
        if (t + 3 < count) v.w = norm(row, t + 3, v.w);
      }
      *reinterpret_cast<float4*>(&buf[b][row][c4]) = v;
    }
  };
  auto store_tile = [&](size_t tile, int b) {
    const int mt = tid - 64;
    float4 v[SC_PER];
#pragma unroll
    for (int q = 0; q < SC_PER; ++q) {
      const int idx = mt + SC_MOVERS * q;
      v[q] = *reinterpret_cast<const float4*>(&buf[b][idx / (SC_T / 4)][(idx % (SC_T / 4)) * 4]);
    }
#pragma unroll
    for (int q = 0; q < SC_PER; ++q) {
      const int idx = mt + SC_MOVERS * q;
      const int row = idx / (SC_T / 4);
      const int c4 = (idx % (SC_T / 4)) * 4;
      const size_t t = tile * SC_T + c4;
      if (r0 + row >= n_series) continue;
      float* dst = S + (size_t)(r0 + row) * spitch + t + 1;
      if (s_vec_ok && t + 3 < count) {
        *reinterpret_cast<float4*>(dst) = v[q];
      } else {
        if (t + 0 < count) dst[0] = v[q].x;
        if (t + 1 < count) dst[1] = v[q].y;
        if (t + 2 < count) dst[2] = v[q].z;
        if (t + 3 < count) dst[3] = v[q].w;
      }
    }
  };

  if (tid < SC_ROWS && r0 + tid < n_series) S[(size_t)(r0 + tid) * spitch] = 0.f;
  if (STAGED) __syncthreads();
  if (tid >= 64 && ntiles > 0) {
    load_raw(0);
    commit(0, 0);
    if (ntiles > 1) load_raw(1);
  }
  __syncthreads();
  float run = 0.f;
  for (size_t p = 0; p <= ntiles; ++p) {
    const int b = (int)(p & 1);
    if (tid < 64) {
      const int row = tid >> 3, seg = tid & 7;
      if (p < ntiles && (size_t)SC_T * (p + 1) <= count) {
        // full tile: lane (row, seg) holds samples 16*seg .. 16*seg + 15 of
        // each 128-sample step of its row in registers; two steps in
        // flight (load B, add A, store A, load A', add B ...)
        constexpr int SEG = SC_SEG;
        float4* s4 = reinterpret_cast<float4*>(&buf[b][row][SEG * seg]);
        constexpr int STEP4 = SEG * 8 / 4;  // float4 per step and row
        float4 va[SEG / 4], vb[SEG / 4];
#pragma unroll
        for (int i = 0; i < SEG / 4; ++i) va[i] = s4[i];
        for (int u = 0; u < SC_T / (8 * SEG); u += 2) {
#pragma unroll
          for (int i = 0; i < SEG / 4; ++i) vb[i] = s4[STEP4 * (u + 1) + i];
          scan_segments<SEG>(va, run, seg);
#pragma unroll
          for (int i = 0; i < SEG / 4; ++i) s4[STEP4 * u + i] = va[i];
          if (u + 2 < SC_T / (8 * SEG)) {
#pragma unroll
            for (int i = 0; i < SEG / 4; ++i) va[i] = s4[STEP4 * (u + 2) + i];
          }
          scan_segments<SEG>(vb, run, seg);
#pragma unroll
          for (int i = 0; i < SEG / 4; ++i) s4[STEP4 * (u + 1) + i] = vb[i];
        }
      } else if (p < ntiles && seg == 0) {
        // the short last tile: the segment-0 lane of each row (which holds
        // the row's running sum) adds it alone
        const int ncol = (int)(count - p * SC_T);
        float* rowp = &buf[b][row][0];
        float4* row4 = reinterpret_cast<float4*>(rowp);
        int c = 0;
        for (; c + 4 <= ncol; c += 4) {
          float4 v = row4[c / 4];
          scan4(v, run);
          row4[c / 4] = v;
        }
        for (; c < ncol; ++c) {
          run = run + rowp[c];
          rowp[c] = run;
        }
      }
    } else {
      if (p >= 1) store_tile(p - 1, b ^ 1);
      if (p + 1 < ntiles) commit(p + 1, b ^ 1);
      if (p + 2 < ntiles) load_raw(p + 2);
    }
    __syncthreads();
  }
}

}  // namespace

hipError_t launch_baseline_points(const float* plane, const SeriesShape& sh, int n_dm,
                                  float* pts, hipStream_t s, const GroupArgs& ga) {
  size_t nl_max = 0;
  unsigned nb = sh.nbeams;
  if (ga.gd) {
    nb = ga.nbmax;
    for (int g = 0; g < ga.G; ++g) {
      const SeriesShape& q = ga.host[g].sh;
      if (q.sc) nl_max = std::max(nl_max, q.count_round / pow5_host(q.bl_direct));
    }
  } else if (sh.sc) {
    nl_max = sh.count_round / pow5_host(sh.bl_direct);
  }
  if (nl_max == 0) return hipSuccess;  // no baseline anywhere (sample_count 0)
  const size_t lds = (nl_max + nl_max / 5 + 8) * sizeof(float);
  hipLaunchKernelGGL(k_baseline_points, dim3(nb, n_dm, (unsigned)ga.G), dim3(NB_BS), lds, s,
                     plane, sh, pts, ga.gd, ga.prow);
  return hipGetLastError();
}

namespace {
int rms_levels_host(size_t bc) {
  int L0 = 0;
  while (L0 < 5 && bc >= pow5_host(L0 + 1)) ++L0;
  return L0;
}
// largest partial count and chunk over the groups (or the one shape)
void rms_maxima(const SeriesShape& sh, const GroupArgs& ga, size_t* np_max, size_t* chunk_max) {
  *np_max = 0;
  *chunk_max = 0;
  const int G = ga.gd ? ga.G : 1;
  for (int g = 0; g < G; ++g) {
    const SeriesShape& q = ga.gd ? ga.host[g].sh : sh;
    const size_t chunk = pow5_host(rms_levels_host(q.bc));
    *np_max = std::max(*np_max, q.bc / chunk);
    *chunk_max = std::max(*chunk_max, chunk);
  }
}
}  // namespace

size_t beam_rms_scratch_words(const SeriesShape& sh, int n_dm, const GroupArgs& ga) {
  size_t np_max, chunk_max;
  rms_maxima(sh, ga, &np_max, &chunk_max);
  const size_t nbm = ga.gd ? ga.nbmax : sh.nbeams;
  return (size_t)n_dm * ga.G * nbm * np_max;
}

hipError_t launch_beam_rms(const float* plane, const float* pts, const SeriesShape& sh,
                           int n_dm, float* rms, hipStream_t s, float* scratch,
                           size_t scratch_words, const GroupArgs& ga) {
  size_t np_max, chunk_max;
  rms_maxima(sh, ga, &np_max, &chunk_max);
  const unsigned nbm = ga.gd ? ga.nbmax : sh.nbeams;
  const bool fits = (size_t)n_dm * ga.G * nbm * np_max <= scratch_words &&
                    (np_max + np_max / 5 + 8) * sizeof(float) <= 65536;
  if (scratch && fits && (ga.gd || chunk_max >= 125)) {
    const size_t prow = ga.gd ? ga.prow : (size_t)sh.nbeams * sh.sc;
    const int stage_pts = (pts && prow > 0 && prow + 1 <= 4096) ? 1 : 0;
    const size_t lds =
        (chunk_max + chunk_max / 5 + 8 + (stage_pts ? prow + 1 : 0)) * sizeof(float);
    hipLaunchKernelGGL(k_beam_rms_partials, dim3((unsigned)np_max, n_dm, ga.G * nbm),
                       dim3(NB_BS), lds, s, plane, pts, sh, np_max, stage_pts, scratch, ga.gd,
                       ga.prow, nbm);
    hipLaunchKernelGGL(k_beam_rms_finish, dim3(nbm, n_dm, ga.G), dim3(64),
                       (np_max + np_max / 5 + 8) * sizeof(float), s, scratch, sh, np_max, nbm,
                       rms, ga.gd);
    return hipGetLastError();
  }
  if (ga.gd) return hipErrorInvalidValue;  // batched calls always bring scratch
  const size_t nl = sh.bc / pow5_host(sh.rms_direct);
  const size_t lds = (nl + nl / 5 + 8) * sizeof(float);
  hipLaunchKernelGGL(k_beam_rms, dim3(sh.nbeams, n_dm), dim3(NB_BS), lds, s, plane, pts, sh,
                     rms);
  return hipGetLastError();
}

hipError_t launch_normalise(float* plane, const float* pts, const float* rms,
                            const SeriesShape& sh, int n_dm, hipStream_t s) {
  size_t gx = (sh.nout + 255) / 256;
  if (gx > 1024) gx = 1024;
  hipLaunchKernelGGL(k_normalise, dim3((unsigned)gx, n_dm), dim3(256), 0, s, plane, pts, rms,
                     sh);
  return hipGetLastError();
}

hipError_t launch_scan(const float* z, size_t zpitch, size_t count, int n_series, float* S,
                       size_t spitch, hipStream_t s, const float* pts, const float* rms,
                       const SeriesShape* sh, const GroupArgs& ga) {
  const int s_vec_ok = (spitch % 4 == 0) && (((reinterpret_cast<uintptr_t>(S) + 4) & 15) == 0);
  SeriesShape none{};
  if (rms && !sh && !ga.gd) return hipErrorInvalidValue;
  const SeriesShape one = sh ? *sh : none;
  const int bpg = (n_series + SC_ROWS - 1) / SC_ROWS;
  const dim3 grid((unsigned)(bpg * ga.G));
  const size_t prow = ga.gd ? ga.prow : (size_t)one.nbeams * one.sc;
  const unsigned nbm = ga.gd ? ga.nbmax : one.nbeams;
  const bool staged = rms && (size_t)SC_ROWS * prow + 1 <= (size_t)SC_PMAX &&
                      (size_t)SC_ROWS * nbm <= (size_t)SC_RMAX;
  if (staged)
    hipLaunchKernelGGL(k_scan<true>, grid, dim3(SC_THREADS), 0, s, z, zpitch, count, n_series, S,
                       spitch, s_vec_ok, pts, rms, one, ga.gd, ga.prow, ga.nbmax);
  else
    hipLaunchKernelGGL(k_scan<false>, grid, dim3(SC_THREADS), 0, s, z, zpitch, count, n_series,
                       S, spitch, s_vec_ok, pts, rms, one, ga.gd, ga.prow, ga.nbmax);
  return hipGetLastError();
}

}  // namespace hd
