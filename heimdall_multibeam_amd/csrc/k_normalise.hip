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
//              prefix sum in a fixed three-level order (segments of 16,
//              chunks of 1024, chunks in turn), four waves per series.
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
  // (a 32-bit division when bc fits: the 64-bit one is a long sequence)
  const size_t np =
      sh.bc <= 0xffffffffull ? (size_t)((unsigned)sh.bc / (unsigned)chunk) : sh.bc / chunk;
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

constexpr int SC_PMAX = 1024;                    // staged baseline points per block
constexpr int SC_RMAX = 256;                     // staged beam RMS values per block

// Exclusive scan per series in a fixed three-level order (the oracle's
// orc_scan follows the same order):
//   samples in segments of 16, summed in order inside a segment (P, the
//   inclusive sum up to the sample);
//   segments in chunks of 64 (1024 samples), their totals summed in order
//   inside the chunk (Q, the sum of the chunk's earlier segments);
//   chunks in order along the series (O, the sum of the earlier chunks);
//   S[0] = 0 and S[k + 1] = O + (Q + P) for sample k.
// With few series, one block of SQ_WAVES waves per series; the waves take
// the chunks in turn (wave w: chunks w, w + SQ_WAVES, ...).  Lane j holds
// segment j of the wave's chunk in registers, so P is 16 adds per lane and Q
// a 64-step chain of wave-uniform adds over the lanes' totals (v_readlane).
// Q restarts in every chunk, so the waves' chunks are independent up to O:
// after each round of SQ_WAVES chunks the waves exchange their chunk totals
// in LDS (one barrier) and each forms its chunk's O from the round's carry
// and the earlier chunks' totals, in order.  The dependent chain of a series
// is a few adds per chunk instead of 1024, and with four waves per series the
// kernel is bound by its HBM traffic (the plane read once, S written once);
// each wave loads its next chunk while it sums this one.
//
// When rms is given the loads are normalised on the fly, z = (x -
// baseline)/rms exactly as k_normalise does, so the pipeline never writes the
// normalised plane back to HBM.
// With many series (batches of beams) one wave per series keeps the GPU as
// busy without the barriers: WPS waves per series, SQ_WAVES / WPS series per
// block; the order of the additions is the same either way.
constexpr int SQ_SEG = 16;
constexpr int SQ_CHUNK = 64 * SQ_SEG;
constexpr int SQ_WAVES = 4;  // waves per block
constexpr size_t SQ_WAVE_SERIES = 4096;  // from this many series on, one wave per series

template <bool STAGED, int WPS>
__global__ __launch_bounds__(64 * SQ_WAVES) void k_scan_seg(
    const float* __restrict__ z, size_t zpitch, size_t count_arg, int n_series,
    float* __restrict__ S, size_t spitch, int s_vec_ok, const float* __restrict__ pts,
    const float* __restrict__ rms, SeriesShape sh_arg, const GroupDev* __restrict__ gd,
    size_t prow_arg, unsigned nbm_arg) {
  static_assert(WPS == 1 || WPS == SQ_WAVES, "a block's waves share one series or none");
  constexpr int NSER = SQ_WAVES / WPS;  // series per block
  __shared__ float sP[STAGED ? SC_PMAX + 1 : 1];
  __shared__ float sR[STAGED ? SC_RMAX : 1];
  __shared__ float sT[2][SQ_WAVES];  // chunk totals of a round (two rounds alternate)
  __shared__ __align__(16) float sQ[SQ_WAVES][68];  // each wave's segment totals
  const int tid = threadIdx.x;
  const int wave = tid >> 6, lane = tid & 63;
  const int ls = wave / WPS, lw = wave % WPS;  // the wave's series in the block, its turn
  // blocks: (group, NSER-row slice of its n_series rows); rows are
  // g*n_series + r in z, S, pts and rms
  const int bpg = (n_series + NSER - 1) / NSER;
  const int gz = blockIdx.x / bpg;
  const int r0 = (blockIdx.x % bpg) * NSER;
  const bool valid = r0 + ls < n_series;
  const size_t rg = (size_t)gz * n_series + min(r0 + ls, n_series - 1);
  const SeriesShape sh = gd ? gd[gz].sh : sh_arg;
  const size_t count = valid ? (gd ? sh.nout : count_arg) : 0;
  const size_t prow = gd ? prow_arg : (size_t)sh.nbeams * sh.sc;  // baseline points per row
  const unsigned nbm = gd ? nbm_arg : sh.nbeams;                   // RMS values per row
  const float* zr = z + rg * zpitch;
  float* Sr = S + rg * spitch;
  if (STAGED) {
    // rows r0.. of pts (plus the float after them, which the stretch may
    // touch at the very end, as it would in global memory) and of rms
    const size_t rb = (size_t)gz * n_series + r0;
    const int nrow = min(NSER, n_series - r0);
    for (size_t i = tid; i < (size_t)nrow * prow + 1; i += 64 * SQ_WAVES) sP[i] = pts[rb * prow + i];
    for (size_t i = tid; i < (size_t)nrow * nbm; i += 64 * SQ_WAVES) sR[i] = rms[rb * nbm + i];
    __syncthreads();
  }
  const float* P = STAGED ? sP + (size_t)ls * prow : (rms ? pts + rg * prow : nullptr);
  const float* R = STAGED ? sR + (size_t)ls * nbm : (rms ? rms + rg * nbm : nullptr);
  const bool z_vec = (zpitch % 4 == 0) && ((reinterpret_cast<uintptr_t>(z) & 15) == 0);
  if (tid % (64 * WPS) == 0 && valid) Sr[0] = 0.f;
  const size_t nchunks = (count + SQ_CHUNK - 1) / SQ_CHUNK;
  const size_t nrounds = (nchunks + WPS - 1) / WPS;
  // this lane's segment of chunk c: samples t0 .. t0 + 15
  auto load = [&](size_t c, float4 (&v)[SQ_SEG / 4]) {
    const size_t t0 = c * SQ_CHUNK + (size_t)lane * SQ_SEG;
    if (z_vec && t0 + SQ_SEG <= count) {
#pragma unroll
      for (int i = 0; i < SQ_SEG / 4; ++i)
        v[i] = *reinterpret_cast<const float4*>(zr + t0 + 4 * i);
    } else {
#pragma unroll
      for (int i = 0; i < SQ_SEG / 4; ++i) {
        const size_t t = t0 + 4 * i;
        v[i].x = t + 0 < count ? zr[t + 0] : 0.f;
        v[i].y = t + 1 < count ? zr[t + 1] : 0.f;
        v[i].z = t + 2 < count ? zr[t + 2] : 0.f;
        v[i].w = t + 3 < count ? zr[t + 3] : 0.f;
      }
    }
  };
  float4 va[SQ_SEG / 4], vb[SQ_SEG / 4];
  if ((size_t)lw < nchunks) load(lw, va);
  float O = 0.f;  // the chunks before this round, in order (the same in every wave)
  for (size_t r = 0; r < nrounds; ++r) {
    const size_t c = r * WPS + lw;
    if (c + WPS < nchunks) load(c + WPS, vb);
    const size_t t0 = c * SQ_CHUNK + (size_t)lane * SQ_SEG;
    const int nv = (c < nchunks && t0 < count) ? (int)min((size_t)SQ_SEG, count - t0) : 0;
    float x[SQ_SEG];
#pragma unroll
    for (int i = 0; i < SQ_SEG / 4; ++i) {
      x[4 * i + 0] = va[i].x;
      x[4 * i + 1] = va[i].y;
      x[4 * i + 2] = va[i].z;
      x[4 * i + 3] = va[i].w;
    }
    // P: in order inside the segment (samples past the series are not added;
    // a whole segment, the usual case, without the per-sample tests)
    float p = 0.f;
    if (nv == SQ_SEG) {
#pragma unroll
      for (int i = 0; i < SQ_SEG; ++i) {
        const float v = rms ? normalised_at(x[i], P, R, sh, t0 + i) : x[i];
        p = p + v;
        x[i] = p;
      }
    } else {
#pragma unroll
      for (int i = 0; i < SQ_SEG; ++i) {
        if (i < nv) {
          const float v = rms ? normalised_at(x[i], P, R, sh, t0 + i) : x[i];
          p = p + v;
          x[i] = p;
        }
      }
    }
    // Q: the chunk's segment totals in lane order, one chain of adds that
    // lane 0 runs over the totals staged in LDS; entry s becomes the sum of
    // totals 0 .. s-1 and entry 64 the chunk's total
    float* qb = sQ[wave];
    qb[lane] = p;
    wave_sync();
    if (lane == 0) {
      float acc = 0.f;
#pragma unroll
      for (int s = 0; s < 64; ++s) {
        const float ts = qb[s];
        qb[s] = acc;
        acc = acc + ts;
      }
      qb[64] = acc;
    }
    wave_sync();
    const float mine = qb[lane];
    const float q = qb[64];
    wave_sync();  // qb is rewritten by the next chunk
    // O of this chunk: the round's carry plus the earlier chunks of the round
    // in order; a chunk past the series adds a zero total after every real one
    float o = O;
    if (WPS == 1) {
      O = O + q;
    } else {  // the block's waves share one series and take the same rounds
      if (lane == 0) sT[r & 1][lw] = q;
      __syncthreads();
#pragma unroll
      for (int w = 0; w < WPS; ++w) {
        const float tw = sT[r & 1][w];
        if (w < lw) o = o + tw;
        O = O + tw;
      }
    }
#pragma unroll
    for (int i = 0; i < SQ_SEG; ++i) x[i] = o + (mine + x[i]);
    float* dst = Sr + t0 + 1;
    if (s_vec_ok && nv == SQ_SEG) {
#pragma unroll
      for (int i = 0; i < SQ_SEG / 4; ++i)
        *reinterpret_cast<float4*>(dst + 4 * i) =
            make_float4(x[4 * i], x[4 * i + 1], x[4 * i + 2], x[4 * i + 3]);
    } else {
#pragma unroll
      for (int i = 0; i < SQ_SEG; ++i)
        if (i < nv) dst[i] = x[i];
    }
#pragma unroll
    for (int i = 0; i < SQ_SEG / 4; ++i) va[i] = vb[i];
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
  const size_t prow = ga.gd ? ga.prow : (size_t)one.nbeams * one.sc;
  const unsigned nbm = ga.gd ? ga.nbmax : one.nbeams;
  // enough series to fill the GPU with one wave each: no barriers in the loop
  const bool one_wave = (size_t)n_series * ga.G >= SQ_WAVE_SERIES;
  const int nser = one_wave ? SQ_WAVES : 1;
  const bool staged =
      rms && nser * prow + 1 <= (size_t)SC_PMAX && nser * (size_t)nbm <= (size_t)SC_RMAX;
  const dim3 grid((unsigned)(((n_series + nser - 1) / nser) * ga.G));
  const dim3 block(64 * SQ_WAVES);
#define HD_SCAN_ARGS z, zpitch, count, n_series, S, spitch, s_vec_ok, pts, rms, one, ga.gd, \
                     ga.prow, ga.nbmax
  if (one_wave && staged)
    hipLaunchKernelGGL((k_scan_seg<true, 1>), grid, block, 0, s, HD_SCAN_ARGS);
  else if (one_wave)
    hipLaunchKernelGGL((k_scan_seg<false, 1>), grid, block, 0, s, HD_SCAN_ARGS);
  else if (staged)
    hipLaunchKernelGGL((k_scan_seg<true, SQ_WAVES>), grid, block, 0, s, HD_SCAN_ARGS);
  else
    hipLaunchKernelGGL((k_scan_seg<false, SQ_WAVES>), grid, block, 0, s, HD_SCAN_ARGS);
#undef HD_SCAN_ARGS
  return hipGetLastError();
}

}  // namespace hd
