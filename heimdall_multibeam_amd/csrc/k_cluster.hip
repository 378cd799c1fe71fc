// k_cluster.hip -- candidate clustering and merging.
//
// Clustering (Pipeline/label_candidate_clusters.cu:141-197, are_coincident.cuh:
// 46-68).  The reference relabels every giant i to the minimum label over all
// coincident giants in one parallel pass and then traces label chains; with a
// single thread (the sequential semantics adopted here, since the
// parallel pass is order dependent) that is
//     L[i] = min(i, min{ L[j] : j < i, coincident(i, j) }),
// i.e. L[i] is the smallest index reachable from i along coincident links to
// earlier giants.  Here that fixed point is reached by repeated in-place
// relaxation with pointer jumping (L[i] <- min(L[i], L[L[i]], L[j])): every
// label stays a reachable index and only decreases, so the iteration stops at
// the unique fixed point.  Neighbours are found through a sort of the giants
// by sample index, so each giant only visits giants within the widest time
// tolerance.
//
// Merging (merge_candidates.cu:21-72,102-131): giants sharing a root form a
// group, ordered by root; its fields come from the first (lowest index) giant
// with the maximum S/N and its member count is the sum.  A 64-bit atomic max
// over (orderable S/N, ~index) picks that giant.
#include <algorithm>
#include <vector>

#include "hd_device.hpp"
#include "hd_internal.hpp"

namespace hd {

namespace {

// Clustering coordinate of giant k: its sample index, shifted by its group's
// offset in a batched call so that groups lie further apart than any
// coincidence window.
__device__ __forceinline__ uint64_t tcoord(const uint64_t* ind, const uint32_t* group,
                                           const GroupDev* gd, unsigned k) {
  return group ? gd[group[k]].t_off + ind[k] : ind[k];
}

__global__ void k_sort_init(const uint64_t* __restrict__ ind, const uint32_t* __restrict__ group,
                            const GroupDev* __restrict__ gd, unsigned n, uint64_t* keys,
                            size_t npad) {
  const size_t k = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= npad) return;
  keys[k] = (k < n) ? ((tcoord(ind, group, gd, (unsigned)k) << 32) | (uint64_t)k) : ~0ull;
}

__global__ void k_bitonic_step(uint64_t* keys, size_t npad, size_t j, size_t kk) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= npad) return;
  const size_t l = i ^ j;
  if (l <= i) return;
  const uint64_t a = keys[i], b = keys[l];
  const bool up = ((i & kk) == 0);
  if ((a > b) == up) {
    keys[i] = b;
    keys[l] = a;
  }
}

// Steps j < kBitonicBlock of the k_bitonic_step network on tiles of
// kBitonicBlock keys held in LDS, one block per tile.  kk_lo == 0: every
// stage kk up to the tile size (npad itself when smaller), which sorts each
// tile up or down as the network would; otherwise the steps j =
// kBitonicBlock/2 .. 1 of stage kk_lo, whose pairs i, i^j never leave a
// tile.  Same compare-exchanges (each pair handled by one thread, a barrier
// between steps) and directions (from the global index), so the same order.
constexpr unsigned kBitonicBlock = 4096;
__global__ __launch_bounds__(1024) void k_bitonic_tile(uint64_t* keys, unsigned npad,
                                                       unsigned kk_lo) {
  __shared__ uint64_t K[kBitonicBlock];
  const unsigned base = blockIdx.x * kBitonicBlock;
  const unsigned nt = npad < kBitonicBlock ? npad : kBitonicBlock;
  for (unsigned i = threadIdx.x; i < nt; i += 1024) K[i] = keys[base + i];
  __syncthreads();
  const unsigned k_first = kk_lo ? kk_lo : 2, k_last = kk_lo ? kk_lo : nt;
  for (unsigned kk = k_first; kk <= k_last; kk <<= 1)
    for (unsigned j = (kk_lo ? nt : kk) >> 1; j > 0; j >>= 1) {
      for (unsigned i = threadIdx.x; i < nt; i += 1024) {
        const unsigned l = i ^ j;
        if (l > i) {
          const uint64_t a = K[i], b = K[l];
          const bool up = ((base + i) & kk) == 0;
          if ((a > b) == up) {
            K[i] = b;
            K[l] = a;
          }
        }
      }
      __syncthreads();
    }
  for (unsigned i = threadIdx.x; i < nt; i += 1024) keys[base + i] = K[i];
}

__device__ __forceinline__ unsigned lower_bound_ind(const uint64_t* keys, unsigned n,
                                                    uint64_t v) {
  unsigned lo = 0, hi = n;
  while (lo < hi) {
    const unsigned mid = (lo + hi) >> 1;
    if ((keys[mid] >> 32) < v) lo = mid + 1;
    else hi = mid;
  }
  return lo;
}

__global__ void k_windows(const uint64_t* __restrict__ ind, const uint32_t* __restrict__ group,
                          const GroupDev* __restrict__ gd, unsigned n,
                          const uint64_t* __restrict__ keys, uint64_t span, uint32_t* lo,
                          uint32_t* hi, int32_t* labels) {
  const unsigned i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const uint64_t t = tcoord(ind, group, gd, i);
  lo[i] = lower_bound_ind(keys, n, t > span ? t - span : 0);
  hi[i] = lower_bound_ind(keys, n, t + span + 1);
  labels[i] = (int32_t)i;
}

// are_coincident_beam (are_coincident.cuh:46-68) with the reference's int
// casts: time tolerance scaled by the wider of the two boxcars.
// Giant i's beam as a sample range: ij / nsamps_beam == ii / nsamps_beam
// exactly when ij lies in it, so the 64-bit division is taken once per giant
// instead of once per pair.
struct BeamRange {
  uint64_t lo, hi;
};
__device__ __forceinline__ BeamRange beam_range(uint64_t ii, uint64_t nsamps_beam) {
  const uint64_t lo = ii / nsamps_beam * nsamps_beam;
  return BeamRange{lo, lo + nsamps_beam};
}
__device__ __forceinline__ bool coincident(uint64_t ii, uint64_t ij, uint32_t fi, uint32_t fj,
                                           uint32_t di, uint32_t dj, uint64_t time_tol,
                                           uint64_t filter_tol, uint64_t dm_tol,
                                           const BeamRange& br) {
  if (ij < br.lo || ij >= br.hi) return false;
  const int wi = 1 << (int)fi, wj = 1 << (int)fj;
  const uint64_t tt = time_tol * (uint64_t)(wi > wj ? wi : wj);
  const int dt = abs((int)ij - (int)ii);
  const int df = abs((int)fj - (int)fi);
  const int dd = abs((int)dj - (int)di);
  return (uint64_t)dt <= tt && (uint64_t)df <= filter_tol && (uint64_t)dd <= dm_tol;
}

// The windows only hold giants of i's own group, which share its beam length.
__global__ void k_relax(const uint64_t* __restrict__ ind, const uint32_t* __restrict__ filter,
                        const uint32_t* __restrict__ dm, unsigned n,
                        const uint64_t* __restrict__ keys, const uint32_t* __restrict__ lo,
                        const uint32_t* __restrict__ hi, uint64_t time_tol, uint64_t filter_tol,
                        uint64_t dm_tol, uint64_t nsamps_beam_one,
                        const uint32_t* __restrict__ group, const GroupDev* __restrict__ gd,
                        int32_t* labels, unsigned* flag) {
  const unsigned i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const uint64_t nsamps_beam = group ? gd[group[i]].nsamps_beam : nsamps_beam_one;
  const int32_t cur = __atomic_load_n(&labels[i], __ATOMIC_RELAXED);
  int32_t best = cur;
  const int32_t jump = __atomic_load_n(&labels[cur], __ATOMIC_RELAXED);
  if (jump < best) best = jump;
  const uint64_t ii = ind[i];
  const BeamRange br = beam_range(ii, nsamps_beam);
  const uint32_t fi = filter[i], di = dm[i];
  for (unsigned p = lo[i]; p < hi[i]; ++p) {
    const unsigned j = (unsigned)(keys[p] & 0xffffffffull);
    if (j >= i) continue;
    if (!coincident(ii, ind[j], fi, filter[j], di, dm[j], time_tol, filter_tol, dm_tol,
                    br))
      continue;
    const int32_t lj = __atomic_load_n(&labels[j], __ATOMIC_RELAXED);
    if (lj < best) best = lj;
  }
  if (best < cur) {
    __atomic_store_n(&labels[i], best, __ATOMIC_RELAXED);
    *flag = 1u;
  }
}

// The same relaxation reading the window from copies of ind, filter and dm
// laid out in the sorted order (s_ind, s_fd = dm << 5 | filter).
__global__ void k_relax_sorted(const uint64_t* __restrict__ ind,
                               const uint32_t* __restrict__ filter,
                               const uint32_t* __restrict__ dm, unsigned n,
                               const uint64_t* __restrict__ keys,
                               const uint64_t* __restrict__ s_ind,
                               const uint32_t* __restrict__ s_fd,
                               const uint32_t* __restrict__ lo, const uint32_t* __restrict__ hi,
                               uint64_t time_tol, uint64_t filter_tol, uint64_t dm_tol,
                               uint64_t nsamps_beam_one, const uint32_t* __restrict__ group,
                               const GroupDev* __restrict__ gd, int32_t* labels,
                               unsigned* flag) {
  const unsigned i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const uint64_t nsamps_beam = group ? gd[group[i]].nsamps_beam : nsamps_beam_one;
  const int32_t cur = __atomic_load_n(&labels[i], __ATOMIC_RELAXED);
  int32_t best = cur;
  const int32_t jump = __atomic_load_n(&labels[cur], __ATOMIC_RELAXED);
  if (jump < best) best = jump;
  const uint64_t ii = ind[i];
  const BeamRange br = beam_range(ii, nsamps_beam);
  const uint32_t fi = filter[i], di = dm[i];
  const unsigned p1 = hi[i];
  for (unsigned p = lo[i]; p < p1; ++p) {
    const unsigned j = (unsigned)(keys[p] & 0xffffffffull);
    if (j >= i) continue;
    const uint32_t fd = s_fd[p];
    if (!coincident(ii, s_ind[p], fi, fd & 31u, di, fd >> 5, time_tol, filter_tol, dm_tol,
                    br))
      continue;
    const int32_t lj = __atomic_load_n(&labels[j], __ATOMIC_RELAXED);
    if (lj < best) best = lj;
  }
  if (best < cur) {
    __atomic_store_n(&labels[i], best, __ATOMIC_RELAXED);
    *flag = 1u;
  }
}

__global__ void k_gather_sorted(const uint64_t* __restrict__ keys, unsigned n,
                                const uint64_t* __restrict__ ind,
                                const uint32_t* __restrict__ filter,
                                const uint32_t* __restrict__ dm, uint64_t* __restrict__ s_ind,
                                uint32_t* __restrict__ s_fd) {
  const unsigned p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n) return;
  const unsigned k = (unsigned)(keys[p] & 0xffffffffull);
  s_ind[p] = ind[k];
  s_fd[p] = (dm[k] << 5) | (filter[k] & 31u);
}

// Filter-banded windows: the giants are sorted by (filter, coordinate), and
// giant i scans, for each filter band fb within filter_tol of its own, only
// the coordinates within time_tol * 2^max(fi, fb) -- the exact time window
// of a pair -- instead of the window of the widest filter.
constexpr int kMaxBands = 16;
__global__ void k_sort_init_banded(const uint64_t* __restrict__ ind,
                                   const uint32_t* __restrict__ filter,
                                   const uint32_t* __restrict__ group,
                                   const GroupDev* __restrict__ gd, unsigned n, uint64_t* keys,
                                   size_t npad) {
  const size_t k = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= npad) return;
  keys[k] = (k < n) ? (((uint64_t)filter[k] << 59) |
                       (tcoord(ind, group, gd, (unsigned)k) << 32) | (uint64_t)k)
                    : ~0ull;
}

__device__ __forceinline__ unsigned lower_bound_key(const uint64_t* keys, unsigned n,
                                                    uint64_t v) {
  unsigned lo = 0, hi = n;
  while (lo < hi) {
    const unsigned mid = (lo + hi) >> 1;
    if (keys[mid] < v) lo = mid + 1;
    else hi = mid;
  }
  return lo;
}

__global__ void k_windows_banded(const uint64_t* __restrict__ ind,
                                 const uint32_t* __restrict__ filter,
                                 const uint32_t* __restrict__ group,
                                 const GroupDev* __restrict__ gd, unsigned n,
                                 const uint64_t* __restrict__ keys, uint64_t time_tol,
                                 unsigned filter_tol, unsigned max_filter, uint32_t* lo,
                                 uint32_t* hi, int32_t* labels) {
  const unsigned i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const uint64_t t = tcoord(ind, group, gd, i);
  const unsigned fi = filter[i];
  const unsigned nb = 2 * filter_tol + 1;
  for (unsigned b = 0; b < nb; ++b) {
    const long long fb = (long long)fi - (long long)filter_tol + (long long)b;
    uint32_t l = 0, h = 0;
    if (fb >= 0 && fb <= (long long)max_filter) {
      const unsigned fm = fi > (unsigned)fb ? fi : (unsigned)fb;
      const uint64_t tol = time_tol << fm;
      const uint64_t base = (uint64_t)fb << 59;
      const uint64_t t_lo = t > tol ? t - tol : 0;
      l = lower_bound_key(keys, n, base | (t_lo << 32));
      h = lower_bound_key(keys, n, base | ((t + tol + 1) << 32));
    }
    lo[(size_t)i * nb + b] = l;
    hi[(size_t)i * nb + b] = h;
  }
  labels[i] = (int32_t)i;
}

__global__ void k_relax_banded(const uint64_t* __restrict__ ind,
                               const uint32_t* __restrict__ filter,
                               const uint32_t* __restrict__ dm, unsigned n, unsigned nb,
                               const uint64_t* __restrict__ keys,
                               const uint64_t* __restrict__ s_ind,
                               const uint32_t* __restrict__ s_fd,
                               const uint32_t* __restrict__ lo, const uint32_t* __restrict__ hi,
                               uint64_t time_tol, uint64_t filter_tol, uint64_t dm_tol,
                               uint64_t nsamps_beam_one, const uint32_t* __restrict__ group,
                               const GroupDev* __restrict__ gd, int32_t* labels,
                               unsigned* flag) {
  const unsigned i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const uint64_t nsamps_beam = group ? gd[group[i]].nsamps_beam : nsamps_beam_one;
  const int32_t cur = __atomic_load_n(&labels[i], __ATOMIC_RELAXED);
  int32_t best = cur;
  const int32_t jump = __atomic_load_n(&labels[cur], __ATOMIC_RELAXED);
  if (jump < best) best = jump;
  const uint64_t ii = ind[i];
  const BeamRange br = beam_range(ii, nsamps_beam);
  const uint32_t fi = filter[i], di = dm[i];
  for (unsigned b = 0; b < nb; ++b) {
    const unsigned p1 = hi[(size_t)i * nb + b];
    for (unsigned p = lo[(size_t)i * nb + b]; p < p1; ++p) {
      const unsigned j = (unsigned)(keys[p] & 0xffffffffull);
      if (j >= i) continue;
      const uint32_t fd = s_fd[p];
      if (!coincident(ii, s_ind[p], fi, fd & 31u, di, fd >> 5, time_tol, filter_tol, dm_tol,
                      br))
        continue;
      const int32_t lj = __atomic_load_n(&labels[j], __ATOMIC_RELAXED);
      if (lj < best) best = lj;
    }
  }
  if (best < cur) {
    __atomic_store_n(&labels[i], best, __ATOMIC_RELAXED);
    *flag = 1u;
  }
}

// The banded relaxation with one wave per giant: the lanes share the window
// entries (contiguous reads) and the minimum is reduced across the wave.
__global__ __launch_bounds__(256) void k_relax_banded_wave(
    const uint64_t* __restrict__ ind, const uint32_t* __restrict__ filter,
    const uint32_t* __restrict__ dm, unsigned n, unsigned nb, const uint64_t* __restrict__ keys,
    const uint64_t* __restrict__ s_ind, const uint32_t* __restrict__ s_fd,
    const uint32_t* __restrict__ lo, const uint32_t* __restrict__ hi, uint64_t time_tol,
    uint64_t filter_tol, uint64_t dm_tol, uint64_t nsamps_beam_one,
    const uint32_t* __restrict__ group, const GroupDev* __restrict__ gd, int32_t* labels,
    unsigned* flag) {
  const unsigned i = blockIdx.x * 4 + (threadIdx.x >> 6);
  const unsigned lane = threadIdx.x & 63;
  if (i >= n) return;  // whole wave
  const uint64_t nsamps_beam = group ? gd[group[i]].nsamps_beam : nsamps_beam_one;
  const int32_t cur = __atomic_load_n(&labels[i], __ATOMIC_RELAXED);
  int32_t best = cur;
  if (lane == 0) {
    const int32_t jump = __atomic_load_n(&labels[cur], __ATOMIC_RELAXED);
    if (jump < best) best = jump;
  }
  const uint64_t ii = ind[i];
  const BeamRange br = beam_range(ii, nsamps_beam);
  const uint32_t fi = filter[i], di = dm[i];
  for (unsigned b = 0; b < nb; ++b) {
    const unsigned p1 = hi[(size_t)i * nb + b];
    for (unsigned p = lo[(size_t)i * nb + b] + lane; p < p1; p += 64) {
      const unsigned j = (unsigned)(keys[p] & 0xffffffffull);
      if (j >= i) continue;
      const uint32_t fd = s_fd[p];
      if (!coincident(ii, s_ind[p], fi, fd & 31u, di, fd >> 5, time_tol, filter_tol, dm_tol,
                      br))
        continue;
      const int32_t lj = __atomic_load_n(&labels[j], __ATOMIC_RELAXED);
      if (lj < best) best = lj;
    }
  }
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    const int32_t other = __shfl_xor(best, o);
    best = other < best ? other : best;
  }
  if (lane == 0 && best < cur) {
    __atomic_store_n(&labels[i], best, __ATOMIC_RELAXED);
    *flag = 1u;
  }
}

// Coincidence lists over the banded windows, one wave per giant i: MODE 0
// counts the earlier giants j < i coincident with i into ecount[i]; MODE 1
// writes them to adj[eoff[i] ..] in window order.
template <int MODE>
__global__ __launch_bounds__(256) void k_edges_banded_wave(
    const uint64_t* __restrict__ ind, const uint32_t* __restrict__ filter,
    const uint32_t* __restrict__ dm, unsigned n, unsigned nb, const uint64_t* __restrict__ keys,
    const uint64_t* __restrict__ s_ind, const uint32_t* __restrict__ s_fd,
    const uint32_t* __restrict__ lo, const uint32_t* __restrict__ hi, uint64_t time_tol,
    uint64_t filter_tol, uint64_t dm_tol, uint64_t nsamps_beam_one,
    const uint32_t* __restrict__ group, const GroupDev* __restrict__ gd,
    uint32_t* __restrict__ ecount, const uint32_t* __restrict__ eoff, uint32_t* __restrict__ adj) {
  const unsigned i = blockIdx.x * 4 + (threadIdx.x >> 6);
  const unsigned lane = threadIdx.x & 63;
  if (i >= n) return;  // whole wave
  const uint64_t nsamps_beam = group ? gd[group[i]].nsamps_beam : nsamps_beam_one;
  const uint64_t ii = ind[i];
  const BeamRange br = beam_range(ii, nsamps_beam);
  const uint32_t fi = filter[i], di = dm[i];
  const uint64_t below = lane ? (~0ull >> (64 - lane)) : 0ull;
  const unsigned base = MODE == 1 ? eoff[i] : 0u;
  unsigned count = 0;
  for (unsigned b = 0; b < nb; ++b) {
    const unsigned p0 = lo[(size_t)i * nb + b], p1 = hi[(size_t)i * nb + b];
    for (unsigned pb = p0; pb < p1; pb += 64) {  // the same bounds in every lane
      const unsigned p = pb + lane;
      bool hit = false;
      unsigned j = 0;
      if (p < p1) {
        j = (unsigned)(keys[p] & 0xffffffffull);
        if (j < i) {
          const uint32_t fd = s_fd[p];
          hit = coincident(ii, s_ind[p], fi, fd & 31u, di, fd >> 5, time_tol, filter_tol,
                           dm_tol, br);
        }
      }
      const uint64_t m = __ballot(hit);
      if (MODE == 1 && hit) adj[base + count + (unsigned)__popcll(m & below)] = j;
      count += (unsigned)__popcll(m);
    }
  }
  if (MODE == 0 && lane == 0) ecount[i] = count;
}

// Pointer jumping: from label cur, follows labels[] while it keeps
// decreasing (up to 4 steps).  Every label is an index reachable from its
// giant along a chain of coincident giants with decreasing indices, and so
// is every label on the way (chains compose), so taking the smallest of them
// never goes below the fixed point: the smallest index reachable along such
// decreasing chains (the single-thread result, which can leave a connected
// group split in two; this is not a union-find).  It only shortens the
// passes needed to reach it.
__device__ __forceinline__ int32_t follow_labels(const int32_t* labels, int32_t cur) {
  int32_t r = cur;
#pragma unroll
  for (int t = 0; t < 4; ++t) {
    const int32_t nx = __atomic_load_n(&labels[r], __ATOMIC_RELAXED);
    if (nx >= r) break;
    r = nx;
  }
  return r;
}

// The relaxation over the coincidence lists, 8 lanes per giant.
constexpr int kListLanes = 8;
__global__ __launch_bounds__(256) void k_relax_lists(unsigned n,
                                                     const uint32_t* __restrict__ eoff,
                                                     const uint32_t* __restrict__ adj,
                                                     int32_t* labels, unsigned* flag) {
  const unsigned i = (blockIdx.x * 256 + threadIdx.x) / kListLanes;
  const unsigned sub = threadIdx.x % kListLanes;
  // giants past n still take part in the shuffles (with nothing to read)
  const bool live = i < n;
  const int32_t cur = live ? __atomic_load_n(&labels[i], __ATOMIC_RELAXED) : 0x7fffffff;
  int32_t best = cur;
  if (live && sub == 0) best = follow_labels(labels, cur);
  const unsigned p1 = live ? eoff[i + 1] : 0u;
  for (unsigned p = (live ? eoff[i] : 0u) + sub; p < p1; p += kListLanes) {
    const int32_t lj = __atomic_load_n(&labels[adj[p]], __ATOMIC_RELAXED);
    if (lj < best) best = lj;
  }
#pragma unroll
  for (int o = kListLanes / 2; o > 0; o >>= 1) {
    const int32_t other = __shfl_xor(best, o);
    best = other < best ? other : best;
  }
  if (live && sub == 0 && best < cur) {
    __atomic_store_n(&labels[i], best, __ATOMIC_RELAXED);
    *flag = 1u;
  }
}

__global__ void k_clear_flag(unsigned* flag) { *flag = 0u; }

// Clustering groups (the gulps of a call) lie further apart than any
// coincidence window, so every list entry of a giant is a giant of its own
// group; their giants are contiguous in index order.  Thread i (0 .. n)
// books the groups that start at giant i: gb[2g] = first giant of group g,
// gb[2g + 1] = first list entry of that giant (gb[2G], gb[2G + 1] close the
// last group).  gb[2G + 2] (cleared before) is set when the group column
// decreases or holds a group >= G: the ranges are then not used.
__global__ void k_group_bounds(const uint32_t* __restrict__ group, unsigned n, unsigned G,
                               const uint32_t* __restrict__ eoff, uint32_t* gb) {
  const unsigned i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i > n) return;
  const long long gp = i == 0 ? -1 : (long long)(group ? group[i - 1] : 0u);
  const long long gc = i == n ? (long long)G : (long long)(group ? group[i] : 0u);
  if (gc < gp || (i < n && gc >= (long long)G)) gb[2 * (size_t)G + 2] = 1u;
  const uint32_t e = eoff[i];
  for (long long g = gp + 1; g <= gc && g <= (long long)G; ++g) {
    gb[2 * g] = i;
    gb[2 * g + 1] = e;
  }
}

// The list relaxation of one clustering group with its labels in LDS: one
// 1024-thread block per group (giants gb[2b] .. gb[2b + 2] - 1, at most
// kLdsGroup; labels local to the group).  The giants are taken in index
// order, 128 per round with 8 lanes each.  A giant's list holds earlier
// giants only, so once the rounds before it are final a round depends on
// itself alone: it is updated again until an update changes none of its
// labels (once, when no giant of the round lists another one of it), and
// then it is final too.  The result is the fixed point (each giant
// labelled with the smallest index reachable along decreasing coincident
// chains), which does not depend on the order of the updates.  After k
// updates of a round every giant whose longest decreasing chain inside the
// round has at most k links holds its final label, so at most 128 + 1
// updates run (the last one changes nothing): the bound 128 + 2 is never
// what ends the loop.
constexpr unsigned kLdsGroup = 16384;  // giants per group (a 32-bit label each)
__global__ __launch_bounds__(1024) void k_relax_lists_lds(const uint32_t* __restrict__ gb,
                                                          const uint32_t* __restrict__ eoff,
                                                          const uint32_t* __restrict__ adj,
                                                          int32_t* __restrict__ labels) {
  extern __shared__ int32_t lab[];
  constexpr unsigned kPer = 1024 / kListLanes;  // giants per round
  constexpr int U = 16;                         // list entries in flight per lane
  const unsigned a = gb[2 * blockIdx.x], e = gb[2 * blockIdx.x + 2];
  if (a >= e) return;
  const unsigned n = e - a;
  for (unsigned k = threadIdx.x; k < n; k += 1024) lab[k] = (int32_t)k;
  __syncthreads();
  const unsigned sub = threadIdx.x % kListLanes;
  const unsigned npad = (n + kPer - 1) / kPer * kPer;
  for (unsigned r0 = 0; r0 < npad; r0 += kPer) {
    const unsigned i = r0 + threadIdx.x / kListLanes;
    const bool live = i < n;
    const unsigned p0 = live ? eoff[a + i] : 0u, p1 = live ? eoff[a + i + 1] : 0u;
    for (unsigned it = 0; it < kPer + 2; ++it) {
      const int32_t cur = live ? __atomic_load_n(&lab[i], __ATOMIC_RELAXED) : 0x7fffffff;
      int32_t best = cur;
      if (live && sub == 0) {  // follow_labels over the group's labels
        int32_t r = cur;
#pragma unroll
        for (int t = 0; t < 4; ++t) {
          const int32_t nx = __atomic_load_n(&lab[r], __ATOMIC_RELAXED);
          if (nx >= r) break;
          r = nx;
        }
        best = r;
      }
      bool intra = false;  // a neighbour in this round
      for (unsigned p = p0 + sub; p < p1; p += kListLanes * U) {
        uint32_t jj[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
          const unsigned q = p + (unsigned)u * kListLanes;
          jj[u] = adj[q < p1 ? q : p] - a;
        }
#pragma unroll
        for (int u = 0; u < U; ++u) {
          const int32_t lj = __atomic_load_n(&lab[jj[u]], __ATOMIC_RELAXED);
          if (lj < best) best = lj;
          intra = intra || jj[u] >= r0;
        }
      }
#pragma unroll
      for (int o = kListLanes / 2; o > 0; o >>= 1) {
        const int32_t other = __shfl_xor(best, o);
        best = other < best ? other : best;
      }
      bool changed = false;
      if (live && sub == 0 && best < cur) {
        __atomic_store_n(&lab[i], best, __ATOMIC_RELAXED);
        changed = true;
      }
      // again only when a label of the round changed and a giant of the
      // round reads another one's (the later iterations repeat only the
      // first test)
      if (!__syncthreads_or(changed)) break;
      if (it == 0 && !__syncthreads_or(intra)) break;
    }
  }
  for (unsigned k = threadIdx.x; k < n; k += 1024) labels[a + k] = (int32_t)a + lab[k];
}

__global__ void k_merge_init(unsigned n, const int32_t* __restrict__ labels, uint64_t* best,
                             unsigned* members, unsigned* root_flag) {
  const unsigned i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  best[i] = 0;
  members[i] = 0;
  root_flag[i] = (labels[i] == (int32_t)i) ? 1u : 0u;
}

__global__ void k_merge_accum(unsigned n, const int32_t* __restrict__ labels,
                              const float* __restrict__ peak,
                              const uint64_t* __restrict__ members_in, uint64_t* best,
                              unsigned* members) {
  const unsigned i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const unsigned r = (unsigned)labels[i];
  const uint64_t key = ((uint64_t)float_key(peak[i]) << 32) | (uint64_t)(0xffffffffu - i);
  atomicMax((unsigned long long*)&best[r], (unsigned long long)key);
  atomicAdd(&members[r], members_in ? (unsigned)members_in[i] : 1u);
}

__global__ void k_merge_write(unsigned n, const unsigned* __restrict__ root_flag,
                              const unsigned* __restrict__ root_pos,
                              const uint64_t* __restrict__ best,
                              const unsigned* __restrict__ members, const float* peak,
                              const uint64_t* ind, const uint64_t* begin, const uint64_t* end,
                              const uint32_t* filter, const uint32_t* dm, const uint32_t* group,
                              float* g_peak, uint64_t* g_ind, uint64_t* g_begin, uint64_t* g_end,
                              uint64_t* g_filter, uint64_t* g_dm, uint64_t* g_members,
                              uint32_t* g_group) {
  const unsigned r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= n || !root_flag[r]) return;
  const unsigned g = root_pos[r];
  const unsigned src = 0xffffffffu - (unsigned)(best[r] & 0xffffffffull);
  g_peak[g] = peak[src];
  g_ind[g] = ind[src];
  g_begin[g] = begin[src];
  g_end[g] = end[src];
  g_filter[g] = filter[src];
  g_dm[g] = dm[src];
  g_members[g] = members[r];
  if (g_group) g_group[g] = group[src];
}

}  // namespace

hipError_t launch_cluster(const uint64_t* ind, const uint32_t* filter, const uint32_t* dm,
                          const uint64_t* members_in, unsigned n, uint64_t time_tol,
                          uint64_t filter_tol, uint64_t dm_tol, uint64_t nsamps_beam,
                          unsigned max_filter, ClusterBuffers& cb, hipStream_t s,
                          unsigned* h_flag_pinned, int* iterations, const uint32_t* group,
                          const GroupDev* gd, uint64_t max_coord) {
  (void)members_in;
  *iterations = 0;
  if (n == 0) return hipSuccess;
  const unsigned T = 256;
  hipError_t e = hipSuccess;
  size_t npad = 1;
  while (npad < n) npad <<= 1;
  // banded windows need the sorted copies, a packed filter and coordinates
  // below 2^27 (the key holds filter:5 | coordinate:27 | index:32), and
  // per-giant window arrays of nb entries (lo/hi sized by the caller)
  const unsigned nb = (unsigned)(2 * filter_tol + 1);
  const bool banded = cb.s_ind && cb.s_fd && cb.band_cap >= nb && max_filter < 32 &&
                      max_coord < (1ull << 27);
  if (banded) {
    hipLaunchKernelGGL(k_sort_init_banded, dim3((unsigned)((npad + T - 1) / T)), dim3(T), 0, s,
                       ind, filter, group, gd, n, cb.sort_keys, npad);
  } else {
    hipLaunchKernelGGL(k_sort_init, dim3((unsigned)((npad + T - 1) / T)), dim3(T), 0, s, ind,
                       group, gd, n, cb.sort_keys, npad);
  }
  // the bitonic network: tiles of kBitonicBlock keys sorted in LDS, then per
  // later stage the steps across tiles one launch each and the rest in LDS
  if (npad <= (size_t(1) << 31)) {
    const unsigned ntile = (unsigned)((npad + kBitonicBlock - 1) / kBitonicBlock);
    hipLaunchKernelGGL(k_bitonic_tile, dim3(ntile), dim3(1024), 0, s, cb.sort_keys,
                       (unsigned)npad, 0u);
    for (size_t kk = 2 * kBitonicBlock; kk <= npad; kk <<= 1) {
      for (size_t j = kk >> 1; j >= kBitonicBlock; j >>= 1)
        hipLaunchKernelGGL(k_bitonic_step, dim3((unsigned)((npad + T - 1) / T)), dim3(T), 0, s,
                           cb.sort_keys, npad, j, kk);
      hipLaunchKernelGGL(k_bitonic_tile, dim3(ntile), dim3(1024), 0, s, cb.sort_keys,
                         (unsigned)npad, (unsigned)kk);
    }
  } else {
    for (size_t kk = 2; kk <= npad; kk <<= 1)
      for (size_t j = kk >> 1; j > 0; j >>= 1)
        hipLaunchKernelGGL(k_bitonic_step, dim3((unsigned)((npad + T - 1) / T)), dim3(T), 0, s,
                           cb.sort_keys, npad, j, kk);
  }
  if (banded) {
    hipLaunchKernelGGL(k_windows_banded, dim3((n + T - 1) / T), dim3(T), 0, s, ind, filter,
                       group, gd, n, cb.sort_keys, time_tol, (unsigned)filter_tol, max_filter,
                       cb.lo, cb.hi, cb.labels);
  } else {
    const uint64_t span = time_tol << max_filter;
    hipLaunchKernelGGL(k_windows, dim3((n + T - 1) / T), dim3(T), 0, s, ind, group, gd, n,
                       cb.sort_keys, span, cb.lo, cb.hi, cb.labels);
  }
  e = hipGetLastError();
  if (e != hipSuccess) return e;
  // the sorted copies need filter < 32 and dm < 2^27 to pack
  const bool sorted = cb.s_ind && cb.s_fd && max_filter < 32;
  if (sorted)
    hipLaunchKernelGGL(k_gather_sorted, dim3((n + T - 1) / T), dim3(T), 0, s, cb.sort_keys, n,
                       ind, filter, dm, cb.s_ind, cb.s_fd);
  // coincidence lists: counted, offsets scanned, sized on the host, written
  // (at most kMaxEdges entries; beyond that the passes scan the windows)
  constexpr size_t kMaxEdges = size_t(1) << 26;
  uint32_t* adj = nullptr;
  // per clustering group (a search without a group column is one group):
  // its range of giants and of list entries, read back with the edge count
  const unsigned G = group ? cb.n_cgroups : 1u;
  std::vector<uint32_t> gbh;
  if (banded && cb.ecount && cb.eoff && cb.adj_alloc) {
    hipLaunchKernelGGL(k_edges_banded_wave<0>, dim3((n + 3) / 4), dim3(256), 0, s, ind, filter,
                       dm, n, nb, cb.sort_keys, cb.s_ind, cb.s_fd, cb.lo, cb.hi, time_tol,
                       filter_tol, dm_tol, nsamps_beam, group, gd, cb.ecount, cb.eoff,
                       (uint32_t*)nullptr);
    e = launch_fill_u32(cb.ecount + n, 1, 0u, s);
    if (e == hipSuccess) e = launch_exclusive_scan_u32(cb.ecount, cb.eoff, (size_t)n + 1,
                                                       cb.scan_tmp, s);
    if (e == hipSuccess && cb.gbounds && G >= 1) {
      e = launch_fill_u32(cb.gbounds + 2 * (size_t)G + 2, 1, 0u, s);
      hipLaunchKernelGGL(k_group_bounds, dim3((n + 1 + T - 1) / T), dim3(T), 0, s, group, n, G,
                         cb.eoff, cb.gbounds);
      gbh.resize(2 * (size_t)G + 3);
      if (e == hipSuccess)
        e = hipMemcpyAsync(gbh.data(), cb.gbounds, gbh.size() * sizeof(uint32_t),
                           hipMemcpyDeviceToHost, s);
    }
    if (e == hipSuccess)
      e = hipMemcpyAsync(h_flag_pinned, cb.eoff + n, sizeof(unsigned), hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (e != hipSuccess) return e;
    const size_t edges = *h_flag_pinned;
    if (edges <= kMaxEdges) adj = cb.adj_alloc(edges);
    if (!adj) (void)hipGetLastError();  // a failed allocation leaves the windows path
    if (adj)
      hipLaunchKernelGGL(k_edges_banded_wave<1>, dim3((n + 3) / 4), dim3(256), 0, s, ind,
                         filter, dm, n, nb, cb.sort_keys, cb.s_ind, cb.s_fd, cb.lo, cb.hi,
                         time_tol, filter_tol, dm_tol, nsamps_beam, group, gd, cb.ecount,
                         cb.eoff, adj);
  }
  // One block per clustering group with its labels in LDS, when the groups'
  // ranges are sound and every group has at most kLdsGroup giants.
  if (adj && !gbh.empty()) {
    bool fits = gbh[2 * (size_t)G + 2] == 0 && gbh[0] == 0 && gbh[2 * (size_t)G] == n;
    uint32_t most = 0;
    for (unsigned g = 0; fits && g < G; ++g) {
      if (gbh[2 * g + 2] < gbh[2 * g] || gbh[2 * g + 3] < gbh[2 * g + 1]) fits = false;
      else most = std::max(most, gbh[2 * g + 2] - gbh[2 * g]);
    }
    if (fits && most <= kLdsGroup) {
      const size_t lds = std::max<size_t>(most, 1) * sizeof(int32_t);
      hipLaunchKernelGGL(k_relax_lists_lds, dim3(G), dim3(1024), lds, s, cb.gbounds, cb.eoff,
                         adj, cb.labels);
      return hipGetLastError();
    }
  }
  // relaxation passes run in batches between host checks of the change flag
  // (list passes are cheap next to a check, window passes are not).  At most
  // n + 1 passes are needed (see k_relax_lists_lds), so a flag still set
  // after n + 2 passes means the labels are wrong: that is reported.
  const int kBatch = adj ? 2 : 4;
  const size_t max_passes = (size_t)n + 2;
  bool converged = false;
  for (size_t it = 0; it < max_passes; it += kBatch) {
    hipLaunchKernelGGL(k_clear_flag, dim3(1), dim3(1), 0, s, cb.flag);
    for (int b = 0; b < kBatch; ++b) {
      if (adj)
        hipLaunchKernelGGL(k_relax_lists, dim3((n + 256 / kListLanes - 1) / (256 / kListLanes)),
                           dim3(256), 0, s, n, cb.eoff, adj, cb.labels, cb.flag);
      else if (banded)
        hipLaunchKernelGGL(k_relax_banded_wave, dim3((n + 3) / 4), dim3(256), 0, s, ind, filter,
                           dm, n, nb, cb.sort_keys, cb.s_ind, cb.s_fd, cb.lo, cb.hi, time_tol,
                           filter_tol, dm_tol, nsamps_beam, group, gd, cb.labels, cb.flag);
      else if (sorted)
        hipLaunchKernelGGL(k_relax_sorted, dim3((n + T - 1) / T), dim3(T), 0, s, ind, filter, dm,
                           n, cb.sort_keys, cb.s_ind, cb.s_fd, cb.lo, cb.hi, time_tol,
                           filter_tol, dm_tol, nsamps_beam, group, gd, cb.labels, cb.flag);
      else
        hipLaunchKernelGGL(k_relax, dim3((n + T - 1) / T), dim3(T), 0, s, ind, filter, dm, n,
                           cb.sort_keys, cb.lo, cb.hi, time_tol, filter_tol, dm_tol,
                           nsamps_beam, group, gd, cb.labels, cb.flag);
    }
    e = hipMemcpyAsync(h_flag_pinned, cb.flag, sizeof(unsigned), hipMemcpyDeviceToHost, s);
    if (e != hipSuccess) return e;
    e = hipStreamSynchronize(s);
    if (e != hipSuccess) return e;
    *iterations = (int)std::min<size_t>(it + kBatch, 0x7fffffff);
    if (*h_flag_pinned == 0) {
      converged = true;
      break;
    }
  }
  if (!converged) return hipErrorUnknown;  // cannot happen: see the pass bound above
  return hipGetLastError();
}

hipError_t launch_merge(const float* peak, const uint64_t* ind, const uint64_t* begin,
                        const uint64_t* end, const uint32_t* filter, const uint32_t* dm,
                        const uint64_t* members_in, unsigned n, ClusterBuffers& cb,
                        float* g_peak, uint64_t* g_ind, uint64_t* g_begin, uint64_t* g_end,
                        uint64_t* g_filter, uint64_t* g_dm, uint64_t* g_members,
                        hipStream_t s, const uint32_t* group, uint32_t* g_group) {
  if (n == 0) return hipSuccess;
  const unsigned T = 256;
  const dim3 grid((n + T - 1) / T);
  hipLaunchKernelGGL(k_merge_init, grid, dim3(T), 0, s, n, cb.labels, cb.best, cb.members,
                     cb.root_flag);
  hipLaunchKernelGGL(k_merge_accum, grid, dim3(T), 0, s, n, cb.labels, peak, members_in, cb.best,
                     cb.members);
  hipError_t e = launch_exclusive_scan_u32(cb.root_flag, cb.root_pos, n, cb.scan_tmp, s);
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(k_merge_write, grid, dim3(T), 0, s, n, cb.root_flag, cb.root_pos, cb.best,
                     cb.members, peak, ind, begin, end, filter, dm, group, g_peak, g_ind,
                     g_begin, g_end, g_filter, g_dm, g_members, g_group);
  return hipGetLastError();
}

}  // namespace hd
