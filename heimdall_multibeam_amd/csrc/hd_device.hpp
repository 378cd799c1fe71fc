// hd_device.hpp -- device helpers shared by the search kernels (gfx950).
//
// All arithmetic here mirrors the reference's float/int conversions one for
// one; the library is compiled with -ffp-contract=off so that no expression
// is fused into an FMA (the reference's host build has none).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace hd {

// ---------------------------------------------------------------------------
// Medians.  The reference's comparison networks (Pipeline/median_filter.cu:
// 21-75) return the order statistic (median4: mean of the middle two).
// Written here as min/max networks: the extreme elements of a pair of pairs
// can never be the median, which leaves a median of three, one v_med3_f32.
// The values searched are finite with no negative zero (integer sums, and
// absolute deviations), where every median network gives the same bits.
__device__ __forceinline__ float med3(float a, float b, float c) {
  return __builtin_amdgcn_fmed3f(a, b, c);
}
__device__ __forceinline__ float med4(float a, float b, float c, float d) {
  const float lo = fmaxf(fminf(a, b), fminf(c, d));
  const float hi = fminf(fmaxf(a, b), fmaxf(c, d));
  return 0.5f * (lo + hi);
}
__device__ __forceinline__ float med5(float a, float b, float c, float d, float e) {
  const float lo = fmaxf(fminf(a, b), fminf(c, d));
  const float hi = fminf(fmaxf(a, b), fmaxf(c, d));
  return med3(e, lo, hi);
}
// Tail of a median-of-5 pyramid once fewer than five values remain
// (Pipeline/median_filter.cu:217-235,295-311).
__device__ __forceinline__ float med_tail(const float* v, int n) {
  if (n == 1) return v[0];
  if (n == 2) return 0.5f * (v[0] + v[1]);
  if (n == 3) return med3(v[0], v[1], v[2]);
  return med4(v[0], v[1], v[2], v[3]);
}

// (double)x > 1e-5 for a float x is the same test as x > 1e-5f: 1e-5f is the
// float just below 1e-5 and no float lies between the two.
__device__ __forceinline__ bool frac_gt(float x) { return x > 1e-5f; }

// linear_stretch_functor (Pipeline/median_filter.cu:451-464) for output i;
// in_end is the element one past the input, read when rounding lands the
// last output at the end of the input.
// Both forms below are written without branches: the second element is read
// at j + 1 only when the fraction counts (as in the reference), else element j
// is read twice, and the result is selected.
struct Stretch {
  float step;  // float(in_count-1)/(out_count-1)
  __device__ __forceinline__ float operator()(const float* in, unsigned i) const {
    const float x = (float)i * step;
    const unsigned j = (unsigned)x;
    const float fr = x - (float)j;
    const bool lerp = frac_gt(fr);
    const float a = in[j];
    const float b = in[lerp ? j + 1 : j];
    const float p = fr * (b - a);
    return lerp ? a + p : a + 0.f;
  }
};

// linear_stretch_beam_functor (Pipeline/median_filter.cu:481-499): the index
// is carried through float across beams.
struct StretchBeam {
  float step;              // float(in_beam-1)/out_beam
  unsigned out_count;      // outputs per beam
  __device__ __forceinline__ float operator()(const float* in, unsigned i) const {
    const unsigned beam = i / out_count;
    const float x = (float)i * step;
    const float xb = x + (float)beam;
    const unsigned k = (unsigned)xb;
    const float o = (x - (float)k) + (float)beam;
    const bool lerp = frac_gt(o);
    const float a = in[k];
    const float b = in[lerp ? k + 1 : k];
    const float p = o * (b - a);
    return lerp ? a + p : a + 0.f;
  }
};

// reblock_abs_beam_functor index (Pipeline/get_rms.cu:214-229):
// (size_t)((float)beam*delta + (float)o), beam = o/length.
// (o < 2^32: the quotient is 0 below the first period, and a 32-bit
// division gives it when the length fits 32 bits; a 64-bit division is a
// long instruction sequence.)
__device__ __forceinline__ size_t reblock_pos(unsigned o, size_t length, float delta) {
  const size_t beam =
      (size_t)o < length ? 0 : (length > 0xffffffffull ? 0 : (size_t)(o / (unsigned)length));
  const float p = (float)beam * delta;
  const float q = p + (float)o;
  return (size_t)q;
}

// bfloat16 bits of a float, rounded to nearest with ties to even (finite
// inputs), and back.
__device__ __forceinline__ uint16_t f2bf(float f) {
  uint32_t u = __float_as_uint(f);
  u += 0x7fffu + ((u >> 16) & 1u);
  return (uint16_t)(u >> 16);
}
__device__ __forceinline__ float bf2f(uint16_t h) { return __uint_as_float((uint32_t)h << 16); }

// Orderable key of a float: unsigned compare of keys matches float compare
// (for non-NaN values).
__device__ __forceinline__ uint32_t float_key(float f) {
  const uint32_t u = __float_as_uint(f);
  return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}
__device__ __forceinline__ float key_float(uint32_t k) {
  const uint32_t u = (k & 0x80000000u) ? (k & 0x7fffffffu) : ~k;
  return __uint_as_float(u);
}

// ---------------------------------------------------------------------------
// XCD-contiguous block order.  The MI355X hands workgroups to its 8 XCDs
// round robin by linear id, so blocks next to each other in (x, y) order
// land on different XCDs and each L2 fetches the halos that neighbouring
// blocks share.  This returns the (x, y) block to work on such that the XCD
// running linear id l (l mod 8) walks one contiguous range of the x-fastest
// order instead: consecutive chunks of consecutive rows.  A bijection on
// the gridDim.x * gridDim.y blocks of each z plane.
__device__ __forceinline__ uint2 xcd_contiguous_block() {
  const unsigned nbx = gridDim.x, nblk = gridDim.x * gridDim.y;
  const unsigned lin = blockIdx.y * nbx + blockIdx.x;
  const unsigned q = nblk / 8, r = nblk % 8, x8 = lin % 8;
  const unsigned l = x8 * q + (x8 < r ? x8 : r) + lin / 8;
  return make_uint2(l % nbx, l / nbx);
}

// ---------------------------------------------------------------------------
// Block-wide helpers.  The block size is a compile-time constant BS and
// every thread calls these.
//
// Every helper that takes a pointer into LDS is __forceinline__.  A helper the
// compiler leaves as a called function receives its LDS buffers as generic
// pointers and reaches them with flat loads and stores, and in such an
// outlined copy of lds_pyramid the barrier that opens each pyramid level was
// emitted with no wait for the previous level's last flat store: waves could
// read a level before it was complete, and the baseline points of a few
// trials changed from run to run.  Inlined, the accesses are ds_read/ds_write
// with their waits before each barrier.  tests/test_build.py checks that no
// kernel of the library calls a device function.

// Orders one wave's LDS writes before its later reads (the wave's LDS
// operations execute in order; this keeps the compiler from moving them).
__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// `levels` median-of-5 levels over src[0 .. n) inside one wave, ping-ponging
// between src and dst; returns the buffer holding the last level.
__device__ __forceinline__ float* wave_pyramid(float* src, float* dst, size_t n, int levels) {
  const unsigned lane = threadIdx.x & 63;
  for (int l = 0; l < levels; ++l) {
    wave_sync();
    const size_t m = n / 5;
    for (size_t q = lane; q < m; q += 64) {
      const float* p = src + 5 * q;
      dst[q] = med5(p[0], p[1], p[2], p[3], p[4]);
    }
    float* t = src;
    src = dst;
    dst = t;
    n = m;
  }
  wave_sync();
  return src;
}

// Reduces `n` values in `src` by median-of-5 groups `levels` times, ping-
// ponging with `dst`; returns a pointer to the buffer holding the result
// (n / 5^levels values).  Requires n >= 5^levels.
template <int BS>
__device__ __forceinline__ float* lds_pyramid(float* src, float* dst, int n, int levels) {
  for (int l = 0; l < levels; ++l) {
    const int m = n / 5;
    __syncthreads();
    for (int j = threadIdx.x; j < m; j += BS) {
      const float* p = src + 5 * j;
      dst[j] = med5(p[0], p[1], p[2], p[3], p[4]);
    }
    float* t = src;
    src = dst;
    dst = t;
    n = m;
  }
  __syncthreads();
  return src;
}

// Finishes a median-of-5 pyramid held in A (n values) with the reference's
// loop `for (size = n; size > 1; size /= 5)` (get_rms.cu:83-87, 119-123):
// groups of five while size > 4, then the small-count rule; no value (n = 0)
// gives 0, the zeroed buffer's first entry.  Returns the final value to
// every thread of the block.
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

// Median of 5^L consecutive leaves starting at `first` (L in 0..3), leaves
// produced by the functor `leaf(k)`.
template <typename Leaf>
__device__ __forceinline__ float tree_value(const Leaf& leaf, size_t first, int L) {
  if (L == 0) return leaf(first);
  if (L == 1) return med5(leaf(first), leaf(first + 1), leaf(first + 2), leaf(first + 3),
                          leaf(first + 4));
  if (L == 2) {
    float m[5];
#pragma unroll
    for (int g = 0; g < 5; ++g) {
      const size_t b = first + 5 * g;
      m[g] = med5(leaf(b), leaf(b + 1), leaf(b + 2), leaf(b + 3), leaf(b + 4));
    }
    return med5(m[0], m[1], m[2], m[3], m[4]);
  }
  float n2[5];
  for (int h = 0; h < 5; ++h) {
    float m[5];
#pragma unroll
    for (int g = 0; g < 5; ++g) {
      const size_t b = first + 25 * h + 5 * g;
      m[g] = med5(leaf(b), leaf(b + 1), leaf(b + 2), leaf(b + 3), leaf(b + 4));
    }
    n2[h] = med5(m[0], m[1], m[2], m[3], m[4]);
  }
  return med5(n2[0], n2[1], n2[2], n2[3], n2[4]);
}

}  // namespace hd
