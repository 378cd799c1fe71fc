// k_stages.hip -- the reference's stage functions as extern "C" calls over
// raw device memory (declared in include/hd/stages.h).
//
// hd_execute fuses these stages; these entry points run one stage at a time
// on caller-provided buffers, with the reference's argument meaning, so that
// callers of the reference's stage classes (and the parity tests) can use
// them directly.  They reuse the pipeline's kernels: a series is a DM series,
// and a batch of `n_series` of them is one launch.
#include <algorithm>
#include <cmath>
#include <mutex>
#include <vector>

#include "hd/stages.h"
#include "hd_device.hpp"
#include "hd_internal.hpp"

namespace hd {
namespace {

constexpr int ST_BS = 256;

hd_error to_hd(hipError_t e) {
  if (e == hipSuccess) return HD_NO_ERROR;
  if (e == hipErrorOutOfMemory || e == hipErrorMemoryAllocation) return HD_MEM_ALLOC_FAILED;
  return HD_INTERNAL_GPU_ERROR;
}

#define ST_CHECK(x)                                   \
  do {                                                \
    hipError_t _e = (x);                              \
    if (_e != hipSuccess) return to_hd(_e);           \
  } while (0)

// Baseline geometry for (beam_stride, beam_count, smooth_radius, nbeams)
// (remove_baseline.cu:30-46).
bool baseline_shape(size_t stride, size_t bc, size_t radius, size_t nbeams, size_t series_stride,
                    SeriesShape& sh) {
  sh = SeriesShape{};
  sh.nbeams = (unsigned)nbeams;
  sh.stride = stride;
  sh.bc = bc;
  sh.nout = stride * nbeams - (stride - bc);
  sh.ppitch = series_stride;
  sh.spitch = series_stride;
  const float oversample = 2;
  sh.sc = radius ? (size_t)((double)(oversample * (float)stride / (float)(2 * radius)) + 0.5) : 0;
  if (sh.sc) {
    sh.nscr = (int)(size_t)(std::log((double)(stride / sh.sc)) / std::log(5.));
    sh.count_round = (size_t)(std::pow(5., (double)sh.nscr) * (double)sh.sc);
    if (sh.count_round < 2) return false;
    int Ld = 0;
    while (Ld < std::min(sh.nscr, 3) && sh.count_round / pow5_host(Ld) > 12288) ++Ld;
    if (sh.count_round / pow5_host(Ld) > 12288) return false;
    sh.bl_direct = Ld;
    sh.step_rs_beam = (float)(stride - 1) / (float)sh.count_round;
    sh.step_rs_last = (float)(bc - 1) / (float)(sh.count_round - 1);
    sh.step_bl_beam = (float)(sh.sc - 1) / (float)stride;
    sh.step_bl_last = (float)(sh.sc - 1) / (float)(bc - 1);
    for (size_t i = bc; i-- > 0;) {  // stale pyramid entry read: refused
      const float x = (float)(unsigned)i * sh.step_bl_last;
      const unsigned j = (unsigned)x;
      if ((size_t)j + 1 < sh.sc) break;
      if (x - (float)j > 1e-5f) return false;
    }
  }
  return true;
}

bool rms_shape(size_t stride, size_t bc, size_t nbeams, SeriesShape& sh) {
  sh.nbeams = (unsigned)nbeams;
  sh.stride = stride;
  sh.bc = bc;
  int Ld = 0;
  while (sh.bc / pow5_host(Ld) > 12288 && sh.bc >= pow5_host(Ld + 1) && Ld < 3) ++Ld;
  if (sh.bc / pow5_host(Ld) > 12288) return false;
  sh.rms_direct = Ld;
  sh.rms_delta = (float)(stride - bc);
  return true;
}

__global__ void k_fill(float* p, size_t n, float v) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

// |d| pooled pyramid chunks of a materialised series (get_rms.cu:104-123)
template <int BS>
__device__ __forceinline__ float* pyr(float* A, float* B, int n, int levels) {
  return lds_pyramid<BS>(A, B, n, levels);
}

__global__ __launch_bounds__(ST_BS) void k_block_rms_direct(const float* __restrict__ d,
                                                            size_t sstride, BlockRmsShape bs,
                                                            float* __restrict__ partial) {
  extern __shared__ float lds[];
  const size_t c = blockIdx.x;
  const size_t s = blockIdx.y;
  const float* row = d + s * sstride;
  const size_t k0 = c * bs.chunk;
  auto leaf = [&](size_t k) -> float {
    return fabsf(row[reblock_pos((unsigned)k, bs.bc, bs.delta)]);
  };
  float value;
  const size_t n1 = bs.chunk / 5;
  if (bs.L0 == 0) {
    value = leaf(k0);
  } else {
    float* A = lds;
    float* B = lds + n1;
    for (size_t q = threadIdx.x; q < n1; q += ST_BS) {
      const size_t b = k0 + 5 * q;
      A[q] = med5(leaf(b), leaf(b + 1), leaf(b + 2), leaf(b + 3), leaf(b + 4));
    }
    value = pyr<ST_BS>(A, B, (int)n1, bs.L0 - 1)[0];
  }
  if (threadIdx.x == 0) partial[s * bs.np + c] = value;
}

__global__ void k_scale_block(float* d, size_t sstride, size_t count, const float* rms) {
  const size_t s = blockIdx.y;
  const float inv = 1.0f / rms[s];  // get_rms.cu:195
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < count;
       i += (size_t)gridDim.x * blockDim.x)
    d[s * sstride + i] = d[s * sstride + i] * inv;
}

__global__ void k_boxcar(const float* __restrict__ S, size_t sstride, size_t n_out, unsigned off,
                         unsigned ahead, unsigned behind, unsigned ts, float* __restrict__ out,
                         size_t ostride) {
  const size_t s = blockIdx.y;
  for (size_t k = (size_t)blockIdx.x * blockDim.x + threadIdx.x; k < n_out;
       k += (size_t)gridDim.x * blockDim.x) {
    const size_t b = k * ts;
    const float a = S[s * sstride + off + ahead + b];
    const float c = S[s * sstride + off - behind + b];
    out[s * ostride + k] = a - c;
  }
}

// GiantFinder's selection on a materialised series (find_giants.cu:47-60)
__global__ __launch_bounds__(ST_BS) void k_threshold_direct(const float* __restrict__ d,
                                                            size_t count, float thresh,
                                                            size_t stride, size_t bcount,
                                                            GiantBuffers gb) {
  constexpr int PER = 16;
  const size_t c = blockIdx.x;
  const size_t ib = c * (ST_BS * PER) + (size_t)threadIdx.x * PER;
  float vals[PER];
  unsigned mask = 0;
#pragma unroll
  for (int q = 0; q < PER; ++q) {
    const size_t i = ib + q;
    vals[q] = 0.f;
    if (i < count) {
      const float v = d[i];
      vals[q] = v;
      if ((size_t)(unsigned)i % stride < bcount && v > thresh) mask |= 1u << q;
    }
  }
  __shared__ unsigned s[ST_BS];
  const unsigned cnt = __popc(mask);
  s[threadIdx.x] = cnt;
  __syncthreads();
  for (int o = 1; o < ST_BS; o <<= 1) {
    const unsigned t = threadIdx.x >= (unsigned)o ? s[threadIdx.x - o] : 0u;
    __syncthreads();
    s[threadIdx.x] += t;
    __syncthreads();
  }
  const unsigned total = s[ST_BS - 1];
  const unsigned rank = s[threadIdx.x] - cnt;
  __shared__ unsigned sbase;
  if (threadIdx.x == 0) {
    const unsigned base = total ? atomicAdd(gb.counter, total) : 0u;
    sbase = base;
    gb.ccount[c] = total;
    gb.coff[c] = base;
  }
  __syncthreads();
  if (!mask || (size_t)sbase + total > gb.scap) return;
  unsigned r = rank;
#pragma unroll
  for (int q = 0; q < PER; ++q)
    if (mask & (1u << q)) {
      gb.skey[sbase + r] = (uint64_t)(ib + q);
      gb.sval[sbase + r] = vals[q];
      ++r;
    }
}

__global__ void k_u64_to_u32(const uint64_t* in, uint32_t* out, size_t n) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[i] = (uint32_t)in[i];
}
__global__ void k_u64_to_i32(const uint64_t* in, int32_t* out, size_t n) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[i] = (int32_t)in[i];
}
__global__ void k_i32_to_u64(const int32_t* in, uint64_t* out, size_t n) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[i] = (uint64_t)in[i];
}
__global__ void k_roots(const int32_t* labels, unsigned n, unsigned* flags) {
  const unsigned i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) flags[i] = labels[i] == (int32_t)i ? 1u : 0u;
}

struct Scratch {  // stream-ordered temporaries
  std::vector<void*> ptrs;
  hipStream_t s;
  explicit Scratch(hipStream_t st) : s(st) {}
  template <typename T>
  hipError_t get(T** p, size_t n) {
    void* q = nullptr;
    hipError_t e = hipMallocAsync(&q, std::max<size_t>(n, 1) * sizeof(T), s);
    if (e == hipSuccess) ptrs.push_back(q);
    *p = static_cast<T*>(q);
    return e;
  }
  ~Scratch() {
    for (void* q : ptrs) (void)hipFreeAsync(q, s);
  }
};

dim3 grid1(size_t n, unsigned bs = 256) { return dim3((unsigned)((n + bs - 1) / bs)); }

}  // namespace
}  // namespace hd

using namespace hd;

extern "C" hd_error hd_remove_baseline(hd_float* d_data, hd_size beam_stride, hd_size beam_count,
                                       hd_size smooth_radius, hd_size nbeams, hd_size n_series,
                                       hd_size series_stride, void* stream) {
  if (!d_data) return HD_INVALID_POINTER;
  if (!nbeams || beam_count > beam_stride || !beam_count) return HD_INVALID_STRIDE;
  hipStream_t s = (hipStream_t)stream;
  SeriesShape sh;
  if (!baseline_shape(beam_stride, beam_count, smooth_radius, nbeams, series_stride, sh))
    return HD_INVALID_STRIDE;
  // the last beam's stretch may read d[nout] (remove_baseline.cu:74-75 with
  // median_filter.cu:460-462): it must lie inside the series' stride
  if (series_stride < sh.nout + 1) return HD_INVALID_STRIDE;
  if (sh.sc == 0 || n_series == 0) return HD_NO_ERROR;  // remove_baseline.cu:39-42
  Scratch tmp(s);
  float *pts, *ones;
  ST_CHECK(tmp.get(&pts, n_series * nbeams * sh.sc + 64));
  ST_CHECK(tmp.get(&ones, n_series * nbeams));
  hipLaunchKernelGGL(k_fill, grid1(n_series * nbeams), dim3(256), 0, s, ones, n_series * nbeams,
                     1.0f);
  ST_CHECK(launch_baseline_points(d_data, sh, (int)n_series, pts, s));
  // subtract: (x - baseline) / 1 is exactly x - baseline
  ST_CHECK(launch_normalise(d_data, pts, ones, sh, (int)n_series, s));
  return HD_NO_ERROR;
}

extern "C" hd_error hd_get_rms_multibeam(const hd_float* d_data, hd_float* d_rms,
                                         hd_size beam_stride, hd_size beam_count, hd_size nbeams,
                                         hd_size n_series, hd_size series_stride, void* stream) {
  if (!d_data || !d_rms) return HD_INVALID_POINTER;
  if (!nbeams || !beam_count || beam_count > beam_stride) return HD_INVALID_STRIDE;
  SeriesShape sh{};
  sh.ppitch = series_stride;
  if (!rms_shape(beam_stride, beam_count, nbeams, sh)) return HD_INVALID_STRIDE;
  sh.sc = 0;  // no baseline: the leaves are |d|
  if (n_series == 0) return HD_NO_ERROR;
  ST_CHECK(launch_beam_rms(d_data, nullptr, sh, (int)n_series, d_rms, (hipStream_t)stream));
  return HD_NO_ERROR;
}

extern "C" hd_error hd_normalise_multibeam(hd_float* d_data, const hd_float* d_rms,
                                           hd_size beam_stride, hd_size cur_nsamps,
                                           hd_size nbeams, hd_size n_series,
                                           hd_size series_stride, void* stream) {
  if (!d_data || !d_rms) return HD_INVALID_POINTER;
  if (!nbeams || !beam_stride || cur_nsamps > series_stride) return HD_INVALID_STRIDE;
  SeriesShape sh{};
  sh.nbeams = (unsigned)nbeams;
  sh.stride = beam_stride;
  sh.nout = cur_nsamps;
  sh.ppitch = series_stride;
  sh.sc = 0;
  if (n_series == 0) return HD_NO_ERROR;
  ST_CHECK(launch_normalise(d_data, nullptr, d_rms, sh, (int)n_series, (hipStream_t)stream));
  return HD_NO_ERROR;
}

namespace hd {
namespace {

// Pooled block-RMS geometry (get_rms.cu:104-123): n = nbeams*beam_count leaves,
// chunks of 5^L0 leaves (L0 <= 5) reduced to np partial medians.
BlockRmsShape block_rms_shape(size_t beam_stride, size_t beam_count, size_t nbeams) {
  BlockRmsShape bs{};
  bs.bc = beam_count;
  bs.n = nbeams * beam_count;
  bs.delta = (float)(beam_stride - beam_count);
  while (bs.L0 < 5 && pow5_host(bs.L0 + 1) <= bs.n) ++bs.L0;
  bs.chunk = pow5_host(bs.L0);
  bs.np = bs.n / bs.chunk;
  return bs;
}

// The two launches of the block RMS: partial medians (n_series x np floats)
// and reciprocals (n_series floats) in caller-given device scratch.
hipError_t block_rms_launch(const float* d_data, float* d_block_rms, size_t series_stride,
                            size_t n_series, const BlockRmsShape& bs, float* partial,
                            float* winv, hipStream_t s) {
  const size_t n1 = bs.chunk / 5;
  const size_t lds = (n1 + n1 / 5 + 8) * sizeof(float);
  hipLaunchKernelGGL(k_block_rms_direct, dim3((unsigned)bs.np, (unsigned)n_series), dim3(ST_BS),
                     lds, s, d_data, series_stride, bs, partial);
  const hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  return launch_block_rms_finish(partial, 1, 0, 1, (int)n_series, d_block_rms, winv, s,
                                 GroupArgs{}, &bs);
}

}  // namespace
}  // namespace hd

extern "C" hd_error hd_block_rms(const hd_float* d_data, hd_float* d_block_rms,
                                 hd_size beam_stride, hd_size beam_count, hd_size nbeams,
                                 hd_size n_series, hd_size series_stride, void* stream) {
  if (!d_data || !d_block_rms) return HD_INVALID_POINTER;
  if (!nbeams || !beam_count || beam_count > beam_stride) return HD_INVALID_STRIDE;
  hipStream_t s = (hipStream_t)stream;
  const BlockRmsShape bs = block_rms_shape(beam_stride, beam_count, nbeams);
  if (n_series == 0) return HD_NO_ERROR;
  Scratch tmp(s);
  float *partial, *winv;
  ST_CHECK(tmp.get(&partial, n_series * bs.np));
  ST_CHECK(tmp.get(&winv, n_series));
  ST_CHECK(block_rms_launch(d_data, d_block_rms, series_stride, n_series, bs, partial, winv, s));
  return HD_NO_ERROR;
}

// ---- hd_block_rms_noalloc: the block RMS with its scratch in a fixed device
// array instead of stream-ordered allocations (for callers that cannot
// allocate: code under HIP graph capture, serving loops).  Each GPU holds its
// own copy of `scratch` until the process ends; its address on a device is
// taken once, with that device current, by hipGetSymbolAddress (the array's
// name in host code is not a device address).  Calls on one GPU share the
// array, so calls that overlap on one GPU race; nothing in it needs clearing
// (every word a call reads, it has written first).
static __device__ __align__(256) float scratch[HD_NOALLOC_SCRATCH_WORDS];

namespace hd {
namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
float* g_scratch_ptr[kMaxDevices] = {};

hipError_t scratch_on_current_device(float** out) {
  int dev = -1;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return e;
  if (dev < 0 || dev >= kMaxDevices) return hipErrorInvalidDevice;
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void* p = nullptr;
    e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return e;
    if (!p) return hipErrorInvalidSymbol;
    g_scratch_ptr[dev] = static_cast<float*>(p);
  }
  *out = g_scratch_ptr[dev];
  return hipSuccess;
}
}  // namespace
}  // namespace hd

extern "C" hd_error hd_block_rms_noalloc(const hd_float* d_data, hd_float* d_block_rms,
                                         hd_size beam_stride, hd_size beam_count,
                                         hd_size nbeams, hd_size n_series,
                                         hd_size series_stride, void* stream) {
  if (!d_data || !d_block_rms) return HD_INVALID_POINTER;
  if (!nbeams || !beam_count || beam_count > beam_stride) return HD_INVALID_STRIDE;
  const BlockRmsShape bs = block_rms_shape(beam_stride, beam_count, nbeams);
  if (n_series == 0) return HD_NO_ERROR;
  // partials first, the reciprocals 256-byte aligned after them
  const size_t partial_words = (n_series * bs.np + 63) / 64 * 64;
  if (n_series > HD_NOALLOC_SCRATCH_WORDS || bs.np > HD_NOALLOC_SCRATCH_WORDS ||
      partial_words + n_series > HD_NOALLOC_SCRATCH_WORDS)
    return HD_MEM_ALLOC_FAILED;  // beyond the fixed scratch: nothing is launched
  float* ptr = nullptr;
  ST_CHECK(scratch_on_current_device(&ptr));
  ST_CHECK(block_rms_launch(d_data, d_block_rms, series_stride, n_series, bs, ptr,
                            ptr + partial_words, (hipStream_t)stream));
  return HD_NO_ERROR;
}

extern "C" hd_error hd_normalise_block(hd_float* d_data, const hd_float* d_block_rms,
                                       hd_size count, hd_size n_series, hd_size series_stride,
                                       void* stream) {
  if (!d_data || !d_block_rms) return HD_INVALID_POINTER;
  if (count > series_stride && n_series > 1) return HD_INVALID_STRIDE;
  if (n_series == 0 || count == 0) return HD_NO_ERROR;
  const unsigned gx = (unsigned)std::min<size_t>((count + 255) / 256, 1024);
  hipLaunchKernelGGL(k_scale_block, dim3(gx, (unsigned)n_series), dim3(256), 0,
                     (hipStream_t)stream, d_data, series_stride, count, d_block_rms);
  ST_CHECK(hipGetLastError());
  return HD_NO_ERROR;
}

extern "C" hd_error hd_matched_filter_prep(const hd_float* d_in, hd_size count,
                                           hd_float* d_scanned, hd_size n_series,
                                           hd_size series_stride, hd_size scanned_stride,
                                           void* stream) {
  if (!d_in || !d_scanned) return HD_INVALID_POINTER;
  if (n_series > 1 && (count > series_stride || count + 1 > scanned_stride))
    return HD_INVALID_STRIDE;
  if (n_series == 0) return HD_NO_ERROR;
  ST_CHECK(launch_scan(d_in, series_stride, count, (int)n_series, d_scanned, scanned_stride,
                       (hipStream_t)stream));
  return HD_NO_ERROR;
}

extern "C" hd_error hd_matched_filter_exec(const hd_float* d_scanned, hd_size scanned_count,
                                           hd_size max_width, hd_size width, hd_size tscrunch,
                                           hd_float* d_out, hd_size n_series,
                                           hd_size scanned_stride, hd_size out_stride,
                                           void* stream) {
  if (!d_scanned || !d_out) return HD_INVALID_POINTER;
  if (!width || !tscrunch || width > max_width || scanned_count <= max_width)
    return HD_INVALID_STRIDE;
  const size_t out_count = scanned_count - max_width;
  const size_t n = (out_count + tscrunch - 1) / tscrunch;
  if (n_series > 1 && n > out_stride) return HD_INVALID_STRIDE;
  if (n_series == 0) return HD_NO_ERROR;
  const unsigned gx = (unsigned)std::min<size_t>((n + 255) / 256, 4096);
  hipLaunchKernelGGL(k_boxcar, dim3(gx, (unsigned)n_series), dim3(256), 0, (hipStream_t)stream,
                     d_scanned, scanned_stride, n, (unsigned)(max_width / 2),
                     (unsigned)((width - 1) / 2 + 1), (unsigned)(width / 2), (unsigned)tscrunch,
                     d_out, out_stride);
  ST_CHECK(hipGetLastError());
  return HD_NO_ERROR;
}

extern "C" hd_error hd_find_giants(const hd_float* d_data, hd_size count, hd_size nbeams,
                                   hd_float thresh, hd_size merge_dist, hd_size beam_stride,
                                   hd_size beam_count, hd_giants* out, void* stream) {
  (void)nbeams;
  if (!d_data || !out) return HD_INVALID_POINTER;
  if (!beam_stride || count >= (1ull << 32)) return HD_INVALID_STRIDE;
  out->count = 0;
  if (count == 0) return HD_NO_ERROR;
  hipStream_t s = (hipStream_t)stream;
  Scratch tmp(s);
  GiantBuffers gb{};
  const size_t chunk = ST_BS * 16;
  const size_t nchunks = (count + chunk - 1) / chunk;
  unsigned* h = nullptr;
  ST_CHECK(hipHostMalloc((void**)&h, 16, hipHostMallocDefault));
  struct HostFree { unsigned* p; ~HostFree() { (void)hipHostFree(p); } } hf{h};
  size_t cap = std::min<size_t>(count, 1u << 22);
  ST_CHECK(tmp.get(&gb.counter, 4));
  ST_CHECK(tmp.get(&gb.ccount, nchunks));
  ST_CHECK(tmp.get(&gb.coff, nchunks));
  ST_CHECK(tmp.get(&gb.cpos, nchunks));
  ST_CHECK(tmp.get(&gb.scan_tmp, exclusive_scan_tmp_words(std::max<size_t>(count, nchunks))));
  unsigned nsamp = 0;
  for (int attempt = 0; attempt < 2; ++attempt) {
    ST_CHECK(tmp.get(&gb.skey, cap));
    ST_CHECK(tmp.get(&gb.sval, cap));
    gb.scap = cap;
    ST_CHECK(hipMemsetAsync(gb.counter, 0, sizeof(unsigned), s));
    hipLaunchKernelGGL(k_threshold_direct, dim3((unsigned)nchunks), dim3(ST_BS), 0, s, d_data,
                       count, thresh, beam_stride, beam_count, gb);
    ST_CHECK(hipGetLastError());
    ST_CHECK(hipMemcpyAsync(h, gb.counter, sizeof(unsigned), hipMemcpyDeviceToHost, s));
    ST_CHECK(hipStreamSynchronize(s));
    nsamp = h[0];
    if (nsamp <= cap) break;
    cap = nsamp;
  }
  if (nsamp == 0) return HD_NO_ERROR;
  ST_CHECK(tmp.get(&gb.okey, nsamp));
  ST_CHECK(tmp.get(&gb.oval, nsamp));
  ST_CHECK(tmp.get(&gb.head, nsamp));
  ST_CHECK(tmp.get(&gb.segid, nsamp));
  gb.nchunks_per_w = nchunks;
  ST_CHECK(launch_exclusive_scan_u32(gb.ccount, gb.cpos, nchunks, gb.scan_tmp, s));
  ST_CHECK(launch_order_samples(gb, nchunks, s));
  WidthTable wt{};
  wt.nw = 1;
  wt.off = 0;
  wt.wi[0].w = 1;
  wt.wi[0].ts = 1;
  wt.wi[0].cs = 1;
  wt.wi[0].merge_dist = (unsigned)merge_dist;
  ST_CHECK(launch_segment_heads(gb, wt, nsamp, s));
  ST_CHECK(launch_exclusive_scan_u32(gb.head, gb.segid, nsamp, gb.scan_tmp, s));
  ST_CHECK(hipMemcpyAsync(h, gb.segid + (nsamp - 1), 4, hipMemcpyDeviceToHost, s));
  ST_CHECK(hipMemcpyAsync(h + 1, gb.head + (nsamp - 1), 4, hipMemcpyDeviceToHost, s));
  ST_CHECK(hipStreamSynchronize(s));
  const size_t ng = (size_t)h[0] + h[1];
  GiantOut go{};
  ST_CHECK(tmp.get(&go.peak, ng));
  ST_CHECK(tmp.get(&go.ind, ng));
  ST_CHECK(tmp.get(&go.begin, ng));
  ST_CHECK(tmp.get(&go.end, ng));
  ST_CHECK(tmp.get(&go.filter, ng));
  ST_CHECK(tmp.get(&go.dm, ng));
  uint64_t* best;
  ST_CHECK(tmp.get(&best, ng));
  ST_CHECK(launch_segment_reduce(gb, wt, nsamp, (unsigned)ng, 1, go, best, s));
  const size_t k = std::min<size_t>(ng, out->capacity);
  if (k) {
    if (out->peaks) ST_CHECK(hipMemcpyAsync(out->peaks, go.peak, k * 4, hipMemcpyDeviceToDevice, s));
    if (out->inds) ST_CHECK(hipMemcpyAsync(out->inds, go.ind, k * 8, hipMemcpyDeviceToDevice, s));
    if (out->begins) ST_CHECK(hipMemcpyAsync(out->begins, go.begin, k * 8, hipMemcpyDeviceToDevice, s));
    if (out->ends) ST_CHECK(hipMemcpyAsync(out->ends, go.end, k * 8, hipMemcpyDeviceToDevice, s));
  }
  ST_CHECK(hipStreamSynchronize(s));
  out->count = ng;
  return HD_NO_ERROR;
}

extern "C" hd_error hd_label_candidate_clusters(hd_size count,
                                                const struct ConstRawCandidates* d_cands,
                                                hd_size time_tol, hd_size filter_tol,
                                                hd_size dm_tol, hd_size nsamps_beam,
                                                hd_size* d_labels, hd_size* label_count,
                                                void* stream) {
  if (!d_cands || !d_labels || !label_count) return HD_INVALID_POINTER;
  if (!d_cands->inds || !d_cands->filter_inds || !d_cands->dm_inds) return HD_INVALID_POINTER;
  if (!nsamps_beam || count >= (1ull << 31)) return HD_INVALID_STRIDE;
  *label_count = 0;
  if (count == 0) return HD_NO_ERROR;
  hipStream_t s = (hipStream_t)stream;
  Scratch tmp(s);
  const unsigned n = (unsigned)count;
  size_t npad = 1;
  while (npad < n) npad <<= 1;
  uint32_t *filt, *dm;
  ST_CHECK(tmp.get(&filt, n));
  ST_CHECK(tmp.get(&dm, n));
  hipLaunchKernelGGL(k_u64_to_u32, grid1(n), dim3(256), 0, s, (const uint64_t*)d_cands->filter_inds, filt, (size_t)n);
  hipLaunchKernelGGL(k_u64_to_u32, grid1(n), dim3(256), 0, s, (const uint64_t*)d_cands->dm_inds, dm, (size_t)n);
  ClusterBuffers cb{};
  ST_CHECK(tmp.get(&cb.sort_keys, npad));
  ST_CHECK(tmp.get(&cb.lo, n));
  ST_CHECK(tmp.get(&cb.hi, n));
  ST_CHECK(tmp.get(&cb.labels, n));
  ST_CHECK(tmp.get(&cb.flag, 4));
  ST_CHECK(tmp.get(&cb.root_flag, n));
  ST_CHECK(tmp.get(&cb.root_pos, n));
  ST_CHECK(tmp.get(&cb.scan_tmp, exclusive_scan_tmp_words(n)));
  cb.sort_n = npad;
  unsigned* h = nullptr;
  ST_CHECK(hipHostMalloc((void**)&h, 16, hipHostMallocDefault));
  struct HostFree { unsigned* p; ~HostFree() { (void)hipHostFree(p); } } hf{h};
  unsigned max_filter = 0;
  {
    std::vector<uint32_t> f(n);
    ST_CHECK(hipMemcpyAsync(f.data(), filt, n * 4, hipMemcpyDeviceToHost, s));
    ST_CHECK(hipStreamSynchronize(s));
    for (uint32_t v : f) max_filter = std::max(max_filter, v);
  }
  if (max_filter > 30) return HD_INVALID_STRIDE;
  int iters = 0;
  ST_CHECK(launch_cluster((const uint64_t*)d_cands->inds, filt, dm, nullptr, n, time_tol,
                          filter_tol, dm_tol, nsamps_beam, max_filter, cb, s, h, &iters));
  hipLaunchKernelGGL(k_i32_to_u64, grid1(n), dim3(256), 0, s, cb.labels, (uint64_t*)d_labels,
                     (size_t)n);
  hipLaunchKernelGGL(k_roots, grid1(n), dim3(256), 0, s, cb.labels, n, cb.root_flag);
  ST_CHECK(launch_exclusive_scan_u32(cb.root_flag, cb.root_pos, n, cb.scan_tmp, s));
  ST_CHECK(hipMemcpyAsync(h, cb.root_pos + (n - 1), 4, hipMemcpyDeviceToHost, s));
  ST_CHECK(hipMemcpyAsync(h + 1, cb.root_flag + (n - 1), 4, hipMemcpyDeviceToHost, s));
  ST_CHECK(hipStreamSynchronize(s));
  *label_count = (hd_size)h[0] + h[1];
  return HD_NO_ERROR;
}

extern "C" hd_error hd_merge_candidates(hd_size count, const hd_size* d_labels,
                                        const struct ConstRawCandidates* d_cands,
                                        const struct RawCandidates* d_groups, void* stream) {
  if (!d_labels || !d_cands || !d_groups) return HD_INVALID_POINTER;
  if (count >= (1ull << 31)) return HD_INVALID_STRIDE;
  if (count == 0) return HD_NO_ERROR;
  hipStream_t s = (hipStream_t)stream;
  Scratch tmp(s);
  const unsigned n = (unsigned)count;
  uint32_t *filt, *dm;
  ClusterBuffers cb{};
  ST_CHECK(tmp.get(&filt, n));
  ST_CHECK(tmp.get(&dm, n));
  ST_CHECK(tmp.get(&cb.labels, n));
  ST_CHECK(tmp.get(&cb.best, n));
  ST_CHECK(tmp.get(&cb.members, n));
  ST_CHECK(tmp.get(&cb.root_flag, n));
  ST_CHECK(tmp.get(&cb.root_pos, n));
  ST_CHECK(tmp.get(&cb.scan_tmp, exclusive_scan_tmp_words(n)));
  hipLaunchKernelGGL(k_u64_to_u32, grid1(n), dim3(256), 0, s, (const uint64_t*)d_cands->filter_inds, filt, (size_t)n);
  hipLaunchKernelGGL(k_u64_to_u32, grid1(n), dim3(256), 0, s, (const uint64_t*)d_cands->dm_inds, dm, (size_t)n);
  hipLaunchKernelGGL(k_u64_to_i32, grid1(n), dim3(256), 0, s, (const uint64_t*)d_labels, cb.labels, (size_t)n);
  ST_CHECK(hipGetLastError());
  ST_CHECK(launch_merge(d_cands->peaks, (const uint64_t*)d_cands->inds,
                        (const uint64_t*)d_cands->begins, (const uint64_t*)d_cands->ends, filt, dm,
                        (const uint64_t*)d_cands->members, n, cb, d_groups->peaks,
                        (uint64_t*)d_groups->inds, (uint64_t*)d_groups->begins,
                        (uint64_t*)d_groups->ends, (uint64_t*)d_groups->filter_inds,
                        (uint64_t*)d_groups->dm_inds, (uint64_t*)d_groups->members, s));
  ST_CHECK(hipStreamSynchronize(s));
  return HD_NO_ERROR;
}

extern "C" hd_error hd_clean_filterbank_rfi(hd_size nchans, hd_float f0, hd_float df,
                                            const hd_byte* d_in, hd_size nsamps, hd_size nbits,
                                            hd_byte* d_out, int* h_killmask, hd_float dm,
                                            hd_float dt, hd_float baseline_length,
                                            hd_float rfi_tol, hd_size rfi_min_beams,
                                            hd_size boxcar_max, hd_size nbeams, void* stream) {
  (void)h_killmask;     // not used by the reference either
  (void)rfi_min_beams;
  if (!d_in || !d_out) return HD_INVALID_POINTER;
  if (d_in == d_out) return HD_INVALID_POINTER;  // cleaning is out of place (pipeline.cu:253)
  CleanPlan cp;
  const int e = clean_plan(nsamps, nbits, nchans, nbeams, dm, dt, f0, df, baseline_length,
                           rfi_tol, boxcar_max, cp);
  if (e != HD_NO_ERROR) return (hd_error)e;
  hipStream_t s = (hipStream_t)stream;
  Scratch tmp(s);
  uint8_t* work;
  ST_CHECK(tmp.get(&work, cp.work_bytes));
  ST_CHECK(run_clean(cp, d_in, d_out, work, s));
  // the delay table is copied from cp's host memory, which must outlive the copy
  if (cp.max_delay > 0) ST_CHECK(hipStreamSynchronize(s));
  return HD_NO_ERROR;
}
