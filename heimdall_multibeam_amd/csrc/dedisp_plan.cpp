// dedisp_plan.cpp -- the dedisp plan calls of the reference's tools
// (include/hd/dedisp_plan.h) over the search's unpack and byte-staged
// dedispersion kernels, with adaptive time resolution by repeated halving of
// the unpacked input (k_scrunch2).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <vector>

#include "hd/dedisp_plan.h"
#include "hd_internal.hpp"
#include "hd_plan.hpp"

namespace hd {

bool pick_byte_tile(const std::vector<int>& d, size_t rows, size_t nl, size_t nbits,
                    DedispGeom& dg) {
  static const int tiles[][4] = {{32, 8, 1, 0}, {32, 4, 1, 0}, {16, 4, 0, 0}, {8, 4, 0, 0}};
  for (const auto& t : tiles) {
    const int db = t[0];
    int spread = 0;
    for (size_t m0 = 0; m0 < rows; m0 += db) {
      const size_t ml = std::min(m0 + db, rows) - 1;
      for (size_t l = 0; l < nl; ++l)
        spread = std::max(spread, d[ml * nl + l] - (d[m0 * nl + l] & ~3));
    }
    dg.db = db;
    dg.cc = t[1];
    dg.wide = t[2];
    dg.remap = 0;
    dg.widen8 = nbits <= 4 ? 255 / ((1 << nbits) - 1) : 256;
    dg.split8 = nbits == 8 ? 1 : 0;
    dg.wlen8 = dedisp8_window_dwords(spread);
    dg.wlen = 0;
    if (dedisp_lds_bytes(dg) <= 160 * 1024) return true;
  }
  return false;
}

}  // namespace hd

namespace {

struct Buf {
  void* p = nullptr;
  size_t bytes = 0;
  hipError_t ensure(size_t n) {
    if (n <= bytes && p) return hipSuccess;
    if (p) (void)hipFree(p);
    p = nullptr;
    bytes = 0;
    const hipError_t e = hipMalloc(&p, n ? n : 16);
    if (e == hipSuccess) bytes = n ? n : 16;
    return e;
  }
  void release() {
    if (p) (void)hipFree(p);
    p = nullptr;
    bytes = 0;
  }
  template <typename T> T* as() const { return static_cast<T*>(p); }
};

size_t round_up(size_t x, size_t a) { return (x + a - 1) / a * a; }

hd_error gpu_err(hipError_t e) {
  if (e == hipSuccess) return HD_NO_ERROR;
  if (e == hipErrorOutOfMemory || e == hipErrorMemoryAllocation) return HD_MEM_ALLOC_FAILED;
  return HD_INTERNAL_GPU_ERROR;
}

#define PL_CHECK(expr)                                  \
  do {                                                  \
    const hipError_t _e = (expr);                       \
    if (_e != hipSuccess) return gpu_err(_e);           \
  } while (0)

}  // namespace

struct hd_dedisp_plan_t {
  int device = 0;
  hipStream_t stream = nullptr;
  size_t nchans = 0;
  float dt = 0.f, f0 = 0.f, df = 0.f;
  std::vector<float> dms, kc;
  std::vector<int> killmask;
  bool adaptive = false;
  float pulse_width = 0.f, tol = 0.f;
  std::vector<hd_size> factors;
  size_t max_delay = 0;
  Buf in, x[2], live, delays, plane, out;
};

namespace {

// the trial list changed: scrunch factors and the largest delay
void refresh(hd_dedisp_plan p) {
  std::vector<size_t> s(p->dms.size(), 1);
  if (p->adaptive && !p->dms.empty())
    s = hdplan::scrunch_list(p->dms, p->dt, p->pulse_width, p->f0, p->df, p->nchans, p->tol);
  p->factors.assign(s.begin(), s.end());
  p->max_delay = 0;
  for (float dm : p->dms)
    for (size_t c = 0; c < p->nchans; ++c) {
      const int d = hdplan::delay(dm, p->kc[c]);
      if (d > 0 && (size_t)d > p->max_delay) p->max_delay = (size_t)d;
    }
}

struct DeviceGuard {
  int prev = -1;
  explicit DeviceGuard(int dev) {
    if (hipGetDevice(&prev) != hipSuccess) prev = -1;
    if (prev != dev) (void)hipSetDevice(dev);
  }
  ~DeviceGuard() {
    int cur = -1;
    if (prev >= 0 && hipGetDevice(&cur) == hipSuccess && cur != prev) (void)hipSetDevice(prev);
  }
};


}  // namespace

extern "C" hd_error hd_dedisp_create_plan(hd_dedisp_plan* plan, hd_size nchans, hd_float dt,
                                          hd_float f0, hd_float df) {
  if (!plan) return HD_INVALID_POINTER;
  *plan = nullptr;
  if (nchans == 0 || !(dt > 0.f)) return HD_INVALID_STRIDE;
  hd_dedisp_plan p = new hd_dedisp_plan_t;
  if (hipGetDevice(&p->device) != hipSuccess) {
    delete p;
    return HD_INVALID_DEVICE_INDEX;
  }
  if (hipStreamCreateWithFlags(&p->stream, hipStreamNonBlocking) != hipSuccess) {
    delete p;
    return HD_INTERNAL_GPU_ERROR;
  }
  p->nchans = nchans;
  p->dt = dt;
  p->f0 = f0;
  p->df = df;
  p->kc = hdplan::delay_slopes(nchans, dt, f0, df);
  p->killmask.assign(nchans, 1);
  *plan = p;
  return HD_NO_ERROR;
}

extern "C" void hd_dedisp_destroy_plan(hd_dedisp_plan p) {
  if (!p) return;
  {
    DeviceGuard g(p->device);
    for (Buf* b : {&p->in, &p->x[0], &p->x[1], &p->live, &p->delays, &p->plane, &p->out})
      b->release();
    if (p->stream) (void)hipStreamDestroy(p->stream);
  }
  delete p;
}

extern "C" hd_error hd_dedisp_set_dm_list(hd_dedisp_plan p, const hd_float* dm_list,
                                          hd_size count) {
  if (!p) return HD_INVALID_PIPELINE;
  if (!dm_list && count) return HD_INVALID_POINTER;
  p->dms.assign(dm_list, dm_list + count);
  refresh(p);
  return HD_NO_ERROR;
}

extern "C" hd_error hd_dedisp_generate_dm_list(hd_dedisp_plan p, hd_float dm_start,
                                               hd_float dm_end, hd_float pulse_width_us,
                                               hd_float tol) {
  if (!p) return HD_INVALID_PIPELINE;
  p->dms = hdplan::dm_list(dm_start, dm_end, p->dt, pulse_width_us, p->f0, p->df, p->nchans,
                           tol);
  refresh(p);
  return HD_NO_ERROR;
}

extern "C" hd_error hd_dedisp_set_killmask(hd_dedisp_plan p, const int* killmask) {
  if (!p) return HD_INVALID_PIPELINE;
  if (killmask)
    p->killmask.assign(killmask, killmask + p->nchans);
  else
    p->killmask.assign(p->nchans, 1);
  return HD_NO_ERROR;
}

extern "C" hd_error hd_dedisp_enable_adaptive_dt(hd_dedisp_plan p, hd_float pulse_width_us,
                                                 hd_float tol) {
  if (!p) return HD_INVALID_PIPELINE;
  p->adaptive = true;
  p->pulse_width = pulse_width_us;
  p->tol = tol;
  refresh(p);
  return HD_NO_ERROR;
}

extern "C" hd_error hd_dedisp_disable_adaptive_dt(hd_dedisp_plan p) {
  if (!p) return HD_INVALID_PIPELINE;
  p->adaptive = false;
  refresh(p);
  return HD_NO_ERROR;
}

extern "C" const hd_size* hd_dedisp_get_dt_factors(hd_dedisp_plan p) {
  return p && !p->factors.empty() ? p->factors.data() : nullptr;
}
extern "C" hd_size hd_dedisp_get_max_delay(hd_dedisp_plan p) { return p ? p->max_delay : 0; }
extern "C" hd_size hd_dedisp_get_dm_delay(hd_dedisp_plan p, hd_size dm_idx) {
  if (!p || dm_idx >= p->dms.size()) return 0;
  size_t m = 0;
  for (size_t c = 0; c < p->nchans; ++c) {
    const int d = hdplan::delay(p->dms[dm_idx], p->kc[c]);
    if (d > 0 && (size_t)d > m) m = (size_t)d;
  }
  return m;
}
extern "C" hd_size hd_dedisp_get_dm_count(hd_dedisp_plan p) { return p ? p->dms.size() : 0; }
extern "C" const hd_float* hd_dedisp_get_dm_list(hd_dedisp_plan p) {
  return p && !p->dms.empty() ? p->dms.data() : nullptr;
}
extern "C" hd_size hd_dedisp_get_channel_count(hd_dedisp_plan p) { return p ? p->nchans : 0; }
extern "C" hd_float hd_dedisp_get_dt(hd_dedisp_plan p) { return p ? p->dt : 0.f; }
extern "C" hd_float hd_dedisp_get_f0(hd_dedisp_plan p) { return p ? p->f0 : 0.f; }
extern "C" hd_float hd_dedisp_get_df(hd_dedisp_plan p) { return p ? p->df : 0.f; }

extern "C" hd_error hd_dedisp_execute(hd_dedisp_plan p, hd_size nsamps, const hd_byte* in,
                                      hd_size in_nbits, hd_byte* out, hd_size out_nbits,
                                      unsigned flags) {
  if (!p) return HD_INVALID_PIPELINE;
  if (!in || !out) return HD_INVALID_POINTER;
  if (!(in_nbits == 1 || in_nbits == 2 || in_nbits == 4 || in_nbits == 8)) return HD_INVALID_NBITS;
  if (out_nbits != 32) return HD_INVALID_NBITS;
  if ((p->nchans * in_nbits) % 8) return HD_INVALID_STRIDE;
  const size_t n_dm = p->dms.size();
  if (n_dm == 0) return HD_INVALID_PIPELINE;
  if (nsamps <= p->max_delay) return HD_TOO_FEW_NSAMPS;
  const bool on_device = (flags & HD_DEDISP_DEVICE_POINTERS) != 0;
  DeviceGuard guard(p->device);
  (void)hipGetLastError();
  hipStream_t s = p->stream;
  const size_t stride = p->nchans * in_nbits / 8;
  const size_t nout = nsamps - p->max_delay;
  const int TB = hd::dedisp_tile_time();
  const size_t ppitch = round_up(nout + 1, (size_t)TB);
  // rows long enough for every tile's staged window (tile start <= nout,
  // window <= tile + span + 32 bytes) and for k_scrunch2's loads
  const size_t xpitch = round_up(ppitch + p->max_delay + 2 * (size_t)TB + 256, 256);

  std::vector<int> live;
  for (size_t c = 0; c < p->nchans; ++c)
    if (p->killmask[c]) live.push_back((int)c);
  const size_t nl = std::max<size_t>(live.size(), 1);

  const hd_byte* d_in = in;
  if (!on_device) {
    PL_CHECK(p->in.ensure(nsamps * stride));
    PL_CHECK(hipMemcpyAsync(p->in.p, in, nsamps * stride, hipMemcpyHostToDevice, s));
    d_in = p->in.as<hd_byte>();
  }
  PL_CHECK(p->x[0].ensure(nl * xpitch));
  PL_CHECK(p->plane.ensure(n_dm * ppitch * sizeof(float)));
  PL_CHECK(hipMemsetAsync(p->plane.p, 0, n_dm * ppitch * sizeof(float), s));
  if (live.empty()) {
    // no live channel: every sum is 0
  } else {
    PL_CHECK(p->live.ensure(live.size() * sizeof(int)));
    PL_CHECK(hipMemcpyAsync(p->live.p, live.data(), live.size() * sizeof(int),
                            hipMemcpyHostToDevice, s));
    PL_CHECK(hd::launch_unpack(d_in, stride, (int)in_nbits, nsamps, p->live.as<int>(),
                               (int)live.size(), p->x[0].as<uint8_t>(), xpitch, s,
                               hd::GroupArgs{}, live.size() == p->nchans));
    // delays of every trial in samples of its own resolution: the slope
    // divided by the factor (a power of two, so exactly)
    std::vector<int> d(n_dm * live.size());
    for (size_t m = 0; m < n_dm; ++m) {
      const float inv = 1.f / (float)p->factors[m];
      for (size_t l = 0; l < live.size(); ++l)
        d[m * live.size() + l] = hdplan::delay(p->dms[m], p->kc[live[l]] * inv);
    }
    PL_CHECK(p->delays.ensure(d.size() * sizeof(int)));
    PL_CHECK(hipMemcpyAsync(p->delays.p, d.data(), d.size() * sizeof(int),
                            hipMemcpyHostToDevice, s));
    int cur = 0;
    size_t level = 1, ns = nsamps;
    for (size_t a = 0; a < n_dm;) {
      const size_t f = p->factors[a];
      size_t b = a;
      while (b < n_dm && p->factors[b] == f) ++b;
      while (level < f) {  // halve the input's resolution up to this class's
        PL_CHECK(p->x[1 - cur].ensure(nl * xpitch));
        PL_CHECK(hd::launch_scrunch2(p->x[cur].as<uint8_t>(), (int)live.size(), xpitch, ns,
                                     p->x[1 - cur].as<uint8_t>(), s));
        cur = 1 - cur;
        ns /= 2;
        level *= 2;
      }
      hd::DedispGeom dg;
      std::memset(&dg, 0, sizeof(dg));
      dg.n_dm = (int)(b - a);
      dg.n_live = (int)live.size();
      dg.nout = nout / f;
      dg.xpitch = xpitch;
      dg.ppitch = ppitch;
      const std::vector<int> dc(d.begin() + a * live.size(), d.begin() + b * live.size());
      if (!hd::pick_byte_tile(dc, b - a, live.size(), in_nbits, dg)) return HD_INVALID_STRIDE;
      PL_CHECK(hd::launch_dedisperse(p->x[cur].as<uint8_t>(),
                                     p->delays.as<int>() + a * live.size(), nullptr, dg,
                                     p->plane.as<float>() + a * ppitch, s));
      a = b;
    }
  }
  // rows of nout floats; a scrunched trial's tail past nout / f is the
  // plane's zero fill (its tiles stop at nout / f)
  if (on_device) {
    PL_CHECK(hipMemcpy2DAsync(out, nout * sizeof(float), p->plane.p, ppitch * sizeof(float),
                              nout * sizeof(float), n_dm, hipMemcpyDeviceToDevice, s));
  } else {
    PL_CHECK(hipMemcpy2DAsync(out, nout * sizeof(float), p->plane.p, ppitch * sizeof(float),
                              nout * sizeof(float), n_dm, hipMemcpyDeviceToHost, s));
  }
  PL_CHECK(hipStreamSynchronize(s));
  return HD_NO_ERROR;
}
