// hd_series.hpp -- device helpers shared by the kernels that normalise a
// dedispersed series (baseline, per-beam RMS, prefix sums, RFI mask).
#pragma once

#include "hd_device.hpp"
#include "hd_internal.hpp"

namespace hd {

// The stretched baseline at series position pos (remove_baseline.cu:88-90):
// beams 0..nb-2 through linear_stretch_beam over their full stride, the last
// beam through linear_stretch over bc samples.
__device__ __forceinline__ float baseline_at(const float* P, const SeriesShape& sh, size_t pos) {
  if (sh.sc == 0) return 0.f;
  const size_t split = (size_t)(sh.nbeams - 1) * sh.stride;
  if (pos < split) {
    StretchBeam f{sh.step_bl_beam, (unsigned)sh.stride};
    return f(P, (unsigned)pos);
  }
  Stretch f{sh.step_bl_last};
  return f(P + (size_t)(sh.nbeams - 1) * sh.sc, (unsigned)(pos - split));
}

// normalise_multibeam after the baseline subtraction (remove_baseline.cu:
// 99-102, get_rms.cu:183-184): z = (x - baseline(pos)) / rms[pos / stride].
// P: the series' baseline points, R: its per-beam RMS values.
__device__ __forceinline__ float normalised_at(float x, const float* P, const float* R,
                                              const SeriesShape& sh, size_t pos) {
  if (sh.nbeams == 1) {
    // one beam: every position is in beam 0, stretched over bc samples
    const float bl = (sh.sc == 0) ? 0.f : Stretch{sh.step_bl_last}(P, (unsigned)pos);
    return (x - bl) / R[0];
  }
  const float y = x - baseline_at(P, sh, pos);
  return y / R[(unsigned)((unsigned)pos / sh.stride)];
}

}  // namespace hd
