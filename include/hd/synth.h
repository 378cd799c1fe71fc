/* hd/synth.h -- seeded synthetic filterbanks with injected pulses.
 *
 * No counterpart in the reference.  The generator writes [nsamps][nchans *
 * nbits/8] bytes in the sigproc/STF layout the reference's readers deliver
 * (channel 0 = f0, sub-byte samples low bits first):
 *   noise   N(mean, sigma^2) per sample from a counter-based hash of (seed,
 *           sample, channel), rounded half to even and clipped;
 *   pulses  Gaussian profiles (FWHM = width samples) centred at t0 + delay(dm,
 *           c) on the search plan's integer delay table, with the per-channel
 *           amplitude that gives the requested optimal matched-filter S/N;
 *   tones   sinusoids on chosen channels (narrow-band RFI);
 *   bursts  zero-DM Gaussian pulses in every live channel (broad-band RFI);
 *   dead    channels held at a constant value.
 * The host generator (libhd_synth.so) and hd_synth_filterbank_device (this
 * library, gfx950) compute the same formulas in double precision; they agree
 * except where the two math libraries' last-bit differences move a value
 * across a rounding boundary.
 */
#pragma once

#include <stddef.h>
#include <stdint.h>

#include "hd/types.h"
#include "hd/error.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct hd_synth_pulse {
  double t0;     /* arrival sample at channel 0 */
  float dm;      /* dispersion measure */
  float width;   /* FWHM in samples */
  float snr;     /* optimal matched-filter S/N */
} hd_synth_pulse;

typedef struct hd_synth_tone {
  int chan;
  float amp_sigma;  /* amplitude in units of sigma */
  float period;     /* samples */
  float phase;      /* radians */
} hd_synth_tone;

typedef struct hd_synth_spec {
  size_t nsamps, nchans, nbits;
  float dt, f0, df;
  uint64_t seed;
  float mean, sigma;
  const int* dead;          /* nchans flags (1 = dead) or NULL */
  float dead_value;
  const hd_synth_pulse* pulses;
  size_t npulses;
  const hd_synth_pulse* bursts;  /* zero-DM: dm ignored */
  size_t nbursts;
  const hd_synth_tone* tones;
  size_t ntones;
} hd_synth_spec;

/* Writes the filterbank of `spec` (host arrays) into device memory d_out on
 * device gpu_id. */
hd_error hd_synth_filterbank_device(const hd_synth_spec* spec, hd_byte* d_out, int gpu_id);

#ifdef __cplusplus
}
#endif
