/*
 * hd_oracle.c -- CPU restatement of Heimdall's per-gulp search.
 *
 * TEST INFRASTRUCTURE ONLY: the checker for the HIP implementation.  Only
 * tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg load it.
 *
 * Semantics are the reference's sequential ones (Thrust with a sequential
 * backend, OMP_NUM_THREADS=1).  Each function cites the reference lines it
 * restates; casts are spelled out to reproduce the reference's mix of float,
 * double and size_t arithmetic.  Build with -ffp-contract=off: the reference
 * host build performs no fused multiply-adds.
 *
 * Parity pins (tests/test_oracle.py, tests/test_oracle_pins.py):
 *  - orc_generate_dm_list     : all 1434 trials of Share/dm2000.list (print-exact)
 *  - orc_delay_table/orc_delay : float64 closed form within half a sample,
 *                                zero at DM 0 / channel 0, monotone, and the
 *                                max delays of four setups
 *  - orc_median3/4/5           : order statistics over every tie pattern
 *  - orc_linear_stretch        : closed form on affine inputs, endpoints
 *  - orc_dedisperse            : brute-force triple loop, impulse response,
 *                                DM-0 = channel sum, linearity
 *  - orc_scan / orc_boxcar     : float64 closed forms within rounding, brute force;
 *                                the three-level order by an independent loop
 *  - orc_remove_baseline       : affine and constant series are removed exactly;
 *                                a numpy restatement, bit for bit
 *  - orc_get_rms_multibeam,    : numpy restatement of both pyramids (tails,
 *    orc_block_rms               float reblock index above 2^24), bit for bit
 *  - orc_find_giants           : brute-force run detection
 *  - orc_label_clusters        : brute-force DAG reachability definition
 *  - orc_merge_candidates      : brute-force group reduction
 *  - orc_scrunch_factors,      : equal to the numpy restatement of
 *    orc_dedisperse_adaptive     oracle/adaptive_dt.py, whose invariants
 *                                (doubling criterion in float64, brute-force
 *                                halved sums, all-ones = orc_dedisperse) are
 *                                tested; factor-1 trials of orc_execute with
 *                                scrunching equal the plain search's
 * Parity unpinned (no reference vector exists): whether dedisp rounds delays
 * as int(dm*k + 0.5f) (the closed form fixes the constant and units, not the
 * rounding rule), its output scaling, its adaptive-dt rule and halving
 * (dedisp is absent), and hash32's output values.
 */
#include "hd_oracle.h"

#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* ------------------------------------------------------------------------ */
/* medians: the reference's comparison networks (median_filter.cu:21-75)
 * return the order statistic (median4: mean of the middle two).  The oracle
 * computes that statistic by insertion sort. */
static void sortn(float* v, int n) {
  for (int i = 1; i < n; ++i) {
    float x = v[i];
    int j = i - 1;
    while (j >= 0 && x < v[j]) { v[j + 1] = v[j]; --j; }
    v[j + 1] = x;
  }
}
float orc_median3(float a, float b, float c) {
  float v[3] = {a, b, c};
  sortn(v, 3);
  return v[1];
}
float orc_median4(float a, float b, float c, float d) {
  float v[4] = {a, b, c, d};
  sortn(v, 4);
  return 0.5f * (v[1] + v[2]);
}
float orc_median5(float a, float b, float c, float d, float e) {
  float v[5] = {a, b, c, d, e};
  sortn(v, 5);
  return v[2];
}

/* ------------------------------------------------------------------------ */
/* DM list.  dedisp_generate_dm_list (called at pipeline.cu:178-182) steps
 * trials by the smearing-tolerance recurrence
 *   DM' = (b2*DM + sqrt(-a2*b2*DM^2 + (a2+b2)*(c + tol^2*a2*DM^2))) / (a2+b2)
 * with f the band centre in GHz, a = 8.3*df/f^3, b2 = a^2*nchans^2/16 and
 * c = (dt_us^2 + ti^2)*(tol^2-1).  Trials are stored as float and each step
 * starts from the stored float value; this reproduces Share/dm2000.list. */
size_t orc_generate_dm_list(float* out, size_t cap, float dm_min, float dm_max,
                            float dt_s, float pulse_width_us, float f0,
                            float df, size_t nchans, float tol) {
  double dt_us = (double)dt_s * 1e6;
  double f = ((double)f0 + ((double)(nchans / 2) - 0.5) * (double)df) * 1e-3;
  double tol2 = (double)tol * (double)tol;
  double a = 8.3 * (double)df / (f * f * f);
  double a2 = a * a;
  double b2 = a2 * (double)(nchans * nchans) / 16.0;
  double c = (dt_us * dt_us + (double)pulse_width_us * (double)pulse_width_us) *
             (tol2 - 1.0);
  size_t n = 0;
  float dm = dm_min;
  if (n < cap) out[n] = dm;
  ++n;
  while (dm < dm_max) {
    double prev = (double)dm;
    double prev2 = prev * prev;
    double k = c + tol2 * a2 * prev2;
    double next = (b2 * prev + sqrt(-a2 * b2 * prev2 + (a2 + b2) * k)) / (a2 + b2);
    dm = (float)next;
    if (n < cap) out[n] = dm;
    ++n;
  }
  return n;
}

/* Delay slope per channel, in samples per unit DM, for channel frequency
 * f_c = f0 + c*df (MHz): k_c = 4.15e3/dt * (1/f_c^2 - 1/f0^2), evaluated in
 * double and stored as float.  (The reference's dedisp library is absent;
 * this is the convention shared by the oracle and the product.) */
void orc_delay_table(float* out, size_t nchans, float dt_s, float f0, float df) {
  double inv0 = 1.0 / (double)f0;
  for (size_t c = 0; c < nchans; ++c) {
    double fc = (double)f0 + (double)c * (double)df;
    double inv = 1.0 / fc;
    out[c] = (float)(4.15e3 / (double)dt_s * (inv * inv - inv0 * inv0));
  }
}
int orc_delay(float dm, float kc) {
  float x = dm * kc + 0.5f;
  return (int)x;
}

/* ------------------------------------------------------------------------ */
/* Sample unpacking: channel c lives in byte c/(8/nbits) at bit
 * (c mod (8/nbits))*nbits, low bits first (clean_filterbank_rfi.cu:137-141). */
static unsigned sample_at(const uint8_t* row, size_t c, size_t nbits) {
  if (nbits == 8) return row[c];
  if (nbits == 16) return (unsigned)row[2 * c] | ((unsigned)row[2 * c + 1] << 8);
  size_t per = 8 / nbits;
  unsigned byte = row[c / per];
  unsigned shift = (unsigned)((c % per) * nbits);
  return (byte >> shift) & ((1u << nbits) - 1u);
}

int orc_dedisperse(const uint8_t* in, size_t nsamps, size_t nbits, size_t nchans,
                   const int* killmask, const float* dm_list, size_t n_dm,
                   const float* delay_table, size_t max_delay, float* out,
                   size_t out_stride, int nthreads) {
  if (!(nbits == 1 || nbits == 2 || nbits == 4 || nbits == 8 || nbits == 16))
    return -1;
  if (max_delay >= nsamps) return -2;
  size_t nout = nsamps - max_delay;
  size_t stride = nchans * nbits / 8;
  /* channel-major copy of the live channels */
  size_t nlive = 0;
  size_t* live = (size_t*)malloc(sizeof(size_t) * (nchans ? nchans : 1));
  for (size_t c = 0; c < nchans; ++c)
    if (!killmask || killmask[c]) live[nlive++] = c;
  uint16_t* x = (uint16_t*)malloc(sizeof(uint16_t) * (nlive ? nlive : 1) * nsamps);
  /* transposed in 64 x 64 blocks (a plain copy, blocked for the caches) */
#ifdef _OPENMP
#pragma omp parallel for schedule(static) num_threads(nthreads < 1 ? 1 : nthreads)
#endif
  for (long tb = 0; tb < (long)((nsamps + 63) / 64); ++tb) {
    const size_t t0 = (size_t)tb * 64, t1 = (t0 + 64 < nsamps) ? t0 + 64 : nsamps;
    for (size_t l0 = 0; l0 < nlive; l0 += 64) {
      const size_t l1 = (l0 + 64 < nlive) ? l0 + 64 : nlive;
      for (size_t t = t0; t < t1; ++t) {
        const uint8_t* row = in + t * stride;
        for (size_t l = l0; l < l1; ++l) x[l * nsamps + t] = (uint16_t)sample_at(row, live[l], nbits);
      }
    }
  }
#ifdef _OPENMP
  if (nthreads < 1) nthreads = 1;
#pragma omp parallel num_threads(nthreads)
#endif
  {
    uint32_t* acc = (uint32_t*)malloc(sizeof(uint32_t) * (nout ? nout : 1));
#ifdef _OPENMP
#pragma omp for schedule(dynamic, 1)
#endif
    for (long d = 0; d < (long)n_dm; ++d) {
      /* integer sums: the order of the additions does not matter, so time
       * is walked in tiles that keep the accumulators in the L1 cache */
      enum { TT = 2048 };
      for (size_t t0 = 0; t0 < nout; t0 += TT) {
        const size_t tn = (nout - t0 < TT) ? nout - t0 : TT;
        uint32_t* a = acc + t0;
        memset(a, 0, sizeof(uint32_t) * tn);
        for (size_t l = 0; l < nlive; ++l) {
          int off = orc_delay(dm_list[d], delay_table[live[l]]);
          const uint16_t* xr = x + l * nsamps + off + t0;
          for (size_t t = 0; t < tn; ++t) a[t] += xr[t];
        }
      }
      float* o = out + (size_t)d * out_stride;
      for (size_t t = 0; t < nout; ++t) o[t] = (float)acc[t];
    }
    free(acc);
  }
  free(x);
  free(live);
  return 0;
}

/* ------------------------------------------------------------------------ */
/* Adaptive time resolution: dedisp's smearing rule and halving (dedisp is
 * absent from the reference tree; pipeline.cu:187-196 enables it and
 * pipeline.cu:489-498 consumes its factors).  Parity unpinned: the numpy
 * restatement in oracle/adaptive_dt.py and the invariants of
 * tests/test_adaptive_dt.py are its checks. */
static float smearing(float dt, float width, float f0, size_t nchans, float df, float dm,
                      float ddm) {
  const float bw = (float)nchans * fabsf(df);
  const float fc = f0 - bw / 2.f;
  const float inv_fc3 = 1.f / (fc * fc * fc);
  const float t_dm = 8.3f * bw * dm * inv_fc3;
  const float t_ddm = 8.3f / 4.f * bw * (float)nchans * ddm * inv_fc3;
  return sqrtf(dt * dt + width * width + t_dm * t_dm + t_ddm * t_ddm);
}

void orc_scrunch_factors(const float* dms, size_t n, float dt, float pulse_width_us, float f0,
                         float df, size_t nchans, float tol, size_t* out) {
  if (n == 0) return;
  const float w = pulse_width_us * 1e-6f;
  out[0] = 1;
  for (size_t d = 1; d < n; ++d) {
    const float ddm = dms[d] - dms[d - 1];
    const float cur = (float)out[d - 1] * dt;
    const float a = smearing(cur, w, f0, nchans, df, dms[d], ddm);
    const float b = smearing(2.f * cur, w, f0, nchans, df, dms[d], ddm);
    out[d] = (b / a < tol) ? 2 * out[d - 1] : out[d - 1];
  }
}

int orc_dedisperse_adaptive(const uint8_t* in, size_t nsamps, size_t nbits, size_t nchans,
                            const int* killmask, const float* dm_list, size_t n_dm,
                            const float* delay_table, size_t max_delay, const size_t* factors,
                            float* out, size_t out_stride, int nthreads) {
  if (!(nbits == 1 || nbits == 2 || nbits == 4 || nbits == 8)) return -1;
  if (max_delay >= nsamps) return -2;
  enum { PAD = 4 };
  const size_t nout = nsamps - max_delay;
  const size_t stride = nchans * nbits / 8;
  /* unpacked samples, [time][channel], PAD zero samples after the last */
  uint8_t* x = (uint8_t*)calloc((nsamps + PAD) * (nchans ? nchans : 1), 1);
  float* kcs = (float*)malloc(sizeof(float) * (nchans ? nchans : 1));
  for (size_t t = 0; t < nsamps; ++t)
    for (size_t c = 0; c < nchans; ++c)
      x[t * nchans + c] = (uint8_t)sample_at(in + t * stride, c, nbits);
  size_t level = 1, ns = nsamps;
  int rc = 0;
  for (size_t a = 0; a < n_dm && rc == 0;) {
    const size_t f = factors[a];
    size_t b = a;
    while (b < n_dm && factors[b] == f) ++b;
    while (level < f) { /* halve: pairwise means rounded down */
      const size_t h = ns / 2;
      for (size_t t = 0; t < h; ++t)
        for (size_t c = 0; c < nchans; ++c)
          x[t * nchans + c] =
              (uint8_t)(((unsigned)x[2 * t * nchans + c] + x[(2 * t + 1) * nchans + c]) >> 1);
      memset(x + h * nchans, 0, (ns - h + PAD) * nchans);
      ns = h;
      level *= 2;
    }
    const size_t nout_s = nout / f;
    for (size_t d = a; d < b; ++d) memset(out + d * out_stride, 0, sizeof(float) * nout);
    if (nout_s > 0) {
      /* the slope over a power of two is exact in float */
      for (size_t c = 0; c < nchans; ++c) kcs[c] = delay_table[c] / (float)f;
      rc = orc_dedisperse(x, ns + PAD, 8, nchans, killmask, dm_list + a, b - a, kcs,
                          ns + PAD - nout_s, out + a * out_stride, out_stride, nthreads);
    }
    a = b;
  }
  free(kcs);
  free(x);
  return rc;
}

/* ------------------------------------------------------------------------ */
/* linear_stretch_functor (median_filter.cu:451-464):
 *   step = float(in-1)/(out-1); x = i*step; j = (unsigned)x;
 *   y = in[j] + ((x-j > 1e-5) ? (x-j)*(in[j+1]-in[j]) : 0) */
void orc_linear_stretch(const float* in, size_t in_count, float* out,
                        size_t out_count) {
  float step = (float)(in_count - 1) / (float)(out_count - 1);
  for (size_t ii = 0; ii < out_count; ++ii) {
    unsigned i = (unsigned)ii;
    float x = (float)i * step;
    unsigned j = (unsigned)x;
    float fr = x - (float)j;
    out[ii] = in[j] + (((double)fr > 1e-5) ? fr * (in[j + 1] - in[j]) : 0.f);
  }
}

/* linear_stretch_beam_functor (median_filter.cu:481-499): the index is carried
 * through float across beams:
 *   step = float(in_beam-1)/out_beam; beam = i/out_beam; x = i*step;
 *   k = (unsigned)(x + beam); o = (x - k) + beam;
 *   y = in[k] + ((o > 1e-5) ? o*(in[k+1]-in[k]) : 0) */
void orc_linear_stretch_beam(const float* in, size_t in_count_beam, float* out,
                             size_t out_count_beam, size_t nbeam) {
  size_t total = out_count_beam * nbeam;
  float step = (float)(in_count_beam - 1) / (float)out_count_beam;
  for (size_t ii = 0; ii < total; ++ii) {
    unsigned i = (unsigned)ii;
    size_t beam = i / out_count_beam;
    float x = (float)i * step;
    unsigned k = (unsigned)(x + (float)beam);
    float o = (x - (float)k) + (float)beam;
    out[ii] = in[k] + (((double)o > 1e-5) ? o * (in[k + 1] - in[k]) : 0.f);
  }
}

/* median_scrunch5_beam (median_filter.cu:194-235,365-395) */
static void median_scrunch5_beam(const float* in, size_t size, size_t nbeam,
                                 float* out) {
  if (size > 4) {
    size_t oc = size / 5;
    size_t total = nbeam * oc;
    for (size_t o = 0; o < total; ++o) {
      size_t beam = o / oc;
      size_t i = (o % oc) * 5;
      const float* p = in + beam * size + i;
      out[o] = orc_median5(p[0], p[1], p[2], p[3], p[4]);
    }
  } else {
    for (size_t o = 0; o < nbeam; ++o) {
      const float* p = in + o * size;
      if (size == 1) out[o] = p[0];
      else if (size == 2) out[o] = 0.5f * (p[0] + p[1]);
      else if (size == 3) out[o] = orc_median3(p[0], p[1], p[2]);
      else out[o] = orc_median4(p[0], p[1], p[2], p[3]);
    }
  }
}

/* median_scrunch5 (median_filter.cu:288-322) */
static void median_scrunch5(const float* in, size_t count, float* out) {
  if (count == 1) out[0] = in[0];
  else if (count == 2) out[0] = 0.5f * (in[0] + in[1]);
  else if (count == 3) out[0] = orc_median3(in[0], in[1], in[2]);
  else if (count == 4) out[0] = orc_median4(in[0], in[1], in[2], in[3]);
  else {
    size_t oc = count / 5;
    for (size_t o = 0; o < oc; ++o) {
      const float* p = in + 5 * o;
      out[o] = orc_median5(p[0], p[1], p[2], p[3], p[4]);
    }
  }
}

/* ------------------------------------------------------------------------ */
/* RemoveBaselinePlan keeps buf1/buf2 across calls (remove_baseline.cu:16-19,
 * device_vector::resize keeps old contents), so the ping-pong buffers here
 * persist too.  They are zero-filled on creation. */
struct orc_work {
  float* buf1;
  float* buf2;
  size_t cap1, cap2;
  float* baseline;
  size_t capb;
};

struct orc_work* orc_work_create(void) {
  struct orc_work* w = (struct orc_work*)calloc(1, sizeof(struct orc_work));
  return w;
}
void orc_work_destroy(struct orc_work* w) {
  if (!w) return;
  free(w->buf1);
  free(w->buf2);
  free(w->baseline);
  free(w);
}
static void grow(float** p, size_t* cap, size_t n) {
  n += 64; /* slack: reads one past a level stay inside the allocation */
  if (*cap >= n) return;
  float* q = (float*)calloc(n, sizeof(float));
  if (*p) memcpy(q, *p, *cap * sizeof(float));
  free(*p);
  *p = q;
  *cap = n;
}

int orc_remove_baseline(struct orc_work* w, float* d, size_t beam_stride,
                        size_t beam_count, size_t smooth_radius, size_t nbeams) {
  int own = 0;
  if (!w) { w = orc_work_create(); own = 1; }
  size_t dm_delay = beam_stride - beam_count;               /* :30 */
  size_t beam_samples = beam_stride;                        /* :32 */
  float oversample = 2;                                     /* :36 */
  /* :38  (float*float)/float, then + 0.5 in double */
  size_t sample_count =
      (size_t)((double)(oversample * (float)beam_samples / (float)(2 * smooth_radius)) + 0.5);
  if (sample_count == 0) { if (own) orc_work_destroy(w); return 0; }
  /* :45-46 integer division inside the log */
  size_t nscrunches = (size_t)(log((double)(beam_samples / sample_count)) / log(5.));
  size_t count_round = (size_t)(pow(5., (double)nscrunches) * (double)sample_count);

  grow(&w->buf1, &w->cap1, nbeams * count_round);           /* :56 */
  grow(&w->buf2, &w->cap2, nbeams * (count_round / 5));     /* :57 */
  float* buf1 = w->buf1;
  float* buf2 = w->buf2;
  size_t nbm = nbeams - 1;                                  /* :66 */
  size_t nsamps = beam_stride * nbeams - dm_delay;          /* :67 */
  grow(&w->baseline, &w->capb, beam_stride * nbeams);       /* :68 */
  float* baseline = w->baseline;

  /* :73-75 resample each beam to count_round points */
  orc_linear_stretch_beam(d, beam_stride, buf1, count_round, nbm);
  orc_linear_stretch(d + nbm * beam_stride, beam_count, buf1 + nbm * count_round,
                     count_round);
  /* :78-82 median-of-5 down to sample_count */
  for (size_t size = count_round; size > sample_count; size /= 5) {
    median_scrunch5_beam(buf1, size, nbeams, buf2);
    float* t = buf1; buf1 = buf2; buf2 = t;
  }
  /* :88-90 stretch back */
  orc_linear_stretch_beam(buf1, sample_count, baseline, beam_stride, nbm);
  orc_linear_stretch(buf1 + nbm * sample_count, sample_count,
                     baseline + nbm * beam_stride, beam_count);
  /* :99-102 subtract */
  for (size_t i = 0; i < nsamps; ++i) d[i] = d[i] - baseline[i];
  if (own) orc_work_destroy(w);
  return 0;
}

/* reblock_abs_beam_functor (get_rms.cu:214-229): the source index is
 * (size_t)((float)beam*delta + (float)o) with delta = stride - length. */
static void reblock_abs_beam(const float* in, size_t in_stride, size_t in_length,
                             float* out, size_t nbeam) {
  size_t out_count = in_length * nbeam;
  float delta = (float)(in_stride - in_length);
  for (size_t oo = 0; oo < out_count; ++oo) {
    unsigned o = (unsigned)oo;
    size_t beam = o / in_length;
    size_t i = (size_t)((float)beam * delta + (float)o);
    out[oo] = fabsf(in[i]);
  }
}

int orc_get_rms_multibeam(float* d, float* rms, size_t beam_stride,
                          size_t beam_count, size_t nbeams) {
  size_t count = nbeams * beam_count;                       /* get_rms.cu:73 */
  float* buf1 = (float*)calloc(count + 64, sizeof(float));
  float* buf2 = (float*)calloc(count / 5 + 64, sizeof(float));
  float* b1 = buf1;
  float* b2 = buf2;
  reblock_abs_beam(d, beam_stride, beam_count, b1, nbeams); /* :81 */
  for (size_t size = beam_count; size > 1; size /= 5) {     /* :83-87 */
    median_scrunch5_beam(b1, size, nbeams, b2);
    float* t = b1; b1 = b2; b2 = t;
  }
  for (size_t b = 0; b < nbeams; ++b) rms[b] = b1[b] * 1.4862f; /* :94-97 */
  free(buf1);
  free(buf2);
  return 0;
}

/* normalise_beam_kernel (get_rms.cu:174-186,202-212) */
void orc_normalise_multibeam(float* d, const float* rms, size_t beam_stride,
                             size_t cur_nsamps) {
  for (size_t ii = 0; ii < cur_nsamps; ++ii) {
    unsigned i = (unsigned)ii;
    size_t beam = i / beam_stride;
    d[ii] = d[ii] / rms[beam];
  }
}

float orc_block_rms(const float* d, size_t beam_stride, size_t beam_count,
                    size_t nbeams) {
  size_t count = nbeams * beam_count;                       /* get_rms.cu:109 */
  float* buf1 = (float*)calloc(count + 64, sizeof(float));
  float* buf2 = (float*)calloc(count / 5 + 64, sizeof(float));
  float* b1 = buf1;
  float* b2 = buf2;
  reblock_abs_beam(d, beam_stride, beam_count, b1, nbeams); /* :117 */
  for (size_t size = count; size > 1; size /= 5) {          /* :119-123 */
    median_scrunch5(b1, size, b2);
    float* t = b1; b1 = b2; b2 = t;
  }
  float med = b1[0];                                        /* :127 */
  float rms = (float)((double)med * 1.4862);                /* :128 */
  free(buf1);
  free(buf2);
  return rms;
}

void orc_normalise_block(float* d, float block_rms, size_t count) {
  float inv = 1.0f / block_rms;                             /* get_rms.cu:195 */
  for (size_t i = 0; i < count; ++i) d[i] = d[i] * inv;
}

/* MatchedFilterPlan::prep (matched_filter.cu:21-32): exclusive scan over
 * count+1 inputs; only the first count inputs reach the output.  The
 * reference's thrust::exclusive_scan on the device adds in a parallel order
 * that Thrust does not specify; the order adopted here (and by the search's
 * k_scan_seg) has three levels: samples in segments of 16, summed in order
 * (P); segments in chunks of 64, their totals summed in order inside the
 * chunk (Q, the chunk's earlier segments); chunks in order (O, the earlier
 * chunks); scanned[k + 1] = O + (Q + P) for sample k. */
void orc_scan(const float* x, size_t count, float* scanned) {
  scanned[0] = 0.f;
  float O = 0.f;
  for (size_t c0 = 0; c0 < count; c0 += 1024) {
    const size_t ce = c0 + 1024 < count ? c0 + 1024 : count;
    float Q = 0.f;
    for (size_t s0 = c0; s0 < ce; s0 += 16) {
      const size_t se = s0 + 16 < ce ? s0 + 16 : ce;
      float P = 0.f;
      for (size_t k = s0; k < se; ++k) {
        P = P + x[k];
        scanned[k + 1] = O + (Q + P);
      }
      Q = Q + P;
    }
    O = O + Q;
  }
}

/* MatchedFilterPlan::exec (matched_filter.cu:37-67) with strided_range
 * (strided_range.h:24-64): out[k] = S[o+ahead+k*ts] - S[o-behind+k*ts]. */
size_t orc_boxcar(const float* scanned, size_t scanned_size, size_t max_width,
                  size_t width, size_t tscrunch, float* out) {
  size_t offset = max_width / 2;
  size_t ahead = (width - 1) / 2 + 1;
  size_t behind = width / 2;
  size_t out_count = scanned_size - max_width;
  size_t n = (out_count + (tscrunch - 1)) / tscrunch;
  for (size_t k = 0; k < n; ++k)
    out[k] = scanned[offset + ahead + k * tscrunch] - scanned[offset - behind + k * tscrunch];
  return n;
}

/* ------------------------------------------------------------------------ */
/* GiantFinder::exec (find_giants.cu:101-316).  Samples i < count with
 * (i % beam_stride) < beam_count and value > thresh are grouped: a new giant
 * starts when ind > previous ind + merge_dist (adjacent pairs, :79-87 and
 * :197-210).  Each giant keeps the first maximum (:62-68), begin = first
 * index, end = last index + 1 (:283-303). */
size_t orc_find_giants(const float* d, size_t count, float thresh,
                       size_t merge_dist, size_t beam_stride, size_t beam_count,
                       orc_cand* out, size_t cap) {
  size_t ng = 0;
  int open = 0;
  orc_cand cur;
  size_t last = 0;
  memset(&cur, 0, sizeof(cur));
  for (size_t ii = 0; ii < count; ++ii) {
    unsigned i = (unsigned)ii; /* counting_iterator<unsigned> (:139) */
    size_t samp = (size_t)i % beam_stride;
    float v = d[ii];
    if (!(samp < beam_count && v > thresh)) continue;
    if (open && (size_t)i > last + merge_dist) {
      if (ng < cap) out[ng] = cur;
      ++ng;
      open = 0;
    }
    if (!open) {
      open = 1;
      cur.peak = v;
      cur.ind = i;
      cur.begin = i;
    } else if (!(cur.peak >= v)) {
      cur.peak = v;
      cur.ind = i;
    }
    cur.end = (uint64_t)i + 1;
    last = i;
  }
  if (open) {
    if (ng < cap) out[ng] = cur;
    ++ng;
  }
  return ng;
}

/* are_coincident_beam (are_coincident.cuh:46-68) */
static int coincident(const orc_cand* a, const orc_cand* b, size_t time_tol,
                      size_t filter_tol, size_t dm_tol, size_t nsamps_beam) {
  size_t beam_i = a->ind / nsamps_beam;
  size_t beam_j = b->ind / nsamps_beam;
  int wi = 1 << (int)a->filter;
  int wj = 1 << (int)b->filter;
  size_t tt = time_tol * (size_t)(wi > wj ? wi : wj);
  return beam_i == beam_j &&
         (size_t)abs((int)b->ind - (int)a->ind) <= tt &&
         (size_t)abs((int)b->filter - (int)a->filter) <= filter_tol &&
         (size_t)abs((int)b->dm - (int)a->dm) <= dm_tol;
}

/* cluster_functor (label_candidate_clusters.cu:141-197) run for i = 0..N-1 in
 * order, then trace_equivalency_chain (:101-139) and the root count
 * (:384-391). */
size_t orc_label_clusters(size_t n, const orc_cand* c, size_t time_tol,
                          size_t filter_tol, size_t dm_tol, size_t nsamps_beam,
                          uint64_t* labels) {
  for (size_t i = 0; i < n; ++i) labels[i] = i;
  for (size_t i = 0; i < n; ++i) {
    for (size_t j = 0; j < n; ++j) {
      if (j == i) continue;
      if (coincident(&c[i], &c[j], time_tol, filter_tol, dm_tol, nsamps_beam)) {
        int li = (int)labels[i], lj = (int)labels[j];
        labels[i] = (uint64_t)(li < lj ? li : lj);
      }
    }
  }
  for (size_t i = 0; i < n; ++i) {
    uint64_t cur = i;
    while (labels[cur] != cur) cur = labels[cur];
    labels[i] = cur;
  }
  size_t count = 0;
  for (size_t i = 0; i < n; ++i) count += (labels[i] == i);
  return count;
}

/* merge_candidates (merge_candidates.cu:21-72,102-131): stable sort by label,
 * then per group the first maximum SNR supplies the fields and members sum. */
typedef struct { uint64_t label; size_t idx; } lab_idx;
static int cmp_lab(const void* a, const void* b) {
  const lab_idx* x = (const lab_idx*)a;
  const lab_idx* y = (const lab_idx*)b;
  if (x->label != y->label) return x->label < y->label ? -1 : 1;
  return x->idx < y->idx ? -1 : (x->idx > y->idx ? 1 : 0); /* stable */
}
void orc_merge_candidates(size_t n, const uint64_t* labels, const orc_cand* c,
                          orc_cand* groups) {
  if (n == 0) return;
  lab_idx* p = (lab_idx*)malloc(sizeof(lab_idx) * n);
  for (size_t i = 0; i < n; ++i) { p[i].label = labels[i]; p[i].idx = i; }
  qsort(p, n, sizeof(lab_idx), cmp_lab);
  size_t g = 0;
  orc_cand acc = c[p[0].idx];
  for (size_t k = 1; k <= n; ++k) {
    if (k == n || p[k].label != p[k - 1].label) {
      groups[g++] = acc;
      if (k < n) acc = c[p[k].idx];
      continue;
    }
    const orc_cand* m = &c[p[k].idx];
    uint64_t members = acc.members + m->members;
    if (!(acc.peak >= m->peak)) acc = *m;
    acc.members = members;
  }
  free(p);
}

/* ------------------------------------------------------------------------ */
void orc_default_params(orc_params* p) {
  memset(p, 0, sizeof(*p));
  /* default_params.C:10-66 */
  p->nchans = 1024;
  p->dt = 64e-6f;
  p->f0 = 1581.804688f;
  p->df = -.390625f;
  p->dm_min = 0.0f;
  p->dm_max = 1000.0f;
  p->dm_tol = 1.25f;
  p->dm_pulse_width = 40;
  p->baseline_length = 2.0f;
  p->boxcar_max = 4096;
  p->detect_thresh = 6.0f;
  p->cand_sep_time = 3;
  p->cand_sep_filter = 3;
  p->cand_sep_dm = 200;
  p->max_giant_rate = 0;
  p->min_tscrunch_width = 4096;
  p->nthreads = 1;
  p->rfi_tol = 5.0f;
  p->use_scrunching = 0;
  p->scrunch_tol = 1.15f;
}

/* ------------------------------------------------------------------------ */
/* Zero-DM RFI cleaning (clean_filterbank_rfi.cu). */

/* hash() at clean_filterbank_rfi.cu:33-43, the published 32-bit integer hash
 * of Robert Jenkins (six add/xor-shift rounds). */
uint32_t orc_hash32(uint32_t a) {
  a = (a + 0x7ed55d16u) + (a << 12);
  a = (a ^ 0xc761c23cu) ^ (a >> 19);
  a = (a + 0x165667b1u) + (a << 5);
  a = (a + 0xd3a2646cu) ^ (a << 9);
  a = (a + 0xfd7046c5u) + (a << 3);
  a = (a ^ 0xb55a4f09u) ^ (a >> 16);
  return a;
}

/* minstd_rand (Park-Miller: multiplier 48271, modulus 2^31-1, increment 0);
 * a seed that is 0 mod m becomes 1 (thrust linear_congruential_engine::seed). */
#define ORC_MINSTD_M 2147483647u
uint32_t orc_minstd_seed(uint32_t s) {
  uint32_t x = s % ORC_MINSTD_M;
  return x == 0 ? 1u : x;
}
uint32_t orc_minstd_next(uint32_t* x) {
  *x = (uint32_t)(((uint64_t)*x * 48271u) % ORC_MINSTD_M);
  return *x;
}

/* uniform_int_distribution<unsigned>(lo, hi)(rng): a uniform_real_distribution
 * <double>(lo, hi + 1) value truncated to unsigned; the real draw maps the
 * engine output u (min 1, max m-1) to (u - 1) / (1 + (m - 2)). */
uint32_t orc_uniform_draw(uint32_t* x, uint32_t lo, uint32_t hi) {
  double r = (double)(uint32_t)(orc_minstd_next(x) - 1u);
  r /= (1.0 + (double)(ORC_MINSTD_M - 1u - 1u));
  const double a = (double)lo;
  const double b = (double)hi + 1.0;
  const double v = (r * (b - a)) + a;
  return (uint32_t)v;
}

/* zap_fb_rfi_functor (clean_filterbank_rfi.cu:51-108) over 32-bit words. */
int orc_zap_filterbank_rfi(const int* mask, const uint8_t* in, size_t nsamps, size_t nbits,
                           size_t nchans, size_t max_resample_dist, uint8_t* out) {
  const size_t stride_bytes = nchans * nbits / 8;
  if (stride_bytes % 4 != 0 || nbits == 0 || nbits > 16) return -1;
  if (nsamps == 0 || nsamps - 1 < max_resample_dist) return -2;
  if (nsamps * (stride_bytes / 4) > 0xffffffffull) return -1;  /* 32-bit word index */
  const uint32_t stride = (uint32_t)(stride_bytes / 4);
  const uint32_t n = (uint32_t)nsamps, maxd = (uint32_t)max_resample_dist;
  const uint32_t bitmask = (1u << nbits) - 1u;
  const uint32_t* win = (const uint32_t*)in;
  uint32_t* wout = (uint32_t*)out;
  const uint32_t nwords = n * stride;
  for (uint32_t i = 0; i < nwords; ++i) {
    const uint32_t t = i / stride, c = i % stride;
    uint32_t result;
    if (mask[t]) {
      uint32_t x = orc_minstd_seed(orc_hash32(i));
      result = 0;
      for (uint32_t k = 0; k < 32; k += (uint32_t)nbits) {
        const uint32_t min_t = t > maxd ? t - maxd : 0;
        const uint32_t max_t = t < n - 1 - maxd ? t + maxd : n - 1;
        uint32_t new_t = orc_uniform_draw(&x, min_t, max_t);
        int attempts = 0;
        while (mask[new_t] && ++attempts < 10 + 1) new_t = orc_uniform_draw(&x, min_t, max_t);
        const uint32_t val = (win[(size_t)new_t * stride + c] >> k) & bitmask;
        result |= val << k;
      }
    } else {
      result = win[i];
    }
    wout[i] = result;
  }
  return 0;
}

int orc_clean_filterbank_rfi(const uint8_t* in, size_t nsamps, size_t nbits, size_t nchans,
                             float dm, float dt, float f0, float df, float baseline_length,
                             float rfi_tol, size_t boxcar_max, size_t nbeams, uint8_t* out,
                             int* mask_out) {
  if (nbeams == 0 || boxcar_max == 0) return -1;
  /* single-DM plan (:361-377): adaptive dt off, all channels */
  float* kc = (float*)malloc(sizeof(float) * (nchans ? nchans : 1));
  orc_delay_table(kc, nchans, dt, f0, df);
  size_t max_delay = 0;
  for (size_t c = 0; c < nchans; ++c) {
    int dd = orc_delay(dm, kc[c]);
    if (dd > 0 && (size_t)dd > max_delay) max_delay = (size_t)dd;
  }
  if ((nchans * nbits / 8) % 4 != 0) { free(kc); return -1; }
  if (max_delay >= nsamps) { free(kc); return -2; }
  const size_t beam_stride = nsamps / nbeams;              /* :379 */
  const size_t N = nsamps - max_delay;                     /* :380 nsamps_computed */
  if (max_delay >= beam_stride || N + 1 < boxcar_max) { free(kc); return -2; }
  const size_t beam_nsamps = beam_stride - max_delay;      /* :381 */
  /* (:383-397) dedisperse at dm */
  float* series = (float*)calloc(N + 64, sizeof(float));
  if (orc_dedisperse(in, nsamps, nbits, nchans, NULL, &dm, 1, kc, max_delay, series, N, 1) != 0) {
    free(series); free(kc); return -1;
  }
  free(kc);
  /* (:404-439) baseline, per-beam RMS, normalise */
  const size_t nsamps_smooth = (size_t)(baseline_length / (2 * dt)); /* :405 */
  orc_remove_baseline(NULL, series, beam_stride, beam_nsamps, nsamps_smooth, nbeams);
  float* beam_rms = (float*)calloc(nbeams, sizeof(float));
  orc_get_rms_multibeam(series, beam_rms, beam_stride, beam_nsamps, nbeams);
  orc_normalise_multibeam(series, beam_rms, beam_stride, N);
  free(beam_rms);
  /* (:448-459) sigma cut */
  int* mask = (int*)calloc(N + 1, sizeof(int));
  for (size_t i = 0; i < N; ++i) mask[i] = fabsf(series[i]) > rfi_tol;
  /* (:461-518) boxcars of the series; each width's cut is widened by a
   * boxcar over the cut itself and ORed in */
  const size_t L = N + 1 - boxcar_max;
  const size_t off = boxcar_max / 2;
  float* S = (float*)calloc(N + 1, sizeof(float));
  float* filt = (float*)calloc(L ? L : 1, sizeof(float));
  int* fm = (int*)calloc(N + 1, sizeof(int));
  int* Sm = (int*)calloc(N + 1, sizeof(int));
  orc_scan(series, N, S);
  for (size_t w = 1; w <= boxcar_max; w *= 2) {
    orc_boxcar(S, N + 1, boxcar_max, w, 1, filt);
    const float norm = (float)(1.0 / sqrtf((float)w));   /* :477-478 */
    for (size_t k = 0; k < L; ++k) filt[k] = filt[k] * norm;
    for (size_t k = 0; k < L; ++k) fm[off + k] = fabsf(filt[k]) > rfi_tol; /* :487-489 */
    /* MatchedFilterPlan<int> over the cut (:494-498), written in place */
    Sm[0] = 0;
    for (size_t k = 1; k <= N; ++k) Sm[k] = Sm[k - 1] + fm[k - 1];
    const size_t ahead = (w - 1) / 2 + 1, behind = w / 2;
    for (size_t k = 0; k < L; ++k) fm[off + k] = Sm[off + ahead + k] - Sm[off - behind + k];
    for (size_t i = 0; i < N; ++i) mask[i] = mask[i] || fm[i];  /* :501-504 */
  }
  free(S); free(filt); free(fm); free(Sm); free(series);
  if (mask_out) memcpy(mask_out, mask, sizeof(int) * N);
  /* (:509-518) zap the first N samples */
  const int rc = orc_zap_filterbank_rfi(mask, in, N, nbits, nchans, nsamps_smooth / 4, out);
  free(mask);
  return rc;
}

static unsigned filter_index(unsigned v) { /* get_filter_index, pipeline.cu:126-136 */
  unsigned r = 0;
  while ((1u << (r + 1)) <= v && r < 31) ++r;
  return r;
}

static double now_s(void) {
  struct timespec ts;
  clock_gettime(CLOCK_MONOTONIC, &ts);
  return ts.tv_sec + 1e-9 * ts.tv_nsec;
}

static void push_giants(orc_result* r, size_t* cap, const orc_cand* g, size_t n) {
  if (r->n_giants + n > *cap) {
    size_t nc = (*cap ? *cap * 2 : 1024);
    while (nc < r->n_giants + n) nc *= 2;
    r->giants = (orc_cand*)realloc(r->giants, nc * sizeof(orc_cand));
    *cap = nc;
  }
  memcpy(r->giants + r->n_giants, g, n * sizeof(orc_cand));
  r->n_giants += n;
}

/* Per-thread copies of the reference's persistent buffers: time series,
 * prefix sums, filtered series (pipeline.cu:76-77, zero-filled by resize) and
 * the baseline plan's scrunch buffers. */
struct dm_state {
  float* ts;
  float* scanned;
  float* filt;
  float* beam_rms;
  struct orc_work* work;
  orc_cand* tmp;
  size_t tmpcap;
};
static struct dm_state* dm_state_create(size_t series_stride, size_t nbeams) {
  struct dm_state* s = (struct dm_state*)calloc(1, sizeof(struct dm_state));
  s->ts = (float*)calloc(series_stride + 64, sizeof(float));
  s->scanned = (float*)calloc(series_stride + 64, sizeof(float));
  s->filt = (float*)calloc(series_stride + 64, sizeof(float));
  s->beam_rms = (float*)calloc(nbeams, sizeof(float));
  s->work = orc_work_create();
  s->tmpcap = 1 << 12;
  s->tmp = (orc_cand*)malloc(sizeof(orc_cand) * s->tmpcap);
  return s;
}
static void dm_state_destroy(struct dm_state* s) {
  free(s->ts); free(s->scanned); free(s->filt); free(s->beam_rms); free(s->tmp);
  orc_work_destroy(s->work);
  free(s);
}
struct dm_ctx {
  const orc_params* p;
  size_t nsamps, nbeams, nsamps_computed, max_delay, W;
};

/* One DM trial of the loop at pipeline.cu:488-790 at scrunch factor
 * cur_dm_scrunch (the trial's row holds nsamps_computed / cur_dm_scrunch
 * samples).  Appends its giants to r; returns 1 when the giant-rate limit
 * trips. */
static int process_dm(const struct dm_ctx* c, struct dm_state* st, size_t dm_idx,
                      const float* row, orc_result* r, size_t* gcap, size_t cur_dm_scrunch) {
  const orc_params* p = c->p;
  const size_t W = c->W, nbeams = c->nbeams, nsamps = c->nsamps;
  /* A factor above boxcar_max searches no width (:657-659): no giants */
  if (cur_dm_scrunch > W) return 0;
  size_t cur_nsamps = c->nsamps_computed / cur_dm_scrunch;  /* :490 */
  float cur_dt = p->dt * (float)cur_dm_scrunch;             /* :491 */
  size_t cur_dm_delay = c->max_delay / cur_dm_scrunch;      /* :493-497 */
  size_t cur_beam_stride = nsamps / nbeams / cur_dm_scrunch; /* :498 */
  size_t cur_beam_nsamps = cur_beam_stride - cur_dm_delay;  /* :499 */
  memcpy(st->ts, row, sizeof(float) * cur_nsamps);
  size_t nsamps_smooth = (size_t)(p->baseline_length / (2 * cur_dt)); /* :572-573 */
  orc_remove_baseline(st->work, st->ts, cur_beam_stride, cur_beam_nsamps, nsamps_smooth, nbeams);
  orc_get_rms_multibeam(st->ts, st->beam_rms, cur_beam_stride, cur_beam_nsamps, nbeams);
  orc_normalise_multibeam(st->ts, st->beam_rms, cur_beam_stride, cur_nsamps);
  size_t rel_boxcar_max = W / cur_dm_scrunch;               /* :626 */
  size_t max_nsamps_filtered = cur_nsamps + 1 - rel_boxcar_max;
  size_t cur_filtered_offset = rel_boxcar_max / 2;
  orc_scan(st->ts, cur_nsamps, st->scanned);                /* :643 */
  for (size_t fw = cur_dm_scrunch; fw <= W; fw *= 2) {      /* :657-659 */
    size_t rel_filter_width = fw / cur_dm_scrunch;
    size_t filter_idx = filter_index((unsigned)fw);
    size_t rel_min_tscrunch_width = p->min_tscrunch_width / cur_dm_scrunch;
    if (rel_min_tscrunch_width < 1) rel_min_tscrunch_width = 1;
    size_t rel_tscrunch_width = 2 * rel_filter_width / rel_min_tscrunch_width;
    if (rel_tscrunch_width < 1) rel_tscrunch_width = 1;
    size_t rel_rel_filter_width = rel_filter_width / rel_tscrunch_width;
    orc_boxcar(st->scanned, cur_nsamps + 1, rel_boxcar_max, rel_filter_width,
               rel_tscrunch_width, st->filt);               /* :680-682 */
    size_t cur_nsamps_filtered = ((max_nsamps_filtered - 1) / rel_tscrunch_width + 1);
    size_t cur_scrunch = cur_dm_scrunch * rel_tscrunch_width;
    size_t bc = cur_beam_stride - (cur_dm_delay + rel_boxcar_max); /* :692 */
    float block_rms = orc_block_rms(st->filt, cur_beam_stride, bc, nbeams); /* :709 */
    orc_normalise_block(st->filt, block_rms, cur_nsamps_filtered);         /* :712 */
    size_t ng;
    for (;;) {
      ng = orc_find_giants(st->filt, cur_nsamps_filtered, p->detect_thresh,
                           p->cand_sep_time * rel_rel_filter_width,
                           cur_beam_stride, bc, st->tmp, st->tmpcap);
      if (ng <= st->tmpcap) break;
      st->tmpcap = ng;
      st->tmp = (orc_cand*)realloc(st->tmp, sizeof(orc_cand) * st->tmpcap);
    }
    size_t rel_off = cur_filtered_offset / rel_tscrunch_width; /* :755-756 */
    for (size_t g = 0; g < ng; ++g) {                        /* :759-775 */
      st->tmp[g].ind = (st->tmp[g].ind + rel_off) * cur_scrunch;
      st->tmp[g].begin = (st->tmp[g].begin + rel_off) * cur_scrunch;
      st->tmp[g].end = (st->tmp[g].end + rel_off) * cur_scrunch;
      st->tmp[g].filter = filter_idx;
      st->tmp[g].dm = dm_idx;
      st->tmp[g].members = 1;
    }
    push_giants(r, gcap, st->tmp, ng);
    /* :780-787 */
    float data_length_mins = (float)((double)((float)nsamps * p->dt) / 60.0);
    if (p->max_giant_rate != 0 &&
        ((float)r->n_giants / data_length_mins > p->max_giant_rate))
      return 1;
  }
  return 0;
}

/* hd_execute (pipeline.cu:233-1087) without the file/TCP output; with
 * use_scrunching, at the adaptive time resolution of orc_scrunch_factors. */
int orc_execute(const orc_params* p, const uint8_t* fb, size_t nsamps,
                size_t nbits, size_t first_idx, size_t nbeams, orc_result* r) {
  (void)first_idx;
  double t0 = now_s();
  memset(r, 0, sizeof(*r));
  size_t nchans = p->nchans;
  /* DM list and delays (pipeline.cu:171-185, 341-365) */
  size_t n_dm = orc_generate_dm_list(NULL, 0, p->dm_min, p->dm_max, p->dt,
                                     p->dm_pulse_width, p->f0, p->df, nchans, p->dm_tol);
  float* dm_list = (float*)malloc(sizeof(float) * n_dm);
  orc_generate_dm_list(dm_list, n_dm, p->dm_min, p->dm_max, p->dt,
                       p->dm_pulse_width, p->f0, p->df, nchans, p->dm_tol);
  float* kc = (float*)malloc(sizeof(float) * nchans);
  orc_delay_table(kc, nchans, p->dt, p->f0, p->df);
  size_t max_delay = 0;
  for (size_t c = 0; c < nchans; ++c) {
    int dd = orc_delay(dm_list[n_dm - 1], kc[c]);
    if (dd > 0 && (size_t)dd > max_delay) max_delay = (size_t)dd;
  }
  r->n_dm = n_dm;
  r->max_delay = max_delay;
  size_t W = p->boxcar_max;
  /* :370-372 */
  if (nsamps < nbeams * (W + max_delay)) r->nsamps_processed = 0;
  else r->nsamps_processed = nsamps - nbeams * (W + max_delay);
  if (max_delay >= nsamps) { r->error = 11; free(dm_list); free(kc); return r->error; }

  size_t nsamps_computed = nsamps - max_delay;              /* :366 */
  size_t series_stride = nsamps_computed;
  size_t dm_first = p->dm_first;
  size_t dm_end = p->dm_count ? dm_first + p->dm_count : n_dm;
  if (dm_end > n_dm) dm_end = n_dm;
  size_t step = p->dm_stride ? p->dm_stride : 1;
  size_t n_sub = 0;
  for (size_t d = dm_first; d < dm_end; d += step) ++n_sub;
  size_t* sub_idx = (size_t*)malloc(sizeof(size_t) * (n_sub ? n_sub : 1));
  float* sub_dm = (float*)malloc(sizeof(float) * (n_sub ? n_sub : 1));
  n_sub = 0;
  for (size_t d = dm_first; d < dm_end; d += step) { sub_idx[n_sub] = d; sub_dm[n_sub++] = dm_list[d]; }
  /* scrunch factors of the whole list (pipeline.cu:187-196, 344-345), then
   * of the trials processed */
  size_t* sub_fac = (size_t*)malloc(sizeof(size_t) * (n_sub ? n_sub : 1));
  int adaptive = 0;
  {
    size_t* fac = (size_t*)malloc(sizeof(size_t) * n_dm);
    for (size_t d = 0; d < n_dm; ++d) fac[d] = 1;
    if (p->use_scrunching)
      orc_scrunch_factors(dm_list, n_dm, p->dt, p->dm_pulse_width, p->f0, p->df, nchans,
                          p->scrunch_tol, fac);
    for (size_t k = 0; k < n_sub; ++k) {
      sub_fac[k] = fac[sub_idx[k]];
      if (sub_fac[k] != 1) adaptive = 1;
    }
    free(fac);
  }
  float* plane = (float*)malloc(sizeof(float) * (n_sub ? n_sub : 1) * series_stride);
  /* CLEANING build (pipeline.cu:263-299): zero-DM cleaning, boxcars to 64 */
  uint8_t* cleaned = NULL;
  if (p->cleaning) {
    const size_t nbytes = nsamps * nchans * nbits / 8;
    cleaned = (uint8_t*)malloc(nbytes ? nbytes : 1);
    memcpy(cleaned, fb, nbytes);
    int rc = orc_clean_filterbank_rfi(fb, nsamps, nbits, nchans, 0.f, p->dt, p->f0, p->df,
                                      p->baseline_length, p->rfi_tol, 64, nbeams, cleaned, NULL);
    if (rc != 0) {
      r->error = rc == -1 ? 7 : 11;
      free(cleaned); free(plane); free(dm_list); free(kc); free(sub_idx); free(sub_dm);
      free(sub_fac);
      return r->error;
    }
    fb = cleaned;
  }
  const int drc =
      adaptive ? orc_dedisperse_adaptive(fb, nsamps, nbits, nchans, p->killmask, sub_dm, n_sub,
                                         kc, max_delay, sub_fac, plane, series_stride,
                                         p->nthreads)
               : orc_dedisperse(fb, nsamps, nbits, nchans, p->killmask, sub_dm, n_sub, kc,
                                max_delay, plane, series_stride, p->nthreads);
  if (drc != 0) {
    r->error = 8; free(plane); free(dm_list); free(kc); free(sub_idx); free(sub_dm);
    free(sub_fac); free(cleaned);
    return r->error;
  }
  free(cleaned);

  /* The DM loop breaks before its first trial when the delay leaves no room
   * (pipeline.cu:501-506, 633-638); both tests are the same for every trial. */
  size_t cur_beam_stride = nsamps / nbeams;                 /* :498 */
  int skip = (max_delay >= cur_beam_stride) || (max_delay + W > cur_beam_stride);
  int too_many = 0;
  struct dm_ctx ctx = {p, nsamps, nbeams, nsamps_computed, max_delay, W};
  if (!skip && (p->max_giant_rate != 0 || p->nthreads <= 1)) {
    /* sequential, with the giant-rate test after every width (:780-787) */
    struct dm_state* st = dm_state_create(series_stride, nbeams);
    size_t gcap = 0;
    for (size_t k = 0; k < n_sub && !too_many; ++k) {
      too_many = process_dm(&ctx, st, sub_idx[k], plane + k * series_stride, r, &gcap,
                            sub_fac[k]);
    }
    dm_state_destroy(st);
  } else if (!skip) {
    /* independent trials in parallel; giants concatenated in trial order */
    orc_result* part = (orc_result*)calloc(n_sub ? n_sub : 1, sizeof(orc_result));
#ifdef _OPENMP
#pragma omp parallel num_threads(p->nthreads)
#endif
    {
      struct dm_state* st = dm_state_create(series_stride, nbeams);
#ifdef _OPENMP
#pragma omp for schedule(dynamic, 1)
#endif
      for (long k = 0; k < (long)n_sub; ++k) {
        size_t gcap = 0;
        process_dm(&ctx, st, sub_idx[k], plane + (size_t)k * series_stride, &part[k], &gcap,
                   sub_fac[k]);
      }
      dm_state_destroy(st);
    }
    size_t total = 0;
    for (size_t k = 0; k < n_sub; ++k) total += part[k].n_giants;
    r->giants = (orc_cand*)malloc(sizeof(orc_cand) * (total ? total : 1));
    for (size_t k = 0; k < n_sub; ++k) {
      memcpy(r->giants + r->n_giants, part[k].giants, part[k].n_giants * sizeof(orc_cand));
      r->n_giants += part[k].n_giants;
      free(part[k].giants);
    }
    free(part);
  }
  free(sub_idx);
  free(sub_dm);
  free(sub_fac);

  if (!too_many && r->n_giants > 0) {
    uint64_t* labels = (uint64_t*)malloc(sizeof(uint64_t) * r->n_giants);
    size_t nsamps_beam = nsamps / nbeams;                   /* :829 */
    size_t label_count = orc_label_clusters(r->n_giants, r->giants, p->cand_sep_time,
                                            p->cand_sep_filter, p->cand_sep_dm,
                                            nsamps_beam, labels);
    r->groups = (orc_cand*)malloc(sizeof(orc_cand) * (label_count ? label_count : 1));
    r->group_dms = (float*)malloc(sizeof(float) * (label_count ? label_count : 1));
    orc_merge_candidates(r->n_giants, labels, r->giants, r->groups);
    r->n_groups = label_count;
    for (size_t g = 0; g < label_count; ++g) r->group_dms[g] = dm_list[r->groups[g].dm];
    free(labels);
  }
  r->error = too_many ? 12 : 0;
  free(plane); free(dm_list); free(kc);
  r->seconds = now_s() - t0;
  return r->error;
}

void orc_result_free(orc_result* r) {
  free(r->giants);
  free(r->groups);
  free(r->group_dms);
  memset(r, 0, sizeof(*r));
}

/* .cand lines (pipeline.cu:1012-1040): groups inside each beam's valid
 * window, fields separated by tabs, floats in the default stream format. */
size_t orc_format_cands(const orc_result* r, const orc_params* p, size_t nsamps,
                        size_t nbeams, size_t first_idx, char* buf, size_t cap) {
  unsigned nsamps_beam = (unsigned)(nsamps / nbeams);       /* :903 */
  unsigned nsamps_valid = (unsigned)(r->nsamps_processed / nbeams); /* :905 */
  size_t used = 0;
  char line[512];
  if (cap) buf[0] = 0;
  for (size_t i = 0; i < r->n_groups; ++i) {
    const orc_cand* g = &r->groups[i];
    size_t beam_idx = (size_t)floor((double)((float)g->ind / (float)nsamps_beam));
    size_t beam_start = beam_idx * nsamps_beam + (p->boxcar_max / 2);
    size_t beam_end = beam_start + nsamps_valid;
    if (!(beam_start <= g->ind && g->ind < beam_end)) continue;
    size_t samp_offset = g->ind - ((size_t)nsamps_beam * beam_idx);
    size_t samp_idx = first_idx + samp_offset;
    float snr = isinf(g->peak) ? 0.0f : g->peak;
    float tsec = p->dt * (float)samp_idx;
    int n = snprintf(line, sizeof(line), "%g\t%zu\t%g\t%zu\t%zu\t%g\t%zu\t%zu\t%zu\t%zu\t\n",
                     (double)snr, samp_idx, (double)tsec, (size_t)g->filter,
                     (size_t)g->dm, (double)r->group_dms[i], (size_t)g->members,
                     first_idx + (size_t)(g->begin % nsamps_beam),
                     first_idx + (size_t)(g->end % nsamps_beam),
                     p->beam + beam_idx + 1);
    if (used + (size_t)n < cap) memcpy(buf + used, line, (size_t)n + 1);
    used += (size_t)n;
  }
  return used;
}
