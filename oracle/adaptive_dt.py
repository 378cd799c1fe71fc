"""CPU restatement of dedispersion at adaptive time resolution, as the
reference requests it from the dedisp library (Pipeline/pipeline.cu:187-196,
Applications/candidate_profiler.C:261-327).

TEST INFRASTRUCTURE ONLY: the checker for hd_dedisp_* (include/hd/
dedisp_plan.h); nothing in the product imports it.

dedisp is an external library that the reference tree does not hold, and no
reference output exists for it, so these rules restate its published method
and are parity unpinned:
  smearing, scrunch_list  the effective pulse width and the doubling rule
                          (float32 throughout)
  scrunch2                pairwise means of unpacked samples, rounded down
  dedisperse_adaptive     trial d from the input halved log2(s_d) times, with
                          delays int(dm * k_c / s_d + 0.5) in those samples and
                          (nsamps - max_delay) // s_d outputs
What the tests pin instead: factors are powers of two from 1, never
decreasing, and follow the criterion evaluated independently in float64
except at its boundary; all factors 1 reproduce orc_dedisperse; a constant
input gives n_live * value for every output; a factor-s trial equals a
brute-force sum over the halved input.
"""
import numpy as np

from oracle import oracle as O

f32 = np.float32


def smearing(dt, width, f0, nchans, df, dm, ddm):
    bw = f32(nchans) * f32(abs(f32(df)))
    fc = f32(f0) - bw / f32(2)
    inv_fc3 = f32(1) / (fc * fc * fc)
    t_dm = f32(8.3) * bw * f32(dm) * inv_fc3
    t_ddm = f32(8.3) / f32(4) * bw * f32(nchans) * f32(ddm) * inv_fc3
    dt, width = f32(dt), f32(width)
    return np.sqrt(dt * dt + width * width + t_dm * t_dm + t_ddm * t_ddm, dtype=f32)


def scrunch_list(dms, dt, pulse_width_us, f0, df, nchans, tol):
    dms = np.asarray(dms, f32)
    s = [1]
    w = f32(pulse_width_us) * f32(1e-6)
    for d in range(1, len(dms)):
        ddm = dms[d] - dms[d - 1]
        cur = f32(s[-1]) * f32(dt)
        a = smearing(cur, w, f0, nchans, df, dms[d], ddm)
        b = smearing(f32(2) * cur, w, f0, nchans, df, dms[d], ddm)
        s.append(2 * s[-1] if b / a < f32(tol) else s[-1])
    return np.array(s, np.int64)


def unpack(fb, nsamps, nbits, nchans):
    """[nsamps, nchans] samples; channel c of a byte at bit (c mod 8/nbits)*nbits."""
    fb = np.asarray(fb, np.uint8)
    if nbits == 8:
        return fb[:nsamps * nchans].reshape(nsamps, nchans).copy()
    per = 8 // nbits
    b = fb[:nsamps * nchans // per].reshape(nsamps, nchans // per)
    out = np.zeros((nsamps, nchans), np.uint8)
    for k in range(per):
        out[:, k::per] = (b >> (k * nbits)) & ((1 << nbits) - 1)
    return out


def scrunch2(x):
    """Halve the time resolution of [nsamps, nchans] samples."""
    n = x.shape[0] // 2
    a = x[0:2 * n:2].astype(np.uint16)
    b = x[1:2 * n:2].astype(np.uint16)
    return ((a + b) >> 1).astype(np.uint8)


def dedisperse_adaptive(fb, nsamps, nbits, nchans, dms, kc, max_delay, factors, killmask=None,
                        nthreads=1):
    """Rows of nsamps - max_delay floats; trial d's first (nsamps - max_delay)
    // s_d are its sums, the rest zero."""
    dms = np.asarray(dms, f32)
    kc = np.asarray(kc, f32)
    factors = np.asarray(factors, np.int64)
    nout = nsamps - max_delay
    out = np.zeros((len(dms), nout), f32)
    x = unpack(fb, nsamps, nbits, nchans)
    level = 1
    a = 0
    pad = 4
    while a < len(dms):
        s = int(factors[a])
        b = a
        while b < len(dms) and factors[b] == s:
            b += 1
        while level < s:
            x = scrunch2(x)
            level *= 2
        ns = x.shape[0]
        nout_s = nout // s
        xs = np.concatenate([x, np.zeros((pad, nchans), np.uint8)])
        # kc / s is exact in float32 (s a power of two)
        rows = O.dedisperse(xs.reshape(-1), ns + pad, 8, nchans, dms[a:b], kc / f32(s),
                            ns + pad - nout_s, killmask=killmask, nthreads=nthreads)
        out[a:b, :nout_s] = rows
        a = b
    return out
