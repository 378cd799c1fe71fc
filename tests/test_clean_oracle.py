"""Pins of the oracle's zero-DM RFI cleaning (clean_filterbank_rfi.cu) on the CPU.

The reference ships no vectors for this path.  What fixes it instead:
  * the random engine: minstd_rand's published known answer;
  * the uniform integer draw: exact rational arithmetic of the mapping the
    rocThrust headers define (a double-precision slip or an off-by-one in
    the interval end shows up as a mismatch or an out-of-range draw);
  * the mask: a float64 brute force over the normalised series, windows and
    offsets written from the reference's boxcar definitions;
  * the zap: structural properties that any correct replacement obeys, plus
    the exact first draw wherever no retry can happen.
The hash constants are taken as written at clean_filterbank_rfi.cu:33-43 and
checked only for being a bijection (parity of the hash itself is unpinned).
"""
import math
from fractions import Fraction

import numpy as np
import pytest

from tests._cases import SMALL_128

f32 = np.float32
M = 2147483647


def test_minstd_known_answer(oracle):
    """C++11 [rand.predef]: the 10000th output of minstd_rand seeded with 1 is 399268537."""
    assert oracle.minstd_seed(1) == 1
    assert oracle.minstd_outputs(1, 10000)[-1] == 399268537


def test_minstd_seed_rule(oracle):
    assert oracle.minstd_seed(0) == 1            # 0 mod m -> 1
    assert oracle.minstd_seed(M) == 1
    assert oracle.minstd_seed(M + 5) == 5
    assert oracle.minstd_seed(0xFFFFFFFF) == 0xFFFFFFFF % M


def test_uniform_draw_exact_mapping(oracle):
    """draw = floor(lo + (u - 1) * (hi - lo + 1) / (m - 1)) for the engine output u."""
    rng = np.random.default_rng(3)
    for lo, hi in [(0, 0), (0, 1), (5, 9), (0, 7812), (250000, 262143), (0, 2**31)]:
        for _ in range(300):
            state = int(rng.integers(1, M))
            (u,) = oracle.minstd_outputs(state, 1)
            (d,), _ = oracle.uniform_draws(state, lo, hi, 1)
            exact = lo + Fraction(u - 1, M - 1) * (hi - lo + 1)
            assert lo <= d <= hi
            assert d == math.floor(exact), (lo, hi, u, d, exact)


def test_hash_is_a_bijection_sample(oracle):
    xs = np.arange(0, 1 << 16, dtype=np.uint64)
    hs = {oracle.hash32(int(x)) for x in xs}
    assert len(hs) == len(xs)
    # and it mixes: consecutive inputs differ in about half their bits
    flips = [bin(oracle.hash32(i) ^ oracle.hash32(i + 1)).count("1") for i in range(2000)]
    assert 12 < np.mean(flips) < 20


def test_norm_is_independent_of_sqrt_precision():
    """1.0/sqrt((float)w): float or double sqrt give the same float for w = 2^k."""
    for k in range(13):
        w = 2 ** k
        a = f32(1.0 / float(np.sqrt(f32(w))))
        b = f32(1.0 / math.sqrt(w))
        assert a == b


def _gulp(seed=1):
    from heimdall_multibeam_amd import synth
    return SMALL_128, synth.make_gulp(SMALL_128, seed)


def _normalised_zero_dm(oracle, fb, w, nbeams):
    kc = oracle.delay_table(w.nchans, f32(w.dt), f32(w.f0), f32(w.df))
    x = oracle.dedisperse(fb, w.nsamps, w.nbits, w.nchans, np.zeros(1, np.float32), kc, 0)[0]
    stride = w.nsamps // nbeams
    smooth = int(f32(2.0) / (2 * f32(w.dt)))
    d = oracle.remove_baseline(x, stride, stride, smooth, nbeams)
    rms = oracle.get_rms_multibeam(d, stride, stride, nbeams)
    beam = np.arange(len(d)) // stride
    return (d / rms[beam]).astype(np.float32), smooth


@pytest.mark.parametrize("nbeams", [1, 2])
def test_mask_matches_float64_brute_force(oracle, nbeams):
    w, fb = _gulp()
    B, tol = 64, f32(5.0)
    z, smooth = _normalised_zero_dm(oracle, fb, w, nbeams)
    rc, _, mask = oracle.clean_filterbank_rfi(fb, w.nsamps, w.nbits, w.nchans, 0.0, w.dt, w.f0,
                                              w.df, 2.0, tol, B, nbeams)
    assert rc == 0
    n = w.nsamps
    zd = z.astype(np.float64)
    S = np.concatenate([[0.0], np.cumsum(zd)])
    ref = np.abs(zd) > tol
    near = np.abs(np.abs(zd) - tol) < 1e-3
    h, L = B // 2, n + 1 - B
    width = 1
    while width <= B:
        ahead, behind = (width - 1) // 2 + 1, width // 2
        j = np.arange(h, h + L)
        f = (S[j + ahead] - S[j - behind]) / math.sqrt(width)
        cut = np.zeros(n, bool)
        cut[j] = np.abs(f) > tol
        close = np.zeros(n, bool)
        close[j] = np.abs(np.abs(f) - tol) < 1e-3
        # widened by a boxcar of the same width over [i - behind, i + ahead)
        c = np.concatenate([[0], np.cumsum(cut)])
        cn = np.concatenate([[0], np.cumsum(close)])
        i = np.arange(h, h + L)
        wide = np.zeros(n, bool)
        wide[i] = (c[i + ahead] - c[i - behind]) > 0
        wnear = np.zeros(n, bool)
        wnear[i] = (cn[i + ahead] - cn[i - behind]) > 0
        ref |= wide
        near |= wnear
        width *= 2
    got = mask.astype(bool)
    bad = (got != ref) & ~near
    assert not bad.any(), np.nonzero(bad)[0][:10]
    assert 0 < ref.sum() < n      # the tones and the burst of this gulp trip the cut


def test_zap_properties(oracle):
    rng = np.random.default_rng(7)
    nchans, nbits, n, maxd = 64, 8, 4096, 100
    fb = rng.integers(0, 256, (n, nchans), dtype=np.uint8)
    mask = np.zeros(n, np.int32)
    mask[rng.choice(n, 60, replace=False)] = 1
    mask[2000:2040] = 1
    rc, out = oracle.zap_filterbank_rfi(mask, fb, n, nbits, nchans, maxd)
    assert rc == 0
    out = out.reshape(n, nchans)
    keep = mask == 0
    np.testing.assert_array_equal(out[keep], fb[keep])
    stride = nchans // 4
    for t in np.nonzero(mask)[0]:
        lo, hi = max(0, t - maxd), min(n - 1, t + maxd)
        for c in range(nchans):
            assert out[t, c] in fb[lo:hi + 1, c]
        # with no other masked sample in range, every first draw is accepted
        if mask[lo:hi + 1].sum() == 1:
            for wd in range(stride):
                i = int(t) * stride + wd
                x = oracle.minstd_seed(oracle.hash32(i))
                draws, _ = oracle.uniform_draws(x, lo, hi, 4)
                # one draw per channel of the word, low byte first; a masked
                # draw (t itself) is retried, so compare only accepted ones
                if t not in draws:
                    for k in range(4):
                        assert out[t, 4 * wd + k] == fb[draws[k], 4 * wd + k]


def test_zap_rejects_narrow_data(oracle):
    fb = np.zeros((10, 4), np.uint8)
    rc, _ = oracle.zap_filterbank_rfi(np.zeros(10, np.int32), fb, 10, 8, 4, 20)
    assert rc != 0


def test_cleaning_removes_zero_dm_burst(oracle):
    """A broadband burst is flagged and replaced by ordinary noise samples."""
    w, fb = _gulp(seed=2)
    fb = fb.copy().reshape(w.nsamps, -1)
    fb[10000:10010, :] = 250
    rc, out, mask = oracle.clean_filterbank_rfi(fb.ravel(), w.nsamps, w.nbits, w.nchans, 0.0,
                                                w.dt, w.f0, w.df, 2.0, 5.0, 64, 1)
    assert rc == 0
    assert mask[10000:10010].all()
    out = out.reshape(w.nsamps, -1)
    assert out[10000:10010].mean() < 160
    # a clean execute differs from a dirty one only through the cleaned bytes
    p = oracle.make_params(nchans=w.nchans, dt=w.dt, f0=w.f0, df=w.df, dm_min=w.dm_min,
                           dm_max=w.dm_max, boxcar_max=w.boxcar_max, nthreads=8, cleaning=1)
    r1 = oracle.execute(p, fb.ravel(), w.nsamps, w.nbits, 0, 1, want_text=False)
    p0 = oracle.make_params(nchans=w.nchans, dt=w.dt, f0=w.f0, df=w.df, dm_min=w.dm_min,
                            dm_max=w.dm_max, boxcar_max=w.boxcar_max, nthreads=8)
    r2 = oracle.execute(p0, out.ravel(), w.nsamps, w.nbits, 0, 1, want_text=False)
    assert r1.error == r2.error == 0
    np.testing.assert_array_equal(r1.groups, r2.groups)
