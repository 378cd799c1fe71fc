"""The restatement of Applications/candidate_profiler.C (oracle/
candidate_profiler.py), pinned on the CPU: a hand-worked normalisation, its
scale invariance, and a brute-force word-level binning of the
frequency-time image (candidate_profiler.C:193-220)."""
import numpy as np

from oracle import candidate_profiler as P

f32 = np.float32


def test_normalise_hand_worked():
    # halves [3,1,4,1] and [5,9,2,6]: quarter points 3 and 6, slope 3/4,
    # intercept 1.5; residuals' median |r| is 1.5, so sigma = 1.5 * 1.4826
    x = np.array([[3, 1, 4, 1, 5, 9, 2, 6]], f32)
    got = P.normalise_rows(x.copy())[0]
    r = np.array([1.5, -1.25, 1, -2.75, 0.5, 3.75, -4, -0.75])
    np.testing.assert_allclose(got, r / (1.5 * 1.4826), rtol=1e-6)


def test_normalise_scale_invariant_and_unit_mad():
    rng = np.random.default_rng(2)
    x = rng.normal(10, 3, (4, 256)).astype(f32) + np.arange(256, dtype=f32) * f32(0.05)
    a = P.normalise_rows(x.copy())
    b = P.normalise_rows((x * f32(4)).copy())    # a power of two: exact scaling
    np.testing.assert_allclose(a, b, rtol=0, atol=1e-6)
    for r in a:
        np.testing.assert_allclose(np.sort(np.abs(r))[128], 1 / 1.4826, rtol=1e-5)


def test_freq_time_binning_brute_force():
    rng = np.random.default_rng(5)
    h = dict(nchans=32, nbits=2, tsamp=1e-3, fch1=400.0, foff=-1.0)
    nsamps = 600
    data = rng.integers(0, 256, nsamps * 32 // 4, dtype=np.uint8)
    samp, filt, dm, out_nsamps, fscrunch = 300, 2, 5.0, 16, 4
    got, in_nsamps = P.freq_time(h, data, samp, filt, dm, out_nsamps, fscrunch)
    kc = P._kc(h)
    maxd = P._max_delay([dm], kc)
    tscr = 1 << filt
    if out_nsamps * tscr < maxd:
        d = 2
        while d < maxd:
            d *= 2
        tscr = d // out_nsamps
    assert in_nsamps == out_nsamps * tscr or in_nsamps >= maxd
    first = samp + maxd // 2 - in_nsamps // 2
    words = np.frombuffer(data.tobytes(), np.uint32)
    per_word, wps = 16, 2                      # 2-bit samples: 16 per word, 2 words per sample
    want = np.zeros((8, out_nsamps), f32)
    for t in range(0, in_nsamps, tscr):
        for c in range(0, 32, fscrunch):
            s = f32(0)
            for ts in range(tscr):
                for k in range(fscrunch):
                    w = (c + k) // per_word + (first + t + ts) * wps
                    word = int(words[w]) if 0 <= w < len(words) else 0
                    x = (word >> (((c + k) % per_word) * 2)) & 3
                    s = f32(s + f32(f32(x) / f32(3)) * f32(255))
            if t // tscr < out_nsamps:
                want[c // fscrunch, t // tscr] = f32(s / f32(tscr * fscrunch))
    np.testing.assert_array_equal(got, P.normalise_rows(want))
