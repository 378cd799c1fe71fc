"""The oracle's restatement of adaptive time resolution (oracle/adaptive_dt.py),
pinned by what the mathematics fixes: the doubling rule's invariants and its
criterion evaluated independently in float64, brute-force sums over the
halved input, and the special cases that reduce to orc_dedisperse.  dedisp
itself is absent, so the rule's constants stay parity unpinned."""
import numpy as np
import pytest

from heimdall_multibeam_amd import synth
from oracle import adaptive_dt as A
from oracle import oracle as O

f32 = np.float32


def _delays(dms, kc):
    return (dms[:, None].astype(f32) * kc[None, :].astype(f32) + f32(0.5)).astype(np.int64)


@pytest.mark.parametrize("name", ["full_1024ch", "gated_4096ch", "small_512ch"])
def test_scrunch_list_invariants_and_float64_criterion(name):
    w = synth.CONFIGS[name]
    dms = synth.trial_dms(w).astype(np.float64)
    s = A.scrunch_list(dms, w.dt, 40.0, w.f0, w.df, w.nchans, 1.15)
    assert s[0] == 1 and np.all(np.diff(s) >= 0)
    assert np.all(s & (s - 1) == 0)                       # powers of two
    assert np.all((s[1:] == s[:-1]) | (s[1:] == 2 * s[:-1]))
    # the criterion in float64, written from the formula's terms directly
    bw = w.nchans * abs(w.df)
    fc = w.f0 - bw / 2
    W = 40e-6
    for d in range(1, len(dms)):
        ddm = dms[d] - dms[d - 1]
        tdm = 8.3 * bw * dms[d] / fc ** 3
        tddm = 8.3 / 4 * bw * w.nchans * ddm / fc ** 3
        base = W * W + tdm * tdm + tddm * tddm
        dt0 = s[d - 1] * w.dt
        ratio = np.sqrt((2 * dt0) ** 2 + base) / np.sqrt(dt0 ** 2 + base)
        if abs(ratio - 1.15) > 1e-5:
            assert (s[d] == 2 * s[d - 1]) == (ratio < 1.15), (name, d)


def test_scrunch_list_limits():
    w = synth.CONFIGS["full_1024ch"]
    dms = synth.trial_dms(w)[:20]
    assert np.all(A.scrunch_list(dms, w.dt, 40.0, w.f0, w.df, w.nchans, 1.0) == 1)
    assert np.all(A.scrunch_list(dms, w.dt, 40.0, w.f0, w.df, w.nchans, 10.0) ==
                  2 ** np.arange(20))


def test_scrunch2_and_unpack():
    rng = np.random.default_rng(1)
    x = rng.integers(0, 256, (11, 5), dtype=np.uint8)
    y = A.scrunch2(x)
    assert y.shape == (5, 5)
    for t in range(5):
        for c in range(5):
            assert y[t, c] == (int(x[2 * t, c]) + int(x[2 * t + 1, c])) // 2
    # 2-bit packing: channel c at bit 2 * (c % 4) of byte c // 4
    v = rng.integers(0, 4, (7, 8), dtype=np.uint8)
    packed = np.zeros((7, 2), np.uint8)
    for c in range(8):
        packed[:, c // 4] |= v[:, c] << (2 * (c % 4))
    assert np.array_equal(A.unpack(packed.reshape(-1), 7, 2, 8), v)


def _brute_trial(x, dm, kc, s, nout, live):
    """Trial dm at factor s: halve log2(s) times, then sum with delays in
    those samples; samples past the halved input are zero."""
    for _ in range(int(np.log2(s))):
        x = np.array([[(int(x[2 * t, c]) + int(x[2 * t + 1, c])) // 2 for c in range(x.shape[1])]
                      for t in range(x.shape[0] // 2)], np.int64)
    n = nout // s
    out = np.zeros(n, np.int64)
    for c in live:
        d = int(f32(dm) * (f32(kc[c]) / f32(s)) + f32(0.5))
        for t in range(n):
            if t + d < x.shape[0]:
                out[t] += x[t + d, c]
    return out


def test_adaptive_dedispersion_brute_force():
    rng = np.random.default_rng(3)
    nchans, nsamps, dt, f0, df = 16, 400, 1e-3, 400.0, -4.0
    kc = O.delay_table(nchans, f32(dt), f32(f0), f32(df))
    dms = np.array([0.0, 3.0, 7.5, 12.0, 20.0, 31.0], f32)
    factors = np.array([1, 1, 2, 2, 4, 8])
    maxd = int(_delays(dms, kc).max())
    fb = rng.integers(0, 256, nsamps * nchans, dtype=np.uint8)
    km = np.ones(nchans, np.int32)
    km[[3, 9]] = 0
    got = A.dedisperse_adaptive(fb, nsamps, 8, nchans, dms, kc, maxd, factors, killmask=km)
    x = fb.reshape(nsamps, nchans)
    live = [c for c in range(nchans) if km[c]]
    nout = nsamps - maxd
    for i, (dm, s) in enumerate(zip(dms, factors)):
        want = _brute_trial(x, dm, kc, int(s), nout, live)
        assert np.array_equal(got[i, :nout // s], want.astype(f32)), (i, s)
        assert np.all(got[i, nout // s:] == 0)


def test_adaptive_special_cases():
    rng = np.random.default_rng(4)
    w = synth.CONFIGS["small_512ch"]
    nchans, nsamps = 64, 3000
    kc = O.delay_table(nchans, f32(w.dt), f32(w.f0), f32(w.df * 8))
    dms = np.linspace(0, 400, 9).astype(f32)
    maxd = int(_delays(dms, kc).max())
    fb = rng.integers(0, 256, nsamps * nchans, dtype=np.uint8)
    ones = A.dedisperse_adaptive(fb, nsamps, 8, nchans, dms, kc, maxd, np.ones(9, int))
    assert np.array_equal(ones, O.dedisperse(fb, nsamps, 8, nchans, dms, kc, maxd))
    const = np.full(nsamps * nchans, 77, np.uint8)
    f = np.array([1, 1, 2, 2, 2, 4, 4, 8, 16])
    got = A.dedisperse_adaptive(const, nsamps, 8, nchans, dms, kc, maxd, f)
    nout = nsamps - maxd
    for i, s in enumerate(f):
        assert np.all(got[i, :nout // s] == 77 * nchans), i


# ---- the C restatement (oracle/hd_oracle.c) used by the whole-gulp oracle ----

@pytest.mark.parametrize("name", sorted(synth.CONFIGS))
def test_c_scrunch_factors_equal_numpy(name):
    w = synth.CONFIGS[name]
    dms = synth.trial_dms(w)
    for tol in (1.05, 1.15, 1.5):
        a = A.scrunch_list(dms, w.dt, 40.0, w.f0, w.df, w.nchans, tol)
        b = O.scrunch_factors(dms, w.dt, 40.0, w.f0, w.df, w.nchans, tol)
        assert np.array_equal(a, b), (name, tol)


@pytest.mark.parametrize("nbits", [8, 2])
def test_c_adaptive_dedispersion_equals_numpy(nbits):
    rng = np.random.default_rng(5 + nbits)
    nchans, nsamps, dt, f0, df = 48, 3001, 5e-4, 600.0, -3.0
    kc = O.delay_table(nchans, f32(dt), f32(f0), f32(df))
    dms = np.linspace(0, 60, 14).astype(f32)
    factors = np.array([1, 1, 1, 2, 2, 2, 4, 4, 4, 8, 8, 16, 16, 32])
    maxd = int(_delays(dms, kc).max())
    fb = rng.integers(0, 256, nsamps * nchans * nbits // 8, dtype=np.uint8)
    km = np.ones(nchans, np.int32)
    km[[0, 7, 30]] = 0
    want = A.dedisperse_adaptive(fb, nsamps, nbits, nchans, dms, kc, maxd, factors, killmask=km)
    got = O.dedisperse_adaptive(fb, nsamps, nbits, nchans, dms, kc, maxd, factors, killmask=km,
                                nthreads=4)
    assert np.array_equal(got, want)


def _same_giants(a, b):
    assert len(a) == len(b)
    for k in a.dtype.names:  # field by field (the records have padding bytes)
        assert np.array_equal(a[k], b[k]), k


def _exec(w, seed, **kw):
    from tests._cases import oracle_params_for
    fb = synth.make_gulp(w, seed)
    p = oracle_params_for(O, w, **kw)
    return O.execute(p, fb, w.nsamps, w.nbits, 0, 1)


def test_execute_all_factors_one_is_the_plain_search():
    """tiny_64ch's list never doubles: scrunching on changes nothing."""
    from tests._cases import TINY
    assert np.all(O.scrunch_factors(synth.trial_dms(TINY), TINY.dt, 40.0, TINY.f0, TINY.df,
                                    TINY.nchans, 1.15) == 1)
    a = _exec(TINY, 1, use_scrunching=0)
    b = _exec(TINY, 1, use_scrunching=1)
    assert len(a.giants) > 0
    _same_giants(a.giants, b.giants)
    assert a.text == b.text


@pytest.mark.parametrize("case", ["SCRUNCH_128", "SCRUNCH_256"])
def test_execute_adaptive_trials(case):
    """With scrunching, the trials whose factor stays 1 give exactly the
    plain search's giants (the loop at pipeline.cu:488-790 treats them
    alike); a trial of factor s reports samples of the full resolution
    (multiples of s) and widths from s up (filter >= log2 s); the strongest
    injected pulses are found either way."""
    import tests._cases as C
    w = getattr(C, case)
    dms = synth.trial_dms(w)
    f = O.scrunch_factors(dms, w.dt, 40.0, w.f0, w.df, w.nchans, 1.15)
    assert f.max() >= 8
    seed = {"SCRUNCH_128": 3, "SCRUNCH_256": 5}[case]
    off = _exec(w, seed, use_scrunching=0)
    on = _exec(w, seed, use_scrunching=1)
    one_off = off.giants[f[off.giants["dm"]] == 1]
    one_on = on.giants[f[on.giants["dm"]] == 1]
    assert len(one_on) > 0
    _same_giants(one_on, one_off)
    sc = on.giants[f[on.giants["dm"]] > 1]
    assert len(sc) > 0
    s = f[sc["dm"]].astype(np.uint64)
    for k in ("ind", "begin", "end"):
        assert np.all(sc[k] % s == 0), k
    assert np.all((1 << sc["filter"].astype(np.uint64)) >= s)
    # the brightest candidates of the plain search at least as wide as their
    # trial's factor are detected as giants (clustering may fold them into
    # another group's peak)
    strong = off.groups[(off.groups["peak"] > 12) &
                        ((1 << off.groups["filter"].astype(np.int64)) >= f[off.groups["dm"]])]
    assert len(strong) > 0
    for g in strong:
        near = on.giants[(np.abs(on.giants["ind"].astype(np.int64) - int(g["ind"])) < 64)]
        assert len(near) > 0 and near["peak"].max() > 0.8 * g["peak"], g
