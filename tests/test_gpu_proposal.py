"""The learned proposal's coarse plane, tiles, scores and gated search.

  * coarse plane: bit-exact against the oracle's own stages run on the
    16-sample sums of the channel data (dedisperse with the coarse delay
    table, remove_baseline, get_rms, divide) and rounded to bfloat16, and
    its wide plane (float32 sums of 4 normalised samples, halved);
  * tiles: numpy windows of both planes with zero padding;
  * scores: the oracle network on the extracted tiles (logit tolerance as in
    test_gpu_cnn);
  * gated search: with every unit proposed it equals the full search bit for
    bit; with a partial proposal its giants are exactly the full search's
    giants in the proposed units (same values, same order), and the proposed
    units are exactly those with a tile scoring above the threshold.
"""
import numpy as np
import pytest

from tests._cases import SMALL_128, TINY, params_for
from tests.conftest import gpu_available

pytestmark = pytest.mark.gpu

f32 = np.float32


@pytest.fixture(scope="module")
def H():
    if not gpu_available():
        pytest.skip("no GPU")
    import heimdall_multibeam_amd as H
    return H


def _unpack(fb, nsamps, nbits, nchans):
    b = np.asarray(fb, np.uint8).reshape(nsamps, nchans * nbits // 8)
    if nbits == 8:
        return b.astype(np.int64)
    per = 8 // nbits
    out = np.zeros((nsamps, nchans), np.int64)
    for k in range(per):
        out[:, k::per] = (b >> (k * nbits)) & ((1 << nbits) - 1)
    return out


def oracle_coarse(O, w, fb, nbits, dms):
    """The coarse plane from the oracle's stages (rows x samples, bf16 bits)."""
    from oracle import cnn
    x = _unpack(fb, w.nsamps, nbits, w.nchans)
    ns = w.nsamps // 16
    s = x[:16 * ns].reshape(ns, 16, w.nchans).sum(1)
    xc = ((s + 8) >> 4) if nbits == 8 else s
    assert xc.max() <= 255
    n_dm = len(dms)
    cd = np.array([dms[min(8 * j + 4, n_dm - 1)] for j in range((n_dm + 7) // 8)], np.float32)
    dt16 = f32(w.dt) * f32(16)
    kc = O.delay_table(w.nchans, dt16, f32(w.f0), f32(w.df))
    maxd = max(O.delay(float(cd[-1]), float(kc[c])) for c in range(w.nchans))
    plane = O.dedisperse(xc.astype(np.uint8).ravel(), ns, 8, w.nchans, cd, kc, maxd)
    nout = ns - maxd
    smooth = int(f32(2.0) / (f32(2.0) * dt16))
    rows = []
    for r in plane:
        z = O.remove_baseline(r[:nout].copy(), ns, nout, smooth, 1)
        rms = O.get_rms_multibeam(z, ns, nout, 1)
        rows.append((z / rms[0]).astype(np.float32))
    z = np.array(rows)
    nw = nout // 4
    s = np.zeros((z.shape[0], nw), np.float32)
    for i in range(4):                      # float32, in index order
        s = (s + z[:, i:4 * nw:4]).astype(np.float32)
    wide = (s * np.float32(0.5)).astype(np.float32)
    return cnn.bf16_bits(z), nout, cnn.bf16_bits(wide)


def _device(a):
    from heimdall_multibeam_amd.hip import DeviceBuffer
    a = np.ascontiguousarray(a)
    d = DeviceBuffer(a.nbytes)
    d.upload(a)
    return d


@pytest.mark.parametrize("nbits", [8, 2])
def test_coarse_plane_matches_oracle(H, oracle, nbits):
    from heimdall_multibeam_amd import proposal as P, synth
    import dataclasses
    w = dataclasses.replace(SMALL_128, nbits=nbits, mean=1.5 if nbits == 2 else 128.0,
                            sigma=0.8 if nbits == 2 else 16.0)
    fb = synth.make_gulp(w, 3)
    pl = H.Pipeline(params_for(w))
    pl.set_output(False)
    d = _device(fb)
    nu, nv = P.coarse(pl, d.address, w.nsamps, nbits)
    got = P.plane(pl)
    want, nout, wide = oracle_coarse(oracle, w, fb, nbits, pl.dm_list())
    print("coarse plane", got.shape, "units", nu, "tpos", nv)
    assert got.shape == want.shape
    np.testing.assert_array_equal(got, want)
    np.testing.assert_array_equal(P.plane(pl, 1), wide)
    assert nu == (len(pl.dm_list()) + 127) // 128
    assert nv == (nout + 31) // 32
    pl.close()


def _window(plane, wide, u, v):
    t = np.zeros((64, 64, 2), np.uint16)
    R = plane.shape[0]
    for ch, src, c0 in ((0, plane, 32 * v - 16), (1, wide, 8 * v - 28)):
        Cc = src.shape[1]
        for y in range(64):
            r = 16 * u - 24 + y
            if not 0 <= r < R:
                continue
            lo, hi = max(c0, 0), min(c0 + 64, Cc)
            if lo < hi:
                t[y, lo - c0:hi - c0, ch] = src[r, lo:hi]
    return t


def test_tiles_and_scores(H):
    from heimdall_multibeam_amd import proposal as P, synth
    from heimdall_multibeam_amd.hip import DeviceBuffer
    from oracle import cnn
    w = SMALL_128
    fb = synth.make_gulp(w, 4)
    pl = H.Pipeline(params_for(w))
    pl.set_output(False)
    d = _device(fb)
    nu, nv = P.coarse(pl, d.address, w.nsamps, 8)
    plane, wide = P.plane(pl), P.plane(pl, 1)
    rng = np.random.default_rng(5)
    pos = [(0, 0), (nu - 1, nv - 1), (0, nv - 1), (nu - 1, 0)]
    pos += [(int(rng.integers(nu)), int(rng.integers(nv))) for _ in range(28)]
    tiles = DeviceBuffer(len(pos) * 64 * 64 * 2 * 2)
    P.extract(pl, pos, tiles.address)
    got = tiles.download().view(np.uint16).reshape(len(pos), 64, 64, 2)
    for k, (u, v) in enumerate(pos):
        np.testing.assert_array_equal(got[k], _window(plane, wide, u, v), err_msg=str((u, v)))
    # scores of every tile with a random network against the oracle
    prng = np.random.default_rng(6)
    p = np.concatenate([prng.normal(0, np.sqrt(2.0 / fi), n) if not bias else np.zeros(n)
                        for (name, shp) in cnn.SHAPES
                        for n, fi, bias in [(int(np.prod(shp)),
                                             {"w1": 18, "fc1_w": 512, "fc2_w": 64}.get(name, 288),
                                             name.startswith("b") or name.endswith("_b"))]])
    p = p.astype(np.float32)
    P.set_proposal(pl, p, 0.0)
    sc = P.scores(pl, nu, nv)
    allpos = [(u, v) for u in range(nu) for v in range(nv)]
    x = cnn.from_bits(np.stack([_window(plane, wide, u, v) for u, v in allpos]))
    zo, _ = cnn.forward(p, x, quantize=True)
    err = np.abs(sc.reshape(-1) - zo)
    print("scores", sc.shape, "max err", err.max())
    assert np.all(err <= 2e-3 * (1 + np.abs(zo)))
    pl.close()


def _giants(pl):
    g = pl.giants()
    return np.stack([g["peak"].view(np.uint32).astype(np.uint64), g["ind"], g["begin"],
                     g["end"], g["filter"], g["dm"]], 1) if len(g["peak"]) else np.zeros((0, 6))


def _random_params(seed):
    from oracle import cnn
    rng = np.random.default_rng(seed)
    out = []
    for name, shp in cnn.SHAPES:
        n = int(np.prod(shp))
        fi = {"w1": 18, "fc1_w": 512, "fc2_w": 64}.get(name, 288)
        out.append(np.zeros(n) if (name.startswith("b") or name.endswith("_b"))
                   else rng.normal(0, np.sqrt(2.0 / fi), n))
    return np.concatenate(out).astype(np.float32)


def test_gated_search_every_unit_equals_full(H):
    from heimdall_multibeam_amd import proposal as P, synth
    w = SMALL_128
    fb = synth.make_gulp(w, 7)
    pl = H.Pipeline(params_for(w))
    pl.set_output(False)
    pl.execute(fb, w.nsamps, 8)
    full_c, full_g, full_t = pl.candidates(), _giants(pl), pl.cand_text()
    assert len(full_c["peak"]) > 0
    P.set_proposal(pl, _random_params(8), -1e30)   # every tile passes
    pl.execute(fb, w.nsamps, 8)
    flags = P.last_proposal(pl)
    assert flags.shape == (1, (len(pl.dm_list()) + 127) // 128) and flags.all()
    gated_c = pl.candidates()
    for k in full_c:
        np.testing.assert_array_equal(gated_c[k], full_c[k], err_msg=k)
    np.testing.assert_array_equal(_giants(pl), full_g)
    assert pl.cand_text() == full_t
    assert "proposal" in pl.timings()
    P.set_proposal(pl, None, 0)
    pl.close()


@pytest.mark.parametrize("case", ["single", "batch"])
def test_gated_search_partial_is_exact(H, case):
    from heimdall_multibeam_amd import proposal as P, synth
    w = SMALL_128
    params = _random_params(9)
    pl = H.Pipeline(params_for(w, dm_max=1500.0, dm_tol=1.02))   # 4+ units
    pl.set_output(False)
    n_dm = len(pl.dm_list())
    assert n_dm > 3 * 128
    if case == "single":
        gulps = [(synth.make_gulp(w, 10), w.nsamps)]
    else:
        gulps = [(synth.make_gulp(w, 11 + k, nsamps=ns), ns)
                 for k, ns in enumerate([w.nsamps, 3 * w.nsamps // 4 + 160, w.nsamps + 4096])]
    # full search and the unit scores of each gulp
    full, flags_want = [], []
    for fb, ns in gulps:
        pl.execute(fb, ns, 8)
        full.append(_giants(pl))
    # scores -> threshold at the median unit maximum
    from heimdall_multibeam_amd.hip import DeviceBuffer
    maxes = []
    P.set_proposal(pl, params, 0.0)
    for fb, ns in gulps:
        d = _device(fb)
        nu, nv = P.coarse(pl, d.address, ns, 8)
        maxes.append(P.scores(pl, nu, nv).max(1))
    thr = float(np.median(np.concatenate(maxes)))
    flags_want = [m > thr for m in maxes]
    P.set_proposal(pl, params, thr)
    if case == "single":
        pl.execute(gulps[0][0], gulps[0][1], 8)
        gated = [_giants(pl)]
        flags = P.last_proposal(pl)
    else:
        pl.execute_batch([(fb, ns, 1, 0, 0) for fb, ns in gulps], 8)
        g = pl.giants()
        flags = P.last_proposal(pl)
        # giants come out gulp by gulp; split them by matching each gulp's
        # filtered full-search list in order
        allg = _giants(pl)
        gated, off = [], 0
        for k in range(len(gulps)):
            units = full[k][:, 5].astype(np.int64) // 128
            n = int(flags_want[k][units].sum())
            gated.append(allg[off:off + n])
            off += n
        assert off == len(allg)
    assert flags.shape == (len(gulps), (n_dm + 127) // 128)
    for k in range(len(gulps)):
        np.testing.assert_array_equal(flags[k].astype(bool), flags_want[k])
        units = full[k][:, 5].astype(np.int64) // 128
        keep = flags_want[k][units]
        print("gulp", k, "units", flags_want[k].astype(int), "giants", len(full[k]), "->",
              int(keep.sum()))
        np.testing.assert_array_equal(gated[k], full[k][keep])
    assert 0 < sum(int(f.sum()) for f in flags_want) < sum(f.size for f in flags_want)
    P.set_proposal(pl, None, 0)
    pl.close()


def _learned():
    from heimdall_multibeam_amd import synth
    return synth.CONFIGS["learned_1024ch"]


def test_coarse_plane_learned_1024ch_matches_oracle(H, oracle):
    """The proposal's input on one full gulp of the learned-proposal workload
    (1024 channels, 262144 samples of 64 us, DM 0-2000, 5% tone channels,
    zero-DM bursts): the coarse and wide planes, bit for bit."""
    from heimdall_multibeam_amd import proposal as P, synth
    w = _learned()
    fb = synth.make_gulp(w, 5)
    pl = H.Pipeline(params_for(w, min_tscrunch_width=w.min_tscrunch_width))
    pl.set_output(False)
    d = _device(fb)
    print("case coarse plane learned_1024ch seed 5", flush=True)
    nu, nv = P.coarse(pl, d.address, w.nsamps, 8)
    got = P.plane(pl)
    want, nout, wide = oracle_coarse(oracle, w, fb, 8, pl.dm_list())
    print("coarse plane", got.shape, "units", nu, "tpos", nv, flush=True)
    assert got.shape == want.shape
    np.testing.assert_array_equal(got, want)
    np.testing.assert_array_equal(P.plane(pl, 1), wide)
    assert nu == (len(pl.dm_list()) + 127) // 128
    assert nv == (nout + 31) // 32
    pl.close()


def test_gated_search_learned_1024ch_every_unit_equals_full(H, oracle):
    """On one full gulp of the learned-proposal workload: the full search
    (the proposal's labeller and recall arbiter) against the oracle, then the
    gated search with every unit proposed against the full search, both bit
    for bit."""
    import os
    from heimdall_multibeam_amd import proposal as P, synth
    from tests._cases import oracle_params_for
    from tests.test_gpu_parity import _compare
    w = _learned()
    fb = synth.make_gulp(w, 6)
    pl = H.Pipeline(params_for(w, min_tscrunch_width=w.min_tscrunch_width))
    pl.set_output(False)
    print("case full search learned_1024ch seed 6", flush=True)
    err, processed = pl.execute(fb, w.nsamps, 8)
    full = dict(err=err, processed=processed, giants=pl.giants(), cands=pl.candidates(),
                text=pl.cand_text(), dms=pl.dm_list(), max_delay=pl.max_delay())
    full_g = _giants(pl)
    r = oracle.execute(oracle_params_for(oracle, w, nthreads=min(16, os.cpu_count() or 1),
                                         min_tscrunch_width=w.min_tscrunch_width),
                       fb, w.nsamps, w.nbits, 0, 1)
    assert len(r.groups) > 5
    _compare(full, r, oracle, w)
    print("case gated search learned_1024ch every unit", flush=True)
    P.set_proposal(pl, _random_params(8), -1e30)
    pl.execute(fb, w.nsamps, 8)
    flags = P.last_proposal(pl)
    assert flags.shape == (1, (len(pl.dm_list()) + 127) // 128) and flags.all()
    gated_c = pl.candidates()
    for k in full["cands"]:
        np.testing.assert_array_equal(gated_c[k], full["cands"][k], err_msg=k)
    np.testing.assert_array_equal(_giants(pl), full_g)
    assert pl.cand_text() == full["text"]
    P.set_proposal(pl, None, 0)
    pl.close()
