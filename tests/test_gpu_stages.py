"""Stage-level parity through the stage C ABI (include/hd/stages.h), bit for
bit against the oracle's stage functions on the same inputs.

  * DM plane: hd_dedisperse (dedisp_execute_adv, Pipeline/pipeline.cu:452-462)
    over every trial of the 1024-channel gulp (1399 trials x 225604 samples of
    exact integer sums) against orc_dedisperse; and at every dedispersion tile
    on a small gulp of 8-bit and 2-bit samples.
  * The per-DM chain on oracle plane rows of that gulp (trials 0, 1, the last
    and those of injected pulses), every stage through its own entry point:
    remove_baseline (remove_baseline.cu:21-160), the per-beam RMS and
    normalise (get_rms.cu:66-100, 202-212), the scan and every boxcar width
    (matched_filter.cu:21-67), the block RMS over the filter buffer that keeps
    the previous width's values (get_rms.cu:104-131, pipeline.cu:692-712) and
    its normalise, and GiantFinder::exec (find_giants.cu:101-316).
  * Clustering and merging (label_candidate_clusters.cu:203-394,
    merge_candidates.cu:74-134) on the oracle's giants of the 512-channel gulp
    and on random giant lists large enough to take the multi-launch
    relaxation.
The oracle runs with up to 16 threads."""
import os

import numpy as np
import pytest

from tests._cases import SMALL_128, params_for
from tests.conftest import gpu_available

pytestmark = pytest.mark.gpu

f32 = np.float32


def _threads():
    return min(16, os.cpu_count() or 1)


@pytest.fixture(scope="module")
def H():
    if not gpu_available():
        pytest.skip("no GPU")
    import heimdall_multibeam_amd as H
    return H


def _device(a):
    from heimdall_multibeam_amd.hip import DeviceBuffer
    a = np.ascontiguousarray(a)
    d = DeviceBuffer(max(a.nbytes, 4))
    d.upload(a)
    return d


def _rows(buf, n, pitch, count, dtype=f32):
    """n rows of `count` values, `pitch` values apart, from a device buffer."""
    item = np.dtype(dtype).itemsize
    a = buf.download(n * pitch * item).view(dtype).reshape(n, pitch)
    return a[:, :count]


def _bits_equal(got, want, what):
    got = np.ascontiguousarray(got, f32)
    want = np.ascontiguousarray(want, f32)
    assert got.shape == want.shape, (what, got.shape, want.shape)
    bad = np.nonzero(got.view(np.uint32) != want.view(np.uint32))
    if len(bad[0]):
        at = tuple(int(b[0]) for b in bad)
        raise AssertionError("%s: %d values differ, first at %s: %r vs %r"
                             % (what, len(bad[0]), at, got[at], want[at]))


def _plan(oracle, w, dms):
    kc = oracle.delay_table(w.nchans, f32(w.dt), f32(w.f0), f32(w.df))
    maxd = max(max(oracle.delay(float(dms[-1]), float(kc[c])), 0) for c in range(w.nchans))
    return kc, maxd


@pytest.fixture(scope="module")
def full(H, oracle):
    """The 1024-channel gulp (seed 1) and the oracle's whole DM plane."""
    from heimdall_multibeam_amd import synth
    w = synth.CONFIGS["full_1024ch"]
    fb, truth = synth.make_gulp(w, 1, return_truth=True)
    dms = synth.trial_dms(w)
    kc, maxd = _plan(oracle, w, dms)
    print("oracle plane: %d trials x %d samples" % (len(dms), w.nsamps - maxd), flush=True)
    plane = oracle.dedisperse(fb, w.nsamps, w.nbits, w.nchans, dms, kc, maxd,
                              nthreads=_threads())
    return dict(w=w, fb=fb, truth=truth, dms=dms, maxd=maxd, plane=plane)


def test_dm_plane_1024ch_matches_oracle(H, full):
    from heimdall_multibeam_amd import stages
    w, plane = full["w"], full["plane"]
    pl = H.Pipeline(params_for(w))
    assert pl.max_delay() == full["maxd"]
    np.testing.assert_array_equal(pl.dm_list(), full["dms"])
    n_dm, nout = plane.shape
    d_in = _device(full["fb"])
    from heimdall_multibeam_amd.hip import DeviceBuffer
    d_out = DeviceBuffer(n_dm * nout * 4)
    print("case dm plane full_1024ch: hd_dedisperse", flush=True)
    stages.dedisperse(pl, w.nsamps, d_in.address, w.nbits, d_out.address, nout * 4)
    print("tile", pl.dedisp_tile(), flush=True)
    got = d_out.download().view(f32).reshape(n_dm, nout)
    d_out.free()
    _bits_equal(got, plane, "DM plane")
    pl.close()


def _tiles():
    out = []
    for db in (8, 16, 32):
        for cc in (4, 8, 16):
            for lanes in (0, 1):
                for remap in (0, 1):
                    out.append((db, remap, lanes, cc, 512))
                if lanes and db == 32:
                    out.append((db, 1, lanes, cc, 1024))
    return out


@pytest.mark.parametrize("nbits", [8, 2])
def test_dm_plane_every_tile_matches_oracle(H, oracle, nbits):
    import dataclasses
    from heimdall_multibeam_amd import stages, synth
    from heimdall_multibeam_amd.hip import DeviceBuffer
    w = dataclasses.replace(SMALL_128, nbits=nbits, mean=1.5 if nbits == 2 else 128.0,
                            sigma=0.8 if nbits == 2 else 16.0)
    fb = synth.make_gulp(w, 21)
    pl = H.Pipeline(params_for(w, channel_zaps=[(10, 13)]))   # with a killmask
    dms = pl.dm_list()
    kc, maxd = _plan(oracle, w, dms)
    assert pl.max_delay() == maxd
    km = np.ones(w.nchans, np.int32)
    km[10:13] = 0
    want = oracle.dedisperse(fb, w.nsamps, nbits, w.nchans, dms, kc, maxd, killmask=km,
                             nthreads=_threads())
    n_dm, nout = want.shape
    pitch = nout + 37                         # a row stride wider than the row
    d_in = _device(fb)
    d_out = DeviceBuffer(n_dm * pitch * 4)
    for tile in _tiles():
        print("case dm plane small_128ch nbits %d tile %s" % (nbits, tile), flush=True)
        pl.set_dedisp_tile(*tile)
        stages.dedisperse(pl, w.nsamps, d_in.address, nbits, d_out.address, pitch * 4)
        _bits_equal(_rows(d_out, n_dm, pitch, nout), want, "DM plane at tile %s" % (tile,))
    pl.close()


def _chain_rows(full, oracle):
    """Trials 0 (the zero-DM bursts), 1 and the last one, and the trial of
    each of the 8 strongest injected pulses with its neighbour."""
    dms = full["dms"]
    rows = {0, 1, len(dms) - 1}
    pulses = sorted(full["truth"]["pulses"], key=lambda p: -p[3])[:8]
    for (t0, dm, width, snr) in pulses:
        m = int(np.argmin(np.abs(dms - dm)))
        rows.update({m, min(m + 1, len(dms) - 1)})
    return sorted(rows)


def test_series_stages_1024ch_match_oracle(H, oracle, full):
    """Every per-DM stage of the search, each through its stage entry point, on
    oracle plane rows; each stage's input is the oracle's previous output."""
    from heimdall_multibeam_amd import stages
    from heimdall_multibeam_amd.hip import DeviceBuffer
    w = full["w"]
    O = oracle
    plane = full["plane"]
    rows = _chain_rows(full, oracle)
    n, nout = len(rows), plane.shape[1]
    stride = w.nsamps                       # one beam: the beam stride is the gulp
    bc = stride - full["maxd"]              # cur_beam_nsamps (pipeline.cu:499)
    assert bc == nout
    W = w.boxcar_max
    smooth = int(f32(2.0) / (f32(2.0) * f32(w.dt)))   # nsamps_smooth (pipeline.cu:572-573)
    assert smooth == 15625
    pitch = nout + 64                       # d[nout] of each series is zero
    print("case series stages full_1024ch rows %s" % rows, flush=True)
    host = np.zeros((n, pitch), f32)
    host[:, :nout] = plane[rows]
    d = _device(host)
    # baseline
    stages.remove_baseline(d.address, stride, bc, smooth, 1, n, pitch)
    want_z = np.stack([O.remove_baseline(plane[r].copy(), stride, bc, smooth, 1) for r in rows])
    _bits_equal(_rows(d, n, pitch, nout), want_z, "remove_baseline")
    # per-beam RMS and normalise
    d_rms = DeviceBuffer(n * 4)
    stages.get_rms_multibeam(d.address, d_rms.address, stride, bc, 1, n, pitch)
    want_rms = np.stack([O.get_rms_multibeam(z, stride, bc, 1) for z in want_z])
    _bits_equal(d_rms.download().view(f32).reshape(n, 1), want_rms, "per-beam RMS")
    stages.normalise_multibeam(d.address, d_rms.address, stride, nout, 1, n, pitch)
    want_zn = np.stack([O.normalise_multibeam(z, r, stride, nout)
                        for z, r in zip(want_z, want_rms)])
    _bits_equal(_rows(d, n, pitch, nout), want_zn, "normalise_multibeam")
    # prefix sums
    spitch = nout + 1 + 63
    d_S = DeviceBuffer(n * spitch * 4)
    stages.matched_filter_prep(d.address, nout, d_S.address, n, pitch, spitch)
    want_S = np.stack([O.scan(z) for z in want_zn])
    _bits_equal(_rows(d_S, n, spitch, nout + 1), want_S, "matched_filter_prep")
    # widths in the reference's order; the filter buffer keeps what the
    # previous width left past the fresh outputs (pipeline.cu:77, 692-712)
    fpitch = nout + 64
    d_f = DeviceBuffer(n * fpitch * 4)
    d_f.upload(np.zeros((n, fpitch), f32))
    filt = np.zeros((n, fpitch), f32)
    d_brms = DeviceBuffer(n * 4)
    bc692 = stride - (full["maxd"] + W)
    max_filtered = nout + 1 - W
    n_giants = 0
    w_ = 1
    while w_ <= W:
        ts = max(2 * w_ // w.min_tscrunch_width, 1)
        n_w = (max_filtered - 1) // ts + 1
        stages.matched_filter_exec(d_S.address, nout + 1, W, w_, ts, d_f.address, n, spitch,
                                   fpitch)
        for k in range(n):
            filt[k, :n_w] = O.boxcar(want_S[k], W, w_, ts)
        _bits_equal(_rows(d_f, n, fpitch, fpitch), filt, "boxcar width %d" % w_)
        stages.block_rms(d_f.address, d_brms.address, stride, bc692, 1, n, fpitch)
        want_b = np.array([O.block_rms(filt[k], stride, bc692, 1) for k in range(n)], f32)
        _bits_equal(d_brms.download().view(f32), want_b, "block RMS width %d" % w_)
        stages.normalise_block(d_f.address, d_brms.address, n_w, n, fpitch)
        for k in range(n):
            filt[k] = O.normalise_block(filt[k], want_b[k], n_w)
        _bits_equal(_rows(d_f, n, fpitch, fpitch), filt, "normalise_block width %d" % w_)
        merge = 3 * (w_ // ts)                # cand_sep_time * rel_rel_filter_width
        cap = 1 << 16
        g_peak, g_ind = DeviceBuffer(cap * 4), DeviceBuffer(cap * 8)
        g_beg, g_end = DeviceBuffer(cap * 8), DeviceBuffer(cap * 8)
        for k in range(n):
            want_g = O.find_giants(filt[k, :n_w], 6.0, merge, stride, bc692)
            cnt = stages.find_giants(d_f.address + k * fpitch * 4, n_w, 1, 6.0, merge, stride,
                                     bc692, g_peak.address, g_ind.address, g_beg.address,
                                     g_end.address, cap)
            assert cnt == len(want_g), ("giants width %d row %d" % (w_, k), cnt, len(want_g))
            if cnt:
                _bits_equal(g_peak.download(cnt * 4).view(f32), want_g["peak"],
                            "giant peaks width %d row %d" % (w_, k))
                for name, buf in (("ind", g_ind), ("begin", g_beg), ("end", g_end)):
                    np.testing.assert_array_equal(buf.download(cnt * 8).view(np.uint64),
                                                  want_g[name], err_msg="%s width %d row %d"
                                                  % (name, w_, k))
            n_giants += cnt
        w_ *= 2
    print("giants compared:", n_giants, flush=True)
    assert n_giants > 20


def _cluster_case(H, oracle, cands, nsamps_beam, what):
    from heimdall_multibeam_amd import stages
    from heimdall_multibeam_amd.hip import DeviceBuffer
    n = len(cands)
    print("case clustering %s: %d giants" % (what, n), flush=True)
    cols = {f: _device(np.ascontiguousarray(cands[f])) for f in
            ("peak", "ind", "begin", "end", "filter", "dm", "members")}
    rc = stages.RawCandidates(*[cols[f].address for f in
                                ("peak", "ind", "begin", "end", "filter", "dm", "members")])
    d_lab = DeviceBuffer(n * 8)
    count = stages.label_candidate_clusters(n, rc, 3, 3, 200, nsamps_beam, d_lab.address)
    want_lab, want_count = oracle.label_clusters(cands, 3, 3, 200, nsamps_beam)
    np.testing.assert_array_equal(d_lab.download().view(np.uint64), want_lab,
                                  err_msg="labels " + what)
    assert count == want_count
    outs = {f: DeviceBuffer(max(count, 1) * (4 if f == "peak" else 8)) for f in
            ("peak", "ind", "begin", "end", "filter", "dm", "members")}
    rg = stages.RawCandidates(*[outs[f].address for f in
                                ("peak", "ind", "begin", "end", "filter", "dm", "members")])
    stages.merge_candidates(n, d_lab.address, rc, rg)
    want = oracle.merge_candidates(want_lab, cands, want_count)
    _bits_equal(outs["peak"].download(count * 4).view(f32), want["peak"], "merged peaks " + what)
    for f in ("ind", "begin", "end", "filter", "dm", "members"):
        np.testing.assert_array_equal(outs[f].download(count * 8).view(np.uint64), want[f],
                                      err_msg="merged %s %s" % (f, what))
    return count


def test_clustering_on_oracle_giants_512ch(H, oracle):
    from heimdall_multibeam_amd import synth
    from tests._cases import oracle_params_for
    w = synth.CONFIGS["small_512ch"]
    fb = synth.make_gulp(w, 2)
    r = oracle.execute(oracle_params_for(oracle, w, nthreads=_threads()), fb, w.nsamps, w.nbits,
                       0, 1, want_text=False)
    assert len(r.giants) > 100
    count = _cluster_case(H, oracle, r.giants, w.nsamps, "small_512ch giants")
    assert count == len(r.groups)


@pytest.mark.parametrize("n", [3000, 24000])
def test_clustering_on_random_giants(H, oracle, n):
    """Giants in dense clumps (long coincidence chains) over 2 beams; 40000
    giants take the multi-launch relaxation."""
    from oracle.oracle import CAND_DTYPE
    rng = np.random.default_rng(n)
    c = np.zeros(n, CAND_DTYPE)
    centres = rng.integers(16, 2 * 300000 - 256, n // 40 + 1)
    c["ind"] = (centres[rng.integers(0, len(centres), n)] + rng.integers(0, 200, n)).astype(
        np.uint64)
    c["filter"] = rng.integers(0, 13, n).astype(np.uint64)
    c["dm"] = rng.integers(0, 1400, n).astype(np.uint64)
    c["peak"] = rng.uniform(6, 40, n).astype(f32)
    c["peak"][rng.integers(0, n, n // 50)] = 12.5          # ties in S/N
    c["begin"] = c["ind"] - rng.integers(0, 8, n).astype(np.uint64)
    c["end"] = c["ind"] + rng.integers(1, 8, n).astype(np.uint64)
    c["members"] = 1
    count = _cluster_case(H, oracle, c, 300000, "random %d" % n)
    assert 1 < count < n
