"""End-to-end parity of hd_execute on the GPU against the CPU oracle.

Same seeded filterbank into both; the bar is bit-exact: giants (peak S/N,
sample, begin, end, filter, DM index) in the reference's order, candidate
groups, and the .cand text.
"""
import numpy as np
import pytest

from tests._cases import SMALL_128, TINY, oracle_params_for, params_for

pytestmark = pytest.mark.gpu


def _run_both(O, w, seed, nbeams=1, nbits=8, first_idx=0, gpu_kw=None, orc_kw=None,
              fb=None, nsamps=None, cleaning=False):
    import heimdall_multibeam_amd as H
    from heimdall_multibeam_amd import synth
    gpu_kw = gpu_kw or {}
    orc_kw = orc_kw or {}
    if fb is None:
        beams = [synth.make_gulp(w, seed + b) for b in range(nbeams)]
        fb = np.concatenate(beams)
        nsamps = w.nsamps * nbeams
    print("case", w.name, "seed", seed, "nbeams", nbeams, "nbits", nbits, "cleaning", cleaning,
          flush=True)
    pl = H.Pipeline(params_for(w, **gpu_kw))
    pl.set_output(False)
    pl.set_cleaning(cleaning)
    allow = (H.HD_NO_ERROR, H.HD_TOO_MANY_EVENTS)
    err, processed = pl.execute(fb, nsamps, nbits, first_idx, nbeams, allow=allow)
    g = pl.giants()
    c = pl.candidates()
    text = pl.cand_text()
    dms = pl.dm_list()
    md = pl.max_delay()
    pl.close()
    p = oracle_params_for(O, w, cleaning=1 if cleaning else 0, **orc_kw)
    r = O.execute(p, fb, nsamps, nbits, first_idx, nbeams)
    return dict(err=err, processed=processed, giants=g, cands=c, text=text, dms=dms,
                max_delay=md), r


def _compare(gpu, r, O, w):
    assert gpu["max_delay"] == r.max_delay
    assert len(gpu["dms"]) == r.n_dm
    assert gpu["processed"] == r.nsamps_processed
    og = r.giants
    gg = gpu["giants"]
    print("giants gpu", len(gg["peak"]), "oracle", len(og), "groups", len(r.groups), flush=True)
    assert len(gg["peak"]) == len(og)
    for f in ("ind", "begin", "end", "filter", "dm"):
        np.testing.assert_array_equal(gg[f], og[f].astype(np.uint64), err_msg=f)
    np.testing.assert_array_equal(gg["peak"].view(np.uint32), og["peak"].view(np.uint32))
    gc = gpu["cands"]
    assert len(gc["peak"]) == len(r.groups)
    for f in ("ind", "begin", "end", "filter", "dm", "members"):
        np.testing.assert_array_equal(gc[f], r.groups[f].astype(np.uint64), err_msg=f)
    np.testing.assert_array_equal(gc["peak"].view(np.uint32), r.groups["peak"].view(np.uint32))
    np.testing.assert_array_equal(gc["dms"], r.group_dms)
    assert gpu["text"] == r.text


@pytest.mark.parametrize("seed", [1, 2])
def test_tiny_single_beam(oracle, seed):
    gpu, r = _run_both(oracle, TINY, seed)
    assert len(r.giants) > 0
    _compare(gpu, r, oracle, TINY)


def test_small_with_rfi(oracle):
    gpu, r = _run_both(oracle, SMALL_128, 3)
    assert len(r.groups) > 0
    _compare(gpu, r, oracle, SMALL_128)


@pytest.mark.parametrize("nbeams", [2, 3])
def test_multibeam_stf(oracle, nbeams):
    gpu, r = _run_both(oracle, TINY, 10, nbeams=nbeams)
    _compare(gpu, r, oracle, TINY)


def test_killmask_zaps(oracle):
    zaps = [(3, 9), (40, 41)]
    km = np.ones(TINY.nchans, np.int32)
    for a, b in zaps:
        km[a:b] = 0
    gpu, r = _run_both(oracle, TINY, 5, gpu_kw=dict(channel_zaps=zaps), orc_kw=dict(killmask=km))
    _compare(gpu, r, oracle, TINY)


@pytest.mark.parametrize("nbits", [2, 4])
def test_subbyte_input(oracle, nbits):
    import dataclasses
    from heimdall_multibeam_amd import synth
    mean, sigma = (1.5, 0.9) if nbits == 2 else (7.5, 2.5)
    w = dataclasses.replace(TINY, nbits=nbits, mean=mean, sigma=sigma, name="tiny_%dbit" % nbits)
    gpu, r = _run_both(oracle, w, 7, nbits=nbits)
    _compare(gpu, r, oracle, w)


def test_decimated_widths(oracle):
    # min_tscrunch_width below boxcar_max: the widest widths are decimated and
    # their block RMS reads the previous width's normalised output
    gpu, r = _run_both(oracle, TINY, 11, gpu_kw=dict(min_tscrunch_width=64),
                       orc_kw=dict(min_tscrunch_width=64))
    _compare(gpu, r, oracle, TINY)


def test_giant_rate_limit(oracle):
    import heimdall_multibeam_amd as H
    gpu, r = _run_both(oracle, TINY, 1, gpu_kw=dict(max_giant_rate=1.0),
                       orc_kw=dict(max_giant_rate=1.0))
    assert r.error == H.HD_TOO_MANY_EVENTS
    assert gpu["err"] == H.HD_TOO_MANY_EVENTS
    assert gpu["text"] == r.text == ""


def test_first_idx_and_beam_offset(oracle):
    gpu, r = _run_both(oracle, TINY, 4, first_idx=123456, gpu_kw=dict(beam=2),
                       orc_kw=dict(beam=2))
    _compare(gpu, r, oracle, TINY)


@pytest.mark.parametrize("db", [8, 16, 32])
def test_every_dedisp_tile_gives_the_same_search(oracle, db):
    """The autotuner's candidates (tile height, block order, channels staged
    per step) all compute the same exact sums: fixed to each, the search
    matches the oracle bit for bit."""
    import heimdall_multibeam_amd as H
    from heimdall_multibeam_amd import synth
    w = SMALL_128
    fb = synth.make_gulp(w, 31)
    p = oracle_params_for(oracle, w)
    r = oracle.execute(p, fb, w.nsamps, w.nbits, 0, 1)
    for remap, cc in ((0, 4), (1, 4), (1, 8), (1, 16)):
        print("case tile", db, "remap", remap, "channels per step", cc, flush=True)
        pl = H.Pipeline(params_for(w))
        pl.set_output(False)
        pl.set_dedisp_tile(db, remap, 0, cc)
        err, processed = pl.execute(fb, w.nsamps, w.nbits)
        assert pl.dedisp_tile() == (db, remap, 0, cc, 512)
        gpu = dict(err=err, processed=processed, giants=pl.giants(), cands=pl.candidates(),
                   text=pl.cand_text(), dms=pl.dm_list(), max_delay=pl.max_delay())
        pl.close()
        _compare(gpu, r, oracle, w)


@pytest.mark.parametrize("nbits,level", [(1, "mid"), (2, "mid"), (4, "mid"), (8, "mid"),
                                         (8, "saturated")])
def test_byte_lane_dedispersion_matches_oracle(oracle, nbits, level):
    """The byte-staged kernel (8-bit lanes for sub-byte input, even/odd
    16-bit-lane sums for 8-bit input), fixed for every tile height, matches
    the oracle bit for bit (and so the 16-bit kernel).  "saturated" puts most
    8-bit samples at 255, the largest sums the lanes must hold."""
    import dataclasses
    import heimdall_multibeam_amd as H
    from heimdall_multibeam_amd import synth
    vmax = (1 << nbits) - 1
    mean, sigma = vmax / 2.0, max(vmax / 3.0, 0.5)
    if level == "saturated":
        mean, sigma = 250.0, 12.0
    w = dataclasses.replace(SMALL_128, nbits=nbits, mean=mean, sigma=sigma)
    fb = synth.make_gulp(w, 40 + nbits)
    p = oracle_params_for(oracle, w)
    r = oracle.execute(p, fb, w.nsamps, nbits, 0, 1)
    for db, cc, threads in ((8, 4, 512), (16, 4, 512), (32, 4, 512), (16, 8, 512), (32, 8, 512),
                            (32, 16, 512), (32, 8, 1024), (32, 16, 1024)):
        print("case byte lanes", nbits, "bits, tile", db, "channels per step", cc, "threads",
              threads, flush=True)
        pl = H.Pipeline(params_for(w))
        pl.set_output(False)
        pl.set_dedisp_tile(db, 1, 1, cc, threads)
        err, processed = pl.execute(fb, w.nsamps, nbits)
        assert pl.dedisp_tile() == (db, 1, 1, cc, threads)
        gpu = dict(err=err, processed=processed, giants=pl.giants(), cands=pl.candidates(),
                   text=pl.cand_text(), dms=pl.dm_list(), max_delay=pl.max_delay())
        pl.close()
        _compare(gpu, r, oracle, w)

