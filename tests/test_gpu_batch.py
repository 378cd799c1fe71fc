"""hd_execute_batch on the GPU: every gulp of a ragged batch must give exactly
what a separate hd_execute call gives, i.e. the oracle run on that gulp alone
(giants, candidate groups, .cand text, processed counts and statuses)."""
import dataclasses

import numpy as np
import pytest

from tests._cases import TINY
from tests.test_gpu_parity import _compare

pytestmark = pytest.mark.gpu

CAND = ("ind", "begin", "end", "filter", "dm", "members")


def _params(w, **kw):
    import heimdall_multibeam_amd as H
    d = dict(nchans=w.nchans, dt=w.dt, f0=w.f0, df=w.df, dm_min=w.dm_min, dm_max=w.dm_max,
             boxcar_max=w.boxcar_max, detect_thresh=w.detect_thresh, use_scrunching=False)
    d.update(kw)
    return H.default_params(**d)


def _oracle(O, w, fb, ns, nb, first_idx, beam, **kw):
    p = O.make_params(nchans=w.nchans, dt=w.dt, f0=w.f0, df=w.df, dm_min=w.dm_min,
                      dm_max=w.dm_max, boxcar_max=w.boxcar_max, detect_thresh=w.detect_thresh,
                      nthreads=8, beam=beam, **kw)
    return O.execute(p, fb, ns, w.nbits, first_idx, nb)


def _run_batch(O, w, gulps, device=False, **kw):
    """gulps: list of (fb, nsamps, nbeams, first_idx, beam)."""
    import heimdall_multibeam_amd as H
    print("batch", w.name, "gulps", [(g[1], g[2]) for g in gulps], flush=True)
    pl = H.Pipeline(_params(w, **kw))
    pl.set_output(False)
    allow = (H.HD_NO_ERROR, H.HD_TOO_MANY_EVENTS, H.HD_TOO_FEW_NSAMPS)
    err, processed, errs = pl.execute_batch(gulps, w.nbits, allow=allow)
    cands = pl.candidates()
    gidx = pl.candidate_gulps()
    giants = pl.giants()
    text = pl.cand_text()
    pl.close()
    return err, processed, errs, cands, gidx, giants, text


def _check(O, w, gulps, out, **kw):
    err, processed, errs, cands, gidx, giants, text = out
    texts, ref_giants = [], []
    for k, (fb, ns, nb, fi, beam) in enumerate(gulps):
        r = _oracle(O, w, fb, ns, nb, fi, beam, **kw)
        assert processed[k] == r.nsamps_processed, k
        if r.error != 0:
            assert errs[k] == r.error, k
            continue
        assert errs[k] == 0, k
        sel = gidx == k
        assert sel.sum() == len(r.groups), (k, sel.sum(), len(r.groups))
        for f in CAND:
            np.testing.assert_array_equal(cands[f][sel], r.groups[f].astype(np.uint64),
                                          err_msg="%s gulp %d" % (f, k))
        np.testing.assert_array_equal(cands["peak"][sel].view(np.uint32),
                                      r.groups["peak"].view(np.uint32))
        np.testing.assert_array_equal(cands["dms"][sel], r.group_dms)
        texts.append(r.text)
        ref_giants.append(r.giants)
    assert text == "".join(texts)
    og = np.concatenate(ref_giants) if ref_giants else np.zeros(0)
    assert len(giants["peak"]) == len(og)
    for f in ("ind", "begin", "end", "filter", "dm"):
        np.testing.assert_array_equal(giants[f], og[f].astype(np.uint64), err_msg=f)
    np.testing.assert_array_equal(giants["peak"].view(np.uint32), og["peak"].view(np.uint32))


def _small_batch_workload():
    from heimdall_multibeam_amd import synth
    # the ragged-batch setup at a size the oracle finishes quickly: 320 ch,
    # 655.36 us, DM 0-1000, boxcars 1..256
    return dataclasses.replace(synth.CONFIGS["batch_320ch"], name="batch_small", dm_max=1000.0,
                               boxcar_max=256, pulse_dm=(0.0, 1000.0))


def test_ragged_batch_matches_separate_calls(oracle):
    from heimdall_multibeam_amd import synth
    w = _small_batch_workload()
    beams = synth.make_batch(w, 5, 7, nsamps_range=(3000, 9000))
    gulps = [(fb, ns, 1, 1000 * k, k) for k, (fb, ns) in enumerate(beams)]
    _check(oracle, w, gulps, _run_batch(oracle, w, gulps))


def test_batch_with_short_and_failing_gulps(oracle):
    from heimdall_multibeam_amd import synth
    w = _small_batch_workload()
    beams = synth.make_batch(w, 9, 4, nsamps_range=(4000, 6000))
    fb0, ns0 = beams[0]
    row = w.nchans * w.nbits // 8
    # max delay 621, widest boxcar 256: 600 samples fail (HD_TOO_FEW_NSAMPS),
    # 800 leave no room for the trial loop (no candidates, no error)
    short = fb0[:600 * row]
    skip = fb0[:800 * row]
    gulps = [(beams[1][0], beams[1][1], 1, 0, 3), (short, 600, 1, 0, 4),
             (beams[2][0], beams[2][1], 1, 77, 5), (skip, 800, 1, 0, 6),
             (beams[3][0], beams[3][1], 1, 5, 7)]
    _check(oracle, w, gulps, _run_batch(oracle, w, gulps))


def test_batch_of_stf_gulps(oracle):
    from heimdall_multibeam_amd import synth
    g1 = np.concatenate([synth.make_gulp(TINY, 21), synth.make_gulp(TINY, 22)])
    g2 = np.concatenate([synth.make_gulp(TINY, 23, nsamps=12000),
                         synth.make_gulp(TINY, 24, nsamps=12000)])
    gulps = [(g1, 2 * TINY.nsamps, 2, 0, 0), (g2, 24000, 2, 4096, 2)]
    _check(oracle, TINY, gulps, _run_batch(oracle, TINY, gulps))


def test_batch_on_device_with_cleaning(oracle):
    import heimdall_multibeam_amd as H
    from heimdall_multibeam_amd import hip, synth
    w = _small_batch_workload()
    beams = synth.make_batch(w, 13, 3, nsamps_range=(5000, 8000))
    bufs = []
    gulps_host, gulps_dev = [], []
    for k, (fb, ns) in enumerate(beams):
        b = hip.DeviceBuffer(fb.nbytes)
        b.upload(fb)
        bufs.append(b)
        gulps_host.append((fb, ns, 1, 0, k))
        gulps_dev.append((b.address, ns, 1, 0, k))
    pl = H.Pipeline(_params(w))
    pl.set_output(False)
    pl.set_cleaning(True)
    print("batch on device with cleaning", flush=True)
    err, processed, errs = pl.execute_batch(gulps_dev, w.nbits, device=True)
    out = (err, processed, errs, pl.candidates(), pl.candidate_gulps(), pl.giants(),
           pl.cand_text())
    pl.close()
    for b in bufs:
        b.free()
    _check(oracle, w, gulps_host, out, cleaning=1)
