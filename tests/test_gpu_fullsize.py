"""Parity at the benchmark workloads' full sizes, bit for bit against the CPU
oracle on the same input:
  * the 1024-channel gulp (262144 samples, 1399 trials, boxcars 1..4096,
    5% tone channels, zero-DM bursts, 50 pulses), one hd_execute call;
  * the ragged batch (352 beams of 4096-16384 samples, 320 channels, 224
    trials, boxcars 1..1024), one hd_execute_batch call against a separate
    oracle run per beam;
  * one beam of the 4096-channel 2-bit setup (the byte-staged kernel at its
    real channel count, with tone channels) on a half-length gulp searched
    to DM 1000.
The oracle runs with up to 16 threads (about 20 s for the gulp and 10 s for
the batch on the MI355X host)."""
import os

import pytest

from tests._cases import oracle_params_for, params_for
from tests.test_gpu_batch import _check, _run_batch
from tests.test_gpu_parity import _compare

pytestmark = pytest.mark.gpu


def _threads():
    return min(16, os.cpu_count() or 1)


def test_full_1024ch_gulp_matches_oracle(oracle):
    import heimdall_multibeam_amd as H
    from heimdall_multibeam_amd import synth
    w = synth.CONFIGS["full_1024ch"]
    fb = synth.make_gulp(w, 1)
    print("case full_1024ch seed 1", flush=True)
    pl = H.Pipeline(params_for(w))
    pl.set_output(False)
    err, processed = pl.execute(fb, w.nsamps, w.nbits)
    gpu = dict(err=err, processed=processed, giants=pl.giants(), cands=pl.candidates(),
               text=pl.cand_text(), dms=pl.dm_list(), max_delay=pl.max_delay())
    print("gpu done, tile", pl.dedisp_tile(), flush=True)
    pl.close()
    r = oracle.execute(oracle_params_for(oracle, w, nthreads=_threads()), fb, w.nsamps,
                       w.nbits, 0, 1)
    assert err == r.error == H.HD_NO_ERROR
    assert len(r.groups) > 20
    _compare(gpu, r, oracle, w)


@pytest.mark.parametrize("seed", [1, 2, 3])
def test_small_512ch_gulp_matches_oracle(oracle, seed):
    """The 512-channel gulp (65536 samples of 256 us, 320 trials to DM 500,
    boxcars 1..256, 10 pulses): one hd_execute call per seed."""
    import heimdall_multibeam_amd as H
    from heimdall_multibeam_amd import synth
    w = synth.CONFIGS["small_512ch"]
    fb = synth.make_gulp(w, seed)
    print("case small_512ch seed %d" % seed, flush=True)
    pl = H.Pipeline(params_for(w))
    pl.set_output(False)
    err, processed = pl.execute(fb, w.nsamps, w.nbits)
    gpu = dict(err=err, processed=processed, giants=pl.giants(), cands=pl.candidates(),
               text=pl.cand_text(), dms=pl.dm_list(), max_delay=pl.max_delay())
    pl.close()
    r = oracle.execute(oracle_params_for(oracle, w, nthreads=_threads()), fb, w.nsamps,
                       w.nbits, 0, 1)
    assert err == r.error == H.HD_NO_ERROR
    assert len(gpu["dms"]) == 320 and gpu["max_delay"] == 1638
    assert len(r.groups) > 5
    _compare(gpu, r, oracle, w)


def test_full_ragged_batch_matches_oracle(oracle):
    from heimdall_multibeam_amd import synth
    w = synth.CONFIGS["batch_320ch"]
    beams = synth.make_batch(w, 1, synth.BATCH_BEAMS)
    gulps = [(fb, ns, 1, 0, k) for k, (fb, ns) in enumerate(beams)]
    print("case batch_320ch seed 1, %d beams" % len(gulps), flush=True)
    _check(oracle, w, gulps, _run_batch(oracle, w, gulps))


def test_4096ch_2bit_beam_matches_oracle(oracle):
    import dataclasses
    import heimdall_multibeam_amd as H
    from heimdall_multibeam_amd import synth
    w = dataclasses.replace(synth.CONFIGS["gated_4096ch"], name="beam_4096ch_dm1000",
                            nsamps=65536, dm_max=1000.0, pulse_dm=(0.0, 1000.0), n_pulses=5)
    fb = synth.make_gulp(w, 3)
    print("case %s seed 3" % w.name, flush=True)
    pl = H.Pipeline(params_for(w, min_tscrunch_width=w.min_tscrunch_width))
    pl.set_output(False)
    err, processed = pl.execute(fb, w.nsamps, w.nbits)
    gpu = dict(err=err, processed=processed, giants=pl.giants(), cands=pl.candidates(),
               text=pl.cand_text(), dms=pl.dm_list(), max_delay=pl.max_delay())
    print("gpu done, tile", pl.dedisp_tile(), flush=True)
    pl.close()
    r = oracle.execute(oracle_params_for(oracle, w, nthreads=_threads(),
                                         min_tscrunch_width=w.min_tscrunch_width),
                       fb, w.nsamps, w.nbits, 0, 1)
    assert err == r.error == H.HD_NO_ERROR
    assert len(r.groups) > 0
    _compare(gpu, r, oracle, w)
