"""Adaptive time resolution (use_scrunching, Pipeline/pipeline.cu:187-196,
488-790) on the GPU against the oracle with the same setting, bit for bit:
giants, candidates and .cand text of gulps whose DM lists reach scrunch
factors 2..16, several beams, sub-byte input with zapped channels, boxcars
narrower than the largest factor with a small min_tscrunch_width, a ragged
batch, and the 1024-channel gulp of the headline benchmark.  The scrunch
rule and halving are the oracle's restatement of dedisp (parity unpinned
against dedisp itself); the rest of the loop restates the reference."""
import dataclasses
import os

import numpy as np
import pytest

from tests._cases import SCRUNCH_128, SCRUNCH_256, oracle_params_for, params_for
from tests.test_gpu_batch import _check, _run_batch
from tests.test_gpu_parity import _compare, _run_both

pytestmark = pytest.mark.gpu

ON_GPU = dict(use_scrunching=True)
ON_ORC = dict(use_scrunching=1)


def _factors(O, w):
    from heimdall_multibeam_amd import synth
    return O.scrunch_factors(synth.trial_dms(w), w.dt, 40.0, w.f0, w.df, w.nchans, 1.15)


@pytest.mark.parametrize("w,seed", [(SCRUNCH_128, 3), (SCRUNCH_256, 5), (SCRUNCH_256, 1)])
def test_adaptive_gulp(oracle, w, seed):
    gpu, r = _run_both(oracle, w, seed, gpu_kw=ON_GPU, orc_kw=ON_ORC)
    f = _factors(oracle, w)
    assert (f[r.giants["dm"]] > 1).sum() > 0   # giants from scrunched trials
    _compare(gpu, r, oracle, w)


def test_adaptive_multibeam(oracle):
    gpu, r = _run_both(oracle, SCRUNCH_128, 11, nbeams=2, gpu_kw=ON_GPU, orc_kw=ON_ORC)
    _compare(gpu, r, oracle, SCRUNCH_128)


def test_adaptive_2bit_zapped(oracle):
    w = dataclasses.replace(SCRUNCH_128, nbits=2, mean=1.5, sigma=0.9, name="scrunch_128ch_2bit")
    zaps = [(5, 12), (90, 91)]
    km = np.ones(w.nchans, np.int32)
    for a, b in zaps:
        km[a:b] = 0
    gpu, r = _run_both(oracle, w, 4, nbits=2, gpu_kw=dict(ON_GPU, channel_zaps=zaps),
                       orc_kw=dict(ON_ORC, killmask=km))
    _compare(gpu, r, oracle, w)


def test_adaptive_decimated_widths(oracle):
    """min_tscrunch_width 64: decimated widths inside the scrunched trials
    (relative minimum 64 / factor)."""
    kw = dict(min_tscrunch_width=64)
    gpu, r = _run_both(oracle, SCRUNCH_256, 5, gpu_kw=dict(ON_GPU, **kw),
                       orc_kw=dict(ON_ORC, **kw))
    assert len(r.giants) > 0
    _compare(gpu, r, oracle, SCRUNCH_256)


def test_adaptive_narrow_boxcars(oracle):
    """boxcar_max 8 with factors up to 16: the factor-16 trials search no
    width, the factor-8 trials one."""
    w = dataclasses.replace(SCRUNCH_256, boxcar_max=8)
    gpu, r = _run_both(oracle, w, 5, gpu_kw=ON_GPU, orc_kw=ON_ORC)
    assert len(r.giants) > 0
    _compare(gpu, r, oracle, w)


def test_adaptive_ragged_batch(oracle):
    from heimdall_multibeam_amd import synth
    w = SCRUNCH_128
    gulps = []
    for k, ns in enumerate((32768, 29001, 24576)):
        fb = synth.make_gulp(w, 20 + k, nsamps=ns)
        gulps.append((fb, ns, 1, 0, k))
    print("case adaptive batch", [g[1] for g in gulps], flush=True)
    _check(oracle, w, gulps, _run_batch(oracle, w, gulps, **ON_GPU), **ON_ORC)


def test_adaptive_full_1024ch_gulp(oracle):
    import heimdall_multibeam_amd as H
    from heimdall_multibeam_amd import synth
    w = synth.CONFIGS["full_1024ch"]
    fb = synth.make_gulp(w, 1)
    print("case full_1024ch seed 1, adaptive dt", flush=True)
    pl = H.Pipeline(params_for(w, **ON_GPU))
    pl.set_output(False)
    err, processed = pl.execute(fb, w.nsamps, w.nbits)
    gpu = dict(err=err, processed=processed, giants=pl.giants(), cands=pl.candidates(),
               text=pl.cand_text(), dms=pl.dm_list(), max_delay=pl.max_delay())
    pl.close()
    r = oracle.execute(oracle_params_for(oracle, w, nthreads=min(16, os.cpu_count() or 1),
                                         **ON_ORC), fb, w.nsamps, w.nbits, 0, 1)
    assert err == r.error == H.HD_NO_ERROR
    assert len(r.groups) > 10
    _compare(gpu, r, oracle, w)


def test_gated_search_refuses_adaptive_dt():
    """The gated search runs at full resolution only: with scrunching on and a
    proposal set, hd_execute returns HD_INVALID_PIPELINE; with the proposal
    off again the same pipeline searches adaptively."""
    import heimdall_multibeam_amd as H
    from heimdall_multibeam_amd import proposal as P, synth
    from tests.test_gpu_proposal import _random_params
    w = SCRUNCH_128
    fb = synth.make_gulp(w, 3)
    pl = H.Pipeline(params_for(w, **ON_GPU))
    pl.set_output(False)
    P.set_proposal(pl, _random_params(8), -1e30)
    err, _ = pl.execute(fb, w.nsamps, w.nbits, allow=(H.HD_INVALID_PIPELINE,))
    assert err == H.HD_INVALID_PIPELINE
    P.set_proposal(pl, None, 0)
    err, _ = pl.execute(fb, w.nsamps, w.nbits)
    assert err == H.HD_NO_ERROR and len(pl.giants()["peak"]) > 0
    pl.close()
