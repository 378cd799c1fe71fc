"""Zero-DM RFI cleaning on the GPU against the CPU oracle (bit-exact bytes).

hd_clean_filterbank_rfi is compared byte for byte with the oracle's
clean_filterbank_rfi restatement, and hd_execute with cleaning on is compared
with the oracle's CLEANING-build execute (giants, groups, .cand text).
"""
import dataclasses

import numpy as np
import pytest

from tests._cases import SMALL_128, TINY
from tests.test_gpu_parity import _compare, _run_both

pytestmark = pytest.mark.gpu


def _clean_both(O, w, fb, nsamps, nbits, dm=0.0, boxcar_max=64, nbeams=1, rfi_tol=5.0):
    import heimdall_multibeam_amd as H
    from heimdall_multibeam_amd import hip
    print("clean case", w.name, "nsamps", nsamps, "nbits", nbits, "dm", dm, "nbeams", nbeams,
          flush=True)
    fb = np.ascontiguousarray(fb, np.uint8)
    din, dout = hip.DeviceBuffer(fb.nbytes), hip.DeviceBuffer(fb.nbytes)
    din.upload(fb)
    dout.upload(fb)  # samples past nsamps - max_delay keep the input's bytes
    H.clean_filterbank_rfi(din.address, dout.address, nsamps, nbits, w.nchans, w.f0, w.df, dm,
                           w.dt, 2.0, rfi_tol, boxcar_max, nbeams)
    hip.synchronize()
    got = dout.download(fb.nbytes)
    din.free()
    dout.free()
    rc, ref, mask = O.clean_filterbank_rfi(fb, nsamps, nbits, w.nchans, dm, w.dt, w.f0, w.df,
                                           2.0, rfi_tol, boxcar_max, nbeams)
    assert rc == 0
    return got, ref, mask


def _burst(fb, nsamps, t0, n):
    """A broadband burst: time samples t0..t0+n-1 set to 250 in every channel."""
    fb = fb.copy().reshape(nsamps, -1)
    fb[t0:t0 + n, :] = 250
    return fb.ravel()


@pytest.mark.parametrize("seed", [1, 2])
def test_clean_bytes_small(oracle, seed):
    from heimdall_multibeam_amd import synth
    fb = _burst(synth.make_gulp(SMALL_128, seed), SMALL_128.nsamps, 9000 + 97 * seed, 12)
    got, ref, mask = _clean_both(oracle, SMALL_128, fb, SMALL_128.nsamps, 8)
    assert mask.sum() > 0
    np.testing.assert_array_equal(got, ref)


def test_clean_bytes_two_beams_wide_boxcar(oracle):
    from heimdall_multibeam_amd import synth
    fb = np.concatenate([synth.make_gulp(TINY, 5), synth.make_gulp(TINY, 6)])
    fb = _burst(fb, 2 * TINY.nsamps, 20000, 8)
    got, ref, mask = _clean_both(oracle, TINY, fb, 2 * TINY.nsamps, 8, boxcar_max=256, nbeams=2)
    assert mask.sum() > 0
    np.testing.assert_array_equal(got, ref)


def test_clean_bytes_two_bit_and_nonzero_dm(oracle):
    from heimdall_multibeam_amd import synth
    w = dataclasses.replace(TINY, name="tiny_2bit", nbits=2)
    fb = synth.make_gulp(w, 3)
    got, ref, _ = _clean_both(oracle, w, fb, w.nsamps, 2)
    np.testing.assert_array_equal(got, ref)
    got, ref, _ = _clean_both(oracle, TINY, synth.make_gulp(TINY, 4), TINY.nsamps, 8, dm=40.0)
    np.testing.assert_array_equal(got, ref)


@pytest.mark.parametrize("seed", [3])
def test_execute_with_cleaning(oracle, seed):
    gpu, r = _run_both(oracle, SMALL_128, seed, cleaning=True)
    assert len(r.groups) > 0
    _compare(gpu, r, oracle, SMALL_128)


def test_execute_with_cleaning_multibeam(oracle):
    gpu, r = _run_both(oracle, TINY, 11, nbeams=2, cleaning=True)
    _compare(gpu, r, oracle, TINY)
