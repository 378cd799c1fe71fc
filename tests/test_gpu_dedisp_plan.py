"""hd_dedisp_* (include/hd/dedisp_plan.h): the dedisp plan calls of the
reference's tools (Applications/candidate_profiler.C:150-164, 261-327;
Pipeline/clean_filterbank_rfi.cu:361-397) on the GPU, against the oracle:

  * the scrunch factors equal the restated rule (oracle/adaptive_dt.py);
  * with adaptive time resolution, every trial's sums over the 1024-channel
    setup (factors 1 to 64) equal the restatement bit for bit, for 8-bit and
    2-bit samples and with dead channels; without it, orc_dedisperse's;
  * device pointers give the same rows as host pointers.
The restated rule is parity unpinned against dedisp (absent)."""
import numpy as np
import pytest

from tests.conftest import gpu_available

pytestmark = pytest.mark.gpu

f32 = np.float32


@pytest.fixture(scope="module")
def D():
    if not gpu_available():
        pytest.skip("no GPU")
    from heimdall_multibeam_amd import dedisp_plan
    return dedisp_plan


def _setup(nbits, seed):
    from heimdall_multibeam_amd import synth
    w = synth.CONFIGS["full_1024ch"]
    dms = synth.trial_dms(w)[::6].astype(f32)        # 234 trials, factors 1 to 64
    rng = np.random.default_rng(seed)
    nsamps = 36540 + 6000
    per = 8 // nbits
    fb = rng.integers(0, 256, nsamps * w.nchans // per, dtype=np.uint8)
    return w, dms, nsamps, fb


@pytest.mark.parametrize("nbits", [8, 2])
def test_adaptive_plane_matches_oracle(D, nbits):
    from oracle import adaptive_dt as A
    from oracle import oracle as O
    w, dms, nsamps, fb = _setup(nbits, 10 + nbits)
    p = D.DedispPlan(w.nchans, w.dt, w.f0, w.df)
    p.set_dm_list(dms)
    km = np.ones(w.nchans, np.int32)
    km[[0, 5, 6, 500, 1023]] = 0
    p.set_killmask(km)
    p.enable_adaptive_dt(40.0, 1.15)
    f = p.dt_factors()
    want_f = A.scrunch_list(dms, w.dt, 40.0, w.f0, w.df, w.nchans, 1.15)
    np.testing.assert_array_equal(f, want_f)
    assert {1, 2, 4, 8, 16, 32} <= set(f.tolist())
    kc = O.delay_table(w.nchans, f32(w.dt), f32(w.f0), f32(w.df))
    maxd = int((dms[:, None] * kc[None, :] + f32(0.5)).astype(np.int64).max())
    assert p.max_delay() == maxd
    print("case adaptive plane nbits %d: %d trials x %d samples" % (nbits, len(dms), nsamps - maxd),
          flush=True)
    got = p.execute(fb, nsamps, nbits)
    want = A.dedisperse_adaptive(fb, nsamps, nbits, w.nchans, dms, kc, maxd, want_f, killmask=km,
                                 nthreads=16)
    bad = np.argwhere(got != want)
    assert len(bad) == 0, "first mismatches (trial, sample): %s" % bad[:5].tolist()
    p.close()


def test_plan_without_adaptive_and_device_pointers(D):
    from oracle import oracle as O
    from heimdall_multibeam_amd.hip import DeviceBuffer
    w, dms, nsamps, fb = _setup(8, 3)
    p = D.DedispPlan(w.nchans, w.dt, w.f0, w.df)
    p.set_dm_list(dms[:80])
    assert np.all(p.dt_factors() == 1)
    kc = O.delay_table(w.nchans, f32(w.dt), f32(w.f0), f32(w.df))
    maxd = p.max_delay()
    print("case plain plan: 80 trials", flush=True)
    host = p.execute(fb, nsamps, 8)
    want = O.dedisperse(fb, nsamps, 8, w.nchans, dms[:80], kc, maxd, nthreads=16)
    np.testing.assert_array_equal(host, want)
    d_in = DeviceBuffer(fb.nbytes)
    d_in.upload(fb)
    d_out = DeviceBuffer(host.nbytes)
    p.execute_device(d_in.address, nsamps, 8, d_out.address)
    dev = d_out.download().view(f32).reshape(host.shape)
    np.testing.assert_array_equal(dev, host)
    d_in.free()
    d_out.free()
    # the delay of one trial is its largest over channels
    assert p.dm_delay(79) == int((dms[79] * kc + f32(0.5)).astype(np.int64).max())
    p.close()
