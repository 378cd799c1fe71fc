"""Shared small search setups for the parity tests."""
from heimdall_multibeam_amd.synth import Workload

# Small gulps the oracle finishes in a second or two.
TINY = Workload("tiny_64ch", 64, 8, 256e-6, 1500.0, -2.0, 16384, 0.0, 300.0, 256, 6.0,
                4, (20.0, 280.0), (1.0, 16.0), (10.0, 30.0))
SMALL_128 = Workload("small_128ch", 128, 8, 128e-6, 1450.0, -1.0, 32768, 0.0, 500.0, 1024, 6.0,
                     6, (20.0, 480.0), (1.0, 64.0), (8.0, 40.0), 0.05, 10.0, 1, 20.0)
# Gulps whose DM lists reach scrunch factors above 1 at scrunch_tol 1.15
# (128 ch: factors 1/2/4/8 over 241 trials; 256 ch: 1..16 over 423 trials).
SCRUNCH_128 = Workload("scrunch_128ch", 128, 8, 64e-6, 1400.0, -2.0, 32768, 0.0, 1000.0, 512,
                       6.0, 8, (100.0, 900.0), (1.0, 64.0), (10.0, 30.0))
SCRUNCH_256 = Workload("scrunch_256ch", 256, 8, 32e-6, 1500.0, -1.0, 40000, 0.0, 1000.0, 1024,
                       6.0, 8, (100.0, 900.0), (1.0, 64.0), (10.0, 30.0), 0.03, 10.0, 1, 20.0)


def params_for(w, **kw):
    import heimdall_multibeam_amd as H
    d = dict(nchans=w.nchans, dt=w.dt, f0=w.f0, df=w.df, dm_min=w.dm_min, dm_max=w.dm_max,
             boxcar_max=w.boxcar_max, detect_thresh=w.detect_thresh, use_scrunching=False)
    d.update(kw)
    return H.default_params(**d)


def oracle_params_for(O, w, **kw):
    d = dict(nchans=w.nchans, dt=w.dt, f0=w.f0, df=w.df, dm_min=w.dm_min, dm_max=w.dm_max,
             boxcar_max=w.boxcar_max, detect_thresh=w.detect_thresh, nthreads=8)
    d.update(kw)
    return O.make_params(**d)
