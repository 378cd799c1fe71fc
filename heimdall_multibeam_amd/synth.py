"""Seeded synthetic filterbanks for the search workloads (host side).

Wraps libhd_synth.so (csrc/synth.cpp).  Every workload is described by a
`Workload` (observing setup + search parameters) and turned into bytes by
`make_gulp`, which draws the injected pulses from numpy's PCG64 with the
given seed and places each at a trial DM so it lines up with the plan's
delay table.
"""
import ctypes as C
import dataclasses
import os
from typing import List, Optional

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
SYNTH_LIB = os.path.join(HERE, "libhd_synth.so")


class _Pulse(C.Structure):
    _fields_ = [("t0", C.c_double), ("dm", C.c_float), ("width", C.c_float), ("snr", C.c_float)]


class _Tone(C.Structure):
    _fields_ = [("chan", C.c_int), ("amp_sigma", C.c_float), ("period", C.c_float),
                ("phase", C.c_float)]


class _Spec(C.Structure):
    _fields_ = [("nsamps", C.c_size_t), ("nchans", C.c_size_t), ("nbits", C.c_size_t),
                ("dt", C.c_float), ("f0", C.c_float), ("df", C.c_float), ("seed", C.c_uint64),
                ("mean", C.c_float), ("sigma", C.c_float), ("dead", C.POINTER(C.c_int)),
                ("dead_value", C.c_float), ("pulses", C.POINTER(_Pulse)),
                ("npulses", C.c_size_t), ("bursts", C.POINTER(_Pulse)), ("nbursts", C.c_size_t),
                ("tones", C.POINTER(_Tone)), ("ntones", C.c_size_t)]


_lib = None


def _synth():
    global _lib
    if _lib is None:
        if not os.path.exists(SYNTH_LIB):
            raise RuntimeError("libhd_synth.so is not built; run __graft_entry__.build()")
        L = C.CDLL(SYNTH_LIB)
        L.hd_synth_filterbank.restype = C.c_int
        L.hd_synth_filterbank.argtypes = [C.POINTER(_Spec), C.c_void_p]
        L.hd_synth_delay.restype = C.c_int
        L.hd_synth_delay.argtypes = [C.c_float, C.c_size_t, C.c_float, C.c_float, C.c_float]
        L.hd_synth_dm_list.restype = C.c_size_t
        L.hd_synth_dm_list.argtypes = [C.POINTER(C.c_float), C.c_size_t, C.c_float, C.c_float,
                                       C.c_float, C.c_float, C.c_float, C.c_float, C.c_size_t,
                                       C.c_float]
        _lib = L
    return _lib


@dataclasses.dataclass
class Workload:
    name: str
    nchans: int
    nbits: int
    dt: float            # seconds
    f0: float            # MHz, channel 0
    df: float            # MHz per channel
    nsamps: int          # per beam
    dm_min: float = 0.0
    dm_max: float = 2000.0
    boxcar_max: int = 4096
    detect_thresh: float = 6.0
    n_pulses: int = 0
    pulse_dm: tuple = (0.0, 2000.0)
    pulse_width: tuple = (1.0, 32.0)
    pulse_snr: tuple = (8.0, 30.0)
    tone_frac: float = 0.0
    tone_sigma: float = 10.0
    n_bursts: int = 0
    burst_snr: float = 20.0
    mean: float = 128.0
    sigma: float = 16.0
    min_tscrunch_width: int = 4096


# The workloads the search is measured and checked on.
CONFIGS = {
    # 1 beam, small: 512 ch, 256 us, gulp 65536, DM 0-500, boxcars 1..256
    "small_512ch": Workload("small_512ch", 512, 8, 256e-6, 1500.0, -0.5, 65536, 0.0, 500.0, 256,
                            6.0, 10, (50.0, 450.0), (1.0, 32.0), (8.0, 30.0)),
    # 1 beam, full size: 1024 ch over 1182-1518 MHz, 64 us, gulp 262144, DM 0-2000,
    # boxcars 1..4096, 5% tone channels at 10 sigma, 3 zero-DM bursts at 20 sigma
    "full_1024ch": Workload("full_1024ch", 1024, 8, 64e-6, 1517.8359375, -0.328125, 262144, 0.0,
                            2000.0, 4096, 6.0, 50, (0.0, 2000.0), (1.0, 512.0), (6.0, 50.0),
                            0.05, 10.0, 3, 20.0),
    # ragged multibeam batch: 320 ch over 835-865 MHz, 655.36 us, per-beam gulps of
    # 4096-16384 samples (nsamps below is the longest), 240-320 live channels with
    # the rest held at the noise median, 0-3 pulses per beam, DM 0-2000, boxcars 1..1024
    "batch_320ch": Workload("batch_320ch", 320, 8, 655.36e-6, 864.953125, -0.09375, 16384, 0.0,
                            2000.0, 1024, 6.0, 3, (0.0, 2000.0), (1.0, 64.0), (6.0, 40.0)),
    # learned proposal training data: the full_1024ch setup with pulses of DM 0-2000,
    # width 1-512 samples and S/N 4-30 (50 per gulp).  The proposal workloads search
    # with -min_tscrunch_width 8192 (every boxcar at full resolution): with the
    # default 4096 the 4096-sample boxcar's RMS is taken over a buffer half filled
    # with the previous width's normalised values (a reference quirk kept by the
    # search), which underestimates it and fills every DM range with spurious
    # high-S/N candidates at that width
    "learned_1024ch": Workload("learned_1024ch", 1024, 8, 64e-6, 1517.8359375, -0.328125, 262144,
                               0.0, 2000.0, 4096, 6.0, 50, (0.0, 2000.0), (1.0, 512.0),
                               (4.0, 30.0), 0.05, 10.0, 3, 20.0, 128.0, 16.0, 8192),
    # gated multibeam: 4096 ch over 1200-1600 MHz, 64 us, 2-bit, gulp 131072, DM 0-3000,
    # boxcars 1..4096, 0-5 pulses per beam (width 1-1024, S/N 6-40), 2% tone channels
    "gated_4096ch": Workload("gated_4096ch", 4096, 2, 64e-6, 1599.951171875, -0.09765625, 131072,
                             0.0, 3000.0, 4096, 6.0, 5, (0.0, 3000.0), (1.0, 1024.0),
                             (6.0, 40.0), 0.02, 10.0, 0, 20.0, 1.5, 1.0, 8192),
}

# beams and per-beam ranges of the ragged batch
BATCH_BEAMS = 352
BATCH_NSAMPS = (4096, 16384)
BATCH_LIVE = (240, 320)


def trial_dms(w: Workload, dm_tol: float = 1.25, pulse_width_us: float = 40.0) -> np.ndarray:
    """The plan's DM trials for a workload (same recurrence as the library)."""
    L = _synth()
    fp = C.POINTER(C.c_float)
    n = L.hd_synth_dm_list(None, 0, w.dm_min, w.dm_max, w.dt, pulse_width_us, w.f0, w.df,
                           w.nchans, dm_tol)
    out = np.zeros(n, np.float32)
    L.hd_synth_dm_list(out.ctypes.data_as(fp), n, w.dm_min, w.dm_max, w.dt, pulse_width_us,
                       w.f0, w.df, w.nchans, dm_tol)
    return out


def make_batch(w: Workload, seed: int, nbeams: int, nsamps_range=BATCH_NSAMPS,
               live_range=BATCH_LIVE, dm_trials=None):
    """Ragged beams: list of (bytes, nsamps) with per-beam gulp lengths drawn
    from nsamps_range, live-channel counts from live_range (the others held at
    the noise median, not masked) and 0..w.n_pulses pulses each."""
    rng = np.random.Generator(np.random.PCG64(seed))
    if dm_trials is None:
        dm_trials = trial_dms(w)
    out = []
    for b in range(nbeams):
        ns = int(rng.integers(nsamps_range[0], nsamps_range[1] + 1))
        n_live = int(rng.integers(live_range[0], live_range[1] + 1))
        dead = np.zeros(w.nchans, np.int32)
        dead[rng.choice(w.nchans, w.nchans - n_live, replace=False)] = 1
        npul = int(rng.integers(0, w.n_pulses + 1))
        fb = make_gulp(w, seed * 100003 + b, nsamps=ns, dm_trials=dm_trials, dead=dead,
                       dead_value=w.mean, n_pulses=npul)
        out.append((fb, ns))
    return out


def delay_samples(w: Workload, dm: float, chan: int) -> int:
    return _synth().hd_synth_delay(dm, chan, w.dt, w.f0, w.df)


def gulp_spec(w: Workload, seed: int, nsamps: Optional[int] = None, dm_trials=None,
              dead: Optional[np.ndarray] = None, dead_value: float = 128.0,
              n_pulses: Optional[int] = None, pulses=None):
    """The generator spec of one beam of workload `w` (pulses, tones and bursts
    drawn from PCG64(seed) unless `pulses` gives (t0, dm, width, snr) tuples).
    Returns (spec, keepalive, truth)."""
    ns = int(nsamps if nsamps is not None else w.nsamps)
    rng = np.random.Generator(np.random.PCG64(seed))
    if dm_trials is None:
        dm_trials = trial_dms(w)
    dm_trials = np.asarray(dm_trials, np.float32)
    max_delay = max(delay_samples(w, float(dm_trials[-1]), w.nchans - 1), 0)
    npul = w.n_pulses if n_pulses is None else n_pulses
    lo_t = w.boxcar_max
    hi_t = max(ns - max_delay - w.boxcar_max, lo_t + 1)
    if pulses is None:
        pulses = []
        for _ in range(npul):
            dm = rng.uniform(*w.pulse_dm)
            dm = float(dm_trials[int(np.argmin(np.abs(dm_trials - dm)))])
            width = float(rng.uniform(*w.pulse_width))
            snr = float(rng.uniform(*w.pulse_snr))
            t0 = float(rng.integers(lo_t, hi_t))
            pulses.append((t0, dm, width, snr))
    tones = []
    if w.tone_frac > 0:
        ntone = int(round(w.tone_frac * w.nchans))
        chans = rng.choice(w.nchans, ntone, replace=False)
        for c in chans:
            tones.append((int(c), w.tone_sigma, float(rng.uniform(20.0, 2000.0)),
                          float(rng.uniform(0, 2 * np.pi))))
    bursts = []
    for _ in range(w.n_bursts):
        bursts.append((float(rng.integers(lo_t, hi_t)), 0.0, float(rng.uniform(1.0, 16.0)),
                       w.burst_snr))
    spec = _Spec()
    spec.nsamps = ns
    spec.nchans = w.nchans
    spec.nbits = w.nbits
    spec.dt, spec.f0, spec.df = w.dt, w.f0, w.df
    spec.seed = seed
    spec.mean, spec.sigma = w.mean, w.sigma
    keep = []
    if dead is not None:
        dead = np.ascontiguousarray(dead, np.int32)
        keep.append(dead)
        spec.dead = dead.ctypes.data_as(C.POINTER(C.c_int))
    spec.dead_value = dead_value
    P = (_Pulse * max(len(pulses), 1))(*[_Pulse(*p) for p in pulses])
    B = (_Pulse * max(len(bursts), 1))(*[_Pulse(*p) for p in bursts])
    T = (_Tone * max(len(tones), 1))(*[_Tone(*t) for t in tones])
    keep += [P, B, T]
    spec.pulses, spec.npulses = P, len(pulses)
    spec.bursts, spec.nbursts = B, len(bursts)
    spec.tones, spec.ntones = T, len(tones)
    return spec, keep, {"pulses": pulses, "tones": tones, "bursts": bursts}


def make_gulp(w: Workload, seed: int, nsamps: Optional[int] = None, dm_trials=None,
              dead: Optional[np.ndarray] = None, dead_value: float = 128.0,
              n_pulses: Optional[int] = None, return_truth: bool = False, pulses=None):
    """Bytes [nsamps][nchans*nbits/8] for one beam of workload `w` (host generator)."""
    spec, keep, truth = gulp_spec(w, seed, nsamps, dm_trials, dead, dead_value, n_pulses,
                                  pulses)
    out = np.empty(spec.nsamps * w.nchans * w.nbits // 8, np.uint8)
    rc = _synth().hd_synth_filterbank(C.byref(spec), out.ctypes.data)
    if rc != 0:
        raise ValueError("hd_synth_filterbank failed")
    if return_truth:
        return out, truth
    return out


def make_gulp_device(w: Workload, seed: int, d_out, gpu_id: int = 0, **kw):
    """The same gulp rendered on the GPU into device memory at address d_out
    (libheimdall_mi355x's hd_synth_filterbank_device); returns the truth."""
    from . import lib, HdError
    L = lib()
    L.hd_synth_filterbank_device.argtypes = [C.POINTER(_Spec), C.c_void_p, C.c_int]
    L.hd_synth_filterbank_device.restype = C.c_int
    spec, keep, truth = gulp_spec(w, seed, **kw)
    rc = L.hd_synth_filterbank_device(C.byref(spec), C.c_void_p(int(d_out)), gpu_id)
    if rc != 0:
        raise HdError(rc, "hd_synth_filterbank_device")
    return truth
