"""ctypes wrapper of the CPU oracle (oracle/libhd_oracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py.  The product package never imports it.
"""
import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libhd_oracle.so")


def build():
    subprocess.check_call(["make", "-s", "-C", HERE])


class Cand(C.Structure):
    _fields_ = [("peak", C.c_float), ("ind", C.c_uint64), ("begin", C.c_uint64),
                ("end", C.c_uint64), ("filter", C.c_uint64), ("dm", C.c_uint64),
                ("members", C.c_uint64)]


CAND_DTYPE = np.dtype([("peak", "<f4"), ("ind", "<u8"), ("begin", "<u8"), ("end", "<u8"),
                       ("filter", "<u8"), ("dm", "<u8"), ("members", "<u8")], align=True)
assert CAND_DTYPE.itemsize == C.sizeof(Cand)


class Params(C.Structure):
    _fields_ = [("nchans", C.c_size_t), ("dt", C.c_float), ("f0", C.c_float), ("df", C.c_float),
                ("dm_min", C.c_float), ("dm_max", C.c_float), ("dm_tol", C.c_float),
                ("dm_pulse_width", C.c_float), ("baseline_length", C.c_float),
                ("boxcar_max", C.c_size_t), ("detect_thresh", C.c_float),
                ("cand_sep_time", C.c_size_t), ("cand_sep_filter", C.c_size_t),
                ("cand_sep_dm", C.c_size_t), ("max_giant_rate", C.c_float),
                ("min_tscrunch_width", C.c_size_t), ("beam", C.c_size_t),
                ("killmask", C.POINTER(C.c_int)), ("nthreads", C.c_int),
                ("dm_first", C.c_size_t), ("dm_count", C.c_size_t), ("dm_stride", C.c_size_t),
                ("cleaning", C.c_int), ("rfi_tol", C.c_float),
                ("use_scrunching", C.c_int), ("scrunch_tol", C.c_float)]


class Result(C.Structure):
    _fields_ = [("error", C.c_int), ("n_dm", C.c_size_t), ("max_delay", C.c_size_t),
                ("nsamps_processed", C.c_size_t), ("n_giants", C.c_size_t),
                ("giants", C.POINTER(Cand)), ("n_groups", C.c_size_t),
                ("groups", C.POINTER(Cand)), ("group_dms", C.POINTER(C.c_float)),
                ("seconds", C.c_double)]


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = C.CDLL(LIB_PATH)
        fp = C.POINTER(C.c_float)
        sz = C.c_size_t
        L.orc_median3.restype = C.c_float
        L.orc_median3.argtypes = [C.c_float] * 3
        L.orc_median4.restype = C.c_float
        L.orc_median4.argtypes = [C.c_float] * 4
        L.orc_median5.restype = C.c_float
        L.orc_median5.argtypes = [C.c_float] * 5
        L.orc_generate_dm_list.restype = sz
        L.orc_generate_dm_list.argtypes = [fp, sz, C.c_float, C.c_float, C.c_float, C.c_float,
                                           C.c_float, C.c_float, sz, C.c_float]
        L.orc_delay_table.restype = None
        L.orc_delay_table.argtypes = [fp, sz, C.c_float, C.c_float, C.c_float]
        L.orc_delay.restype = C.c_int
        L.orc_delay.argtypes = [C.c_float, C.c_float]
        L.orc_dedisperse.restype = C.c_int
        L.orc_dedisperse.argtypes = [C.c_void_p, sz, sz, sz, C.c_void_p, fp, sz, fp, sz, fp, sz,
                                     C.c_int]
        L.orc_linear_stretch.restype = None
        L.orc_linear_stretch.argtypes = [fp, sz, fp, sz]
        L.orc_linear_stretch_beam.restype = None
        L.orc_linear_stretch_beam.argtypes = [fp, sz, fp, sz, sz]
        L.orc_work_create.restype = C.c_void_p
        L.orc_work_destroy.argtypes = [C.c_void_p]
        L.orc_remove_baseline.restype = C.c_int
        L.orc_remove_baseline.argtypes = [C.c_void_p, fp, sz, sz, sz, sz]
        L.orc_get_rms_multibeam.restype = C.c_int
        L.orc_get_rms_multibeam.argtypes = [fp, fp, sz, sz, sz]
        L.orc_normalise_multibeam.restype = None
        L.orc_normalise_multibeam.argtypes = [fp, fp, sz, sz]
        L.orc_block_rms.restype = C.c_float
        L.orc_block_rms.argtypes = [fp, sz, sz, sz]
        L.orc_normalise_block.restype = None
        L.orc_normalise_block.argtypes = [fp, C.c_float, sz]
        L.orc_scan.restype = None
        L.orc_scan.argtypes = [fp, sz, fp]
        L.orc_boxcar.restype = sz
        L.orc_boxcar.argtypes = [fp, sz, sz, sz, sz, fp]
        L.orc_find_giants.restype = sz
        L.orc_find_giants.argtypes = [fp, sz, C.c_float, sz, sz, sz, C.POINTER(Cand), sz]
        L.orc_label_clusters.restype = sz
        L.orc_label_clusters.argtypes = [sz, C.POINTER(Cand), sz, sz, sz, sz,
                                         C.POINTER(C.c_uint64)]
        L.orc_merge_candidates.restype = None
        L.orc_merge_candidates.argtypes = [sz, C.POINTER(C.c_uint64), C.POINTER(Cand),
                                           C.POINTER(Cand)]
        L.orc_hash32.restype = C.c_uint32
        L.orc_hash32.argtypes = [C.c_uint32]
        L.orc_minstd_seed.restype = C.c_uint32
        L.orc_minstd_seed.argtypes = [C.c_uint32]
        L.orc_minstd_next.restype = C.c_uint32
        L.orc_minstd_next.argtypes = [C.POINTER(C.c_uint32)]
        L.orc_uniform_draw.restype = C.c_uint32
        L.orc_uniform_draw.argtypes = [C.POINTER(C.c_uint32), C.c_uint32, C.c_uint32]
        L.orc_zap_filterbank_rfi.restype = C.c_int
        L.orc_zap_filterbank_rfi.argtypes = [C.c_void_p, C.c_void_p, sz, sz, sz, sz, C.c_void_p]
        L.orc_clean_filterbank_rfi.restype = C.c_int
        L.orc_clean_filterbank_rfi.argtypes = [C.c_void_p, sz, sz, sz, C.c_float, C.c_float,
                                               C.c_float, C.c_float, C.c_float, C.c_float, sz,
                                               sz, C.c_void_p, C.c_void_p]
        L.orc_scrunch_factors.restype = None
        L.orc_scrunch_factors.argtypes = [fp, sz, C.c_float, C.c_float, C.c_float, C.c_float,
                                          sz, C.c_float, C.POINTER(C.c_size_t)]
        L.orc_dedisperse_adaptive.restype = C.c_int
        L.orc_dedisperse_adaptive.argtypes = [C.c_void_p, sz, sz, sz, C.c_void_p, fp, sz, fp, sz,
                                              C.POINTER(C.c_size_t), fp, sz, C.c_int]
        L.orc_default_params.restype = None
        L.orc_default_params.argtypes = [C.POINTER(Params)]
        L.orc_execute.restype = C.c_int
        L.orc_execute.argtypes = [C.POINTER(Params), C.c_void_p, sz, sz, sz, sz,
                                  C.POINTER(Result)]
        L.orc_result_free.restype = None
        L.orc_result_free.argtypes = [C.POINTER(Result)]
        L.orc_format_cands.restype = sz
        L.orc_format_cands.argtypes = [C.POINTER(Result), C.POINTER(Params), sz, sz, sz,
                                       C.c_char_p, sz]
        _lib = L
    return _lib


def _fp(a):
    return a.ctypes.data_as(C.POINTER(C.c_float))


def median3(*v):
    return lib().orc_median3(*v)


def median4(*v):
    return lib().orc_median4(*v)


def median5(*v):
    return lib().orc_median5(*v)


def dm_list(dm_min, dm_max, dt, pulse_width, f0, df, nchans, tol):
    L = lib()
    n = L.orc_generate_dm_list(None, 0, dm_min, dm_max, dt, pulse_width, f0, df, nchans, tol)
    out = np.zeros(n, np.float32)
    L.orc_generate_dm_list(_fp(out), n, dm_min, dm_max, dt, pulse_width, f0, df, nchans, tol)
    return out


def delay_table(nchans, dt, f0, df):
    out = np.zeros(nchans, np.float32)
    lib().orc_delay_table(_fp(out), nchans, dt, f0, df)
    return out


def delay(dm, kc):
    return lib().orc_delay(dm, kc)


def dedisperse(fb, nsamps, nbits, nchans, dms, kc, max_delay, killmask=None, nthreads=1):
    fb = np.ascontiguousarray(fb, dtype=np.uint8)
    dms = np.ascontiguousarray(dms, np.float32)
    kc = np.ascontiguousarray(kc, np.float32)
    nout = nsamps - max_delay
    out = np.zeros((len(dms), nout), np.float32)
    km = None
    if killmask is not None:
        killmask = np.ascontiguousarray(killmask, np.int32)
        km = killmask.ctypes.data
    rc = lib().orc_dedisperse(fb.ctypes.data, nsamps, nbits, nchans, km, _fp(dms), len(dms),
                              _fp(kc), max_delay, _fp(out), nout, nthreads)
    if rc != 0:
        raise ValueError("orc_dedisperse failed %d" % rc)
    return out


def scrunch_factors(dms, dt, pulse_width_us, f0, df, nchans, tol):
    dms = np.ascontiguousarray(dms, np.float32)
    out = np.zeros(len(dms), np.uintp)
    lib().orc_scrunch_factors(_fp(dms), len(dms), dt, pulse_width_us, f0, df, nchans, tol,
                              out.ctypes.data_as(C.POINTER(C.c_size_t)))
    return out.astype(np.int64)


def dedisperse_adaptive(fb, nsamps, nbits, nchans, dms, kc, max_delay, factors, killmask=None,
                        nthreads=1):
    fb = np.ascontiguousarray(fb, dtype=np.uint8)
    dms = np.ascontiguousarray(dms, np.float32)
    kc = np.ascontiguousarray(kc, np.float32)
    fac = np.ascontiguousarray(factors, np.uintp)
    nout = nsamps - max_delay
    out = np.zeros((len(dms), nout), np.float32)
    km = None
    if killmask is not None:
        killmask = np.ascontiguousarray(killmask, np.int32)
        km = killmask.ctypes.data
    rc = lib().orc_dedisperse_adaptive(fb.ctypes.data, nsamps, nbits, nchans, km, _fp(dms),
                                       len(dms), _fp(kc), max_delay,
                                       fac.ctypes.data_as(C.POINTER(C.c_size_t)), _fp(out), nout,
                                       nthreads)
    if rc != 0:
        raise ValueError("orc_dedisperse_adaptive failed %d" % rc)
    return out


def linear_stretch(x, out_count):
    x = np.concatenate([np.asarray(x, np.float32), np.zeros(1, np.float32)])
    out = np.zeros(out_count, np.float32)
    lib().orc_linear_stretch(_fp(x), len(x) - 1, _fp(out), out_count)
    return out


def remove_baseline(d, beam_stride, beam_count, smooth_radius, nbeams):
    buf = np.zeros(len(d) + 64, np.float32)
    buf[:len(d)] = d
    lib().orc_remove_baseline(None, _fp(buf), beam_stride, beam_count, smooth_radius, nbeams)
    return buf[:len(d)].copy()


def get_rms_multibeam(d, beam_stride, beam_count, nbeams):
    d = np.ascontiguousarray(d, np.float32)
    rms = np.zeros(nbeams, np.float32)
    lib().orc_get_rms_multibeam(_fp(d), _fp(rms), beam_stride, beam_count, nbeams)
    return rms


def normalise_multibeam(d, rms, beam_stride, cur_nsamps):
    d = np.array(d, np.float32)
    rms = np.ascontiguousarray(rms, np.float32)
    lib().orc_normalise_multibeam(_fp(d), _fp(rms), beam_stride, cur_nsamps)
    return d


def normalise_block(d, rms, count):
    d = np.array(d, np.float32)
    lib().orc_normalise_block(_fp(d), rms, count)
    return d


def block_rms(d, beam_stride, beam_count, nbeams):
    d = np.ascontiguousarray(d, np.float32)
    return np.float32(lib().orc_block_rms(_fp(d), beam_stride, beam_count, nbeams))


def scan(x):
    x = np.ascontiguousarray(x, np.float32)
    out = np.zeros(len(x) + 1, np.float32)
    lib().orc_scan(_fp(x), len(x), _fp(out))
    return out


def boxcar(scanned, max_width, width, tscrunch):
    scanned = np.ascontiguousarray(scanned, np.float32)
    n_out = len(scanned) - max_width
    out = np.zeros((n_out + tscrunch - 1) // tscrunch, np.float32)
    n = lib().orc_boxcar(_fp(scanned), len(scanned), max_width, width, tscrunch, _fp(out))
    return out[:n]


def find_giants(d, thresh, merge_dist, beam_stride, beam_count):
    d = np.ascontiguousarray(d, np.float32)
    cap = 1024
    while True:
        out = np.zeros(cap, CAND_DTYPE)
        n = lib().orc_find_giants(_fp(d), len(d), thresh, merge_dist, beam_stride, beam_count,
                                  out.ctypes.data_as(C.POINTER(Cand)), cap)
        if n <= cap:
            return out[:n]
        cap = n


def label_clusters(cands, time_tol, filter_tol, dm_tol, nsamps_beam):
    cands = np.ascontiguousarray(cands, CAND_DTYPE)
    labels = np.zeros(len(cands), np.uint64)
    n = lib().orc_label_clusters(len(cands), cands.ctypes.data_as(C.POINTER(Cand)), time_tol,
                                 filter_tol, dm_tol, nsamps_beam,
                                 labels.ctypes.data_as(C.POINTER(C.c_uint64)))
    return labels, n


def merge_candidates(labels, cands, n_groups):
    cands = np.ascontiguousarray(cands, CAND_DTYPE)
    labels = np.ascontiguousarray(labels, np.uint64)
    groups = np.zeros(max(n_groups, 1), CAND_DTYPE)
    lib().orc_merge_candidates(len(cands), labels.ctypes.data_as(C.POINTER(C.c_uint64)),
                               cands.ctypes.data_as(C.POINTER(Cand)),
                               groups.ctypes.data_as(C.POINTER(Cand)))
    return groups[:n_groups]


def hash32(a):
    return lib().orc_hash32(a & 0xffffffff)


def minstd_seed(s):
    return lib().orc_minstd_seed(s & 0xffffffff)


def minstd_outputs(state, n):
    """The next n outputs of the minstd engine from `state`."""
    x = C.c_uint32(state)
    return [lib().orc_minstd_next(C.byref(x)) for _ in range(n)]


def uniform_draws(seed_state, lo, hi, n):
    """n successive uniform_int_distribution(lo, hi) draws from engine state
    seed_state; returns (draws, final state)."""
    x = C.c_uint32(seed_state)
    out = [lib().orc_uniform_draw(C.byref(x), lo, hi) for _ in range(n)]
    return out, x.value


def zap_filterbank_rfi(mask, fb, nsamps, nbits, nchans, max_resample_dist):
    mask = np.ascontiguousarray(mask, np.int32)
    fb = np.ascontiguousarray(fb, np.uint8)
    out = np.zeros_like(fb)
    rc = lib().orc_zap_filterbank_rfi(mask.ctypes.data, fb.ctypes.data, nsamps, nbits, nchans,
                                      max_resample_dist, out.ctypes.data)
    return rc, out


def clean_filterbank_rfi(fb, nsamps, nbits, nchans, dm, dt, f0, df, baseline_length, rfi_tol,
                         boxcar_max, nbeams):
    """Returns (rc, cleaned filterbank, mask).  Samples past nsamps -
    max_delay(dm) keep the input's bytes."""
    fb = np.ascontiguousarray(fb, np.uint8)
    out = fb.copy()
    mask = np.zeros(nsamps, np.int32)
    rc = lib().orc_clean_filterbank_rfi(fb.ctypes.data, nsamps, nbits, nchans, dm, dt, f0, df,
                                        baseline_length, rfi_tol, boxcar_max, nbeams,
                                        out.ctypes.data, mask.ctypes.data)
    return rc, out, mask


def default_params():
    p = Params()
    lib().orc_default_params(C.byref(p))
    return p


def make_params(**kw):
    p = default_params()
    km = kw.pop("killmask", None)
    for k, v in kw.items():
        setattr(p, k, v)
    if km is not None:
        km = np.ascontiguousarray(km, np.int32)
        p.killmask = km.ctypes.data_as(C.POINTER(C.c_int))
        p._km_ref = km
    return p


class ExecResult:
    pass


def execute(params, fb, nsamps, nbits, first_idx, nbeams, want_text=True):
    """Run the whole-gulp oracle; returns an ExecResult with numpy arrays."""
    fb = np.ascontiguousarray(fb, np.uint8)
    r = Result()
    lib().orc_execute(C.byref(params), fb.ctypes.data, nsamps, nbits, first_idx, nbeams,
                      C.byref(r))
    out = ExecResult()
    out.error = r.error
    out.n_dm = r.n_dm
    out.max_delay = r.max_delay
    out.nsamps_processed = r.nsamps_processed
    out.seconds = r.seconds
    ng = r.n_giants
    out.giants = (np.ctypeslib.as_array(C.cast(r.giants, C.POINTER(C.c_uint8)),
                                        (ng * CAND_DTYPE.itemsize,)).view(CAND_DTYPE).copy()
                  if ng else np.zeros(0, CAND_DTYPE))
    ngr = r.n_groups
    out.groups = (np.ctypeslib.as_array(C.cast(r.groups, C.POINTER(C.c_uint8)),
                                        (ngr * CAND_DTYPE.itemsize,)).view(CAND_DTYPE).copy()
                  if ngr else np.zeros(0, CAND_DTYPE))
    out.group_dms = (np.ctypeslib.as_array(r.group_dms, (ngr,)).copy()
                     if ngr else np.zeros(0, np.float32))
    if want_text:
        need = lib().orc_format_cands(C.byref(r), C.byref(params), nsamps, nbeams, first_idx,
                                      None, 0)
        buf = C.create_string_buffer(need + 1)
        lib().orc_format_cands(C.byref(r), C.byref(params), nsamps, nbeams, first_idx, buf,
                               need + 1)
        out.text = buf.value.decode()
    lib().orc_result_free(C.byref(r))
    return out
