"""MI355X-native Heimdall per-gulp search.

The product is the C-ABI library libheimdall_mi355x.so (include/hd/*.h), a
drop-in for the reference's Pipeline library.  This module is a thin ctypes
mirror of that interface for Python callers (tests, benchmarks): the same
entry points (`hd_create_pipeline`, `hd_preallocate`, `hd_get_max_overlap`,
`hd_execute`, `hd_destroy_pipeline`), the same `hd_params` struct and error
codes, plus accessors for the candidates the reference only writes to disk.

There is no CPU fallback: if the library is missing, importing the search
raises.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libheimdall_mi355x.so")

# hd_error values (include/hd/error.h)
HD_NO_ERROR = 0
HD_MEM_ALLOC_FAILED = 1
HD_MEM_COPY_FAILED = 2
HD_INVALID_DEVICE_INDEX = 3
HD_DEVICE_ALREADY_SET = 4
HD_INVALID_PIPELINE = 5
HD_INVALID_POINTER = 6
HD_INVALID_STRIDE = 7
HD_INVALID_NBITS = 8
HD_PRIOR_GPU_ERROR = 9
HD_INTERNAL_GPU_ERROR = 10
HD_TOO_FEW_NSAMPS = 11
HD_TOO_MANY_EVENTS = 12
HD_UNKNOWN_ERROR = 13


class HdRange(C.Structure):
    _fields_ = [("start", C.c_size_t), ("end", C.c_size_t)]


class HdParams(C.Structure):
    """struct hd_params (include/hd/params.h, reference Pipeline/hd/params.h:19-78)."""
    _fields_ = [
        ("verbosity", C.c_int), ("dada_id", C.c_int), ("sigproc_file", C.c_char_p),
        ("yield_cpu", C.c_bool), ("nsamps_gulp", C.c_size_t), ("dm_gulp_size", C.c_size_t),
        ("baseline_length", C.c_float), ("beam", C.c_size_t), ("nbeams", C.c_size_t),
        ("override_beam", C.c_bool), ("nchans", C.c_size_t), ("dt", C.c_float),
        ("f0", C.c_float), ("df", C.c_float), ("dm_min", C.c_float), ("dm_max", C.c_float),
        ("dm_tol", C.c_float), ("dm_pulse_width", C.c_float), ("dm_nbits", C.c_size_t),
        ("use_scrunching", C.c_bool), ("scrunch_tol", C.c_float), ("rfi_tol", C.c_float),
        ("rfi_min_beams", C.c_size_t), ("boxcar_max", C.c_size_t),
        ("detect_thresh", C.c_float), ("cand_sep_time", C.c_size_t),
        ("cand_sep_filter", C.c_size_t), ("cand_sep_dm", C.c_size_t),
        ("cand_rfi_dm_cut", C.c_size_t), ("max_giant_rate", C.c_float),
        ("min_tscrunch_width", C.c_size_t), ("coincidencer_host", C.c_char_p),
        ("coincidencer_port", C.c_int), ("num_channel_zaps", C.c_uint),
        ("channel_zaps", C.POINTER(HdRange)), ("beam_count", C.c_size_t),
        ("gpu_id", C.c_int), ("utc_start", C.c_long), ("spectra_per_second", C.c_size_t),
        ("output_dir", C.c_char_p),
    ]


class HdGulpDesc(C.Structure):
    """hd_gulp_desc (include/hd/mi355x.h): one gulp of hd_execute_batch."""
    _fields_ = [("filterbank", C.c_void_p), ("nsamps", C.c_size_t), ("nbeams", C.c_size_t),
                ("first_idx", C.c_size_t), ("beam", C.c_size_t)]


_lib = None


def lib():
    """The loaded product library (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError("libheimdall_mi355x.so is missing: run __graft_entry__.build() "
                           "(there is no CPU fallback)")
    L = C.CDLL(LIB_PATH)
    sz, fp, szp = C.c_size_t, C.POINTER(C.c_float), C.POINTER(C.c_size_t)
    vp = C.c_void_p
    L.hd_set_default_params.argtypes = [C.POINTER(HdParams)]
    L.hd_set_default_params.restype = None
    L.hd_get_error_string.argtypes = [C.c_int]
    L.hd_get_error_string.restype = C.c_char_p
    L.hd_create_pipeline.argtypes = [C.POINTER(vp), HdParams]
    L.hd_create_pipeline.restype = C.c_int
    L.hd_preallocate.argtypes = [vp, sz, sz]
    L.hd_preallocate.restype = C.c_int
    L.hd_get_max_overlap.argtypes = [vp, sz]
    L.hd_get_max_overlap.restype = sz
    L.hd_execute.argtypes = [vp, vp, sz, sz, sz, sz, szp]
    L.hd_execute.restype = C.c_int
    L.hd_execute_device.argtypes = [vp, vp, sz, sz, sz, sz, szp]
    L.hd_execute_device.restype = C.c_int
    L.hd_destroy_pipeline.argtypes = [vp]
    L.hd_destroy_pipeline.restype = None
    L.hd_get_last_candidates.argtypes = [vp, sz, fp, szp, szp, szp, szp, szp, szp, fp]
    L.hd_get_last_candidates.restype = sz
    L.hd_get_last_giants.argtypes = [vp, sz, fp, szp, szp, szp, szp, szp]
    L.hd_get_last_giants.restype = sz
    L.hd_get_last_cand_text.argtypes = [vp, C.c_char_p, sz]
    L.hd_get_last_cand_text.restype = sz
    L.hd_get_dm_count.argtypes = [vp]
    L.hd_get_dm_count.restype = sz
    L.hd_get_max_delay.argtypes = [vp]
    L.hd_get_max_delay.restype = sz
    L.hd_get_dm_list.argtypes = [vp, fp, sz]
    L.hd_get_dm_list.restype = sz
    L.hd_get_last_timings.argtypes = [vp, C.POINTER(C.c_double), sz]
    L.hd_get_last_timings.restype = sz
    L.hd_set_output_enabled.argtypes = [vp, C.c_int]
    L.hd_set_output_enabled.restype = None
    L.hd_execute_batch.argtypes = [vp, C.POINTER(HdGulpDesc), sz, sz, C.c_int,
                                   C.POINTER(C.c_size_t), C.POINTER(C.c_int)]
    L.hd_execute_batch.restype = C.c_int
    L.hd_get_last_candidate_gulps.argtypes = [vp, C.POINTER(C.c_size_t), sz]
    L.hd_get_last_candidate_gulps.restype = sz
    L.hd_set_cleaning.argtypes = [vp, C.c_int]
    L.hd_set_cleaning.restype = None
    L.hd_get_dedisp_tile.argtypes = [vp, C.POINTER(C.c_int), C.POINTER(C.c_int),
                                     C.POINTER(C.c_int), C.POINTER(C.c_int), C.POINTER(C.c_int)]
    L.hd_get_dedisp_tile.restype = None
    L.hd_set_graph.argtypes = [vp, C.c_int]
    L.hd_set_graph.restype = None
    L.hd_get_graph_stats.argtypes = [vp, C.POINTER(C.c_size_t)]
    L.hd_get_graph_stats.restype = sz
    L.hd_set_dedisp_tile.argtypes = [vp, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int]
    L.hd_set_dedisp_tile.restype = C.c_int
    L.hd_clean_filterbank_rfi.argtypes = [sz, C.c_float, C.c_float, vp, sz, sz, vp,
                                          C.POINTER(C.c_int), C.c_float, C.c_float, C.c_float,
                                          C.c_float, sz, sz, sz, vp]
    L.hd_clean_filterbank_rfi.restype = C.c_int
    L.hd_generate_dm_list.argtypes = [fp, sz, C.c_float, C.c_float, C.c_float, C.c_float,
                                      C.c_float, C.c_float, sz, C.c_float]
    L.hd_generate_dm_list.restype = sz
    _lib = L
    return L


def default_params(**overrides):
    """hd_set_default_params (reference Pipeline/default_params.C:10-66) plus overrides."""
    p = HdParams()
    lib().hd_set_default_params(C.byref(p))
    keep = []
    for k, v in overrides.items():
        if k == "channel_zaps":
            arr = (HdRange * max(len(v), 1))(*[HdRange(a, b) for a, b in v])
            keep.append(arr)
            p.channel_zaps = C.cast(arr, C.POINTER(HdRange))
            p.num_channel_zaps = len(v)
        elif isinstance(v, str):
            b = v.encode()
            keep.append(b)
            setattr(p, k, b)
        else:
            setattr(p, k, v)
    p._keep = keep
    return p


def error_string(err):
    return lib().hd_get_error_string(err).decode()


class HdError(RuntimeError):
    def __init__(self, code, what):
        super().__init__("%s: %s (%d)" % (what, error_string(code), code))
        self.code = code


CAND_FIELDS = ("peak", "ind", "begin", "end", "filter", "dm", "members")


class Pipeline:
    """One hd_pipeline (reference Pipeline/hd/pipeline.h)."""

    def __init__(self, params: HdParams):
        self._params = params
        self.handle = C.c_void_p()
        err = lib().hd_create_pipeline(C.byref(self.handle), params)
        if err != HD_NO_ERROR:
            raise HdError(err, "hd_create_pipeline")

    def close(self):
        if self.handle:
            lib().hd_destroy_pipeline(self.handle)
            self.handle = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def preallocate(self, nsamps, nbeams):
        err = lib().hd_preallocate(self.handle, nsamps, nbeams)
        if err != HD_NO_ERROR:
            raise HdError(err, "hd_preallocate")

    def max_overlap(self, nbeams):
        return lib().hd_get_max_overlap(self.handle, nbeams)

    def set_output(self, enabled):
        lib().hd_set_output_enabled(self.handle, 1 if enabled else 0)

    def set_graph(self, enabled):
        """Stage chain as a HIP graph (default on) or launched kernel by kernel."""
        lib().hd_set_graph(self.handle, 1 if enabled else 0)

    def graph_stats(self):
        """(graphs captured, graph replays) of this pipeline."""
        n = C.c_size_t(0)
        c = lib().hd_get_graph_stats(self.handle, C.byref(n))
        return c, n.value

    def dedisp_tile(self):
        """(trials per block, XCD-aware order, byte staging, channels per step, threads
        per block) of the last call's dedispersion."""
        v = [C.c_int(0) for _ in range(5)]
        lib().hd_get_dedisp_tile(self.handle, *[C.byref(x) for x in v])
        return tuple(x.value for x in v)

    def set_dedisp_tile(self, trials_per_block, xcd_remap=1, byte_lanes=0, channels_per_step=4,
                        block_threads=512):
        """Fix the dedispersion tile (8, 16 or 32 trials per block; byte staging;
        4, 8 or 16 channels staged per step; 512 or 1024 threads, the latter for
        byte staging at 32 trials), or 0 for the autotuner."""
        err = lib().hd_set_dedisp_tile(self.handle, trials_per_block, xcd_remap, byte_lanes,
                                       channels_per_step, block_threads)
        if err != HD_NO_ERROR:
            raise HdError(err, "hd_set_dedisp_tile")

    def set_cleaning(self, enabled):
        """Zero-DM RFI cleaning before each search (the reference's CLEANING build)."""
        lib().hd_set_cleaning(self.handle, 1 if enabled else 0)

    def execute(self, filterbank, nsamps, nbits, first_idx=0, nbeams=1, device=False,
                allow=(HD_NO_ERROR,)):
        """hd_execute on a host array (or a device pointer with device=True).
        Returns (error, nsamps_processed)."""
        processed = C.c_size_t(0)
        if device:
            ptr = C.c_void_p(int(filterbank))
            err = lib().hd_execute_device(self.handle, ptr, nsamps, nbits, first_idx, nbeams,
                                          C.byref(processed))
        else:
            fb = np.ascontiguousarray(filterbank, dtype=np.uint8)
            err = lib().hd_execute(self.handle, fb.ctypes.data, nsamps, nbits, first_idx,
                                   nbeams, C.byref(processed))
        if err not in allow:
            raise HdError(err, "hd_execute")
        return err, processed.value

    def execute_batch(self, gulps, nbits, device=False, allow=(HD_NO_ERROR,)):
        """hd_execute_batch.  gulps: sequence of (filterbank, nsamps, nbeams,
        first_idx, beam), filterbank a host array (or a device address with
        device=True).  Returns (error, processed list, per-gulp errors)."""
        n = len(gulps)
        descs = (HdGulpDesc * max(n, 1))()
        keep = []
        for k, (fb, ns, nb, fi, beam) in enumerate(gulps):
            if device:
                addr = int(fb)
            else:
                a = np.ascontiguousarray(fb, dtype=np.uint8)
                keep.append(a)
                addr = a.ctypes.data
            descs[k] = HdGulpDesc(addr, ns, nb, fi, beam)
        processed = (C.c_size_t * max(n, 1))()
        errs = (C.c_int * max(n, 1))()
        err = lib().hd_execute_batch(self.handle, descs, n, nbits, 1 if device else 0, processed,
                                     errs)
        if err not in allow:
            raise HdError(err, "hd_execute_batch")
        return err, list(processed)[:n], list(errs)[:n]

    def candidate_gulps(self):
        """Gulp index (in the last batch) of each candidate."""
        L = lib()
        n = L.hd_get_last_candidate_gulps(self.handle, None, 0)
        out = np.zeros(n, np.uint64)
        if n:
            L.hd_get_last_candidate_gulps(self.handle,
                                          out.ctypes.data_as(C.POINTER(C.c_size_t)), n)
        return out

    def dm_list(self):
        n = lib().hd_get_dm_count(self.handle)
        out = np.zeros(n, np.float32)
        lib().hd_get_dm_list(self.handle, out.ctypes.data_as(C.POINTER(C.c_float)), n)
        return out

    def max_delay(self):
        return lib().hd_get_max_delay(self.handle)

    def candidates(self):
        L = lib()
        n = L.hd_get_last_candidates(self.handle, 0, None, None, None, None, None, None, None,
                                     None)
        cols = {f: np.zeros(n, np.float32 if f == "peak" else np.uint64) for f in CAND_FIELDS}
        dms = np.zeros(n, np.float32)
        if n:
            fp = C.POINTER(C.c_float)
            szp = C.POINTER(C.c_size_t)
            L.hd_get_last_candidates(self.handle, n, cols["peak"].ctypes.data_as(fp),
                                     *[cols[f].ctypes.data_as(szp) for f in CAND_FIELDS[1:]],
                                     dms.ctypes.data_as(fp))
        cols["dms"] = dms
        return cols

    def giants(self):
        L = lib()
        n = L.hd_get_last_giants(self.handle, 0, None, None, None, None, None, None)
        fields = CAND_FIELDS[:6]
        cols = {f: np.zeros(n, np.float32 if f == "peak" else np.uint64) for f in fields}
        if n:
            fp = C.POINTER(C.c_float)
            szp = C.POINTER(C.c_size_t)
            L.hd_get_last_giants(self.handle, n, cols["peak"].ctypes.data_as(fp),
                                 *[cols[f].ctypes.data_as(szp) for f in fields[1:]])
        return cols

    def cand_text(self):
        L = lib()
        n = L.hd_get_last_cand_text(self.handle, None, 0)
        buf = C.create_string_buffer(n + 1)
        L.hd_get_last_cand_text(self.handle, buf, n + 1)
        return buf.value.decode()

    STAGES = ("unpack", "dedisperse", "baseline", "rms", "scan", "block_rms", "giants",
              "cluster", "total", "clean", "proposal")

    def timings(self):
        out = (C.c_double * len(self.STAGES))()
        n = lib().hd_get_last_timings(self.handle, out, len(self.STAGES))
        return dict(zip(self.STAGES[:n], list(out)[:n]))


def clean_filterbank_rfi(d_in, d_out, nsamps, nbits, nchans, f0, df, dm, dt, baseline_length,
                         rfi_tol, boxcar_max, nbeams, stream=None):
    """hd_clean_filterbank_rfi over device buffers (addresses); raises HdError."""
    rc = lib().hd_clean_filterbank_rfi(nchans, f0, df, d_in, nsamps, nbits, d_out, None, dm, dt,
                                       baseline_length, rfi_tol, 0, boxcar_max, nbeams, stream)
    if rc != HD_NO_ERROR:
        raise HdError(rc, "hd_clean_filterbank_rfi")


def generate_dm_list(dm_min, dm_max, dt, pulse_width_us, f0, df, nchans, tol):
    """hd_generate_dm_list (the reference's dedisp_generate_dm_list call)."""
    L = lib()
    n = L.hd_generate_dm_list(None, 0, dm_min, dm_max, dt, pulse_width_us, f0, df, nchans, tol)
    out = np.zeros(n, np.float32)
    L.hd_generate_dm_list(out.ctypes.data_as(C.POINTER(C.c_float)), n, dm_min, dm_max, dt,
                          pulse_width_us, f0, df, nchans, tol)
    return out
