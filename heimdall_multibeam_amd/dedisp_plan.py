"""ctypes mirror of include/hd/dedisp_plan.h: a dedispersion plan with the
dedisp calls the reference's tools make (dedisp_create_plan, set_dm_list,
set_killmask, enable/disable_adaptive_dt, get_dt_factors, get_max_delay,
execute).  Raises HdError when the library returns an error."""
import ctypes as C

import numpy as np

from . import HD_NO_ERROR, HdError, lib as _pipeline_lib

sz = C.c_size_t
vp = C.c_void_p
fl = C.c_float
HOST_POINTERS = 0
DEVICE_POINTERS = 1

_bound = False


def lib():
    global _bound
    L = _pipeline_lib()
    if not _bound:
        L.hd_dedisp_create_plan.argtypes = [C.POINTER(vp), sz, fl, fl, fl]
        L.hd_dedisp_destroy_plan.argtypes = [vp]
        L.hd_dedisp_destroy_plan.restype = None
        L.hd_dedisp_set_dm_list.argtypes = [vp, C.POINTER(fl), sz]
        L.hd_dedisp_generate_dm_list.argtypes = [vp, fl, fl, fl, fl]
        L.hd_dedisp_set_killmask.argtypes = [vp, C.POINTER(C.c_int)]
        L.hd_dedisp_enable_adaptive_dt.argtypes = [vp, fl, fl]
        L.hd_dedisp_disable_adaptive_dt.argtypes = [vp]
        L.hd_dedisp_get_dt_factors.argtypes = [vp]
        L.hd_dedisp_get_dt_factors.restype = C.POINTER(sz)
        L.hd_dedisp_get_dm_list.argtypes = [vp]
        L.hd_dedisp_get_dm_list.restype = C.POINTER(fl)
        for name in ("hd_dedisp_get_max_delay", "hd_dedisp_get_dm_count",
                     "hd_dedisp_get_channel_count"):
            getattr(L, name).argtypes = [vp]
            getattr(L, name).restype = sz
        L.hd_dedisp_get_dm_delay.argtypes = [vp, sz]
        L.hd_dedisp_get_dm_delay.restype = sz
        L.hd_dedisp_execute.argtypes = [vp, sz, vp, sz, vp, sz, C.c_uint]
        for name in ("hd_dedisp_create_plan", "hd_dedisp_set_dm_list",
                     "hd_dedisp_generate_dm_list", "hd_dedisp_set_killmask",
                     "hd_dedisp_enable_adaptive_dt", "hd_dedisp_disable_adaptive_dt",
                     "hd_dedisp_execute"):
            getattr(L, name).restype = C.c_int
        _bound = True
    return L


def _check(rc, what):
    if rc != HD_NO_ERROR:
        raise HdError(rc, what)


class DedispPlan:
    """dedisp_create_plan(nchans, dt [s], f0, df [MHz]) on the current device."""

    def __init__(self, nchans, dt, f0, df):
        self.handle = vp()
        self.nchans = nchans
        _check(lib().hd_dedisp_create_plan(C.byref(self.handle), nchans, dt, f0, df),
               "hd_dedisp_create_plan")

    def set_dm_list(self, dms):
        a = np.ascontiguousarray(dms, np.float32)
        _check(lib().hd_dedisp_set_dm_list(self.handle, a.ctypes.data_as(C.POINTER(fl)), len(a)),
               "hd_dedisp_set_dm_list")

    def generate_dm_list(self, dm_start, dm_end, pulse_width_us, tol):
        _check(lib().hd_dedisp_generate_dm_list(self.handle, dm_start, dm_end, pulse_width_us,
                                                tol), "hd_dedisp_generate_dm_list")

    def set_killmask(self, killmask):
        a = np.ascontiguousarray(killmask, np.int32)
        _check(lib().hd_dedisp_set_killmask(self.handle, a.ctypes.data_as(C.POINTER(C.c_int))),
               "hd_dedisp_set_killmask")

    def enable_adaptive_dt(self, pulse_width_us, tol):
        _check(lib().hd_dedisp_enable_adaptive_dt(self.handle, pulse_width_us, tol),
               "hd_dedisp_enable_adaptive_dt")

    def disable_adaptive_dt(self):
        _check(lib().hd_dedisp_disable_adaptive_dt(self.handle), "hd_dedisp_disable_adaptive_dt")

    def dm_count(self):
        return lib().hd_dedisp_get_dm_count(self.handle)

    def dm_list(self):
        n = self.dm_count()
        p = lib().hd_dedisp_get_dm_list(self.handle)
        return np.ctypeslib.as_array(p, (n,)).copy() if n else np.zeros(0, np.float32)

    def dt_factors(self):
        n = self.dm_count()
        p = lib().hd_dedisp_get_dt_factors(self.handle)
        return np.ctypeslib.as_array(p, (n,)).astype(np.int64) if n else np.zeros(0, np.int64)

    def max_delay(self):
        return lib().hd_dedisp_get_max_delay(self.handle)

    def dm_delay(self, i):
        return lib().hd_dedisp_get_dm_delay(self.handle, i)

    def execute(self, fb, nsamps, nbits):
        """Host filterbank bytes in, [dm_count, nsamps - max_delay] float32 out."""
        fb = np.ascontiguousarray(fb, np.uint8)
        out = np.zeros((self.dm_count(), nsamps - self.max_delay()), np.float32)
        _check(lib().hd_dedisp_execute(self.handle, nsamps, fb.ctypes.data, nbits,
                                       out.ctypes.data, 32, HOST_POINTERS), "hd_dedisp_execute")
        return out

    def execute_device(self, d_in, nsamps, nbits, d_out):
        _check(lib().hd_dedisp_execute(self.handle, nsamps, d_in, nbits, d_out, 32,
                                       DEVICE_POINTERS), "hd_dedisp_execute")

    def close(self):
        if self.handle:
            lib().hd_dedisp_destroy_plan(self.handle)
            self.handle = vp()
