"""ctypes mirror of the learned proposal (include/hd/proposal.h).

`Cnn` wraps one hd_cnn: the network's parameters, its forward and backward
kernels and the Adam step, all on the GPU.  `set_proposal`, `coarse`,
`extract`, `plane` and `scores` drive the coarse DM-time plane and gating of
a `Pipeline`.
"""
import ctypes as C

import numpy as np

from . import HdError, HD_NO_ERROR, lib as _lib

TILE = 64
TIME_DECIM = 16
DM_DECIM = 8
UNIT = 128
TSTEP = 32

_bound = False


def lib():
    global _bound
    L = _lib()
    if _bound:
        return L
    sz, fp, vp = C.c_size_t, C.POINTER(C.c_float), C.c_void_p
    L.hd_cnn_param_count.argtypes = []
    L.hd_cnn_param_count.restype = sz
    L.hd_cnn_create.argtypes = [C.POINTER(vp), C.c_int, sz]
    L.hd_cnn_create.restype = C.c_int
    L.hd_cnn_destroy.argtypes = [vp]
    L.hd_cnn_destroy.restype = None
    L.hd_cnn_init_params.argtypes = [vp, C.c_ulonglong]
    L.hd_cnn_init_params.restype = C.c_int
    L.hd_cnn_set_params.argtypes = [vp, fp]
    L.hd_cnn_set_params.restype = C.c_int
    L.hd_cnn_get_params.argtypes = [vp, fp]
    L.hd_cnn_get_params.restype = C.c_int
    L.hd_cnn_forward.argtypes = [vp, vp, sz, vp]
    L.hd_cnn_forward.restype = C.c_int
    L.hd_cnn_get_logits.argtypes = [vp, fp, sz]
    L.hd_cnn_get_logits.restype = C.c_int
    L.hd_cnn_backward.argtypes = [vp, vp, vp, fp]
    L.hd_cnn_backward.restype = C.c_int
    L.hd_cnn_get_grads.argtypes = [vp, fp]
    L.hd_cnn_get_grads.restype = C.c_int
    L.hd_cnn_adam.argtypes = [vp, C.c_float, C.c_float, C.c_float, C.c_float, C.c_float]
    L.hd_cnn_adam.restype = C.c_int
    L.hd_set_proposal.argtypes = [vp, fp, C.c_float]
    L.hd_set_proposal.restype = C.c_int
    L.hd_proposal_coarse.argtypes = [vp, vp, sz, sz, C.POINTER(sz), C.POINTER(sz)]
    L.hd_proposal_coarse.restype = C.c_int
    L.hd_proposal_extract.argtypes = [vp, C.POINTER(C.c_uint), sz, vp]
    L.hd_proposal_extract.restype = C.c_int
    L.hd_proposal_get_plane.argtypes = [vp, C.c_int, C.POINTER(C.c_uint16), sz, C.POINTER(sz),
                                        C.POINTER(sz)]
    L.hd_proposal_get_plane.restype = sz
    L.hd_proposal_scores.argtypes = [vp, fp, sz]
    L.hd_proposal_scores.restype = sz
    L.hd_get_last_proposal.argtypes = [vp, C.POINTER(C.c_uint8), sz, C.POINTER(sz)]
    L.hd_get_last_proposal.restype = sz
    _bound = True
    return L


def _check(err, what):
    if err != HD_NO_ERROR:
        raise HdError(err, what)


def _fptr(a):
    return a.ctypes.data_as(C.POINTER(C.c_float))


def param_count():
    return lib().hd_cnn_param_count()


class Cnn:
    """One hd_cnn on `gpu_id` for batches of at most `max_batch` tiles."""

    def __init__(self, max_batch, gpu_id=0, seed=1):
        self.handle = C.c_void_p()
        _check(lib().hd_cnn_create(C.byref(self.handle), gpu_id, max_batch), "hd_cnn_create")
        self.max_batch = max_batch
        if seed != 1:
            _check(lib().hd_cnn_init_params(self.handle, seed), "hd_cnn_init_params")

    def close(self):
        if self.handle:
            lib().hd_cnn_destroy(self.handle)
            self.handle = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_params(self, p):
        p = np.ascontiguousarray(p, np.float32)
        assert p.size == param_count()
        _check(lib().hd_cnn_set_params(self.handle, _fptr(p)), "hd_cnn_set_params")

    def params(self):
        out = np.zeros(param_count(), np.float32)
        _check(lib().hd_cnn_get_params(self.handle, _fptr(out)), "hd_cnn_get_params")
        return out

    def forward(self, d_tiles, n):
        """Logits of n device tiles (address of bfloat16 [n][64][64][2]) -> numpy."""
        _check(lib().hd_cnn_forward(self.handle, C.c_void_p(int(d_tiles)), n, None),
               "hd_cnn_forward")
        out = np.zeros(n, np.float32)
        _check(lib().hd_cnn_get_logits(self.handle, _fptr(out), n), "hd_cnn_get_logits")
        return out

    def forward_device(self, d_tiles, n):
        _check(lib().hd_cnn_forward(self.handle, C.c_void_p(int(d_tiles)), n, None),
               "hd_cnn_forward")

    def backward(self, d_labels, d_weights):
        loss = C.c_float(0)
        _check(lib().hd_cnn_backward(self.handle, C.c_void_p(int(d_labels)),
                                     C.c_void_p(int(d_weights)), C.byref(loss)),
               "hd_cnn_backward")
        return loss.value

    def grads(self):
        out = np.zeros(param_count(), np.float32)
        _check(lib().hd_cnn_get_grads(self.handle, _fptr(out)), "hd_cnn_get_grads")
        return out

    def adam(self, lr, beta1=0.9, beta2=0.999, eps=1e-8, weight_decay=0.0):
        _check(lib().hd_cnn_adam(self.handle, lr, beta1, beta2, eps, weight_decay),
               "hd_cnn_adam")


def set_proposal(pipeline, params, threshold):
    """Gating on (params: float32 network parameters) or off (params None)."""
    if params is None:
        _check(lib().hd_set_proposal(pipeline.handle, None, 0.0), "hd_set_proposal")
        return
    p = np.ascontiguousarray(params, np.float32)
    pipeline._proposal_params = p
    _check(lib().hd_set_proposal(pipeline.handle, _fptr(p), float(threshold)),
           "hd_set_proposal")


def coarse(pipeline, d_fb, nsamps, nbits):
    """Coarse plane of one single-beam device gulp; returns (n_units, n_tpos)."""
    nu, nv = C.c_size_t(0), C.c_size_t(0)
    _check(lib().hd_proposal_coarse(pipeline.handle, C.c_void_p(int(d_fb)), nsamps, nbits,
                                    C.byref(nu), C.byref(nv)), "hd_proposal_coarse")
    return nu.value, nv.value


def extract(pipeline, positions, d_tiles):
    """Tiles at (u, v) positions of the last coarse plane into device memory."""
    pos = np.ascontiguousarray(np.asarray(positions, np.uint32).reshape(-1))
    n = pos.size // 2
    _check(lib().hd_proposal_extract(pipeline.handle, pos.ctypes.data_as(C.POINTER(C.c_uint)),
                                     n, C.c_void_p(int(d_tiles))), "hd_proposal_extract")
    return n


def plane(pipeline, channel=0):
    """The last coarse plane (channel 0) or its wide plane (channel 1) as
    bfloat16 bits [rows, cols] (uint16)."""
    L = lib()
    rows, cols = C.c_size_t(0), C.c_size_t(0)
    n = L.hd_proposal_get_plane(pipeline.handle, channel, None, 0, C.byref(rows), C.byref(cols))
    out = np.zeros(n, np.uint16)
    if n:
        L.hd_proposal_get_plane(pipeline.handle, channel,
                                out.ctypes.data_as(C.POINTER(C.c_uint16)), n, C.byref(rows),
                                C.byref(cols))
    return out.reshape(rows.value, cols.value)


def scores(pipeline, n_units, n_tpos):
    """Logits of every tile of the last coarse plane, [n_units, n_tpos]."""
    out = np.zeros(n_units * n_tpos, np.float32)
    n = lib().hd_proposal_scores(pipeline.handle, _fptr(out), out.size)
    if n != out.size:
        raise RuntimeError("hd_proposal_scores returned %d of %d" % (n, out.size))
    return out.reshape(n_units, n_tpos)


def last_proposal(pipeline):
    """Units searched by the last gated call: uint8 [gulps, units_per_gulp]."""
    L = lib()
    per = C.c_size_t(0)
    n = L.hd_get_last_proposal(pipeline.handle, None, 0, C.byref(per))
    out = np.zeros(n, np.uint8)
    if n:
        L.hd_get_last_proposal(pipeline.handle, out.ctypes.data_as(C.POINTER(C.c_uint8)), n,
                               C.byref(per))
    return out.reshape(-1, per.value) if per.value else out.reshape(0, 0)
