"""ctypes mirror of include/hd/stages.h: the search's stage functions over
device memory, one call per reference stage (dedisp_execute_adv,
RemoveBaselinePlan::exec, GetRMSPlanMB::exec_multibeam, normalise_multibeam,
MatchedFilterPlan::prep/exec, GiantFinder::exec, label_candidate_clusters,
merge_candidates).

Every function takes device addresses (ints, e.g. DeviceBuffer.address) and
raises HdError when the library returns an error.  `stream` is a HIP stream
handle or None for the default stream.
"""
import ctypes as C

from . import HD_NO_ERROR, HdError, lib as _pipeline_lib

sz = C.c_size_t
vp = C.c_void_p


class HdGiants(C.Structure):
    """hd_giants (include/hd/stages.h): device arrays, capacity, count."""
    _fields_ = [("peaks", vp), ("inds", vp), ("begins", vp), ("ends", vp), ("capacity", sz),
                ("count", sz)]


class RawCandidates(C.Structure):
    """RawCandidates / ConstRawCandidates (include/hd/types.h): device arrays."""
    _fields_ = [("peaks", vp), ("inds", vp), ("begins", vp), ("ends", vp),
                ("filter_inds", vp), ("dm_inds", vp), ("members", vp)]


_bound = False


def lib():
    global _bound
    L = _pipeline_lib()
    if not _bound:
        f = C.c_float
        L.hd_dedisperse.argtypes = [vp, sz, vp, sz, sz, vp, sz, sz, C.c_uint]
        L.hd_remove_baseline.argtypes = [vp, sz, sz, sz, sz, sz, sz, vp]
        L.hd_get_rms_multibeam.argtypes = [vp, vp, sz, sz, sz, sz, sz, vp]
        L.hd_normalise_multibeam.argtypes = [vp, vp, sz, sz, sz, sz, sz, vp]
        L.hd_block_rms.argtypes = [vp, vp, sz, sz, sz, sz, sz, vp]
        L.hd_normalise_block.argtypes = [vp, vp, sz, sz, sz, vp]
        L.hd_matched_filter_prep.argtypes = [vp, sz, vp, sz, sz, sz, vp]
        L.hd_matched_filter_exec.argtypes = [vp, sz, sz, sz, sz, vp, sz, sz, sz, vp]
        L.hd_find_giants.argtypes = [vp, sz, sz, f, sz, sz, sz, C.POINTER(HdGiants), vp]
        L.hd_label_candidate_clusters.argtypes = [sz, C.POINTER(RawCandidates), sz, sz, sz, sz,
                                                  vp, C.POINTER(sz), vp]
        L.hd_merge_candidates.argtypes = [sz, vp, C.POINTER(RawCandidates),
                                          C.POINTER(RawCandidates), vp]
        for name in ("hd_dedisperse", "hd_remove_baseline", "hd_get_rms_multibeam",
                     "hd_normalise_multibeam", "hd_block_rms", "hd_normalise_block",
                     "hd_matched_filter_prep", "hd_matched_filter_exec", "hd_find_giants",
                     "hd_label_candidate_clusters", "hd_merge_candidates"):
            getattr(L, name).restype = C.c_int
        _bound = True
    return L


def _check(rc, what):
    if rc != HD_NO_ERROR:
        raise HdError(rc, what)


def dedisperse(pipeline, nsamps, d_in, nbits, d_out, out_stride_bytes):
    """hd_dedisperse: the pipeline's trials over nsamps samples at d_in."""
    p = pipeline._params
    _check(lib().hd_dedisperse(pipeline.handle, nsamps, d_in, nbits, p.nchans * nbits // 8,
                               d_out, 32, out_stride_bytes, 0), "hd_dedisperse")


def remove_baseline(d, beam_stride, beam_count, smooth_radius, nbeams, n_series,
                    series_stride, stream=None):
    _check(lib().hd_remove_baseline(d, beam_stride, beam_count, smooth_radius, nbeams,
                                    n_series, series_stride, stream), "hd_remove_baseline")


def get_rms_multibeam(d, d_rms, beam_stride, beam_count, nbeams, n_series, series_stride,
                      stream=None):
    _check(lib().hd_get_rms_multibeam(d, d_rms, beam_stride, beam_count, nbeams, n_series,
                                      series_stride, stream), "hd_get_rms_multibeam")


def normalise_multibeam(d, d_rms, beam_stride, cur_nsamps, nbeams, n_series, series_stride,
                        stream=None):
    _check(lib().hd_normalise_multibeam(d, d_rms, beam_stride, cur_nsamps, nbeams, n_series,
                                        series_stride, stream), "hd_normalise_multibeam")


def block_rms(d, d_rms, beam_stride, beam_count, nbeams, n_series, series_stride,
              stream=None):
    _check(lib().hd_block_rms(d, d_rms, beam_stride, beam_count, nbeams, n_series,
                              series_stride, stream), "hd_block_rms")


def normalise_block(d, d_rms, count, n_series, series_stride, stream=None):
    _check(lib().hd_normalise_block(d, d_rms, count, n_series, series_stride, stream),
           "hd_normalise_block")


def matched_filter_prep(d_in, count, d_scanned, n_series, series_stride, scanned_stride,
                        stream=None):
    _check(lib().hd_matched_filter_prep(d_in, count, d_scanned, n_series, series_stride,
                                        scanned_stride, stream), "hd_matched_filter_prep")


def matched_filter_exec(d_scanned, scanned_count, max_width, width, tscrunch, d_out, n_series,
                        scanned_stride, out_stride, stream=None):
    _check(lib().hd_matched_filter_exec(d_scanned, scanned_count, max_width, width, tscrunch,
                                        d_out, n_series, scanned_stride, out_stride, stream),
           "hd_matched_filter_exec")


def find_giants(d, count, nbeams, thresh, merge_dist, beam_stride, beam_count, peaks, inds,
                begins, ends, capacity, stream=None):
    """Returns the number of giants found (only `capacity` are stored)."""
    g = HdGiants(peaks, inds, begins, ends, capacity, 0)
    _check(lib().hd_find_giants(d, count, nbeams, thresh, merge_dist, beam_stride, beam_count,
                                C.byref(g), stream), "hd_find_giants")
    return g.count


def label_candidate_clusters(count, cands, time_tol, filter_tol, dm_tol, nsamps_beam,
                             d_labels, stream=None):
    """cands: RawCandidates of device arrays.  Returns label_count."""
    n = sz(0)
    _check(lib().hd_label_candidate_clusters(count, C.byref(cands), time_tol, filter_tol,
                                             dm_tol, nsamps_beam, d_labels, C.byref(n), stream),
           "hd_label_candidate_clusters")
    return n.value


def merge_candidates(count, d_labels, cands, groups, stream=None):
    _check(lib().hd_merge_candidates(count, d_labels, C.byref(cands), C.byref(groups), stream),
           "hd_merge_candidates")


__all__ = ["HdGiants", "RawCandidates", "dedisperse", "remove_baseline", "get_rms_multibeam",
           "normalise_multibeam", "block_rms", "normalise_block", "matched_filter_prep",
           "matched_filter_exec", "find_giants", "label_candidate_clusters",
           "merge_candidates"]
