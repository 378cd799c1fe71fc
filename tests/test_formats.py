"""The sigproc reader, the MJD conversion and the DM-list tool (CPU only)."""
import ctypes as C
import os
import struct
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "heimdall_multibeam_amd", "bin")
GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


def sigproc_header(fields):
    """Bytes of a sigproc header: length-prefixed keywords, each followed by
    its value (int32, float64 or a length-prefixed string)."""
    def s(x):
        b = x.encode()
        return struct.pack("<i", len(b)) + b
    out = s("HEADER_START")
    for k, v in fields:
        out += s(k)
        if isinstance(v, str):
            out += s(v)
        elif isinstance(v, float):
            out += struct.pack("<d", v)
        else:
            out += struct.pack("<i", v)
    return out + s("HEADER_END")


class Hdr(C.Structure):
    _fields_ = [("tstart", C.c_double), ("tsamp", C.c_double), ("fch1", C.c_double),
                ("foff", C.c_double), ("nchans", C.c_int), ("nbits", C.c_int), ("nifs", C.c_int),
                ("ibeam", C.c_int), ("nbeams", C.c_int), ("nsamples", C.c_int),
                ("header_size", C.c_int), ("source_name", C.c_char * 80)]


def _lib():
    import heimdall_multibeam_amd as H
    L = H.lib()
    L.hd_sigproc_read_header.argtypes = [C.c_char_p, C.POINTER(Hdr)]
    L.hd_sigproc_read_header.restype = C.c_int
    L.hd_mjd_to_utc.argtypes = [C.c_double]
    L.hd_mjd_to_utc.restype = C.c_long
    return L


def test_header_fields_and_size(tmp_path):
    fields = [("source_name", "J0000+0000"), ("telescope_id", 4), ("nchans", 64), ("nbits", 8),
              ("tsamp", 256e-6), ("fch1", 1500.0), ("foff", -2.0), ("tstart", 60000.25),
              ("ibeam", 3), ("nbeams", 13), ("nifs", 1), ("az_start", 1.5), ("signed", "")]
    # "signed" carries one byte; write it by hand
    head = sigproc_header(fields[:-1])[:-len(struct.pack("<i", 10) + b"HEADER_END")]
    head += struct.pack("<i", 6) + b"signed" + b"\x00"
    head += struct.pack("<i", 10) + b"HEADER_END"
    data = np.arange(100 * 64, dtype=np.uint8)
    p = tmp_path / "a.fil"
    p.write_bytes(head + data.tobytes())
    h = Hdr()
    assert _lib().hd_sigproc_read_header(str(p).encode(), C.byref(h)) == 0
    assert (h.nchans, h.nbits, h.ibeam, h.nbeams, h.nifs) == (64, 8, 3, 13, 1)
    assert (h.tsamp, h.fch1, h.foff, h.tstart) == (256e-6, 1500.0, -2.0, 60000.25)
    assert h.source_name == b"J0000+0000"
    assert h.header_size == len(head)
    # nsamples from the file size, header bytes included (header.h:189-195)
    assert h.nsamples == (len(head) + data.nbytes) // 64 * 8 // 8


def test_header_rejects_non_sigproc(tmp_path):
    p = tmp_path / "b.fil"
    p.write_bytes(b"\x00" * 64)
    h = Hdr()
    assert _lib().hd_sigproc_read_header(str(p).encode(), C.byref(h)) == -2
    assert _lib().hd_sigproc_read_header(b"/nonexistent/x.fil", C.byref(h)) == -1


def test_mjd_conversion_matches_reference_rounding():
    """DataSource.h:49-80 in UTC: the second is always incremented."""
    import time
    old = os.environ.get("TZ")
    os.environ["TZ"] = "UTC"
    time.tzset()
    try:
        L = _lib()
        # MJD 55000.5 = 2009-06-18 12:00:00 UTC, plus the reference's extra second
        assert L.hd_mjd_to_utc(55000.5) == 1245326400 + 1
        assert L.hd_mjd_to_utc(40587.0) == 0 + 1   # the Unix epoch
    finally:
        if old is None:
            del os.environ["TZ"]
        else:
            os.environ["TZ"] = old
        time.tzset()


def test_generate_dmlist_tool_matches_fixture():
    exe = os.path.join(BIN, "generate_dmlist")
    if not os.path.exists(exe):
        pytest.skip("tools not built")
    out = subprocess.run([exe, "-f0", "1581.804688", "-df", "-0.390625", "-nchan", "1024",
                          "-dt", "64e-6", "-dm", "0", "2000"], capture_output=True, text=True,
                         check=True).stdout.split()
    ref = open(os.path.join(GOLDEN, "dm2000.list")).read().split()
    assert out[0] == "0"          # trial 0 (the fixture starts at trial 1)
    assert out[1:] == ref
