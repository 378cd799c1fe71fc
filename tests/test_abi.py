"""The drop-in boundary on the CPU: the library loads, exports every entry
point the headers declare, and its host-side behaviour (params, defaults,
errors, command line, DM list, kill masks) matches the reference interface.
No GPU compute is called here."""
import ctypes as C
import glob
import os
import re
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
f32 = np.float32


def _declared_functions():
    names = set()
    for h in glob.glob(os.path.join(ROOT, "include", "hd", "*.h")):
        src = open(h).read()
        src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
        for m in re.finditer(r"^\s*[A-Za-z_][\w\s\*]*?\b(hd_\w+|throw_error)\s*\(", src, re.M):
            names.add(m.group(1))
    return names


def test_library_exports_every_declared_function():
    import heimdall_multibeam_amd as H
    L = H.lib()
    names = _declared_functions()
    assert {"hd_create_pipeline", "hd_execute", "hd_destroy_pipeline", "hd_find_giants",
            "hd_remove_baseline"} <= names
    missing = [n for n in sorted(names) if not hasattr(L, n)]
    assert not missing, missing


def test_params_layout_matches_c_header(tmp_path):
    """sizeof/offsetof of struct hd_params from the C header, with and without
    HAVE_PSRDADA, against the ctypes mirror."""
    import heimdall_multibeam_amd as H
    fields = [f for f, _ in H.HdParams._fields_ if f != "dada_id"]
    prog = tmp_path / "layout.c"
    body = "\n".join('  printf("%s %%zu\\n", offsetof(struct hd_params, %s));' % (f, f)
                     for f in fields)
    prog.write_text('#include <stdio.h>\n#include <stddef.h>\n#include "hd/params.h"\n'
                    "int main(void){\n" + body +
                    '\n  printf("sizeof %zu\\n", sizeof(struct hd_params));\n  return 0;\n}\n')
    for defs in ([], ["-DHAVE_PSRDADA"]):
        exe = tmp_path / ("layout%d" % len(defs))
        subprocess.check_call(["gcc", "-std=gnu11"] + defs + ["-I", os.path.join(ROOT, "include"),
                                                           str(prog), "-o", str(exe)])
        out = dict(l.split() for l in subprocess.check_output([str(exe)]).decode().splitlines())
        assert int(out["sizeof"]) == C.sizeof(H.HdParams)
        for f in fields:
            assert int(out[f]) == getattr(H.HdParams, f).offset, f


def test_default_params_match_reference():
    import heimdall_multibeam_amd as H
    p = H.default_params()
    # Pipeline/default_params.C:10-66
    assert p.nsamps_gulp == 262144 and p.baseline_length == f32(2.0)
    assert p.nchans == 1024 and p.dt == f32(64e-6) and p.f0 == f32(1581.804688)
    assert p.df == f32(-0.390625) and p.dm_max == 1000.0 and p.dm_tol == f32(1.25)
    assert p.dm_pulse_width == 40 and p.dm_nbits == 32 and p.use_scrunching
    assert p.boxcar_max == 4096 and p.detect_thresh == 6.0
    assert (p.cand_sep_time, p.cand_sep_filter, p.cand_sep_dm) == (3, 3, 200)
    assert p.min_tscrunch_width == 4096 and p.max_giant_rate == 0
    assert p.coincidencer_port == -1 and p.output_dir == b"."
    assert p.gpu_id == 0 and p.spectra_per_second == 0 and p.beam_count == 13


def test_error_strings():
    import heimdall_multibeam_amd as H
    assert H.error_string(H.HD_NO_ERROR) == "No error"
    assert H.error_string(H.HD_TOO_MANY_EVENTS) == "Too many events"
    assert H.error_string(H.HD_TOO_FEW_NSAMPS) == "No. samples < maximum delay"
    assert H.error_string(99) == "Invalid error code"
    assert H.HD_UNKNOWN_ERROR == 13


def test_parse_command_line():
    import heimdall_multibeam_amd as H
    L = H.lib()
    L.hd_parse_command_line.argtypes = [C.c_int, C.POINTER(C.c_char_p), C.POINTER(H.HdParams)]
    L.hd_parse_command_line.restype = C.c_int
    p = H.default_params()
    args = ["heimdall", "-f", "x.fil", "-dm", "0", "2000", "-boxcar_max", "1024", "-V",
            "-zap_chans", "10", "20", "-cand_sep_dm_trial", "17", "-beam", "3",
            "-coincidencer", "host.example:51041", "-min_tscrunch_width", "512"]
    argv = (C.c_char_p * len(args))(*[a.encode() for a in args])
    assert L.hd_parse_command_line(len(args), argv, C.byref(p)) == 0
    assert p.sigproc_file == b"x.fil" and p.dm_max == 2000.0 and p.boxcar_max == 1024
    assert p.verbosity == 2 and p.num_channel_zaps == 1
    assert (p.channel_zaps[0].start, p.channel_zaps[0].end) == (10, 20)
    assert p.cand_sep_dm == 17 and p.beam == 2 and p.override_beam
    assert p.coincidencer_host == b"host.example" and p.coincidencer_port == 51041
    assert p.min_tscrunch_width == 512


def test_dm_list_equals_oracle(oracle):
    import heimdall_multibeam_amd as H
    for args in [(0, 2000, 64e-6, 40, 1581.804688, -0.390625, 1024, 1.25),
                 (0, 2000, 64e-6, 40, 1517.8359375, -0.328125, 1024, 1.25),
                 (0, 2000, 655.36e-6, 40, 864.953125, -0.09375, 320, 1.25),
                 (0, 3000, 64e-6, 40, 1599.951171875, -0.09765625, 4096, 1.25),
                 (0, 500, 256e-6, 40, 1500.0, -0.5, 512, 1.25)]:
        a = [f32(v) if isinstance(v, float) or i < 6 or i == 7 else v for i, v in enumerate(args)]
        got = H.generate_dm_list(*a)
        ref = oracle.dm_list(*a)
        np.testing.assert_array_equal(got, ref)


def test_config_trial_counts():
    """DM trial counts of the measured 512-, 1024- and 320-channel workloads."""
    import heimdall_multibeam_amd as H
    cases = [((0, 500, 256e-6, 40, 1500.0, -0.5, 512, 1.25), 320),
             ((0, 2000, 64e-6, 40, 1517.8359375, -0.328125, 1024, 1.25), 1399),
             ((0, 2000, 655.36e-6, 40, 864.953125, -0.09375, 320, 1.25), 224)]
    for args, n in cases:
        assert len(H.generate_dm_list(*args)) == n


def test_manual_killmask_end_exclusive():
    import heimdall_multibeam_amd as H
    L = H.lib()
    L.hd_apply_manual_killmasks.argtypes = [C.POINTER(C.c_int), C.c_size_t, C.c_uint,
                                            C.POINTER(H.HdRange)]
    km = (C.c_int * 16)(*([1] * 16))
    zaps = (H.HdRange * 2)(H.HdRange(2, 5), H.HdRange(14, 40))
    assert L.hd_apply_manual_killmasks(km, 16, 2, zaps) == 0
    assert list(km) == [1, 1, 0, 0, 0, 1, 1, 1, 1, 1, 1, 1, 1, 1, 0, 0]


def test_create_pipeline_without_gpu_fails_loudly():
    """No silent CPU path: without a device the pipeline refuses to exist."""
    import heimdall_multibeam_amd as H
    from tests.conftest import gpu_available
    if gpu_available():
        pytest.skip("a GPU is present")
    with pytest.raises(H.HdError) as e:
        H.Pipeline(H.default_params())
    assert e.value.code == H.HD_INVALID_DEVICE_INDEX
