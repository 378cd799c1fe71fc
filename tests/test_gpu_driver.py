"""The gulp driver (bin/heimdall) on the GPU: a sigproc file searched gulp by
gulp with overlap-save must write, for every gulp, the .cand file of the
oracle run on that gulp's samples."""
import calendar
import os
import subprocess
import time

import numpy as np
import pytest

from tests._cases import TINY
from tests.test_formats import sigproc_header

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "heimdall_multibeam_amd", "bin", "heimdall")


def test_driver_overlap_save_matches_per_gulp_oracle(oracle, tmp_path):
    from heimdall_multibeam_amd import synth
    w = TINY
    G = 8192
    total = 3 * G + 3000
    data = synth.make_gulp(w, 31, nsamps=total).reshape(total, -1)
    tstart = 55000.5
    head = sigproc_header([("nchans", w.nchans), ("nbits", w.nbits), ("tsamp", w.dt),
                           ("fch1", w.f0), ("foff", w.df), ("tstart", tstart), ("ibeam", 2),
                           ("nifs", 1)])
    fil = tmp_path / "in.fil"
    fil.write_bytes(head + data.tobytes())
    out = tmp_path / "cands"
    out.mkdir()
    env = dict(os.environ, TZ="UTC")
    cmd = [EXE, "-f", str(fil), "-nsamps_gulp", str(G), "-dm", str(w.dm_min), str(w.dm_max),
           "-boxcar_max", str(w.boxcar_max), "-output_dir", str(out)]
    print("running", " ".join(cmd), flush=True)
    r = subprocess.run(cmd, env=env, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    # the same loop on the oracle (heimdall.C:176-292, one beam)
    p = oracle.make_params(nchans=w.nchans, dt=w.dt, f0=w.f0, df=w.df, dm_min=w.dm_min,
                           dm_max=w.dm_max, boxcar_max=w.boxcar_max, nthreads=8, beam=1)
    utc = calendar.timegm((2009, 6, 18, 12, 0, 0)) + 1  # tstart, with the extra second
    sps = int(1.0 / float(np.float32(w.dt)))
    start, overlap, expected = 0, 0, {}
    while start + overlap + G <= total:
        n = G + overlap
        fb = np.ascontiguousarray(data[start:start + n]).ravel()
        res = oracle.execute(p, fb, n, w.nbits, start, 1)
        name = time.strftime("%Y-%m-%d-%H:%M:%S", time.gmtime(utc + start // sps)) + "_02.cand"
        expected[name] = res.text
        overlap = n - res.nsamps_processed
        start += res.nsamps_processed
    got = {f: open(out / f).read() for f in os.listdir(out)}
    assert sorted(got) == sorted(expected)
    for k in expected:
        assert got[k] == expected[k], k
    assert sum(len(t) for t in expected.values()) > 0
