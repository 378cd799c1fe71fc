"""bin/candidate_profiler (Applications/candidate_profiler.C) on a synthetic
sigproc file with a dispersed pulse: its four output files equal the CPU
restatement (oracle/candidate_profiler.py) text for text, binned and
boxcar-filtered.  The DM window spans several time-scrunch factors, and the
frequency-time window is widened to the pulse's delay."""
import os
import subprocess

import numpy as np
import pytest

from tests.conftest import gpu_available
from tests.test_formats import sigproc_header

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "heimdall_multibeam_amd", "bin", "candidate_profiler")
f32 = np.float32


def _file(tmp_path):
    from oracle import oracle as O
    h = dict(nchans=256, nbits=8, tsamp=32e-6, fch1=1500.0, foff=-1.0)
    nsamps = 40000
    rng = np.random.default_rng(21)
    x = np.clip(rng.normal(120, 18, (nsamps, h["nchans"])), 0, 255)
    dms = O.dm_list(0.0, 1000.0, f32(h["tsamp"]), 40.0, f32(h["fch1"]), f32(h["foff"]),
                    h["nchans"], 1.25)
    kc = O.delay_table(h["nchans"], f32(h["tsamp"]), f32(h["fch1"]), f32(h["foff"]))
    i = len(dms) * 2 // 3
    t0 = 5000
    d = (dms[i] * kc + f32(0.5)).astype(np.int64)
    for c in range(h["nchans"]):
        x[t0 + d[c]:t0 + d[c] + 6, c] += 60
    data = np.clip(x, 0, 255).astype(np.uint8).reshape(-1)
    fields = [("source_name", "FAKE"), ("nchans", h["nchans"]), ("nbits", 8),
              ("tsamp", h["tsamp"]), ("fch1", h["fch1"]), ("foff", h["foff"]),
              ("tstart", 60000.0), ("nifs", 1)]
    path = tmp_path / "cand.fil"
    path.write_bytes(sigproc_header(fields) + data.tobytes())
    (tmp_path / "dms.txt").write_text("".join("%.9g\n" % v for v in dms))
    return h, data, dms, i, t0


@pytest.mark.parametrize("do_filter", [0, 1])
def test_profiler_outputs_match_restatement(tmp_path, do_filter):
    if not gpu_available():
        pytest.skip("no GPU")
    from oracle import candidate_profiler as P
    from oracle import adaptive_dt as A
    assert os.path.exists(BIN), "bin/candidate_profiler not built"
    h, data, dms, i, t0 = _file(tmp_path)
    filt, out_nsamps, out_dm_count, fscrunch = 3, 64, 24, 4
    f = A.scrunch_list(dms[i - 12:i + 12], f32(h["tsamp"]), 40.0, f32(h["fch1"]),
                       f32(h["foff"]), h["nchans"], 1.15)
    assert len(set(f.tolist())) >= 2, f           # the window spans scrunch factors
    args = [BIN, "cand.fil", "dms.txt", str(t0), str(filt), str(i), str(out_nsamps),
            str(out_dm_count), str(fscrunch), str(do_filter)]
    print("case candidate_profiler do_filter %d" % do_filter, flush=True)
    r = subprocess.run(args, cwd=tmp_path, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr[-2000:]
    want = P.outputs(h, data, dms, t0, filt, i, out_nsamps, out_dm_count, fscrunch,
                     do_filter=bool(do_filter))
    assert "in_nsamps=" in r.stdout
    for name, text in want.items():
        got = (tmp_path / name).read_text()
        if got != text:
            g, w = got.splitlines(), text.splitlines()
            bad = [k for k in range(min(len(g), len(w))) if g[k] != w[k]]
            raise AssertionError("%s differs (%d vs %d lines); first line %s:\n%s\n%s" % (
                name, len(g), len(w), bad[:1], g[bad[0]][:300] if bad else "",
                w[bad[0]][:300] if bad else ""))
    # the pulse peaks in S/N against DM at its own trial
    rows = [ln.split("\t") for ln in (tmp_path / "snr_dm.dat").read_text().splitlines()[1:]]
    best = max(rows, key=lambda r: float(r[2]))
    assert abs(int(best[0]) - i) <= 2 and float(best[2]) > 8, best
