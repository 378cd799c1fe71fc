"""The multibeam coincidencer (bin/coincidencer, replacing Applications/
coincidencer.C + Candidates.C) against the CPU restatement oracle/
coincidence.py: in file mode over .cand files and in network mode over the
TCP messages heimdall sends (Pipeline/pipeline.cu:907-995).  The restatement
itself is pinned by hand-worked cases of the reference's rules."""
import os
import socket
import subprocess
import time

import numpy as np
import pytest

from oracle import coincidence as K

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "heimdall_multibeam_amd", "bin", "coincidencer")


def _line(snr, samp, filt, dmi, beam, members=5, dt=64e-6, dm=100.0):
    # heimdall's .cand line (pipeline.cu:1026-1038), trailing tab included
    return "%g\t%d\t%g\t%d\t%d\t%g\t%d\t%d\t%d\t%d\t\n" % (snr, samp, samp * dt, filt, dmi, dm,
                                                             members, samp - 2, samp + 2, beam)


def test_restatement_rules_by_hand():
    # tolerance sep_time * 2^max(filter): 3 * 2^2 = 12 samples
    a = K.parse_line(_line(10.0, 1000, 2, 50, 1))
    assert K.is_coincident(a, K.parse_line(_line(10.0, 1012, 1, 50, 2)), 3, 4)
    assert not K.is_coincident(a, K.parse_line(_line(10.0, 1013, 1, 50, 2)), 3, 4)
    # filters more than sep_filter apart never coincide
    assert not K.is_coincident(a, K.parse_line(_line(10.0, 1000, 7, 50, 2)), 3, 4)
    # relative S/N difference: |15 - 6| / 21 = 0.43 > 0.4; |14 - 6| / 20 = 0.4
    assert not K.is_coincident(K.parse_line(_line(6.0, 1000, 2, 50, 1)),
                               K.parse_line(_line(15.0, 1000, 2, 50, 2)), 3, 4)
    assert K.is_coincident(K.parse_line(_line(6.0, 1000, 2, 50, 1)),
                           K.parse_line(_line(14.0, 1000, 2, 50, 2)), 3, 4)
    # only the first coincident candidate of each beam counts: beam 2's
    # second (brighter) one does not become the primary
    beams = [[K.parse_line(_line(10.0, 1000, 2, 50, 1))],
             [K.parse_line(_line(11.0, 1001, 2, 50, 2)),
              K.parse_line(_line(13.0, 1002, 2, 50, 2))],
             [K.parse_line(_line(12.0, 999, 2, 50, 3))]]
    K.compute(beams)
    c = beams[0][0]
    assert c["nbeams"] == 3 and c["primary_beam"] == 3 and c["max_snr"] == np.float32(12.0)
    assert K.format_line(c) == "10\t1000\t0.064\t2\t50\t100\t5\t998\t1002\t3\t\t3\t12\t1"


def _random_lines(rng, nbeams, per_beam):
    lines = []
    centres = rng.integers(10000, 200000, 12)
    for b in range(1, nbeams + 1):
        for _ in range(per_beam):
            s = int(rng.choice(centres) + rng.integers(-40, 40))
            lines.append(_line(float(np.float32(rng.uniform(6, 30))), s, int(rng.integers(0, 10)),
                               int(rng.integers(0, 1400)), b, int(rng.integers(1, 400)),
                               dm=float(np.float32(rng.uniform(0, 2000)))))
    return lines


@pytest.fixture(scope="module")
def tool():
    if not os.path.exists(BIN):
        from heimdall_multibeam_amd import build
        build.build()
    return BIN


def test_file_mode_matches_restatement(tool, tmp_path):
    rng = np.random.default_rng(4)
    nb = 3
    lines = _random_lines(rng, nb, 60)
    utc = "2026-12-04-10:00:00"
    names = []
    for b in range(1, nb + 1):
        name = "%s_%02d.cand" % (utc, b)
        with open(tmp_path / name, "w") as f:
            f.write("".join(ln for ln in lines if ln.rstrip("\t\n").endswith("\t%d" % b)))
        names.append(name)
    r = subprocess.run([tool, "-n", str(nb)] + names, cwd=tmp_path, capture_output=True,
                       text=True, timeout=60)
    assert r.returncode == 0, r.stderr
    got = (tmp_path / ("%s_all.cand" % utc)).read_text()
    want = K.coincide_lines(lines, nb)
    assert got == want
    assert "\t\t2\t" in got or "\t\t3\t" in got     # some coincidences found


def _free_port():
    s = socket.socket()
    s.bind(("127.0.0.1", 0))
    p = s.getsockname()[1]
    s.close()
    return p


def _send(port, header, lines):
    with socket.create_connection(("127.0.0.1", port), timeout=10) as s:
        s.sendall((header + "\n").encode())
        s.sendall("".join(ln.rstrip("\t\n") + "\n" for ln in lines).encode())


def test_network_mode_matches_restatement(tool, tmp_path):
    """Two beam groups of one gulp arrive as heimdall's TCP messages; the
    coincidencer writes <utc_start>/<first_sample_utc>_all.cand."""
    rng = np.random.default_rng(5)
    nb = 2
    lines = _random_lines(rng, nb, 40)
    utc_start, first_utc = "2026-12-04-09:00:00", "2026-12-04-09:00:16"
    (tmp_path / utc_start).mkdir()
    port = _free_port()
    proc = subprocess.Popen([tool, "-p", str(port), "-a", "127.0.0.1", "-n", str(nb)],
                            cwd=tmp_path, stderr=subprocess.PIPE, text=True)
    try:
        for _ in range(100):                     # wait until it listens
            try:
                socket.create_connection(("127.0.0.1", port), timeout=1).close()
                break
            except OSError:
                time.sleep(0.05)
        # the probe connection above carried no header: ignored by the tool
        for b in range(1, nb + 1):
            mine = [ln for ln in lines if ln.rstrip("\t\n").endswith("\t%d" % b)]
            _send(port, "%s %s %d %d 1 %d" % (utc_start, first_utc, 262144, b - 1, len(mine)),
                  mine)
        out = tmp_path / utc_start / ("%s_all.cand" % first_utc)
        for _ in range(200):
            if out.exists() and out.stat().st_size > 0:
                break
            time.sleep(0.05)
        time.sleep(0.2)
        got = out.read_text()
    finally:
        proc.terminate()
        proc.wait(timeout=20)
    assert got == K.coincide_lines(lines, nb)
