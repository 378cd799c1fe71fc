"""The candidate wire path: hd_execute's TCP messages to a coincidencer
(Pipeline/pipeline.cu:907-995) and the coincidencer fed by them.

  * the header "utc_start first_sample_utc first_idx beam nbeams n_events" and
    the event lines received on a listening socket equal the oracle's
    candidate lines for the same gulp (the .cand lines without their trailing
    tab);
  * two pipelines (beams 1 and 2 of one gulp) send to bin/coincidencer, whose
    <utc_start>/<first_sample_utc>_all.cand equals the CPU restatement
    (oracle/coincidence.py) of the oracle's candidates of both beams."""
import os
import socket
import subprocess
import threading
import time

import pytest

from tests._cases import SMALL_128, oracle_params_for, params_for
from tests.conftest import gpu_available

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "heimdall_multibeam_amd", "bin", "coincidencer")
UTC_START = 1796374800          # 2026-12-04-09:00:00 UTC


def _utc(t):
    return time.strftime("%Y-%m-%d-%H:%M:%S", time.gmtime(t))


def _params(w, port, beam):
    sps = int(1.0 / w.dt)
    return params_for(w, coincidencer_host="127.0.0.1", coincidencer_port=port,
                      utc_start=UTC_START, spectra_per_second=sps, beam=beam)


def _oracle_lines(oracle, w, fb, first_idx, beam):
    r = oracle.execute(oracle_params_for(oracle, w, beam=beam), fb, w.nsamps, w.nbits, first_idx,
                       1)
    return [ln.rstrip("\t") for ln in r.text.splitlines()]


@pytest.fixture(scope="module")
def H():
    if not gpu_available():
        pytest.skip("no GPU")
    import heimdall_multibeam_amd as H
    return H


def test_tcp_messages_match_oracle_lines(H, oracle):
    from heimdall_multibeam_amd import synth
    w = SMALL_128
    fb = synth.make_gulp(w, 3)
    srv = socket.socket()
    srv.bind(("127.0.0.1", 0))
    srv.listen(4)
    srv.settimeout(60)
    port = srv.getsockname()[1]
    got = []

    def serve():
        c, _ = srv.accept()
        data = b""
        while True:
            b = c.recv(65536)
            if not b:
                break
            data += b
        c.close()
        got.append(data.decode())

    t = threading.Thread(target=serve)
    t.start()
    sps = int(1.0 / w.dt)
    first_idx = 10 * sps
    print("case tcp small_128ch seed 3 port %d" % port, flush=True)
    pl = H.Pipeline(_params(w, port, 0))
    pl.set_output(True)
    pl.execute(fb, w.nsamps, w.nbits, first_idx, 1)
    pl.close()
    t.join(60)
    srv.close()
    assert got, "no connection"
    lines = got[0].splitlines()
    want = _oracle_lines(oracle, w, fb, first_idx, 0)
    assert len(want) > 0
    assert lines[0].split() == [_utc(UTC_START), _utc(UTC_START + first_idx // sps),
                                str(first_idx), "0", "1", str(len(want))]
    assert lines[1:] == want


def test_heimdall_to_coincidencer(H, oracle, tmp_path):
    from heimdall_multibeam_amd import synth
    from oracle import coincidence as K
    if not os.path.exists(BIN):
        pytest.skip("bin/coincidencer not built")
    w = SMALL_128
    (tmp_path / _utc(UTC_START)).mkdir()
    s = socket.socket()
    s.bind(("127.0.0.1", 0))
    port = s.getsockname()[1]
    s.close()
    proc = subprocess.Popen([BIN, "-p", str(port), "-a", "127.0.0.1", "-n", "2"], cwd=tmp_path,
                            stderr=subprocess.PIPE, text=True)
    try:
        for _ in range(100):
            try:
                socket.create_connection(("127.0.0.1", port), timeout=1).close()
                break
            except OSError:
                time.sleep(0.05)
        sps = int(1.0 / w.dt)
        first_idx = 5 * sps
        lines = []
        # the same sky in both beams (one gulp rendered twice with different
        # noise), so that their candidates coincide
        for beam, seed in ((0, 11), (1, 12)):
            fb = synth.make_gulp(w, seed, pulses=[(9000.0, 200.0, 8.0, 25.0),
                                                  (20000.0, 350.0, 3.0, 15.0)])
            print("case heimdall -> coincidencer beam %d" % (beam + 1), flush=True)
            pl = H.Pipeline(_params(w, port, beam))
            pl.set_output(True)
            pl.execute(fb, w.nsamps, w.nbits, first_idx, 1)
            pl.close()
            lines += _oracle_lines(oracle, w, fb, first_idx, beam)
        out = tmp_path / _utc(UTC_START) / ("%s_all.cand" % _utc(UTC_START + first_idx // sps))
        for _ in range(200):
            if out.exists() and out.stat().st_size > 0:
                break
            time.sleep(0.05)
        time.sleep(0.2)
        got = out.read_text()
    finally:
        proc.terminate()
        proc.wait(timeout=20)
    want = K.coincide_lines([ln + "\n" for ln in lines], 2)
    assert got == want
    assert "\t2\t\t" in got                     # candidates seen in both beams
