"""The stage chain as a HIP graph: every replay equals the eager launches.

Each check replays a captured graph several times with eager work and
synchronisations of another pipeline in between, and compares every replay
with the one-kernel-at-a-time result bit for bit (giants, candidates, .cand
text).  A replay on new input of the same shape must give the new input's
result (the graph reads the unpacked input, it holds no data).
"""
import numpy as np
import pytest

from tests._cases import SMALL_128, TINY, params_for
from tests.conftest import gpu_available

pytestmark = pytest.mark.gpu


def _result(pl):
    g, c = pl.giants(), pl.candidates()
    return ({k: v.copy() for k, v in g.items()}, {k: v.copy() for k, v in c.items()},
            pl.cand_text())


def _same(a, b):
    for da, db in zip(a[:2], b[:2]):
        assert da.keys() == db.keys()
        for k in da:
            np.testing.assert_array_equal(da[k], db[k], err_msg=k)
    assert a[2] == b[2]


def test_graph_replays_equal_eager():
    if not gpu_available():
        pytest.skip("no GPU")
    import heimdall_multibeam_amd as H
    from heimdall_multibeam_amd import synth
    w = SMALL_128
    fbs = [synth.make_gulp(w, 60 + k) for k in range(3)]
    eager = H.Pipeline(params_for(w))
    eager.set_output(False)
    eager.set_graph(False)
    want = []
    for fb in fbs:
        eager.execute(fb, w.nsamps, 8)
        want.append(_result(eager))
    assert len(want[0][1]["peak"]) > 0
    pl = H.Pipeline(params_for(w))
    pl.set_output(False)
    pl.set_graph(True)
    order = [0, 0, 1, 2, 1, 0]
    for step, k in enumerate(order):
        print("case graph replay", step, "gulp", k, flush=True)
        pl.execute(fbs[k], w.nsamps, 8)
        _same(_result(pl), want[k])
        eager.execute(fbs[(k + 1) % 3], w.nsamps, 8)   # eager work in between
    captures, launches = pl.graph_stats()
    print("captures", captures, "launches", launches)
    assert captures == 1 and launches == len(order)
    assert eager.graph_stats() == (0, 0)
    pl.close()
    eager.close()


def test_graph_batch_replays_equal_eager():
    if not gpu_available():
        pytest.skip("no GPU")
    import heimdall_multibeam_amd as H
    from heimdall_multibeam_amd import synth
    w = TINY
    gulps = [(synth.make_gulp(w, 70 + k, nsamps=ns), ns, 1, 0, k)
             for k, ns in enumerate([w.nsamps, w.nsamps - 1500, w.nsamps + 2048])]
    eager = H.Pipeline(params_for(w))
    eager.set_output(False)
    eager.set_graph(False)
    eager.execute_batch(gulps, 8)
    want = _result(eager)
    pl = H.Pipeline(params_for(w))
    pl.set_output(False)
    for rep in range(4):
        pl.execute_batch(gulps, 8)
        _same(_result(pl), want)
        eager.execute(gulps[rep % 3][0], gulps[rep % 3][1], 8)
    assert pl.graph_stats() == (1, 4)
    pl.close()
    eager.close()
