"""What the trained proposal network gains over the method without it.

On gated_4096ch beams the training never saw, every unit (128 trials) gets
the network's score (its best tile) and the score of the coarse-plane method
without a network (the largest normalised boxcar of 1..64 coarse samples in
the unit's rows).  For each recall target the threshold is the largest that
keeps that share of the full search's S/N >= 8 candidates in proposed units;
the test compares the share of units each method then proposes (the work
the gated search does).  The network must propose fewer units on average
over the targets 0.97, 0.98 and 0.99.
"""
import numpy as np
import pytest

from tests.conftest import gpu_available

pytestmark = pytest.mark.gpu


def test_network_proposes_fewer_units_than_coarse_boxcar():
    if not gpu_available():
        pytest.skip("no GPU")
    import os
    from heimdall_multibeam_amd import train_proposal as T
    path = os.path.join(os.path.dirname(T.__file__), "data", "proposal.npz")
    z = np.load(path, allow_pickle=False)
    st = T.Setup("gated_4096ch")
    n = 60
    base = 1000 * 1000003 + 2 * 100000 + 50000   # disjoint from the training seeds
    seeds = [base + i for i in range(n)]
    kinds = ["noise" if i % 10 == 0 else ("rfi" if i % 10 == 1 else "pulses") for i in range(n)]
    ev = T.evaluate(st, seeds, kinds, z["params"])
    ncand = sum(len(e["units"]) for e in ev)
    print("held-out beams", n, "candidates >= 8:", ncand, "ideal unit fraction",
          round(T.ideal_fraction(ev), 4))
    fn, fh = [], []
    for r in (0.97, 0.98, 0.99):
        a = T.threshold_for(ev, "net", r)
        b = T.threshold_for(ev, "heur", r)
        print("recall target %.2f: network %.3f of units (recall %.3f), coarse boxcar %.3f "
              "(recall %.3f)" % (r, a[2], a[1], b[2], b[1]))
        assert a[1] >= r and b[1] >= r
        fn.append(a[2])
        fh.append(b[2])
    assert ncand >= 50
    assert np.mean(fn) < np.mean(fh)
