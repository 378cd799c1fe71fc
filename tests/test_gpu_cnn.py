"""The proposal network's HIP kernels (MFMA forward, backward, Adam) against
the CPU oracle (oracle/cnn.py), which rounds to bfloat16 at the same points.

Tolerances: float32 accumulation order differs between the MFMA tiles and the
oracle (float64 sums rounded to float32), and a bfloat16 activation that sits
on a rounding boundary can round the other way; both are far below these
bounds:
  logits            |dz| <= 2e-3 * (1 + |z|)
  weight gradients  relative L2 error <= 1e-2 per tensor, cosine >= 0.9995
  Adam              1e-6 relative against the update written out in numpy
"""
import ctypes as C

import numpy as np
import pytest

from tests.conftest import gpu_available

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def P():
    if not gpu_available():
        pytest.skip("no GPU")
    from heimdall_multibeam_amd import proposal
    return proposal


def _params(seed, P):
    from oracle import cnn
    rng = np.random.default_rng(seed)
    d = {}
    for name, shp in cnn.SHAPES:
        fan_in = {"w1": 18, "fc1_w": 512, "fc2_w": 64}.get(name, 288)
        if name.startswith("b") or name.endswith("_b"):
            d[name] = rng.normal(0, 0.05, shp)
        else:
            d[name] = rng.normal(0, 1.3 * np.sqrt(2.0 / fan_in), shp)
    p = cnn.flatten(d).astype(np.float32)
    assert p.size == P.param_count()
    return p


def _tiles(seed, n):
    from oracle import cnn
    rng = np.random.default_rng(seed)
    x = rng.normal(0, 1, (n, 64, 64, 2)).astype(np.float32)
    for b in range(0, n, 2):
        y, t = rng.integers(8, 56, 2)
        x[b, y - 3:y + 3, t - 2:t + 2, 0] += 3.0
        x[b, y - 3:y + 3, t // 4 + 20:t // 4 + 22, 1] += 2.0
    return cnn.bf16(x), cnn.bf16_bits(x)


def _dev(a):
    from heimdall_multibeam_amd.hip import DeviceBuffer
    a = np.ascontiguousarray(a)
    d = DeviceBuffer(a.nbytes)
    d.upload(a)
    return d


def test_forward_matches_oracle(P):
    from oracle import cnn
    n = 96
    p = _params(11, P)
    xf, xb = _tiles(12, n)
    net = P.Cnn(max_batch=128)
    net.set_params(p)
    np.testing.assert_array_equal(net.params(), p)
    dx = _dev(xb)
    z = net.forward(dx.address, n)
    zo, _ = cnn.forward(p, xf, quantize=True)
    err = np.abs(z - zo)
    print("logit max err", err.max(), "range", zo.min(), zo.max())
    assert np.all(err <= 2e-3 * (1 + np.abs(zo)))
    # partial batches use the same kernels with a ragged last tile row
    z2 = net.forward(dx.address, 37)
    np.testing.assert_array_equal(z2, z[:37])
    net.close()


def test_backward_matches_oracle(P):
    from oracle import cnn
    n = 80
    p = _params(21, P)
    xf, xb = _tiles(22, n)
    rng = np.random.default_rng(23)
    lab = (rng.random(n) < 0.4).astype(np.float32)
    wt = rng.uniform(0.5, 2.0, n).astype(np.float32)
    net = P.Cnn(max_batch=n)
    net.set_params(p)
    dx, dl, dw = _dev(xb), _dev(lab), _dev(wt)
    net.forward(dx.address, n)
    loss = net.backward(dl.address, dw.address)
    g = net.grads()
    _, cache = cnn.forward(p, xf, quantize=True)
    go, loss_o = cnn.backward(cache, lab, wt)
    assert abs(loss - loss_o) <= 1e-3 * (1 + abs(loss_o))
    off = 0
    for name, shp in cnn.SHAPES:
        k = int(np.prod(shp))
        a, b = g[off:off + k].astype(np.float64), go[off:off + k]
        off += k
        nb = np.linalg.norm(b)
        rel = np.linalg.norm(a - b) / max(nb, 1e-30)
        cos = float(a @ b / max(np.linalg.norm(a) * nb, 1e-30))
        print(name, "rel", rel, "cos", cos, "norm", nb)
        assert nb > 0, name
        assert rel <= 1e-2 and cos >= 0.9995, name
    net.close()


def test_adam_step(P):
    n = 32
    p = _params(31, P)
    _, xb = _tiles(32, n)
    lab = np.ones(n, np.float32)
    wt = np.ones(n, np.float32)
    net = P.Cnn(max_batch=n)
    net.set_params(p)
    dx, dl, dw = _dev(xb), _dev(lab), _dev(wt)
    m = np.zeros_like(p, dtype=np.float64)
    v = np.zeros_like(p, dtype=np.float64)
    q = p.astype(np.float64)
    lr, b1, b2, eps, wd = 1e-3, 0.9, 0.999, 1e-8, 1e-4
    for step in (1, 2):
        net.forward(dx.address, n)
        net.backward(dl.address, dw.address)
        g = net.grads().astype(np.float64)
        net.adam(lr, b1, b2, eps, wd)
        m = b1 * m + (1 - b1) * g
        v = b2 * v + (1 - b2) * g * g
        upd = (m / (1 - b1 ** step)) / (np.sqrt(v / (1 - b2 ** step)) + eps)
        q = q - lr * (upd + wd * q)
        got = net.params()
        np.testing.assert_allclose(got, q, rtol=1e-5, atol=1e-7)
        q = got.astype(np.float64)
    net.close()


def test_training_reduces_loss(P):
    """A few hundred Adam steps of the project's kernels separate tiles that
    hold a bright blob from noise tiles."""
    from oracle import cnn
    rng = np.random.default_rng(41)
    n = 256
    x = rng.normal(0, 1, (n, 64, 64, 2)).astype(np.float32)
    lab = (np.arange(n) % 2).astype(np.float32)
    for b in np.nonzero(lab)[0]:
        y, t = rng.integers(4, 60, 2)
        x[b, y - 2:y + 2, t - 2:t + 2, 0] += 2.5
    xb = cnn.bf16_bits(x)
    net = P.Cnn(max_batch=n, seed=5)
    dx, dl, dw = _dev(xb), _dev(lab), _dev(np.ones(n, np.float32))
    losses = []
    for it in range(150):
        net.forward(dx.address, n)
        losses.append(net.backward(dl.address, dw.address))
        net.adam(2e-3)
    z = net.forward(dx.address, n)
    acc = float(np.mean((z > 0) == (lab > 0.5)))
    print("loss", losses[0], "->", losses[-1], "accuracy", acc)
    assert losses[-1] < 0.25 * losses[0]
    assert acc >= 0.95
    net.close()
