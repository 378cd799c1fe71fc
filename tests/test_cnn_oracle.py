"""Pins of the proposal network's CPU oracle (oracle/cnn.py).

What fixes the oracle without trusting its own formulas:
  * bfloat16 rounding: hand-checked bit patterns (ties to even, carries);
  * forward: one output of every layer recomputed by a direct loop over the
    3x3 window at stride 2 with zero padding (an off-by-one in the padding,
    a swapped kernel axis or a transposed weight fails it);
  * backward: central finite differences of the float64 loss, for a sample
    of weights of every tensor and every bias (a dropped ReLU mask, a wrong
    transpose in the data gradient or a mis-indexed col2im fails it).
"""
import numpy as np
import pytest

from oracle import cnn


def test_bf16_rounding_bits():
    cases = {
        0x3F800000: 0x3F80,          # 1.0
        0x3F808000: 0x3F80,          # halfway, even stays
        0x3F818000: 0x3F82,          # halfway, odd rounds up
        0x3F808001: 0x3F81,          # above halfway
        0x3F7FFFFF: 0x3F80,          # carry into the exponent
        0xBF808001: 0xBF81,          # sign kept
        0x00000000: 0x0000,
    }
    for u, want in cases.items():
        f = np.array([u], np.uint32).view(np.float32)
        assert int(cnn.bf16_bits(f)[0]) == want, hex(u)
        assert cnn.from_bits(cnn.bf16_bits(f))[0] == cnn.bf16(f)[0]


def _params(seed, scale=1.0):
    rng = np.random.default_rng(seed)
    d = {}
    for name, shp in cnn.SHAPES:
        fan_in = {"w1": 18, "fc1_w": 512, "fc2_w": 64}.get(name, 288)
        if name.startswith("b") or name.endswith("_b"):
            d[name] = rng.normal(0, 0.05, shp)
        else:
            d[name] = rng.normal(0, scale * np.sqrt(2.0 / fan_in), shp)
    return cnn.flatten(d)


def _direct_conv(x, w, b, oy, ox, co):
    """One output of a 3x3 / stride 2 / pad 1 conv, by a direct loop."""
    H, W, Ci = x.shape
    s = 0.0
    for ky in range(3):
        for kx in range(3):
            iy, ix = 2 * oy - 1 + ky, 2 * ox - 1 + kx
            if 0 <= iy < H and 0 <= ix < W:
                for ci in range(Ci):
                    s += x[iy, ix, ci] * w[co, ky * 3 + kx, ci]
    return max(s + b[co], 0.0)


def test_forward_matches_direct_loops():
    p = _params(1)
    rng = np.random.default_rng(2)
    x = rng.normal(0, 1, (2, 64, 64, 2))
    z, cache = cnn.forward(p, x, quantize=False)
    P = cnn.unflatten(p)
    acts = cache["acts"]
    for l in range(1, 5):
        w = P["w%d" % l]
        Ho = acts[l].shape[1]
        for (b, oy, ox, co) in [(0, 0, 0, 0), (1, Ho - 1, Ho - 1, 31), (0, Ho // 2, 1, 7),
                                (1, 0, Ho - 1, 12)]:
            want = _direct_conv(acts[l - 1][b], w, P["b%d" % l], oy, ox, co)
            assert abs(acts[l][b, oy, ox, co] - want) < 1e-9 * max(1, abs(want))
    f = acts[4].reshape(2, -1)
    # flatten order (y*4 + x)*32 + c
    assert f[1, (2 * 4 + 3) * 32 + 5] == acts[4][1, 2, 3, 5]
    h = np.maximum(f @ P["fc1_w"].T + P["fc1_b"], 0)
    np.testing.assert_allclose(z, h @ P["fc2_w"] + P["fc2_b"][0], rtol=1e-12)


def test_backward_matches_finite_differences():
    p = _params(3, scale=1.5)
    rng = np.random.default_rng(4)
    x = rng.normal(0, 1, (3, 64, 64, 2))
    x[0, 20:24, 30:34] += 4.0
    labels = np.array([1.0, 0.0, 1.0])
    weights = np.array([1.0, 0.5, 2.0])
    z, cache = cnn.forward(p, x, quantize=False)
    g, loss = cnn.backward(cache, labels, weights)
    assert abs(loss - cnn.loss_only(p, x, labels, weights)) < 1e-12
    # a sample of every tensor plus all of the small ones
    offs, o = {}, 0
    for name, shp in cnn.SHAPES:
        offs[name] = (o, int(np.prod(shp)))
        o += int(np.prod(shp))
    idx = []
    for name, (o, n) in offs.items():
        k = n if n <= 64 else 12
        idx += list(o + rng.choice(n, k, replace=False))
    eps = 1e-6    # small enough that no first-layer ReLU crosses its kink
    bad = 0
    for i in idx:
        pp, pm = p.copy(), p.copy()
        pp[i] += eps
        pm[i] -= eps
        fd = (cnn.loss_only(pp, x, labels, weights) - cnn.loss_only(pm, x, labels, weights)) / (2 * eps)
        err = abs(fd - g[i])
        if err > 1e-6 + 1e-4 * abs(fd):
            bad += 1
    # a ReLU kink inside +-eps can spoil a difference quotient: allow one
    assert bad <= 1, bad


def test_quantized_forward_close_to_exact():
    """bfloat16 storage perturbs the logits by a few parts in a thousand."""
    p = _params(5)
    x = np.random.default_rng(6).normal(0, 1, (4, 64, 64, 2))
    zq, _ = cnn.forward(p, x, quantize=True)
    ze, _ = cnn.forward(p, x, quantize=False)
    assert np.max(np.abs(zq - ze)) < 0.05 * (1 + np.max(np.abs(ze)))
