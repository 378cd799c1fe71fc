"""CPU restatement of the proposal network (numpy).

TEST INFRASTRUCTURE ONLY: imported by tests/ and __graft_entry__.smoke() as the
checker of the HIP forward and backward kernels; the product never imports it.

The reference has no learned stage, so nothing here follows a reference file.
The network and its quantisation points are defined by the product's kernels:
  input  x: 64x64 tile of 2 channels (coarse plane, wide plane), bfloat16
  conv_l (l = 1..4): 3x3, pad 1, stride 2, 32 output channels, ReLU;
         64 -> 32 -> 16 -> 8 -> 4 pixels, activations stored as bfloat16 (NHWC)
  fc1: 512 -> 64, ReLU, stored as bfloat16;  fc2: 64 -> 1 logit (float32)
Every product of the matrix stages multiplies bfloat16 operands (weights are
rounded from the float32 master copy) and accumulates in float32; biases are
float32.  Backward stores the gradient of each activation, masked by its ReLU,
as bfloat16, and accumulates weight gradients in float32.

`quantize=False` runs the same formulas in float64 with no rounding: the
finite-difference check of the backward pass uses that mode.

Flat parameter layout (float32), in order:
  w1[32][9][2]  b1[32]  w2[32][9][32]  b2[32]  w3 ..  b3  w4 ..  b4
  fc1_w[64][512]  fc1_b[64]  fc2_w[64]  fc2_b[1]
with conv weights indexed [out channel][ky*3 + kx][in channel] and the fc1
input flattened as (y*4 + x)*32 + channel.
"""
import numpy as np

TILE = 64
C = 32
HID = 64
CIN = 2
SHAPES = [("w1", (C, 9, CIN)), ("b1", (C,)),
          ("w2", (C, 9, C)), ("b2", (C,)),
          ("w3", (C, 9, C)), ("b3", (C,)),
          ("w4", (C, 9, C)), ("b4", (C,)),
          ("fc1_w", (HID, 16 * C)), ("fc1_b", (HID,)),
          ("fc2_w", (HID,)), ("fc2_b", (1,))]
N_PARAMS = sum(int(np.prod(s)) for _, s in SHAPES)


def bf16(a):
    """Round float32 values to bfloat16 (round to nearest, ties to even)."""
    a = np.ascontiguousarray(a, dtype=np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    return r.astype(np.uint32).view(np.float32)


def bf16_bits(a):
    """bfloat16 bit patterns (uint16) of float32 values."""
    return (bf16(a).view(np.uint32) >> 16).astype(np.uint16)


def from_bits(h):
    return (np.asarray(h, np.uint16).astype(np.uint32) << 16).view(np.float32)


def unflatten(p):
    out, o = {}, 0
    for name, shp in SHAPES:
        n = int(np.prod(shp))
        out[name] = np.asarray(p[o:o + n]).reshape(shp)
        o += n
    assert o == len(p)
    return out


def flatten(d):
    return np.concatenate([np.asarray(d[name]).reshape(-1) for name, _ in SHAPES])


def _im2col(x):
    """x [B,H,W,Ci] -> patches [B,H/2,W/2,9,Ci] of the 3x3 / stride 2 / pad 1 conv."""
    B, H, W, Ci = x.shape
    xp = np.zeros((B, H + 2, W + 2, Ci), x.dtype)
    xp[:, 1:H + 1, 1:W + 1] = x
    Ho, Wo = H // 2, W // 2
    cols = np.empty((B, Ho, Wo, 9, Ci), x.dtype)
    for ky in range(3):
        for kx in range(3):
            cols[:, :, :, ky * 3 + kx] = xp[:, ky:ky + 2 * Ho:2, kx:kx + 2 * Wo:2]
    return cols


def _col2im(g, H, W):
    """Adjoint of _im2col: g [B,Ho,Wo,9,Ci] -> [B,H,W,Ci]."""
    B, Ho, Wo, _, Ci = g.shape
    xp = np.zeros((B, H + 2, W + 2, Ci), g.dtype)
    for ky in range(3):
        for kx in range(3):
            xp[:, ky:ky + 2 * Ho:2, kx:kx + 2 * Wo:2] += g[:, :, :, ky * 3 + kx]
    return xp[:, 1:H + 1, 1:W + 1]


def forward(p, x, quantize=True):
    """p: flat parameters; x: [B,64,64,2] tiles.  Returns (logits [B], cache)."""
    P = unflatten(np.asarray(p, np.float64 if not quantize else np.float32))
    if quantize:
        q = bf16
        dt = np.float64   # accumulate wide; float32 rounding of the sums is below bf16's
    else:
        q = lambda a: np.asarray(a, np.float64)
        dt = np.float64
    a = q(np.asarray(x, np.float32)).astype(dt)                 # [B,64,64,2]
    acts = [a]
    for l in range(1, 5):
        wq = q(P["w%d" % l]).astype(dt)                         # [32,9,Ci]
        cols = _im2col(acts[-1])                                 # [B,Ho,Wo,9,Ci]
        v = np.tensordot(cols, wq, axes=([3, 4], [1, 2]))        # [B,Ho,Wo,32]
        if quantize:
            v = v.astype(np.float32).astype(dt)
        v = v + P["b%d" % l].astype(dt)
        if quantize:
            v = v.astype(np.float32).astype(dt)
        acts.append(q(np.maximum(v, 0).astype(np.float32) if quantize else np.maximum(v, 0))
                    .astype(dt))
    B = x.shape[0]
    f = acts[-1].reshape(B, 16 * C)
    v = f @ q(P["fc1_w"]).astype(dt).T
    if quantize:
        v = v.astype(np.float32).astype(dt)
    v = v + P["fc1_b"].astype(dt)
    if quantize:
        v = v.astype(np.float32).astype(dt)
    h = q(np.maximum(v, 0).astype(np.float32) if quantize else np.maximum(v, 0)).astype(dt)
    z = h @ q(P["fc2_w"]).astype(dt) + P["fc2_b"].astype(dt)[0]
    return z, {"acts": acts, "h": h, "P": P, "q": q, "dt": dt, "quantize": quantize}


def loss_and_dz(z, labels, weights):
    """Weighted mean binary cross-entropy of logits z and its gradient."""
    z = np.asarray(z, np.float64)
    y = np.asarray(labels, np.float64)
    w = np.asarray(weights, np.float64)
    W = w.sum()
    sp = np.maximum(z, 0) + np.log1p(np.exp(-np.abs(z)))
    loss = float((w * (sp - y * z)).sum() / W)
    sig = 1.0 / (1.0 + np.exp(-z))
    return loss, w * (sig - y) / W


def backward(cache, labels, weights):
    """Gradients (flat float64 array, SHAPES order) and the loss."""
    P, q, dt, quant = cache["P"], cache["q"], cache["dt"], cache["quantize"]
    acts, h = cache["acts"], cache["h"]
    B = h.shape[0]
    hidden = h @ q(P["fc2_w"]).astype(dt) + P["fc2_b"].astype(dt)[0]
    loss, dz = loss_and_dz(hidden, labels, weights)
    if quant:
        dz = dz.astype(np.float32).astype(dt)
    G = {}
    G["fc2_w"] = dz @ h
    G["fc2_b"] = np.array([dz.sum()])
    dh = np.outer(dz, q(P["fc2_w"]).astype(dt)) * (h > 0)
    dh = q(dh.astype(np.float32) if quant else dh).astype(dt)
    f = acts[-1].reshape(B, 16 * C)
    G["fc1_w"] = dh.T @ f
    G["fc1_b"] = dh.sum(0)
    df = dh @ q(P["fc1_w"]).astype(dt)
    d = (df * (f > 0)).reshape(acts[-1].shape)
    d = q(d.astype(np.float32) if quant else d).astype(dt)
    for l in range(4, 0, -1):
        x_in = acts[l - 1]
        cols = _im2col(x_in)                                     # [B,Ho,Wo,9,Ci]
        gw = np.tensordot(d, cols, axes=([0, 1, 2], [0, 1, 2]))  # [32,9,Ci]
        G["w%d" % l] = gw
        G["b%d" % l] = d.sum((0, 1, 2))
        if l == 1:
            break
        w = q(P["w%d" % l]).astype(dt)                           # [32,9,Ci]
        gcols = np.tensordot(d, w, axes=([3], [0]))              # [B,Ho,Wo,9,Ci]
        dx = _col2im(gcols, x_in.shape[1], x_in.shape[2])
        dx = dx * (x_in > 0)
        d = q(dx.astype(np.float32) if quant else dx).astype(dt)
    return flatten(G), loss


def loss_only(p, x, labels, weights, quantize=False):
    z, _ = forward(p, x, quantize)
    return loss_and_dz(z, labels, weights)[0]
