"""CPU oracle of the BP warm-start network (test infrastructure only: used by
tests/ as the checker of the GPU kernels; never imported by the engine).

The network has no counterpart in the reference: it replaces the start of the
reference's BP solve (src/rotamer.cpp:1110-1133, beliefs = probabilities,
messages = 1) by a learned correction of the first pass's beliefs.  This file
restates, in float64, exactly what the GPU kernels compute:

  forward   x_b = bf16(x);  h1 = relu(bf16(W1) x_b + b1);  h1_b = bf16(h1)
            h2 = relu(bf16(W2) h1_b + b2);  h2_b = bf16(h2)
            y  = bf16(W3) h2_b + b3
  loss      z_r = f[6+r] + y_r over the row's states (mask f[18+r]),
            loss = mean_rows( -sum_r t_r log softmax(z)_r )
  backward  dy = (softmax(z) - t) / n_rows (0 off the states);  dy_b = bf16(dy)
            dh2 = (bf16(W3)^T dy_b) * [h2 > 0];  dh2_b = bf16(dh2)
            dh1 = (bf16(W2)^T dh2_b) * [h1 > 0];  dh1_b = bf16(dh1)
            dW3 = dy_b^T h2_b   dW2 = dh2_b^T h1_b   dW1 = dh1_b^T x_b
            db3 = sum dy_b      db2 = sum dh2_b      db1 = sum dh1_b

bf16 rounding is round-to-nearest-even on the float32 value.  With
quantize=False every bf16() is the identity and the gradient is the exact
gradient of the float64 loss (pinned by central differences in
tests/test_oracle.py).

Parameter vector layout (float32 on the C ABI): W1 [128][32], b1 [128],
W2 [128][128], b2 [128], W3 [16][128], b3 [16].
"""
import numpy as np

N_IN, N_H, N_OUT, N_USED = 32, 128, 16, 6
SHAPES = [("W1", (N_H, N_IN)), ("b1", (N_H,)), ("W2", (N_H, N_H)), ("b2", (N_H,)),
          ("W3", (N_OUT, N_H)), ("b3", (N_OUT,))]
N_PARAM = sum(int(np.prod(s)) for _, s in SHAPES)


def unpack(p):
    out, k = {}, 0
    for name, shape in SHAPES:
        n = int(np.prod(shape))
        out[name] = np.asarray(p[k:k + n], np.float64).reshape(shape)
        k += n
    return out


def pack(d):
    return np.concatenate([np.asarray(d[name], np.float64).ravel() for name, _ in SHAPES])


def bf16(x):
    """Round to bfloat16 (nearest even) through float32, returned as float64."""
    u = np.ascontiguousarray(np.asarray(x, np.float32)).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def forward(p, feat, quantize=True):
    q = bf16 if quantize else (lambda a: np.asarray(a, np.float64))
    w = unpack(p)
    x = q(feat)
    h1 = np.maximum(x @ q(w["W1"]).T + w["b1"], 0.0)
    h1b = q(h1)
    h2 = np.maximum(h1b @ q(w["W2"]).T + w["b2"], 0.0)
    h2b = q(h2)
    y = h2b @ q(w["W3"]).T + w["b3"]
    return dict(x=x, h1=h1, h1b=h1b, h2=h2, h2b=h2b, y=y)


def loss_and_dy(feat, target, y):
    feat = np.asarray(feat, np.float64)
    n = len(feat)
    mask = feat[:, 18:18 + N_USED] > 0.5
    z = feat[:, 6:6 + N_USED] + y[:, :N_USED]
    zm = np.where(mask, z, -np.inf)
    mx = zm.max(axis=1, keepdims=True)
    e = np.where(mask, np.exp(zm - mx), 0.0)
    s = e.sum(axis=1, keepdims=True)
    lse = mx + np.log(s)
    t = np.asarray(target, np.float64)[:, :N_USED]
    loss = -np.where(mask, t * (z - lse), 0.0).sum() / n
    dy = np.zeros((n, N_OUT))
    dy[:, :N_USED] = np.where(mask, e / s - t, 0.0) / n
    return loss, dy


def loss_grad(p, feat, target, quantize=True):
    """Mean loss and its gradient (as a parameter vector) over all rows."""
    q = bf16 if quantize else (lambda a: np.asarray(a, np.float64))
    w = unpack(p)
    f = forward(p, feat, quantize)
    loss, dy = loss_and_dy(feat, target, f["y"])
    dyb = q(dy)
    dh2 = (dyb @ q(w["W3"])) * (f["h2"] > 0)
    dh2b = q(dh2)
    dh1 = (dh2b @ q(w["W2"])) * (f["h1"] > 0)
    dh1b = q(dh1)
    g = dict(W3=dyb.T @ f["h2b"], b3=dyb.sum(0), W2=dh2b.T @ f["h1b"], b2=dh2b.sum(0),
             W1=dh1b.T @ f["x"], b1=dh1b.sum(0))
    return loss, pack(g)


def loss_only(p, feat, target, quantize=False):
    f = forward(p, feat, quantize)
    return loss_and_dy(feat, target, f["y"])[0]
