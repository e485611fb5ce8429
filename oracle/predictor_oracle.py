"""CPU oracle of the learned field predictor's building blocks.

TEST INFRASTRUCTURE ONLY.  Used by the U-Net oracle (oracle/unet_oracle.py,
the checker of ``lcode3d_amd/csrc/unet.hip``) and by ``tests/``.  The
reference has no learned predictor (its predictor pass is
``lcode.py:841-881``), so there is no reference line to follow: each
function restates a textbook operation, in FP64 NumPy.

* ``conv3x3`` - 3 x 3 cross-correlation with zero padding ("same" size),
  out[co](i, j) = b[co] + sum_ci sum_{di,dj} w[co, ci, di, dj]
  in[ci](i + di - 1, j + dj - 1).
* ``conv3x3_backward`` - its adjoints: the input gradient is the full
  correlation of the output gradient with the flipped kernel, the weight
  gradient the correlation of the input with the output gradient.
* ``forward``/``backward`` - a plain stack of such layers (with ReLU), the
  smallest composition the adjoints are pinned through.
* ``adam`` - Kingma & Ba, "Adam: a method for stochastic optimization"
  (2015), Algorithm 1, with bias-corrected moments.

Parity pinning: ``conv3x3`` against a brute-force loop on tiny planes and
against ``scipy.signal.correlate``; the backward pass against central
finite differences of the forward pass (tests/test_predictor_oracle.py).
"""

import numpy as np


def conv3x3(x, w, b):
    """x [Cin, H, W], w [Cout, Cin, 3, 3], b [Cout] -> [Cout, H, W]."""
    cin, H, W = x.shape
    xp = np.pad(x, ((0, 0), (1, 1), (1, 1)))
    out = np.broadcast_to(b[:, None, None], (w.shape[0], H, W)).astype(np.float64).copy()
    for di in range(3):
        for dj in range(3):
            out += np.einsum("oc,chw->ohw", w[:, :, di, dj], xp[:, di:di + H, dj:dj + W])
    return out


def conv3x3_backward(x, w, gout):
    """Gradients of sum(gout * conv3x3(x, w, b)) with respect to x, w, b."""
    cin, H, W = x.shape
    xp = np.pad(x, ((0, 0), (1, 1), (1, 1)))
    gp = np.pad(gout, ((0, 0), (1, 1), (1, 1)))
    gw = np.empty_like(w, dtype=np.float64)
    gx = np.zeros((cin, H, W))
    for di in range(3):
        for dj in range(3):
            gw[:, :, di, dj] = np.einsum("ohw,chw->oc", gout, xp[:, di:di + H, dj:dj + W])
            # x(i, j) feeds out(i - di + 1, j - dj + 1) with weight w[., ., di, dj]
            gx += np.einsum("oc,ohw->chw", w[:, :, di, dj],
                            gp[:, 2 - di:2 - di + H, 2 - dj:2 - dj + W])
    return gx, gw, gout.sum(axis=(1, 2))


def forward(layers, x):
    """layers: [(w, b, relu)]; returns the activations [x, a1, ..., y]."""
    acts = [x]
    for w, b, relu in layers:
        y = conv3x3(acts[-1], w, b)
        acts.append(np.maximum(y, 0) if relu else y)
    return acts


def backward(layers, acts, gy):
    """Parameter gradients [(gw, gb)] of sum(gy * y) and the input gradient."""
    grads = [None] * len(layers)
    g = gy
    for li in range(len(layers) - 1, -1, -1):
        w, b, relu = layers[li]
        gx, gw, gb = conv3x3_backward(acts[li], w, g)
        grads[li] = (gw, gb)
        if li > 0 and layers[li - 1][2]:
            gx = gx * (acts[li] > 0)     # ReLU of the previous layer
        g = gx
    return grads, g


def mse(y, t):
    """mean((y - t)^2) and its gradient 2 (y - t) / n."""
    d = y - t
    return float((d * d).mean()), 2 * d / d.size


def adam(p, g, m, v, t, lr, b1=0.9, b2=0.999, eps=1e-8):
    """One Adam step t >= 1; returns new (p, m, v)."""
    m = b1 * m + (1 - b1) * g
    v = b2 * v + (1 - b2) * g * g
    mhat = m / (1 - b1 ** t)
    vhat = v / (1 - b2 ** t)
    return p - lr * mhat / (np.sqrt(vhat) + eps), m, v
