"""CPU oracle of the learned field predictor: a 4-level convolutional
encoder-decoder with 32 channels per level.

TEST INFRASTRUCTURE ONLY.  The checker for the U-Net kernels of
``lcode3d_amd/csrc/unet.hip``; only ``tests/`` import it.  The reference has
no learned predictor (its predictor pass is ``lcode.py:841-881``), so there
is no reference line to follow: each function restates a textbook
operation in FP64 NumPy, on [C, H, W] arrays.

Network (input planes at level-0 size S0, levels S_{l+1} = (S_l + 1) // 2):

    e0 = relu(C0(x))                      14 -> 32   level 0
    e1 = relu(C1(pool(e0)))               32 -> 32   level 1
    e2 = relu(C2(pool(e1)))               32 -> 32   level 2
    e3 = relu(C3(pool(e2)))               32 -> 32   level 3
    d2 = relu(C4(cat(up(e3), e2)))        64 -> 32   level 2
    d1 = relu(C5(cat(up(d2), e1)))        64 -> 32   level 1
    d0 = relu(C6(cat(up(d1), e0)))        64 -> 32   level 0
    y  = C7(d0)                           32 -> 6    level 0

Ck are 3 x 3 cross-correlations with zero padding and a bias
(``predictor_oracle.conv3x3``); pool is the 2 x 2 mean with zeros beyond
the edge; up is nearest-neighbour upsampling, y(i, j) = x(i // 2, j // 2).

Mixed precision: with ``q = bf16`` every tensor the kernels store in
bfloat16 (the input planes, the weights, the activations e*, d*, their
pooled and upsampled copies, and the gradients of the backward pass) is
rounded to bfloat16 at the same point; sums are exact (FP64) here and FP32
on the device.  With ``q = None`` (identity) the arithmetic is exact FP64
for the finite-difference pins.

Parity pinning (tests/test_unet_oracle.py): pool and up against brute-force
loops and their adjoints against the dot-product identity
<A x, y> = <x, A^T y>; the whole backward pass against central finite
differences of the forward pass in exact arithmetic; bf16 against hand
examples (round to nearest even).
"""

import numpy as np

from .predictor_oracle import conv3x3, conv3x3_backward

C = 32
CIN = 14
COUT = 6
#: (cin, cout) of C0 .. C7
LAYERS = ((CIN, C), (C, C), (C, C), (C, C), (2 * C, C), (2 * C, C), (2 * C, C),
          (C, COUT))


def bf16(a):
    """Round to the nearest bfloat16 (ties to even), as float64; finite
    inputs."""
    f = np.asarray(a, dtype=np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def _ident(a):
    return a


def level_sizes(s0, levels=4):
    s = [s0]
    for _ in range(levels - 1):
        s.append((s[-1] + 1) // 2)
    return s


def pool(x):
    """2 x 2 mean, zeros beyond the edge: [C, H, W] -> [C, (H+1)//2, (W+1)//2]."""
    c, h, w = x.shape
    xp = np.zeros((c, h + h % 2, w + w % 2))
    xp[:, :h, :w] = x
    return (xp[:, 0::2, 0::2] + xp[:, 1::2, 0::2] + xp[:, 0::2, 1::2]
            + xp[:, 1::2, 1::2]) / 4


def pool_T(gy, h, w):
    """Adjoint of pool: g(i, j) = gy(i // 2, j // 2) / 4 on [C, h, w]."""
    return np.repeat(np.repeat(gy, 2, axis=1), 2, axis=2)[:, :h, :w] / 4


def up(x, h, w):
    """Nearest-neighbour upsampling to [C, h, w]: y(i, j) = x(i // 2, j // 2)."""
    return np.repeat(np.repeat(x, 2, axis=1), 2, axis=2)[:, :h, :w]


def up_T(gy, hs, ws):
    """Adjoint of up: g(I, J) = sum of gy(i, j) over i // 2 = I, j // 2 = J."""
    c, h, w = gy.shape
    gp = np.zeros((c, 2 * hs, 2 * ws))
    gp[:, :h, :w] = gy
    return (gp[:, 0::2, 0::2] + gp[:, 1::2, 0::2] + gp[:, 0::2, 1::2]
            + gp[:, 1::2, 1::2])


def init_params(seed=0):
    """He-normal weights (the output layer scaled by 0.1), zero biases."""
    rng = np.random.default_rng(seed)
    params = []
    for k, (cin, cout) in enumerate(LAYERS):
        std = np.sqrt(2.0 / (cin * 9)) * (0.1 if k == len(LAYERS) - 1 else 1.0)
        params.append((rng.standard_normal((cout, cin, 3, 3)) * std, np.zeros(cout)))
    return params


def forward(params, x, q=None):
    """Activations of the network on input planes x [14, S0, S0]."""
    q = q or _ident
    qw = [(q(w), b) for w, b in params]
    s = level_sizes(x.shape[1])
    a = {"x": q(x)}

    def conv(k, inp, relu=True):
        y = conv3x3(inp, *qw[k])
        return q(np.maximum(y, 0)) if relu else y
    a["e0"] = conv(0, a["x"])
    for l in (1, 2, 3):
        a[f"p{l - 1}"] = q(pool(a[f"e{l - 1}"]))
        a[f"e{l}"] = conv(l, a[f"p{l - 1}"])
    low = "e3"
    for k, l in ((4, 2), (5, 1), (6, 0)):
        a[f"u{l + 1}"] = q(up(a[low], s[l], s[l]))
        a[f"d{l}"] = conv(k, np.concatenate([a[f"u{l + 1}"], a[f"e{l}"]]))
        low = f"d{l}"
    a["y"] = conv(7, a["d0"], relu=False)
    return a


def backward(params, a, gy, q=None):
    """Weight and bias gradients [(gw, gb)] of sum(gy * y), in the order
    and with the storage roundings of the kernels."""
    q = q or _ident
    qw = [(q(w), b) for w, b in params]
    s = level_sizes(a["x"].shape[1])
    grads = [None] * len(params)

    def conv_T(k, inp, g):
        gx, gw, gb = conv3x3_backward(inp, qw[k][0], g)
        grads[k] = (gw, gb)
        return gx
    g = q(gy)
    gx = conv_T(7, a["d0"], g)
    g = q(gx * (a["d0"] > 0))                      # gradient of C6's output
    skip = {}
    for k, l in ((6, 0), (5, 1), (4, 2)):
        inp = np.concatenate([a[f"u{l + 1}"], a[f"e{l}"]])
        gx = conv_T(k, inp, g)
        skip[l] = gx[C:]
        low = a["e3"] if l == 2 else a[f"d{l + 1}"]
        g = q(q(up_T(q(gx[:C]), s[l + 1], s[l + 1])) * (low > 0))
    for l in (3, 2, 1):
        gx = conv_T(l, a[f"p{l - 1}"], g)
        g = q((pool_T(q(gx), s[l - 1], s[l - 1]) + q(skip[l - 1])) * (a[f"e{l - 1}"] > 0))
    conv_T(0, a["x"], g)
    return grads


def mse_grad(y, t):
    """mean((y - t)^2) over all values, and its gradient 2 (y - t) / n."""
    d = y - t
    return float((d * d).mean()), 2 * d / d.size
