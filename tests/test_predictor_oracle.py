"""Pins of the predictor oracle (CPU): the convolution against a brute-force
loop and scipy.signal.correlate, its backward pass against central finite
differences, Adam against a hand-worked first step."""

import numpy as np
import pytest
import scipy.signal

from oracle import predictor_oracle as po


def _brute(x, w, b):
    cin, H, W = x.shape
    out = np.zeros((w.shape[0], H, W))
    for o in range(w.shape[0]):
        for i in range(H):
            for j in range(W):
                s = b[o]
                for c in range(cin):
                    for di in range(3):
                        for dj in range(3):
                            ii, jj = i + di - 1, j + dj - 1
                            if 0 <= ii < H and 0 <= jj < W:
                                s += w[o, c, di, dj] * x[c, ii, jj]
                out[o, i, j] = s
    return out


def test_conv_brute_force():
    rng = np.random.default_rng(0)
    x = rng.standard_normal((3, 5, 7))
    w = rng.standard_normal((2, 3, 3, 3))
    b = rng.standard_normal(2)
    assert np.allclose(po.conv3x3(x, w, b), _brute(x, w, b), rtol=0, atol=1e-12)


def test_conv_matches_scipy_correlate():
    rng = np.random.default_rng(1)
    x = rng.standard_normal((4, 11, 9))
    w = rng.standard_normal((3, 4, 3, 3))
    b = np.zeros(3)
    want = np.stack([sum(scipy.signal.correlate(x[c], w[o, c], mode="same")
                         for c in range(4)) for o in range(3)])
    assert np.allclose(po.conv3x3(x, w, b), want, rtol=0, atol=1e-12)


def test_backward_finite_differences():
    rng = np.random.default_rng(2)
    layers = [(rng.standard_normal((4, 3, 3, 3)) * .5, rng.standard_normal(4) * .1, True),
              (rng.standard_normal((2, 4, 3, 3)) * .5, rng.standard_normal(2) * .1, False)]
    x = rng.standard_normal((3, 6, 5))
    t = rng.standard_normal((2, 6, 5))

    def loss(layers_, x_):
        return po.mse(po.forward(layers_, x_)[-1], t)[0]

    acts = po.forward(layers, x)
    _, gy = po.mse(acts[-1], t)
    grads, gx = po.backward(layers, acts, gy)
    h = 1e-6
    # a few weight, bias and input entries of each layer
    for li, (w, b, relu) in enumerate(layers):
        for idx in [(0, 0, 0, 0), (1, 2, 1, 2), (w.shape[0] - 1, w.shape[1] - 1, 2, 1)]:
            wp, wm = w.copy(), w.copy()
            wp[idx] += h
            wm[idx] -= h
            lp = loss([(wp, b, relu) if k == li else L for k, L in enumerate(layers)], x)
            lm = loss([(wm, b, relu) if k == li else L for k, L in enumerate(layers)], x)
            assert abs((lp - lm) / (2 * h) - grads[li][0][idx]) < 1e-6
        bp, bm = b.copy(), b.copy()
        bp[0] += h
        bm[0] -= h
        lp = loss([(w, bp, relu) if k == li else L for k, L in enumerate(layers)], x)
        lm = loss([(w, bm, relu) if k == li else L for k, L in enumerate(layers)], x)
        assert abs((lp - lm) / (2 * h) - grads[li][1][0]) < 1e-6
    for idx in [(0, 0, 0), (2, 5, 4), (1, 3, 2)]:
        xp, xm = x.copy(), x.copy()
        xp[idx] += h
        xm[idx] -= h
        assert abs((loss(layers, xp) - loss(layers, xm)) / (2 * h) - gx[idx]) < 1e-6


def test_adam_first_step():
    # at t = 1 the bias-corrected step is lr * g / (|g| + eps) = lr * sign(g)
    p = np.array([1.0, -2.0, 3.0])
    g = np.array([0.5, -4.0, 1e-3])
    p1, m, v = po.adam(p, g, np.zeros(3), np.zeros(3), 1, lr=0.1, eps=0)
    assert np.allclose(p1, p - 0.1 * np.sign(g), rtol=0, atol=1e-15)
    assert np.allclose(m, 0.1 * g) and np.allclose(v, 0.001 * g * g)


@pytest.mark.parametrize("shape", [(1, 1, 1), (2, 3, 1), (1, 1, 4)])
def test_conv_degenerate_planes(shape):
    rng = np.random.default_rng(3)
    x = rng.standard_normal(shape)
    w = rng.standard_normal((2, shape[0], 3, 3))
    b = rng.standard_normal(2)
    assert np.allclose(po.conv3x3(x, w, b), _brute(x, w, b), rtol=0, atol=1e-12)
