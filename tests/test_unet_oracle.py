"""Pins of the U-Net predictor oracle (CPU): bfloat16 rounding by hand
examples, pool / up against brute-force loops and their adjoints against
the dot-product identity, and the whole backward pass against central
finite differences of the forward pass in exact arithmetic (a dropped skip
term, a wrong mask or a transposed kernel fails these)."""

import numpy as np
import pytest

from oracle import unet_oracle as uo


def test_bf16_round_to_nearest_even():
    one = 1.0
    assert uo.bf16(one) == 1.0
    assert uo.bf16(1 + 2.0 ** -8) == 1.0                  # tie -> even (down)
    assert uo.bf16(1 + 3 * 2.0 ** -8) == 1 + 2.0 ** -6    # tie -> even (up)
    assert uo.bf16(1 + 2.0 ** -8 + 2.0 ** -12) == 1 + 2.0 ** -7
    assert uo.bf16(-2.5) == -2.5
    assert uo.bf16(np.float32(3.0e-3)) == pytest.approx(3.0e-3, rel=2 ** -8)


def test_pool_and_up_brute_force():
    rng = np.random.default_rng(0)
    x = rng.standard_normal((2, 7, 5))
    p = uo.pool(x)
    assert p.shape == (2, 4, 3)
    for c in range(2):
        for i in range(4):
            for j in range(3):
                s = sum(x[c, a, b] for a in (2 * i, 2 * i + 1) for b in (2 * j, 2 * j + 1)
                        if a < 7 and b < 5)
                assert p[c, i, j] == pytest.approx(s / 4, abs=1e-15)
    u = uo.up(p, 7, 5)
    for c in range(2):
        for i in range(7):
            for j in range(5):
                assert u[c, i, j] == p[c, i // 2, j // 2]


@pytest.mark.parametrize("h,w", [(7, 5), (8, 6), (11, 11)])
def test_pool_and_up_adjoints(h, w):
    rng = np.random.default_rng(h * w)
    x = rng.standard_normal((3, h, w))
    hs, ws = (h + 1) // 2, (w + 1) // 2
    y = rng.standard_normal((3, hs, ws))
    assert np.sum(uo.pool(x) * y) == pytest.approx(np.sum(x * uo.pool_T(y, h, w)), rel=1e-13)
    z = rng.standard_normal((3, h, w))
    assert np.sum(uo.up(y, h, w) * z) == pytest.approx(np.sum(y * uo.up_T(z, hs, ws)), rel=1e-13)


def test_level_sizes():
    assert uo.level_sizes(161) == [161, 81, 41, 21]
    assert uo.level_sizes(11) == [11, 6, 3, 2]


def test_backward_matches_finite_differences():
    rng = np.random.default_rng(3)
    s0 = 11
    params = uo.init_params(seed=1)
    # non-zero biases so every bias gradient is exercised
    params = [(w, rng.standard_normal(b.shape) * .1) for w, b in params]
    x = rng.standard_normal((uo.CIN, s0, s0))
    gy = rng.standard_normal((uo.COUT, s0, s0))
    a = uo.forward(params, x)
    grads = uo.backward(params, a, gy)

    def loss(ps):
        return float(np.sum(uo.forward(ps, x)["y"] * gy))
    eps = 1e-6
    for k in range(len(params)):
        w, b = params[k]
        gw, gb = grads[k]
        for _ in range(4):
            idx = tuple(rng.integers(0, n) for n in w.shape)
            wp, wm = w.copy(), w.copy()
            wp[idx] += eps
            wm[idx] -= eps
            psp = list(params)
            psp[k] = (wp, b)
            psm = list(params)
            psm[k] = (wm, b)
            fd = (loss(psp) - loss(psm)) / (2 * eps)
            assert gw[idx] == pytest.approx(fd, rel=1e-5, abs=1e-7), (k, idx)
        o = int(rng.integers(0, b.shape[0]))
        bp, bm = b.copy(), b.copy()
        bp[o] += eps
        bm[o] -= eps
        psp = list(params)
        psp[k] = (w, bp)
        psm = list(params)
        psm[k] = (w, bm)
        fd = (loss(psp) - loss(psm)) / (2 * eps)
        assert gb[o] == pytest.approx(fd, rel=1e-5, abs=1e-7), (k, "bias")


def test_bf16_forward_close_to_exact():
    rng = np.random.default_rng(4)
    params = uo.init_params(seed=2)
    x = rng.standard_normal((uo.CIN, 11, 11))
    y = uo.forward(params, x)["y"]
    yq = uo.forward(params, x, q=uo.bf16)["y"]
    assert np.linalg.norm(yq - y) / np.linalg.norm(y) < 3e-2
