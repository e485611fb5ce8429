"""The learned field predictor's HIP kernels against the NumPy oracle
(oracle/predictor_oracle.py), and what the trained network gains.

* forward activations, parameter gradients and input gradients of a small
  network at 17 x 19 planes: relative error <= 1e-5 (FP32 kernels against
  an FP64 oracle);
* three Adam steps: <= 1e-6 relative;
* plane conversions (inputs, targets, guess) against their formulas;
* at 129^2 (grid_step_size 0.05, xi_step_size 0.05) a network trained on
  three plain runs (amplitude and sigma drawn from default_rng(0)) guesses
  the next slice's fields of a fourth, held-out run better than the
  quadratic extrapolation it starts from (a zero network); the test prints
  its error beside the linear and quadratic extrapolations'.
"""

import types

import numpy as np
import pytest
import torch

from gpu_helpers import host, say

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def pred_mod():
    import lcode3d_amd.predictor as p
    assert torch.cuda.is_available(), "no HIP device visible"
    return p


def _rel(a, b):
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def _net(pred_mod, N, hidden, seed):
    net = pred_mod.ConvPredictor(N, hidden=hidden, seed=seed)
    # non-trivial biases and last layer for the check
    rng = np.random.default_rng(seed + 1)
    p = net.params.cpu().numpy()
    for cin, cout, wo, bo, relu in net.layers:
        p[bo:bo + cout] = rng.standard_normal(cout) * 0.1
        if not relu:
            p[wo:bo] *= 10
    net.params.copy_(torch.as_tensor(p))
    return net


def test_forward_and_gradients_match_oracle(pred_mod, po):
    say("predictor gradients 17x17")
    N = 17
    net = _net(pred_mod, N, (5, 4), seed=3)
    rng = np.random.default_rng(4)
    x = rng.standard_normal((14, N, N)).astype(np.float32)
    t = rng.standard_normal((6, N, N)).astype(np.float32)
    layers = net.layer_arrays()
    acts = po.forward(layers, x.astype(np.float64))
    loss, gy = po.mse(acts[-1], t.astype(np.float64))
    grads, gx = po.backward(layers, acts, gy)

    y = net.forward(torch.as_tensor(x, device="cuda"))
    assert _rel(host(y), acts[-1]) < 1e-5
    for k in range(1, len(acts) - 1):
        assert _rel(host(net._acts[k]), acts[k]) < 1e-5, k
    gx_dev = net.input_grad(torch.as_tensor(gy.astype(np.float32), device="cuda"))
    g = net.grad.cpu().numpy()
    for li, (cin, cout, wo, bo, relu) in enumerate(net.layers):
        assert _rel(g[wo:bo].reshape(cout, cin, 3, 3), grads[li][0]) < 1e-5, li
        assert _rel(g[bo:bo + cout], grads[li][1]) < 1e-5, li
    assert _rel(host(gx_dev), gx) < 1e-5


def test_train_step_loss_and_adam_match_oracle(pred_mod, po):
    say("predictor Adam steps")
    N = 17
    net = _net(pred_mod, N, (4, 4), seed=5)
    rng = np.random.default_rng(6)
    x = rng.standard_normal((14, N, N)).astype(np.float32)
    t = rng.standard_normal((6, N, N)).astype(np.float32)
    p = net.params.cpu().numpy().astype(np.float64)
    m = np.zeros_like(p)
    v = np.zeros_like(p)
    for step in range(1, 4):
        layers = []
        for cin, cout, wo, bo, relu in net.layers:
            layers.append((p[wo:bo].reshape(cout, cin, 3, 3), p[bo:bo + cout], relu))
        acts = po.forward(layers, x.astype(np.float64))
        loss, gy = po.mse(acts[-1], t.astype(np.float64))
        grads, _ = po.backward(layers, acts, gy)
        g = np.concatenate([np.concatenate([gw.ravel(), gb]) for gw, gb in grads])
        got = net.train_step(torch.as_tensor(x, device="cuda"),
                             torch.as_tensor(t, device="cuda")).item()
        assert abs(got - loss) <= 1e-5 * loss
        p, m, v = po.adam(p, g, m, v, step, lr=net.lr)
        assert _rel(net.params.cpu().numpy(), p) < 1e-5, step


def test_plane_conversions(pred_mod):
    say("predictor plane conversions")
    N = 9
    net = pred_mod.ConvPredictor(N, hidden=(3, 3))
    rng = np.random.default_rng(7)
    net.scale[:] = rng.uniform(.5, 2, 14)
    net.tscale[:] = rng.uniform(.5, 2, 6)

    def state():
        return types.SimpleNamespace(**{n: torch.as_tensor(rng.standard_normal((N, N)),
                                                           device="cuda")
                                        for n in pred_mod.FIELDS + ("ro",)})
    s1, s2, s3, s0 = state(), state(), state(), state()
    beam = torch.as_tensor(rng.standard_normal((N, N)), device="cuda")

    def quad(n):
        return 3 * host(getattr(s1, n)) - 3 * host(getattr(s2, n)) + host(getattr(s3, n))
    x = host(net.inputs(s1, s2, s3, beam))
    for f, n in enumerate(pred_mod.FIELDS):
        assert np.allclose(x[f], quad(n) / net.scale[f], rtol=1e-6)
        assert np.allclose(x[6 + f], (host(getattr(s1, n)) - host(getattr(s2, n)))
                           / net.scale[6 + f], rtol=1e-6)
    assert np.allclose(x[12], host(beam) / net.scale[12], rtol=1e-6)
    assert np.allclose(x[13], host(s1.ro) / net.scale[13], rtol=1e-6)
    t = host(net.targets(s1, s2, s3, s0))
    for f, n in enumerate(pred_mod.FIELDS):
        want = (host(getattr(s0, n)) - quad(n)) / net.tscale[f]
        assert np.allclose(t[f], want, rtol=1e-5, atol=1e-6)
    # the guess of a zero network is the quadratic extrapolation
    net.params.zero_()
    g = net.guess(s1, s2, s3, beam)
    for f, n in enumerate(pred_mod.FIELDS):
        assert np.allclose(host(g[f]), quad(n), rtol=0, atol=1e-12)


def _cfg129():
    return types.SimpleNamespace(
        grid_steps=129, grid_step_size=.05, xi_step_size=.05,
        field_solver_subtraction_trick=1, field_solver_variant_A=True,
        reflect_padding_steps=5, plasma_padding_steps=10,
        plasma_coarseness=2, plasma_fineness=2)


def _run(lc, cfg, amplitude, sigma, slices):
    from lcode3d_amd.beams import cosine_gaussian
    xs, ys, const, vp, st = lc.init(cfg)
    states, beams = [], []
    for i in range(slices):
        beam = cosine_gaussian(cfg, i, amplitude=amplitude, sigma=sigma)
        st = lc.step(cfg, const, vp, st, beam)
        states.append(st)
        beams.append(beam)
    return states, beams


def test_trained_network_beats_extrapolation(pred_mod):
    import lcode3d_amd as lc
    say("predictor training at 129^2")
    cfg = _cfg129()
    rng = np.random.default_rng(0)
    params = [(rng.uniform(.01, .1), rng.uniform(.5, 2)) for _ in range(4)]
    runs = [_run(lc, cfg, a, s, 120) for a, s in params]
    net = pred_mod.ConvPredictor(129, hidden=(16, 16), seed=0, lr=2e-3)
    net.fit_scales([states[20:] for states, _ in runs[:3]])
    samples = []
    for states, beams in runs[:3]:
        for n in range(23, 120):
            s1, s2, s3 = states[n - 1], states[n - 2], states[n - 3]
            samples.append((net.inputs(s1, s2, s3, beams[n]).clone(),
                            net.targets(s1, s2, s3, states[n]).clone()))
    hist = net.fit(samples, epochs=8)
    say("  loss by epoch:", [f"{h:.3g}" for h in hist])
    states, beams = runs[3]
    err = {"learned": [], "linear": [], "quadratic": []}
    for n in range(23, 120):
        s1, s2, s3, s0 = states[n - 1], states[n - 2], states[n - 3], states[n]
        g = net.guess(s1, s2, s3, beams[n])
        num = {"learned": 0., "linear": 0., "quadratic": 0.}
        den = 0.
        for f, name in enumerate(pred_mod.FIELDS):
            F0, F1, F2, F3 = (getattr(s, name) for s in (s0, s1, s2, s3))
            num["learned"] += torch.sum((g[f] - F0) ** 2).item()
            num["linear"] += torch.sum((2 * F1 - F2 - F0) ** 2).item()
            num["quadratic"] += torch.sum((3 * F1 - 3 * F2 + F3 - F0) ** 2).item()
            den += torch.sum(F0 ** 2).item()
        for k in err:
            err[k].append((num[k] / den) ** .5)
    mean = {k: float(np.mean(v)) for k, v in err.items()}
    say("  held-out guess rel L2 (mean over 97 slices):",
        {k: f"{v:.3e}" for k, v in mean.items()})
    assert hist[-1] < hist[0]
    # measured: learned 2.1e-4, quadratic 3.3e-4, linear 2.9e-3
    assert mean["learned"] < 0.9 * mean["quadratic"]
