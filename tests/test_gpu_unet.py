"""The learned predictor's kernels (BF16 MFMA convolutions, pooling,
upsampling, their adjoints, the loss, Adam, the plane conversions) against
the U-Net oracle (oracle/unet_oracle.py, pinned by finite differences),
through the C ABI; then the trained network on the 641^2 workload.

Tolerances: the oracle rounds to bfloat16 at the points the kernels store
bfloat16, so the two differ only by FP32 against FP64 summation and the
rare bfloat16 rounding flip that follows from it: forward activations and
output within relative L2 1e-2, weight and bias gradients within 3e-2
(measured values are printed)."""

import types

import numpy as np
import pytest
import torch

from gpu_helpers import host, rel_l2, say

pytestmark = pytest.mark.gpu

S0 = 23                       # levels 23, 12, 6, 3
GRID = 4 * S0 - 3             # level-0 size (N + 3) / 4 = S0


@pytest.fixture(scope="module")
def lc():
    import lcode3d_amd
    assert torch.cuda.is_available(), "no HIP device visible"
    return lcode3d_amd


def _bf_to_chw(t):
    """int16 bfloat16 bits [s, s, 32] -> float64 [32, s, s]."""
    u = t.cpu().numpy().view(np.uint16).astype(np.uint32) << 16
    return u.view(np.float32).astype(np.float64).transpose(2, 0, 1)


def _chw_to_bf(x, s):
    """float64 [c, s, s] (c <= 32) -> int16 bfloat16 bits [s, s, 32] device."""
    hw = np.zeros((s, s, 32), dtype=np.float32)
    hw[..., :x.shape[0]] = x.transpose(1, 2, 0)
    return torch.as_tensor(hw, device="cuda").to(torch.bfloat16).view(torch.int16)


# enum order of unet.hip's tensors and their levels
TENSORS = {"e0": (0, 0), "p0": (1, 1), "e1": (2, 1), "p1": (3, 2), "e2": (4, 2),
           "p2": (5, 3), "e3": (6, 3), "u3": (7, 2), "d2": (8, 2), "u2": (9, 1),
           "d1": (10, 1), "u1": (11, 0), "d0": (12, 0)}


def test_forward_backward_match_oracle(lc):
    from lcode3d_amd.unet import UNetPredictor
    from oracle import unet_oracle as uo
    say(f"unet forward/backward vs oracle, S0={S0}")
    rng = np.random.default_rng(0)
    net = UNetPredictor(GRID, seed=0)
    assert net.S == S0
    params = uo.init_params(seed=5)
    params = [(w, rng.standard_normal(b.shape) * .05) for w, b in params]
    net.set_params(params)
    x = uo.bf16(rng.standard_normal((uo.CIN, S0, S0)))
    t = rng.standard_normal((uo.COUT, S0, S0)) * .3
    xd = _chw_to_bf(x, S0)
    y = net.forward(xd)
    ref = uo.forward(params, x, q=uo.bf16)
    errs = {}
    for name, (idx, lvl) in TENSORS.items():
        errs[name] = rel_l2(_bf_to_chw(net.tensor(idx, lvl)), ref[name])
    errs["y"] = rel_l2(host(y).transpose(2, 0, 1), ref["y"])
    say("forward rel L2:", {k: f"{v:.1e}" for k, v in errs.items()})
    for k, v in errs.items():
        assert v < 1e-2, (k, v)
    td = torch.as_tensor(t.transpose(1, 2, 0).astype(np.float32), device="cuda").contiguous()
    net.backward(xd, td)
    loss, gy = uo.mse_grad(ref["y"], t)
    grads = uo.backward(params, ref, gy, q=uo.bf16)
    got = net.layer_arrays(net.grad)
    gerr = []
    for k, ((gw, gb), (rw, rb)) in enumerate(zip(got, grads)):
        gerr.append((rel_l2(gw, rw), rel_l2(gb, rb)))
    say("gradient rel L2 (w, b) per layer:", [f"{a:.1e}/{b:.1e}" for a, b in gerr])
    assert abs(net.loss.item() - loss) < 1e-3 * loss
    for k, (ew, eb) in enumerate(gerr):
        assert ew < 3e-2 and eb < 3e-2, (k, ew, eb)


def test_adam_step_matches_oracle(lc):
    from lcode3d_amd.unet import UNetPredictor
    from oracle import predictor_oracle as po
    from oracle import unet_oracle as uo
    say("unet Adam step")
    rng = np.random.default_rng(1)
    net = UNetPredictor(GRID, seed=3, lr=1e-3)
    p0 = net.params.double().cpu().numpy()
    x = _chw_to_bf(uo.bf16(rng.standard_normal((uo.CIN, S0, S0))), S0)
    t = torch.as_tensor(rng.standard_normal((S0, S0, 6)).astype(np.float32), device="cuda")
    net.train_step(x, t)
    g = net.grad.double().cpu().numpy()
    want, _, _ = po.adam(p0, g, np.zeros_like(p0), np.zeros_like(p0), 1, 1e-3)
    got = net.params.double().cpu().numpy()
    assert np.abs(got - want).max() <= 1e-6 * max(np.abs(want).max(), 1)


def _state(rng, n, dtype):
    st = types.SimpleNamespace()
    for k in ("Ex", "Ey", "Ez", "Bx", "By", "Bz", "ro"):
        setattr(st, k, torch.as_tensor(rng.standard_normal((n, n)), device="cuda").to(dtype))
    return st


def _mean4(a, s):
    n = a.shape[0]
    p = np.zeros((4 * s, 4 * s))
    p[:n, :n] = a
    return p.reshape(s, 4, s, 4).sum(axis=(1, 3)) / 16


@pytest.mark.parametrize("base", ["quadratic", "alternating"])
@pytest.mark.parametrize("dtype", [torch.float64, torch.float32])
def test_io_planes_targets_guess(lc, dtype, base):
    from lcode3d_amd.unet import FIELDS, UNetPredictor
    say(f"unet io planes, {dtype}, {base} base")
    rng = np.random.default_rng(2)
    net = UNetPredictor(GRID, seed=0)
    net.base = base
    s1, s2, s3, s4 = (_state(rng, GRID, dtype) for _ in range(4))

    def base_of(name):
        a = [h(getattr(s, name)) for s in (s1, s2, s3, s4)]
        if base == "alternating":
            return 2 * a[0] - 2 * a[2] + a[3]
        return 3 * a[0] - 3 * a[1] + a[2]
    beam = torch.as_tensor(rng.standard_normal((GRID, GRID)), device="cuda")
    raw = host(net.raw_inputs(s1, s2, beam))
    h = lambda t: t.double().cpu().numpy()  # noqa: E731
    for f, name in enumerate(FIELDS):
        a1, a2 = h(getattr(s1, name)), h(getattr(s2, name))
        assert np.allclose(raw[..., f], _mean4(a1, S0), rtol=1e-6, atol=1e-6)
        assert np.allclose(raw[..., 6 + f], _mean4(a1 - a2, S0), rtol=1e-6, atol=1e-6)
    assert np.allclose(raw[..., 12], _mean4(h(beam), S0), rtol=1e-6, atol=1e-6)
    assert np.allclose(raw[..., 13], _mean4(h(s1.ro), S0), rtol=1e-6, atol=1e-6)
    nxt = [torch.as_tensor(rng.standard_normal((GRID, GRID)), device="cuda").to(dtype)
           for _ in FIELDS]
    tg = host(net.raw_targets(s1, s2, s3, nxt, s4))
    for f, name in enumerate(FIELDS):
        q = base_of(name)
        assert np.allclose(tg[..., f], _mean4(h(nxt[f]) - q, S0), rtol=1e-5, atol=1e-5)
    net.tscale = np.linspace(.5, 2, 6).astype(np.float32)
    g = net.guess(s1, s2, s3, beam, s4)
    y = host(net.y.view(S0, S0, 16))
    for f, name in enumerate(FIELDS):
        q = base_of(name)
        up = np.repeat(np.repeat(y[..., f].astype(np.float64), 4, 0), 4, 1)[:GRID, :GRID]
        want = q + float(net.tscale[f]) * up
        tol = 1e-12 if dtype == torch.float64 else 1e-5
        assert rel_l2(h(g[f]), want) < tol, name


def test_infer_equals_forward(lc):
    """The inference form (pooling and upsampling inside the convolutions'
    staging) gives the training forward's output bit for bit."""
    from lcode3d_amd.unet import UNetPredictor
    from oracle import unet_oracle as uo
    say("unet infer vs forward")
    rng = np.random.default_rng(7)
    net = UNetPredictor(GRID, seed=4)
    x = _chw_to_bf(uo.bf16(rng.standard_normal((uo.CIN, S0, S0))), S0)
    y1 = net.forward(x).clone()
    y2 = net.infer(x).clone()
    assert torch.equal(y1, y2)
