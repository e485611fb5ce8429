"""The learned predictor on the 641^2 workload (grid_step_size 0.025,
xi_step_size 0.005, coarseness 3, fineness 2): the U-Net trained on the GPU
by this package's own kernels on slices of the plain FP64 method, against
the quadratic extrapolation it refines.

Training: two plain runs (amplitude, sigma from default_rng(0) in
U[0.01, 0.1] x U[0.5, 2]) of 600 slices, every 3rd slice (~400 samples),
4 epochs of Adam.  Held out: the amplitude-0.05, sigma-1 beam.

What is checked (measured values printed):
* the learned guess of the predictor pass's fields F1 is closer to them
  than the extrapolation, averaged over slices 150..550 of the held-out run
  (relative L2 over the six fields);
* the mixed-precision step with the learned guess (GuidedStepper, corrector
  residual threshold 2e-3) keeps Ez(xi, 0, 0) within relative L2 1e-4 of
  the plain FP64 step over 600 slices and accepts most guesses."""

import types

import numpy as np
import pytest
import torch

from gpu_helpers import say

pytestmark = pytest.mark.gpu

FIELDS = ("Ex", "Ey", "Ez", "Bx", "By", "Bz")


@pytest.fixture(scope="module")
def lc():
    import lcode3d_amd
    assert torch.cuda.is_available(), "no HIP device visible"
    return lcode3d_amd


def _cfg(precision="fp64"):
    return types.SimpleNamespace(
        grid_steps=641, grid_step_size=.025, xi_step_size=.005,
        field_solver_subtraction_trick=1, field_solver_variant_A=True,
        reflect_padding_steps=5, plasma_padding_steps=10,
        plasma_coarseness=3, plasma_fineness=2, field_solver_precision=precision)


@pytest.fixture(scope="module")
def net(lc):
    from lcode3d_amd import unet
    say("learned predictor: collecting and training at 641^2")
    rng = np.random.default_rng(0)
    runs = [(rng.uniform(.01, .1), rng.uniform(.5, 2)) for _ in range(2)]
    n = unet.UNetPredictor(641, seed=0, lr=1e-3)
    samples, sx, st = unet.collect_training_data(lc, n, _cfg(), runs, 600, every=3)
    hist = unet.train(n, samples, (sx, st), 4)
    say(f"  {len(samples)} samples, loss by epoch {[f'{h:.3g}' for h in hist]}")
    assert hist[-1] < hist[0]
    return n


def _rel6(a, b):
    num = sum(torch.linalg.norm(a[k].double() - b[k].double()) ** 2 for k in range(6))
    den = sum(torch.linalg.norm(b[k].double()) ** 2 for k in range(6))
    return (num / den).sqrt().item()


def test_learned_guess_beats_extrapolation(lc, net):
    from lcode3d_amd.beams import cosine_gaussian
    c = _cfg()
    _, _, const, vp, s = lc.init(c)
    hist, eq, eg = [], [], []
    for i in range(551):
        beam = cosine_gaussian(c, i)
        if i >= 150 and i % 50 == 0:
            f1 = lc.predictor_fields(c, const, vp, hist[0], beam)
            q = [3 * getattr(hist[0], k) - 3 * getattr(hist[1], k) + getattr(hist[2], k)
                 for k in FIELDS]
            net.refresh_every = 1
            g = net.guess(hist[0], hist[1], hist[2], beam)
            eq.append(_rel6(q, f1))
            eg.append(_rel6(g, f1))
        s = lc.step(c, const, vp, s, beam)
        hist.insert(0, s)
        del hist[3:]
    say(f"guess vs F1 over slices 150..550: extrapolation {np.mean(eq):.2e}, "
        f"learned {np.mean(eg):.2e}")
    assert np.mean(eg) < np.mean(eq)


def test_mixed_learned_step_641(lc, net):
    from lcode3d_amd.beams import cosine_gaussian
    c64, cm = _cfg(), _cfg("mixed")
    _, _, const, vp, sp = lc.init(c64)
    _, _, constm, vpm, sm = lc.init(cm)
    net.refresh_every = 4
    gs = lc.GuidedStepper(cm, constm, vpm, 2e-3, predictor=net.guess)
    ezp, ezm = [], []
    for i in range(600):
        b = cosine_gaussian(c64, i)
        sp = lc.step(c64, const, vp, sp, b)
        sm = gs.step(sm, b)
        ezp.append(sp.Ez[320, 320])
        ezm.append(sm.Ez[320, 320])
    gs.drain(wait=True)
    ezp, ezm = torch.stack(ezp), torch.stack(ezm).double()
    axis = (torch.linalg.norm(ezm - ezp) / torch.linalg.norm(ezp)).item()
    full = (torch.linalg.norm(sm.Ez.double() - sp.Ez) / torch.linalg.norm(sp.Ez)).item()
    say(f"mixed+learned 641^2 x 600: Ez axis {axis:.1e}, final Ez {full:.1e}; "
        f"{gs.accepted} of {gs.guided} accepted, {gs.plain} plain")
    assert axis < 1e-4 and full < 1e-4
    assert gs.accepted > gs.guided // 2
