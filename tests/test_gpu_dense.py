"""The dense form of the field solves (N x N sine/cosine matrices on MFMA)
against the FP64 oracle.

* FP64 MFMA (field_solver_transform='dense', precision 'fp64'): the same
  arithmetic as the oracle up to summation order, so single solves keep
  relative L2 <= 1e-12 and the step keeps 1e-10 over 100 slices of the
  129^2 workload, like the FFT path.
* BF16x3 (precision 'bf16x3'): both operands split into hi + lo BF16
  (|x - hi - lo| <= 2^-18 |x|), three products accumulated in FP32, so a
  single transform carries a relative error of a few 1e-6; a solve (four
  transforms) keeps relative L2 <= 1e-4 and the step keeps Ez(xi, 0, 0)
  within 1e-3 of its peak over 200 slices.  The measured errors are printed.
"""

import types

import numpy as np
import pytest
import torch

from gpu_helpers import STATE, dev, host, rel_l2, say, smooth_field

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def lc():
    import lcode3d_amd
    assert torch.cuda.is_available(), "no HIP device visible"
    return lcode3d_amd


def _cfg(N, precision, **kw):
    base = dict(grid_steps=N, grid_step_size=3.2 / N, xi_step_size=.05,
                field_solver_subtraction_trick=1, field_solver_variant_A=True,
                reflect_padding_steps=5, plasma_padding_steps=10,
                plasma_coarseness=2, plasma_fineness=2,
                field_solver_transform="dense",
                field_solver_precision=precision)
    base.update(kw)
    return types.SimpleNamespace(**base)


BOUND = {"fp64": 1e-12, "bf16x3": 1e-4}


@pytest.mark.parametrize("precision", ["fp64", "bf16x3"])
@pytest.mark.parametrize("N", [17, 25, 129, 641])
def test_dense_solves(lc, oracle, N, precision):
    say(f"dense {precision} solves N={N}")
    rng = np.random.default_rng(N + 100)
    cfg = _cfg(N, precision)
    jx, jy = smooth_field(rng, N), smooth_field(rng, N)
    errs = {}
    ez = host(lc.calculate_Ez(cfg, dev(jx), dev(jy)))
    errs["Ez"] = rel_l2(ez, oracle.calculate_Ez(cfg, jx, jy))
    bz = host(lc.calculate_Bz(cfg, dev(jx), dev(jy)))
    errs["Bz"] = rel_l2(bz, oracle.calculate_Bz(cfg, jx, jy))
    for trick in (1, 0):
        cfg.field_solver_subtraction_trick = trick
        args = [smooth_field(rng, N) for _ in range(4)]
        beam = smooth_field(rng, N, .05)
        dens = [smooth_field(rng, N) for _ in range(6)]
        got = lc.calculate_Ex_Ey_Bx_By(cfg, *[dev(a) for a in args], dev(beam),
                                       *[dev(d) for d in dens])
        want = oracle.calculate_Ex_Ey_Bx_By(cfg, *args, beam, *dens)
        for name, g, w in zip(("Ex", "Ey", "Bx", "By"), got, want):
            errs[f"{name} trick={trick}"] = rel_l2(host(g), w)
    say(f"  rel L2: {{{', '.join(f'{k}: {v:.1e}' for k, v in errs.items())}}}")
    for name, e in errs.items():
        assert e < BOUND[precision], (name, e)
    # boundary rows/columns the solvers leave at zero stay exact zeros
    assert np.all(ez[0] == 0) and np.all(ez[:, -1] == 0)
    # the Bz mean is removed from the spectra's analytic column sums
    assert abs(bz.mean()) < (1e-13 if precision == "fp64" else 1e-6) * np.abs(bz).max()


@pytest.mark.parametrize("precision,slices", [("fp64", 100), ("bf16x3", 200)])
def test_dense_step_129(lc, oracle, precision, slices):
    say(f"dense {precision} step 129^2, {slices} slices")
    cfg_o = _cfg(129, "fp64", grid_step_size=.05)
    cfg_g = _cfg(129, precision, grid_step_size=.05)
    xs, ys, const, vp, st = oracle.init(cfg_o)
    _, _, gconst, gvp, gst = lc.init(cfg_g)
    worst, ez_o, ez_g = {}, [], []
    for xi_i in range(slices):
        beam = oracle.beam(cfg_o, xi_i, xs, ys)
        st = oracle.step(cfg_o, const, vp, st, beam)
        gst = lc.step(cfg_g, gconst, gvp, gst, beam)
        ez_o.append(st.Ez[64, 64])
        ez_g.append(gst.Ez[64, 64].item())
        if (xi_i + 1) % 50 == 0:
            for name in STATE:
                floor = np.linalg.norm(st.Bx) if name == "Bz" else 0.
                e = rel_l2(host(getattr(gst, name)), getattr(st, name), floor)
                worst[name] = max(worst.get(name, 0.), e)
    ez_o, ez_g = np.array(ez_o), np.array(ez_g)
    ez_err = np.abs(ez_g - ez_o).max() / np.abs(ez_o).max()
    say(f"  worst rel L2: {{{', '.join(f'{k}: {v:.1e}' for k, v in worst.items())}}}"
        f", Ez axis max {ez_err:.1e}")
    if precision == "fp64":
        for name, e in worst.items():
            assert e < 1e-10, (name, e)
        assert ez_err < 1e-10
    else:
        assert ez_err < 1e-3
    assert lc.fault_count(cfg_g, const.m.shape[0]) == 0


def test_dense_guided_fallback_matches_plain(lc):
    """The gated dense solve: a rejected guess (zeros, threshold 1e-5) runs
    the plain predictor, so the result equals the plain dense step."""
    from lcode3d_amd.beams import cosine_gaussian
    say("dense guided fallback")
    cfg = _cfg(129, "fp64", grid_step_size=.05)
    xs, ys, const, vp, st = lc.init(cfg)
    for i in range(40):
        st = lc.step(cfg, const, vp, st, cosine_gaussian(cfg, i))
    beam = cosine_gaussian(cfg, 40)
    plain = lc.step(cfg, const, vp, st, beam)
    zeros = tuple(torch.zeros_like(st.Ex) for _ in range(6))
    guided = lc.step(cfg, const, vp, st, beam, guess=[(1.0, zeros)],
                     threshold=1e-5)
    assert lc.guided_residual(cfg, const.m.shape[0], vp.fine_grid.numel()) > 1e-5
    for name in STATE:
        a, b = host(getattr(guided, name)), host(getattr(plain, name))
        floor = np.linalg.norm(host(plain.Bx)) if name == "Bz" else 0.
        assert rel_l2(a, b, floor) < 1e-12, name
