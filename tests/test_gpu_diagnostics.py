"""The on-device zn diagnostic against scipy.ndimage.gaussian_filter (the
library call of lcode.py:983-987) and the run driver of lcode.py:1032-1051
against a plain step loop."""

import types

import numpy as np
import pytest
import scipy.ndimage
import torch

from gpu_helpers import dev, say, smooth_field

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def lc():
    import lcode3d_amd
    assert torch.cuda.is_available(), "no HIP device visible"
    return lcode3d_amd


def _cfg(N, h, **kw):
    base = dict(grid_steps=N, grid_step_size=h, xi_step_size=.05,
                field_solver_subtraction_trick=1, field_solver_variant_A=True,
                reflect_padding_steps=5, plasma_padding_steps=10,
                plasma_coarseness=2, plasma_fineness=2)
    base.update(kw)
    return types.SimpleNamespace(**base)


def _zn_scipy(config, ro):
    sigma = 0.25 / config.grid_step_size
    hf = ro - scipy.ndimage.gaussian_filter(ro, sigma=sigma)
    return np.abs(hf).mean() / 4.23045376e-04


# h = 0.1 puts the blur radius (10 cells) above half the 17-point grid, so
# the reflection wraps more than once
@pytest.mark.parametrize("N,h", [(17, .1), (129, .05), (641, .025)])
def test_zn_matches_scipy(lc, N, h):
    from lcode3d_amd.diagnostics import zn
    say(f"zn N={N} h={h}")
    cfg = _cfg(N, h)
    rng = np.random.default_rng(N)
    ro = smooth_field(rng, N, .1, zero_rim=0) + 1e-3 * rng.standard_normal((N, N))
    got = zn(cfg, dev(ro)).item()
    want = _zn_scipy(cfg, ro)
    assert abs(got - want) <= 1e-12 * abs(want)


def test_main_driver_matches_step_loop(lc, capsys):
    from lcode3d_amd.beams import cosine_gaussian
    from lcode3d_amd.main import main
    import lcode3d_amd.diagnostics as diags
    say("main driver 129^2, 30 slices")
    cfg = _cfg(129, .05, xi_steps=30, diagnostics_each_N_steps=10,
               gpu_index=0)
    cfg.beam = lambda xi_i, xs, ys: cosine_gaussian(cfg, xi_i)
    diags.max_zn = 0.0
    st = main(cfg, write_png=False)
    lines = [ln for ln in capsys.readouterr().out.splitlines()
             if ln.startswith("xi=")]
    assert len(lines) == 4   # slices 0, 10, 20 and the last
    # the same run as a plain loop
    xs, ys, const, vp, ref = lc.init(cfg)
    ez = []
    for xi_i in range(30):
        ref = lc.step(cfg, const, vp, ref, cosine_gaussian(cfg, xi_i))
        ez.append(ref.Ez[64, 64].item())
    # FP64 atomics make runs agree to rounding, not bit for bit
    assert torch.allclose(st.Ez, ref.Ez, rtol=0, atol=1e-12 * ref.Ez.abs().max().item())
    assert lines[-1].startswith(f"xi={-29 * .05:+.4f} {ez[-1]:+.4e}|")
    assert lines[-1].split("|")[1] == diags.diags_peak_msg(ez)
    zn_last = diags.zn(cfg, ref.ro).item()
    assert float(lines[-1].split("zn=")[1]) >= round(zn_last, 3) - 1e-3
