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


def test_checkpoint_resume_continues_the_run(lc, tmp_path):
    """save_state/load_state mid-run: resuming from the file reproduces the
    uninterrupted run (to 1e-10 of its peak: FP64 atomics round in any order)."""
    from lcode3d_amd import state_io
    from lcode3d_amd.beams import cosine_gaussian
    say("checkpoint/resume 129^2")
    cfg = _cfg(129, .05)
    xs, ys, const, vp, st = lc.init(cfg)
    for i in range(30):
        st = lc.step(cfg, const, vp, st, cosine_gaussian(cfg, i))
    path = tmp_path / "slice_29.npz"
    state_io.save_state(path, st, cfg, xi_i=29, const=const, virt_params=vp)
    ref = st
    for i in range(30, 50):
        ref = lc.step(cfg, const, vp, ref, cosine_gaussian(cfg, i))
    header, st2, const2, vp2 = state_io.load_state(path)
    assert header["xi_i"] == 29
    cfg2 = _cfg(129, .05)
    cfg2.reflect_boundary = header["config"]["reflect_boundary"]
    for i in range(header["xi_i"] + 1, 50):
        st2 = lc.step(cfg2, const2, vp2, st2, cosine_gaussian(cfg2, i))
    for name in ("Ez", "Ex", "ro", "x_offt", "pz"):
        a, b = getattr(st2, name), getattr(ref, name)
        assert torch.allclose(a, b, rtol=0, atol=1e-10 * b.abs().max().item()), name
