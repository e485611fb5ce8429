"""The FP32-transform (mixed precision) field solves against the FP64
oracle.  Transforms, spectra and eigenvalue tables are FP32; the RHS, the
Bz zero mode and mean, Variant A and the averages stay FP64.  Tolerances
are stated per check: a single solve keeps relative L2 <= 1e-5 (the Bz
mean is zero to 1e-7 of max |Bz|); the step
keeps Ez(xi, 0, 0) and every field within 1e-4 relative L2 over 200 slices
of the 129^2 workload."""

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


def _cfg(N, precision="fp32", **kw):
    base = dict(grid_steps=N, grid_step_size=3.2 / N, xi_step_size=.05,
                field_solver_subtraction_trick=1, field_solver_variant_A=True,
                reflect_padding_steps=5, plasma_padding_steps=10,
                plasma_coarseness=2, plasma_fineness=2,
                field_solver_precision=precision)
    base.update(kw)
    return types.SimpleNamespace(**base)


@pytest.mark.parametrize("N", [33, 41, 129, 641, 1281])
def test_fp32_solves(lc, oracle, N):
    say(f"fp32 solves N={N}")
    rng = np.random.default_rng(N)
    cfg = _cfg(N)
    jx, jy = smooth_field(rng, N), smooth_field(rng, N)
    assert rel_l2(host(lc.calculate_Ez(cfg, dev(jx), dev(jy))),
                  oracle.calculate_Ez(cfg, jx, jy)) < 1e-5
    bz = host(lc.calculate_Bz(cfg, dev(jx), dev(jy)))
    assert rel_l2(bz, oracle.calculate_Bz(cfg, jx, jy)) < 1e-5
    # the mean is removed in FP64 from the analytic column sums of the FP32
    # spectra: what is left is the FP32 rounding of the last transform
    assert abs(bz.mean()) < 1e-7 * np.abs(bz).max()
    args = [smooth_field(rng, N) for _ in range(4)]
    dens = [smooth_field(rng, N) for _ in range(6)]
    got = lc.calculate_Ex_Ey_Bx_By(cfg, *[dev(a) for a in args], 0,
                                   *[dev(d) for d in dens])
    want = oracle.calculate_Ex_Ey_Bx_By(cfg, *args, 0, *dens)
    for g, w in zip(got, want):
        assert rel_l2(host(g), w) < 1e-5


def test_fp32_step_129_200_slices(lc, oracle):
    cfg_o = _cfg(129, "fp64", grid_step_size=.05)
    cfg_g = _cfg(129, "fp32", grid_step_size=.05)
    xs, ys, const, vp, st = oracle.init(cfg_o)
    _, _, gconst, gvp, gst = lc.init(cfg_g)
    worst, ez_o, ez_g = {}, [], []
    for xi_i in range(200):
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
    say("fp32 129^2 worst rel L2:", {k: f"{v:.1e}" for k, v in worst.items()})
    for name, e in worst.items():
        assert e < 1e-4, (name, e)
    ez_o = np.array(ez_o)
    assert np.abs(np.array(ez_g) - ez_o).max() < 1e-4 * np.abs(ez_o).max()


def test_mixed_step_129_200_slices(lc, oracle):
    """The mixed-precision step (FP32 grids, deposition arithmetic and
    transforms; FP64 particles, RHS arithmetic, Bz mean, ion background)
    against the FP64 oracle over the 129^2 workload: every field within
    relative L2 1e-3, momenta 1e-4, Ez(xi, 0, 0) within 1e-4 of its peak."""
    cfg_o = _cfg(129, "fp64", grid_step_size=.05)
    cfg_g = _cfg(129, "mixed", grid_step_size=.05)
    xs, ys, const, vp, st = oracle.init(cfg_o)
    _, _, gconst, gvp, gst = lc.init(cfg_g)
    assert gst.Ex.dtype == torch.float32 and gst.x_offt.dtype == torch.float64
    worst, ez_o, ez_g = {}, [], []
    for xi_i in range(200):
        beam = oracle.beam(cfg_o, xi_i, xs, ys)
        st = oracle.step(cfg_o, const, vp, st, beam)
        gst = lc.step(cfg_g, gconst, gvp, gst, beam)
        ez_o.append(st.Ez[64, 64])
        ez_g.append(gst.Ez[64, 64].item())
        if (xi_i + 1) % 50 == 0:
            for name in STATE:
                floor = np.linalg.norm(st.Bx) if name == "Bz" else 0.
                e = rel_l2(host(getattr(gst, name)).astype(np.float64),
                           getattr(st, name), floor)
                worst[name] = max(worst.get(name, 0.), e)
    say("mixed 129^2 worst rel L2:", {k: f"{v:.1e}" for k, v in worst.items()})
    for name, e in worst.items():
        bound = 1e-4 if name in ("px", "py", "pz", "x_offt", "y_offt") else 1e-3
        assert e < bound, (name, e)
    ez_o = np.array(ez_o)
    assert np.abs(np.array(ez_g) - ez_o).max() < 1e-4 * np.abs(ez_o).max()
    assert lc.fault_count(cfg_g, gconst.m.shape[0]) == 0


def test_mixed_state_dtypes_checked(lc):
    """A mixed config takes float32 grids: an FP64 state is refused before
    any launch, and the standalone operators stay FP64."""
    cfg = _cfg(33, "mixed")
    _, _, const, vp, st = lc.init(cfg)
    cfg64 = _cfg(33, "fp64")
    _, _, _, _, st64 = lc.init(cfg64)
    with pytest.raises(ValueError, match="float32"):
        lc.step(cfg, const, vp, st64, 0)
    with pytest.raises(ValueError, match="float64"):
        lc.step(cfg64, const, vp, st, 0)
    ez = lc.calculate_Ez(cfg, torch.zeros_like(st64.Ex), torch.zeros_like(st64.Ex))
    assert ez.dtype == torch.float64
