"""Parity of the fused HIP xi step (lcode_step through the C ABI) against
the CPU oracle.

* 129^2 production-shaped run (grid_step_size 0.05, xi_step_size 0.05,
  coarseness 2, fineness 2, Variant A, trick 1, cosine-Gaussian beam of
  amplitude 0.05): 200 slices from init, per-field relative L2 <= 1e-10
  every 50 slices and Ez(xi, 0, 0) on every slice.
* 641^2 (grid_step_size 0.025, xi_step_size 0.005, coarseness 3,
  fineness 2) and 1281^2 (0.0125, 0.0025, coarseness 3, fineness 3): one
  step from a perturbed state, relative L2 <= 1e-11.
* Variant B and the Laplace (trick 0) paths, one step each.
* Particles displaced up to 7.5 cells (beyond the deposition tiles' 6-cell
  halo: the spill path inside the fused step), two steps.
Tolerances are relative L2 over each array; Bz uses the Bx norm as its floor
(Bz is small for a round beam).
"""

import types

import numpy as np
import pytest
import torch

from gpu_helpers import (STATE, dev, host, perturbed_state, rel_l2, say,
                         to_device_state)

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def lc():
    import lcode3d_amd
    assert torch.cuda.is_available(), "no HIP device visible"
    return lcode3d_amd


def _cfg(**kw):
    base = dict(grid_steps=129, grid_step_size=.05, xi_step_size=.05,
                field_solver_subtraction_trick=1, field_solver_variant_A=True,
                reflect_padding_steps=5, plasma_padding_steps=10,
                plasma_coarseness=2, plasma_fineness=2)
    base.update(kw)
    return types.SimpleNamespace(**base)


def _floor(name, ref):
    if name == "Bz":
        return np.linalg.norm(ref.Bx)
    return 0.0


def test_step_129_200_slices(lc, oracle):
    cfg_o, cfg_g = _cfg(), _cfg()
    xs, ys, const, vp, st = oracle.init(cfg_o)
    _, _, gconst, gvp, gst = lc.init(cfg_g)
    n = cfg_o.grid_steps
    worst = {}
    ez_o, ez_g = [], []
    for xi_i in range(200):
        beam = oracle.beam(cfg_o, xi_i, xs, ys)
        st = oracle.step(cfg_o, const, vp, st, beam)
        gst = lc.step(cfg_g, gconst, gvp, gst, beam)
        ez_o.append(st.Ez[n // 2, n // 2])
        ez_g.append(gst.Ez[n // 2, n // 2].item())
        if (xi_i + 1) % 50 == 0:
            say(f"129^2 slice {xi_i + 1}")
            for name in STATE:
                ref = getattr(st, name)
                err = rel_l2(host(getattr(gst, name)), ref, _floor(name, st))
                worst[name] = max(worst.get(name, 0), err)
    say("129^2 worst rel L2:", {k: f"{v:.2e}" for k, v in worst.items()})
    for name, err in worst.items():
        assert err < 1e-10, (name, err)
    ez_o, ez_g = np.array(ez_o), np.array(ez_g)
    assert np.abs(ez_g - ez_o).max() < 1e-10 * np.abs(ez_o).max()
    assert lc.fault_count(cfg_g, const.m.shape[0]) == 0


def _one_step(lc, oracle, cfg_o, cfg_g, seed, beam_scale=.05):
    xs, ys, const, vp, st = perturbed_state(oracle, cfg_o, seed)
    lc.init(cfg_g)  # sets reflect_boundary like the oracle's init
    gconst = lc.GPUArrays(**{k: dev(getattr(const, k)) for k in (
        "m", "q", "x_init", "y_init", "ro_initial")})
    gvp = lc.GPUArrays(**{k: getattr(vp, k) for k in (
        "influence_prev", "influence_next", "indices_prev", "indices_next",
        "fine_grid")})
    gst = to_device_state(lc, st)
    beam = beam_scale * np.exp(-.5 * (xs**2 + ys**2))
    ref = oracle.step(cfg_o, const, vp, st, beam)
    got = lc.step(cfg_g, gconst, gvp, gst, dev(beam))
    errs = {name: rel_l2(host(getattr(got, name)), getattr(ref, name),
                         _floor(name, ref)) for name in STATE}
    # prev stays valid and unmodified (lcode.py:926-931)
    assert np.array_equal(host(gst.Ex), st.Ex)
    assert lc.fault_count(cfg_g, const.m.shape[0]) == 0
    return errs


def test_step_641_one_step_perturbed(lc, oracle):
    say("641^2 one step")
    kw = dict(grid_steps=641, grid_step_size=.025, xi_step_size=.005,
              plasma_coarseness=3, plasma_fineness=2)
    errs = _one_step(lc, oracle, _cfg(**kw), _cfg(**kw), seed=11)
    say("641^2 rel L2:", {k: f"{v:.1e}" for k, v in errs.items()})
    for name, err in errs.items():
        assert err < 1e-11, (name, err)


def test_step_1281_fineness_3_one_step_perturbed(lc, oracle):
    """1281^2 (grid_step_size 0.0125, xi_step_size 0.0025, coarseness 3,
    fineness 3): transforms of 2560 points (one line pair per block,
    twiddles from global memory) and groups of 9 fine particles."""
    say("1281^2 fineness 3 one step")
    kw = dict(grid_steps=1281, grid_step_size=.0125, xi_step_size=.0025,
              plasma_coarseness=3, plasma_fineness=3)
    errs = _one_step(lc, oracle, _cfg(**kw), _cfg(**kw), seed=13, beam_scale=.1)
    say("1281^2 rel L2:", {k: f"{v:.1e}" for k, v in errs.items()})
    for name, err in errs.items():
        assert err < 1e-11, (name, err)


@pytest.mark.parametrize("variant_A,trick", [(False, 1), (True, 0), (False, 0)])
def test_step_variants(lc, oracle, variant_A, trick):
    say(f"variant_A={variant_A} trick={trick}")
    kw = dict(field_solver_variant_A=variant_A,
              field_solver_subtraction_trick=trick)
    errs = _one_step(lc, oracle, _cfg(**kw), _cfg(**kw), seed=5)
    for name, err in errs.items():
        assert err < 1e-11, (name, err)


def test_step_accepts_scalar_and_host_beams(lc, oracle):
    say("scalar/host beams")
    cfg = _cfg()
    xs, ys, gconst, gvp, gst = lc.init(cfg)
    a = lc.step(cfg, gconst, gvp, gst, 0)
    b = lc.step(cfg, gconst, gvp, gst, np.zeros((129, 129)))
    for name in STATE:
        assert torch.allclose(getattr(a, name), getattr(b, name), rtol=0,
                              atol=1e-12), name
    # quiet start stays at rest (to round-off: the deposition's FP64
    # atomics add in a different order than the initial deposition did)
    for name in STATE:
        assert torch.abs(getattr(a, name)).max().item() < 1e-12, name
    view = lc.GPUArraysView(a)
    assert isinstance(view.Ez, np.ndarray) and view.Ez.shape == (129, 129)


def test_step_particles_beyond_the_deposition_halo(lc, oracle):
    """Offsets up to 7.5 cells: the pushes' displacement bound exceeds the
    largest tile halo (6 cells), so the fused step's depositions list the particles
    beyond it and deposit them with global atomics (the spill path), and
    the bound slots hand over between the steps; two steps against the
    oracle, relative L2 <= 1e-11 and 1e-10."""
    say("129^2 two steps, offsets up to 7.5 cells")
    cfg_o, cfg_g = _cfg(), _cfg()
    xs, ys, const, vp, st = perturbed_state(oracle, cfg_o, seed=5,
                                            amp_x=7.5 * cfg_o.grid_step_size)
    lc.init(cfg_g)
    gconst = lc.GPUArrays(**{k: dev(getattr(const, k)) for k in (
        "m", "q", "x_init", "y_init", "ro_initial")})
    gvp = lc.GPUArrays(**{k: getattr(vp, k) for k in (
        "influence_prev", "influence_next", "indices_prev", "indices_next",
        "fine_grid")})
    beam = .05 * np.exp(-.5 * (xs**2 + ys**2))
    ref, got = st, to_device_state(lc, st)
    for k, tol in enumerate((1e-11, 1e-10)):
        ref = oracle.step(cfg_o, const, vp, ref, beam)
        got = lc.step(cfg_g, gconst, gvp, got, dev(beam))
        for name in STATE:
            err = rel_l2(host(getattr(got, name)), getattr(ref, name), _floor(name, ref))
            assert err < tol, (k, name, err)
    disp = max(np.abs(ref.x_offt).max(), np.abs(ref.y_offt).max()) / cfg_o.grid_step_size
    say(f"  max offset after two steps {disp:.2f} cells")
    assert disp > 6.5
    assert lc.fault_count(cfg_g, const.m.shape[0]) == 0


def test_step_641_particles_ten_cells_out(lc, oracle):
    """641^2: offsets up to 10 cells, beyond the 6-cell largest tile halo,
    so every deposition of the step lists the far particles and the
    grid-wide spill kernel deposits them; one step against the oracle,
    relative L2 <= 1e-11."""
    say("641^2 one step, offsets up to 10 cells (spill kernel)")
    kw = dict(grid_steps=641, grid_step_size=.025, xi_step_size=.005,
              plasma_coarseness=3, plasma_fineness=2)
    cfg_o, cfg_g = _cfg(**kw), _cfg(**kw)
    xs, ys, const, vp, st = perturbed_state(oracle, cfg_o, seed=17,
                                            amp_x=10 * cfg_o.grid_step_size)
    lc.init(cfg_g)
    gconst = lc.GPUArrays(**{k: dev(getattr(const, k)) for k in (
        "m", "q", "x_init", "y_init", "ro_initial")})
    gvp = lc.GPUArrays(**{k: getattr(vp, k) for k in (
        "influence_prev", "influence_next", "indices_prev", "indices_next",
        "fine_grid")})
    beam = .05 * np.exp(-.5 * (xs**2 + ys**2))
    ref = oracle.step(cfg_o, const, vp, st, beam)
    got = lc.step(cfg_g, gconst, gvp, to_device_state(lc, st), dev(beam))
    torch.cuda.synchronize()
    errs = {name: rel_l2(host(getattr(got, name)), getattr(ref, name),
                         _floor(name, ref)) for name in STATE}
    say("  rel L2:", {k: f"{v:.1e}" for k, v in errs.items()})
    for name, err in errs.items():
        assert err < 1e-11, (name, err)
    disp = max(np.abs(st.x_offt).max(), np.abs(st.y_offt).max()) / cfg_o.grid_step_size
    assert disp > 9
    assert lc.fault_count(cfg_g, const.m.shape[0]) == 0


@pytest.mark.parametrize("cells", [4.5, 8.0])
def test_step_641_local_displacement(lc, oracle, cells):
    """641^2: offsets that are zero except in two bumps (one pushing
    particles outward, one inward) of up to `cells` cells, so the
    deposition tiles' per-tile halos differ from tile to tile: particles
    leave tiles with large bounds for tiles whose own bound is zero (4.5
    cells), or go beyond the largest halo and take the spill kernel (8
    cells); one step against the oracle, relative L2 <= 1e-11."""
    say(f"641^2 one step, two local displacement bumps of up to {cells} cells")
    kw = dict(grid_steps=641, grid_step_size=.025, xi_step_size=.005,
              plasma_coarseness=3, plasma_fineness=2)
    cfg_o, cfg_g = _cfg(**kw), _cfg(**kw)
    xs, ys, const, vp, st = perturbed_state(oracle, cfg_o, seed=23)
    h, sig = cfg_o.grid_step_size, .3
    x0, y0 = const.x_init, const.y_init
    st.x_offt = np.zeros_like(x0)
    st.y_offt = np.zeros_like(y0)
    # peak |offset| of A r / sig exp(-r^2 / 2 sig^2) is A exp(-1/2)
    amp = cells * h / np.exp(-.5)
    for (cx, cy), sign in (((2.0, -1.3), 1), ((-1.1, 0.7), -1)):
        dx, dy = x0 - cx, y0 - cy
        g = sign * amp / sig * np.exp(-(dx * dx + dy * dy) / (2 * sig * sig))
        st.x_offt += g * dx
        st.y_offt += g * dy
    lc.init(cfg_g)
    gconst = lc.GPUArrays(**{k: dev(getattr(const, k)) for k in (
        "m", "q", "x_init", "y_init", "ro_initial")})
    gvp = lc.GPUArrays(**{k: getattr(vp, k) for k in (
        "influence_prev", "influence_next", "indices_prev", "indices_next",
        "fine_grid")})
    beam = .05 * np.exp(-.5 * (xs**2 + ys**2))
    ref = oracle.step(cfg_o, const, vp, st, beam)
    got = lc.step(cfg_g, gconst, gvp, to_device_state(lc, st), dev(beam))
    torch.cuda.synchronize()
    errs = {name: rel_l2(host(getattr(got, name)), getattr(ref, name),
                         _floor(name, ref)) for name in STATE}
    say("  rel L2:", {k: f"{v:.1e}" for k, v in errs.items()})
    for name, err in errs.items():
        assert err < 1e-11, (name, err)
    disp = max(np.abs(st.x_offt).max(), np.abs(st.y_offt).max()) / h
    assert cells - .1 < disp < cells + .1
    assert lc.fault_count(cfg_g, const.m.shape[0]) == 0


def test_step_nonfinite_particle_is_reported(lc, oracle):
    """A NaN momentum makes the step's depositions write NaN grids (no
    particle is deposited, so nothing grinds on a NaN cell index) and the
    next step() call raises FloatingPointError; a reset workspace then
    steps normally again."""
    say("129^2 NaN momentum")
    cfg = _cfg()
    xs, ys, gconst, gvp, gst = lc.init(cfg)
    bad = lc.GPUArrays(**{k: getattr(gst, k).clone() for k in STATE})
    bad.px[3, 5] = float("nan")
    out = lc.step(cfg, gconst, gvp, bad, 0)
    torch.cuda.synchronize()
    assert torch.isnan(out.ro).all() and torch.isnan(out.jz).all()
    with pytest.raises(FloatingPointError):
        lc.step(cfg, gconst, gvp, gst, 0)
    # the flag was reset by the raise: a clean state steps normally
    ok = lc.step(cfg, gconst, gvp, gst, 0)
    torch.cuda.synchronize()
    assert torch.isfinite(ok.ro).all()
    lc.step(cfg, gconst, gvp, ok, 0)
