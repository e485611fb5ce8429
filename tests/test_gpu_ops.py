"""Parity of each HIP operator against the CPU oracle on seeded inputs
(calls go through the C ABI via lcode3d_amd).  Tolerances: the field
solves, pusher and deposition are FP64 and differ from the oracle only by
summation order (FFT algorithm, atomic order), so relative L2 <= 1e-12."""

import types

import numpy as np
import pytest
import torch

from gpu_helpers import dev, host, rel_l2, say, smooth_field

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def lc():
    import lcode3d_amd
    assert torch.cuda.is_available(), "no HIP device visible"
    return lcode3d_amd


def _cfg(N, h=.05, dxi=.05, trick=1, variant_A=True, c=2, f=2, pad=10):
    return types.SimpleNamespace(
        grid_steps=N, grid_step_size=h, xi_step_size=dxi,
        field_solver_subtraction_trick=trick,
        field_solver_variant_A=variant_A, reflect_padding_steps=5,
        plasma_padding_steps=pad, plasma_coarseness=c, plasma_fineness=f)


SIZES = [17, 25, 33, 41, 129, 641, 1281]   # 25, 41: runtime FFT plans (L = 48, 80);
# 641, 1281: the wave FFT (L = 1280, 2560)


@pytest.mark.parametrize("N", SIZES)
def test_calculate_Ez_and_Bz(lc, oracle, N):
    say(f"Ez/Bz N={N}")
    rng = np.random.default_rng(N)
    cfg = _cfg(N, h=3.2 / N)
    jx, jy = smooth_field(rng, N), smooth_field(rng, N)
    Ez = host(lc.calculate_Ez(cfg, dev(jx), dev(jy)))
    Bz = host(lc.calculate_Bz(cfg, dev(jx), dev(jy)))
    torch.cuda.synchronize()
    ez_ref = oracle.calculate_Ez(cfg, jx, jy)
    bz_ref = oracle.calculate_Bz(cfg, jx, jy)
    assert rel_l2(Ez, ez_ref) < 1e-12
    assert rel_l2(Bz, bz_ref) < 1e-12
    assert np.all(Ez[0] == 0) and np.all(Ez[:, -1] == 0)
    assert abs(Bz.mean()) < 1e-14 * np.abs(Bz).max()


@pytest.mark.parametrize("N", SIZES)
@pytest.mark.parametrize("trick", [1, 0])
def test_calculate_Ex_Ey_Bx_By(lc, oracle, N, trick):
    say(f"transverse N={N} trick={trick}")
    rng = np.random.default_rng(N + 1)
    cfg = _cfg(N, h=3.2 / N, trick=trick)
    args = [smooth_field(rng, N, zero_rim=1) for _ in range(4)]
    beam = smooth_field(rng, N, .05)
    dens = [smooth_field(rng, N) for _ in range(6)]
    got = lc.calculate_Ex_Ey_Bx_By(cfg, *[dev(a) for a in args], dev(beam),
                                   *[dev(d) for d in dens])
    want = oracle.calculate_Ex_Ey_Bx_By(cfg, *args, beam, *dens)
    for name, g, w in zip(("Ex", "Ey", "Bx", "By"), got, want):
        assert rel_l2(host(g), w) < 1e-12, name
    # a scalar beam is broadcast like the reference's `ro + beam_ro`
    got0 = lc.calculate_Ex_Ey_Bx_By(cfg, *[dev(a) for a in args], 0,
                                    *[dev(d) for d in dens])
    want0 = oracle.calculate_Ex_Ey_Bx_By(cfg, *args, 0, *dens)
    for g, w in zip(got0, want0):
        assert rel_l2(host(g), w) < 1e-12


def _particles(oracle, cfg, seed, amp_x, amp_p):
    xs, ys, const, vp, st = oracle.init(cfg)
    rng = np.random.default_rng(seed)
    nc = const.m.shape[0]
    xo = amp_x * rng.uniform(-1, 1, (nc, nc))
    yo = amp_x * rng.uniform(-1, 1, (nc, nc))
    px, py, pz = (amp_p * rng.uniform(-1, 1, (nc, nc)) for _ in range(3))
    return const, vp, xo, yo, px, py, pz


@pytest.mark.parametrize("N,c,f", [(33, 2, 2), (129, 2, 2), (641, 3, 2),
                                   (129, 3, 3)])
def test_move_estimate_and_move_smart(lc, oracle, N, c, f):
    say(f"push N={N} c={c} f={f}")
    cfg = _cfg(N, h=3.2 / N, c=c, f=f)
    const, vp, xo, yo, px, py, pz = _particles(oracle, cfg, N, 3 * cfg.grid_step_size, .3)
    # push some particles across the reflection boundary
    xo[0, :] = -(cfg.reflect_boundary + const.x_init[0, :]) - .01
    yo[:, -1] = (cfg.reflect_boundary - const.y_init[:, -1]) + .01
    ex, ey = oracle.move_estimate_wo_fields(cfg, const.m, const.x_init,
                                            const.y_init, xo, yo, px, py, pz)
    gex, gey = lc.move_estimate_wo_fields(cfg, *[dev(a) for a in (
        const.m, const.x_init, const.y_init, xo, yo, px, py, pz)])
    assert np.abs(host(gex) - ex).max() < 1e-13
    assert np.abs(host(gey) - ey).max() < 1e-13
    rng = np.random.default_rng(N + 7)
    F = [smooth_field(rng, N, .2, zero_rim=1) for _ in range(6)]
    # keep particles well inside for the push
    xo2, yo2 = xo.copy(), yo.copy()
    xo2[0, :] = 0
    yo2[:, -1] = 0
    ex2, ey2 = oracle.move_estimate_wo_fields(cfg, const.m, const.x_init,
                                              const.y_init, xo2, yo2, px, py, pz)
    want = oracle.move_smart(cfg, const.m, const.q, const.x_init, const.y_init,
                             xo2, yo2, ex2, ey2, px, py, pz, *F)
    got = lc.move_smart(cfg, *[dev(a) for a in (
        const.m, const.q, const.x_init, const.y_init, xo2, yo2, ex2, ey2,
        px, py, pz)], *[dev(a) for a in F])
    for name, g, w in zip(("x_offt", "y_offt", "px", "py", "pz"), got, want):
        assert np.abs(host(g) - w).max() < 1e-13 * max(1, np.abs(w).max()), name
    nc = const.m.shape[0]
    assert lc.fault_count(cfg, nc) == 0


@pytest.mark.parametrize("N,c,f,cells", [(33, 2, 2, 2), (129, 2, 2, 2),
                                         (641, 3, 2, 2), (129, 3, 3, 2),
                                         (129, 2, 2, 7), (641, 3, 2, 7)])
def test_deposit(lc, oracle, N, c, f, cells):
    """cells: amplitude of the random particle offsets in cells; at 7 most
    particles leave the tile halo (at most 6 cells) and take the atomic spill path."""
    say(f"deposit N={N} c={c} f={f} offsets<={cells} cells")
    cfg = _cfg(N, h=3.2 / N, c=c, f=f)
    const, vp, xo, yo, px, py, pz = _particles(oracle, cfg, N + 3,
                                               cells * cfg.grid_step_size, .3)
    want = oracle.deposit(cfg, const.ro_initial, xo, yo, const.m, const.q,
                          px, py, pz, vp)
    dvp = lc.GPUArrays(**{k: getattr(vp, k) for k in (
        "influence_prev", "influence_next", "indices_prev", "indices_next",
        "fine_grid")})
    got = lc.deposit(cfg, dev(const.ro_initial), *[dev(a) for a in (
        xo, yo, const.m, const.q, px, py, pz)], dvp)
    for name, g, w in zip(("ro", "jx", "jy", "jz"), got, want):
        assert rel_l2(host(g), w) < 1e-12, name
    # scalar-zero ion background (initial_deposition)
    got0 = lc.deposit(cfg, 0, *[dev(a) for a in (
        xo, yo, const.m, const.q, px, py, pz)], dvp)
    want0 = oracle.deposit(cfg, 0, xo, yo, const.m, const.q, px, py, pz, vp)
    assert rel_l2(host(got0[0]), want0[0]) < 1e-12
    assert lc.fault_count(cfg, const.m.shape[0]) == 0


@pytest.mark.parametrize("N,c,f,pad", [(129, 2, 2, 10), (641, 3, 2, 10),
                                       (129, 3, 3, 10)])
def test_init_matches_oracle(lc, oracle, N, c, f, pad):
    say(f"init N={N} c={c} f={f}")
    cfg_o = _cfg(N, h=3.2 / N, c=c, f=f, pad=pad)
    cfg_g = _cfg(N, h=3.2 / N, c=c, f=f, pad=pad)
    xs, ys, const, vp, st = oracle.init(cfg_o)
    gxs, gys, gconst, gvp, gst = lc.init(cfg_g)
    assert cfg_g.reflect_boundary == cfg_o.reflect_boundary
    assert np.array_equal(gxs, xs) and np.array_equal(gys, ys)
    for k in ("m", "q", "x_init", "y_init"):
        assert np.array_equal(host(getattr(gconst, k)), getattr(const, k)), k
    for k in ("influence_prev", "influence_next", "indices_prev",
              "indices_next", "fine_grid"):
        assert np.array_equal(host(getattr(gvp, k)), getattr(vp, k)), k
    assert rel_l2(host(gconst.ro_initial), const.ro_initial) < 1e-13
    for k in ("Ex", "ro", "x_offt", "pz"):
        assert not np.any(host(getattr(gst, k)))
