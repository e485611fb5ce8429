"""Pins the oracle's particle functions: TSC shape, gather/scatter,
plasma set-up, coarse->fine interpolation, deposition and the pushers
(CPU only)."""

import math
import types

import numpy as np
import pytest


def _bspline2(d):
    """Quadratic B-spline of the signed distance d (in cells)."""
    d = np.abs(d)
    return np.where(d < .5, .75 - d**2, np.where(d < 1.5, .5 * (1.5 - d)**2, 0.))


def test_weights_are_the_quadratic_bspline(oracle):
    rng = np.random.default_rng(0)
    N, h = 33, .1
    x = rng.uniform(-1.2, 1.2, 500)
    y = rng.uniform(-1.2, 1.2, 500)
    i, j, *w = oracle.weights(x, y, N, h)
    for (di, dj), wk in zip(oracle.STENCIL, w):
        cx = (i + di - N // 2) * h     # cell-centre coordinates
        cy = (j + dj - N // 2) * h
        ref = _bspline2((x - cx) / h) * _bspline2((y - cy) / h)
        assert np.allclose(wk, ref, atol=1e-14)
    assert np.allclose(sum(w), 1, atol=1e-14)
    # the centre cell is the nearest one
    assert np.all(np.abs(x - (i - N // 2) * h) <= h / 2 + 1e-12)


def test_interp9_is_exact_on_linear_fields(oracle):
    rng = np.random.default_rng(1)
    N, h = 33, .1
    g = (np.arange(N) - N // 2) * h
    a = 2.5 * g[:, None] - 1.25 * g[None, :] + .3
    x, y = rng.uniform(-1, 1, 300), rng.uniform(-1, 1, 300)
    i, j, *w = oracle.weights(x, y, N, h)
    assert np.allclose(oracle.interp9(a, i, j, *w), 2.5 * x - 1.25 * y + .3,
                       atol=1e-13)


def test_deposit9_matches_per_particle_loop(oracle):
    rng = np.random.default_rng(2)
    N, h = 17, .1
    x, y = rng.uniform(-.5, .5, 40), rng.uniform(-.5, .5, 40)
    val = rng.standard_normal(40)
    i, j, *w = oracle.weights(x, y, N, h)
    out = np.zeros((N, N))
    oracle.deposit9(out, i, j, val, *w)
    ref = np.zeros((N, N))
    for k in range(40):
        for (di, dj), wk in zip(oracle.STENCIL, w):
            ref[i[k] + di, j[k] + dj] += val[k] * wk[k]
    assert np.allclose(out, ref, atol=1e-14)
    assert math.isclose(out.sum(), val.sum(), abs_tol=1e-12)


def test_virt_params_worked_example(oracle):
    """The example written out in the reference's comment
    (lcode.py:527-535)."""
    coarse = np.array([-2., -1., 0., 1., 2.])
    fine = np.array([-2.4, -1.8, -1.2, -.6, 0., .6, 1.2, 1.8, 2.4])
    vp = oracle.virt_params_from_grids(coarse, fine, 1.)
    assert list(vp.indices_next) == [0, 1, 1, 2, 2, 3, 4, 4, 4]
    assert list(vp.indices_prev) == [0, 0, 0, 1, 1, 2, 3, 3, 4]
    # edges use only the existing neighbour
    assert vp.influence_prev[0] == 0 and vp.influence_next[0] == 1
    assert vp.influence_prev[-1] == 1 and vp.influence_next[-1] == 0
    inner = slice(1, -1)
    assert np.allclose(vp.influence_prev[inner] + vp.influence_next[inner], 1)
    # interpolating the coarse positions themselves gives the fine ones
    xs = (vp.influence_prev[inner] * coarse[vp.indices_prev[inner]]
          + vp.influence_next[inner] * coarse[vp.indices_next[inner]])
    assert np.allclose(xs, fine[inner])


@pytest.mark.parametrize("steps,coarseness,fineness,nc,nf",
                         [(109, 2, 2, 53, 216), (621, 3, 2, 205, 1240),
                          (1261, 3, 3, 419, 3779)])
def test_plasma_grid_sizes(oracle, steps, coarseness, fineness, nc, nf):
    c = oracle.make_coarse_plasma_grid(steps, .1, coarseness)
    f = oracle.make_fine_plasma_grid(steps, .1, fineness)
    assert len(c) == nc and len(f) == nf
    assert np.allclose(np.diff(c), .1 * coarseness)
    assert np.allclose(np.diff(f), .1 / fineness)
    if fineness % 2:
        assert 0. in f
    else:
        assert 0. not in f


def _cfg(N=33, h=.1, dxi=.05, coarseness=2, fineness=2):
    cfg = types.SimpleNamespace(
        grid_steps=N, grid_step_size=h, xi_step_size=dxi,
        field_solver_subtraction_trick=1, field_solver_variant_A=True,
        reflect_padding_steps=5, plasma_padding_steps=6,
        plasma_coarseness=coarseness, plasma_fineness=fineness)
    return cfg


def test_coarse_to_fine_zero_offsets(oracle):
    cfg = _cfg()
    (xi, yi, xo, yo, px, py, pz, m, q,
     vp) = oracle.make_plasma(cfg.grid_steps - 2 * cfg.plasma_padding_steps,
                              cfg.grid_step_size, 2, 2)
    s = 1 / 16
    x, y, fm, fq, fpx, fpy, fpz = oracle.coarse_to_fine(
        xo, yo, m, q, px, py, pz, s, vp)
    fg = vp.fine_grid
    assert np.array_equal(x, np.broadcast_to(fg[:, None], x.shape))
    assert np.array_equal(y, np.broadcast_to(fg[None, :], y.shape))
    # interior fine particles carry 1/fineness^2 of a unit charge
    assert np.allclose(fq[2:-2, 2:-2], -1 / 4)
    assert np.allclose(fm[2:-2, 2:-2], 1 / 4)


def test_initial_density_is_uniform_one(oracle):
    cfg = _cfg(N=65, h=.05)
    xs, ys, const, vp, state = oracle.init(cfg)
    N = cfg.grid_steps
    inner = const.ro_initial[15:N - 15, 15:N - 15]
    assert np.allclose(inner, 1, atol=1e-12)
    # every fine particle carries -1/fineness^2 (its four bilinear weights
    # sum to 1 even at the clipped edges), so the ions total Nf^2/fineness^2
    nf = len(vp.fine_grid)
    assert math.isclose(const.ro_initial.sum(), nf**2 / 4, rel_tol=1e-12)
    assert cfg.reflect_boundary == cfg.grid_step_size * (N / 2 - 5)


def test_deposit_charge_identity_and_brute_force(oracle):
    cfg = _cfg(N=33, h=.1)
    xs, ys, const, vp, state = oracle.init(cfg)
    rng = np.random.default_rng(4)
    shape = const.m.shape
    xo, yo = rng.uniform(-.1, .1, shape), rng.uniform(-.1, .1, shape)
    px, py, pz = (rng.uniform(-.3, .3, shape) for _ in range(3))
    ro, jx, jy, jz = oracle.deposit(cfg, const.ro_initial, xo, yo, const.m,
                                    const.q, px, py, pz, vp)
    # dro - djz = q for every fine particle, ro_initial cancels sum(q)
    assert abs((ro - jz).sum()) < 1e-11
    # brute force: one fine particle at a time
    s = 1 / (cfg.plasma_coarseness * cfg.plasma_fineness)**2
    ip, inn = vp.influence_prev, vp.influence_next
    P, Nn, fg = vp.indices_prev, vp.indices_next, vp.fine_grid
    N, h = cfg.grid_steps, cfg.grid_step_size
    ref = np.zeros((4, N, N))
    for fi in range(len(fg)):
        for fj in range(len(fg)):
            wts = (ip[fi] * ip[fj], ip[fi] * inn[fj], inn[fi] * ip[fj], inn[fi] * inn[fj])
            idx = ((P[fi], P[fj]), (P[fi], Nn[fj]), (Nn[fi], P[fj]), (Nn[fi], Nn[fj]))
            mix = lambda c: sum(w * c[a, b] for w, (a, b) in zip(wts, idx))  # noqa
            x = fg[fi] + mix(xo)
            y = fg[fj] + mix(yo)
            fm_, fq_ = s * mix(const.m), s * mix(const.q)
            fpx, fpy, fpz = s * mix(px), s * mix(py), s * mix(pz)
            g = math.sqrt(fm_**2 + fpx**2 + fpy**2 + fpz**2)
            dro = fq_ / (1 - fpz / g)
            vals = (dro, fpx * dro / g, fpy * dro / g, fpz * dro / g)
            ci = int(math.floor(x / h + .5)) + N // 2
            cj = int(math.floor(y / h + .5)) + N // 2
            for di in (-1, 0, 1):
                for dj in (-1, 0, 1):
                    wgt = (_bspline2((x - (ci + di - N // 2) * h) / h) *
                           _bspline2((y - (cj + dj - N // 2) * h) / h))
                    for c in range(4):
                        ref[c, ci + di, cj + dj] += vals[c] * wgt
    ref[0] += const.ro_initial
    for got, want in zip((ro, jx, jy, jz), ref):
        assert np.allclose(got, want, atol=1e-13)


def test_estimate_zero_field_drift_and_reflection(oracle):
    cfg = _cfg()
    cfg.reflect_boundary = 1.0
    x_init = np.array([0., .9, -.9, .5])
    y_init = np.array([0., 0., .95, -.99])
    m = np.ones(4)
    px = np.array([.1, .5, 0., 0.])
    py = np.array([0., 0., .3, -.2])
    pz = np.array([.2, 0., 0., .1])
    xo, yo = oracle.move_estimate_wo_fields(cfg, m, x_init, y_init,
                                            np.zeros(4), np.zeros(4), px, py, pz)
    g = np.sqrt(1 + px**2 + py**2 + pz**2)
    x = x_init + px / (g - pz) * .05
    y = y_init + py / (g - pz) * .05
    x = np.where(np.abs(x) >= 1, np.sign(x) * 2 - x, x)
    y = np.where(np.abs(y) >= 1, np.sign(y) * 2 - y, y)
    assert np.allclose(xo, x - x_init) and np.allclose(yo, y - y_init)
    assert np.all(np.abs(x_init + xo) <= 1) and np.all(np.abs(y_init + yo) <= 1)


def _push_uniform(oracle, cfg, p0, E, B, nsteps):
    """Push one particle nsteps through uniform fields."""
    N = cfg.grid_steps
    F = [np.full((N, N), v) for v in (*E, *B)]
    one = lambda v: np.array([float(v)])  # noqa: E731
    xo, yo = one(0), one(0)
    px, py, pz = (one(v) for v in p0)
    m, q = one(1), one(-1)
    for _ in range(nsteps):
        ex, ey = oracle.move_estimate_wo_fields(cfg, m, one(0), one(0), xo, yo,
                                                px, py, pz)
        xo, yo, px, py, pz = oracle.move_smart(
            cfg, m, q, one(0), one(0), xo, yo, ex, ey, px, py, pz, *F)
    return xo, yo, px, py, pz


def test_pusher_uniform_Ez_matches_exact_invariant(oracle):
    """In a uniform Ez (nothing else) the exact motion keeps
    gamma - pz = m - q Ez s after stepping a distance s; the pusher is
    second order, so its error falls ~4x when the step halves."""
    errs = []
    for k in range(3):
        cfg = _cfg(N=17, dxi=.2 / 2**k)
        cfg.reflect_boundary = 100.
        n = 10 * 2**k
        s = n * cfg.xi_step_size
        _, _, px, py, pz = _push_uniform(oracle, cfg, (0, 0, 0), (0, 0, .5),
                                         (0, 0, 0), n)
        g = np.sqrt(1 + pz**2)
        exact_gmpz = 1 + .5 * s  # m - q Ez s with q = -1
        errs.append(abs((g - pz)[0] - exact_gmpz))
        assert px[0] == 0 and py[0] == 0
    assert errs[0] < 1e-3
    assert 3.0 < errs[0] / errs[1] < 5.0 or errs[1] < 1e-14


def test_pusher_uniform_Bz_rotates(oracle):
    """A uniform Bz rotates p_perp; the magnitude stays put to second order
    and the angle after path s is q Bz s / (gamma - pz) (pz = 0)."""
    res = []
    for k in range(3):
        cfg = _cfg(N=17, dxi=.1 / 2**k)
        cfg.reflect_boundary = 100.
        n = 10 * 2**k
        _, _, px, py, pz = _push_uniform(oracle, cfg, (.2, 0, 0), (0, 0, 0),
                                         (0, 0, .7), n)
        ang = math.atan2(py[0], px[0])
        g = math.sqrt(1 + .2**2)
        # dp/ds = q (v x B) / (1 - vz):  angle rate = -q Bz / gamma
        exact = .7 / g * n * cfg.xi_step_size
        res.append(abs(ang - exact))
    assert res[0] < 1e-3
    assert res[0] / res[1] > 3.0 or res[1] < 1e-13


def test_move_smart_reflection_flips_momentum(oracle):
    cfg = _cfg(N=17)
    cfg.reflect_boundary = .5
    N = cfg.grid_steps
    F = [np.zeros((N, N)) for _ in range(6)]
    x_init = np.array([.49, -.49, 0.])
    y_init = np.array([0., 0., .49])
    z = np.zeros(3)
    px = np.array([1., -1., 0.])
    py = np.array([0., 0., 1.])
    pz = np.zeros(3)
    m, q = np.ones(3), -np.ones(3)
    xo, yo, npx, npy, npz = oracle.move_smart(cfg, m, q, x_init, y_init, z, z,
                                              z, z, px, py, pz, *F)
    g = np.sqrt(2.)
    d = 1 / g * cfg.xi_step_size
    assert np.allclose(x_init[:2] + xo[:2], [1 - (.49 + d), -1 + (.49 + d)])
    assert np.allclose(npx, [-1, 1, 0])
    assert np.allclose(y_init[2] + yo[2], 1 - (.49 + d)) and npy[2] == -1
