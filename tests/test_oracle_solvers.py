"""Pins the oracle's field solves: applying the discrete 5-point operator
with the boundary ghosts each transform implies to the solution must give
back the right-hand side the solver was handed (CPU only)."""

import types

import numpy as np
import pytest


def _cfg(N, h, trick=1):
    return types.SimpleNamespace(grid_steps=N, grid_step_size=h,
                                 xi_step_size=.05,
                                 field_solver_subtraction_trick=trick,
                                 field_solver_variant_A=True)


def _smooth_random(rng, N, zero_rim=2):
    a = rng.standard_normal((N, N))
    k = np.array([.25, .5, .25])
    for _ in range(3):
        a = np.apply_along_axis(lambda r: np.convolve(r, k, 'same'), 0, a)
        a = np.apply_along_axis(lambda r: np.convolve(r, k, 'same'), 1, a)
    a[:zero_rim] = a[-zero_rim:] = 0
    a[:, :zero_rim] = a[:, -zero_rim:] = 0
    return a


def _lap(u, h, ghost_x, ghost_y):
    """-Laplacian with ghosts: 'zero' (Dirichlet: u outside is 0) or
    'even' (u[-1] = u[1], u[N] = u[N-2])."""
    def pad(a, axis, kind):
        a = np.moveaxis(a, axis, 0)
        if kind == 'zero':
            lo, hi = np.zeros_like(a[:1]), np.zeros_like(a[:1])
        else:
            lo, hi = a[1:2], a[-2:-1]
        return np.moveaxis(np.concatenate([lo, a, hi]), 0, axis)
    p = pad(pad(u, 0, ghost_x), 1, ghost_y)
    return (4 * p[1:-1, 1:-1] - p[2:, 1:-1] - p[:-2, 1:-1]
            - p[1:-1, 2:] - p[1:-1, :-2]) / h**2


@pytest.mark.parametrize("N,h", [(17, .1), (33, .05), (65, .05)])
def test_Ez_solves_dirichlet_poisson(oracle, N, h):
    rng = np.random.default_rng(N)
    jx, jy = _smooth_random(rng, N), _smooth_random(rng, N)
    Ez = oracle.calculate_Ez(_cfg(N, h), jx, jy)
    assert np.all(Ez[0] == 0) and np.all(Ez[-1] == 0)
    assert np.all(Ez[:, 0] == 0) and np.all(Ez[:, -1] == 0)
    # -Lap Ez = -(djx/dx + djy/dy) (central differences) on the interior.
    div = ((jx[2:, 1:-1] - jx[:-2, 1:-1]) + (jy[1:-1, 2:] - jy[1:-1, :-2])) / (2 * h)
    got = _lap(Ez[1:-1, 1:-1], h, 'zero', 'zero')
    assert np.allclose(got, -div, rtol=0, atol=1e-10 * abs(div).max())


@pytest.mark.parametrize("N,h,trick", [(17, .1, 1), (33, .05, 1), (33, .05, 0)])
def test_mixed_solve_is_helmholtz_dirichlet_x_neumann_y(oracle, N, h, trick):
    rng = np.random.default_rng(N + trick)
    rhs = rng.standard_normal((N, N))
    u = oracle.mixed_solve(_cfg(N, h, trick), rhs)
    assert u.shape == (N, N)
    assert np.abs(u[0]).max() < 1e-14 and np.abs(u[-1]).max() < 1e-14
    got = _lap(u[1:-1], h, 'zero', 'even') + (1 if trick else 0) * u[1:-1]
    assert np.allclose(got, rhs[1:-1], rtol=0, atol=1e-9 * abs(rhs).max())


@pytest.mark.parametrize("N,h", [(17, .1), (33, .05)])
def test_transverse_solve_transposes_Ex_and_By(oracle, N, h):
    """Ex and By are solved on the transposed RHS: Dirichlet in y, even in x."""
    rng = np.random.default_rng(3)
    cfg = _cfg(N, h)
    f = [_smooth_random(rng, N) for _ in range(11)]
    Ex, Ey, Bx, By = oracle.calculate_Ex_Ey_Bx_By(cfg, *f)
    Ex_rhs, Ey_rhs, Bx_rhs, By_rhs = oracle.transverse_rhs(cfg, *f)
    for u, rhs, transposed in ((Ex, Ex_rhs, True), (Ey, Ey_rhs, False),
                               (Bx, Bx_rhs, False), (By, By_rhs, True)):
        if transposed:
            u, rhs = u.T, rhs.T
        got = _lap(u[1:-1], h, 'zero', 'even') + u[1:-1]
        assert np.allclose(got, rhs[1:-1], rtol=0, atol=1e-9 * abs(rhs).max())


@pytest.mark.parametrize("N,h", [(17, .1), (33, .05), (65, .05)])
def test_Bz_solves_neumann_poisson_minus_zero_mode(oracle, N, h):
    rng = np.random.default_rng(N + 5)
    jx, jy = _smooth_random(rng, N), _smooth_random(rng, N)
    Bz = oracle.calculate_Bz(_cfg(N, h), jx, jy)
    assert abs(Bz.mean()) < 1e-14 * abs(Bz).max()
    curl = np.zeros((N, N))
    curl[1:-1, 1:-1] = ((jx[1:-1, 2:] - jx[1:-1, :-2])
                        - (jy[2:, 1:-1] - jy[:-2, 1:-1])) / (2 * h)
    rhs = -curl
    # the DCT-I zero mode is the trapezoid-weighted mean
    w = np.ones(N)
    w[0] = w[-1] = .5
    c = (w[:, None] * w[None, :] * rhs).sum() / (N - 1)**2
    got = _lap(Bz, h, 'even', 'even')
    assert np.allclose(got, rhs - c, rtol=0, atol=1e-9 * abs(rhs).max())


def test_Ez_of_a_known_potential(oracle):
    """Manufactured solution: jx = dphi/dx numerically makes -Lap Ez equal
    -(div grad phi); checks scale and sign end to end with a smooth phi."""
    N, h = 65, .05
    x = (np.arange(N) - N // 2) * h
    X, Y = np.meshgrid(x, x, indexing='ij')
    L = h * (N - 1)
    # phi vanishes with its normal derivative's stencil support at the rim
    Ez_true = np.sin(np.pi * (X + L / 2) / L) * np.sin(2 * np.pi * (Y + L / 2) / L)
    # choose jx so that div j (central) equals Lap_h Ez_true on the interior
    lap = np.zeros((N, N))
    lap[1:-1, 1:-1] = -_lap(Ez_true[1:-1, 1:-1], h, 'zero', 'zero')
    # jx with d/dx (central) = lap: integrate along x with step 2h
    jx = np.zeros((N, N))
    for i in range(2, N):
        jx[i] = jx[i - 2] + 2 * h * lap[i - 1]
    Ez = oracle.calculate_Ez(_cfg(N, h), jx, np.zeros((N, N)))
    assert np.allclose(Ez[1:-1, 1:-1], Ez_true[1:-1, 1:-1], atol=1e-9)
