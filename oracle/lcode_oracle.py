"""CPU oracle: a NumPy/SciPy restatement of LCODE 3D's per-xi-slice step.

TEST INFRASTRUCTURE ONLY.  This module is the checker for the HIP path and
the CPU baseline timed by ``bench.py``.  Only ``tests/``,
``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of ``bench.py`` may
import it.  The product package ``lcode3d_amd`` never imports it and has no
CPU fallback.

Every function cites the reference line range it restates (``lcode.py:N``
is lotov/lcode3d's single source file).  CuPy arrays become NumPy arrays,
``cp.fft.rfft2`` becomes ``scipy.fft.rfft2``, and the two numba kernels
become vectorised NumPy over all particles at once; the 9-point scatter uses
``np.bincount``, so the oracle is deterministic (the reference's FP64 atomics
are not).

Parity pinning (what checks each function, all under ``tests/``):

* ``dst2d``/``dct2d``/``mix2d`` - against ``scipy.fft.dstn``/``dctn`` (type 1)
  and against self-inversion (T(T(a)) = (2(N-1))^2 a).
* ``calculate_Ez``/``calculate_Bz``/``calculate_Ex_Ey_Bx_By`` solves - the
  5-point discrete operator with the matching boundary ghosts, applied to
  the solution, gives back the right-hand side (Neumann: minus its
  trapezoid-weighted mean); eigen-tables against brute-force eigenvalues of
  the 1D second-difference matrices.
* ``weights``/``interp9``/``deposit9`` - quadratic B-spline closed form,
  partition of unity, exactness on linear fields, brute-force loops.
* ``make_plasma``/``coarse_to_fine`` - hand-worked example from the
  reference's own comment (lcode.py:527-535), zero-offset identity,
  bilinear weights summing to 1.
* ``deposit``/``initial_deposition`` - exact charge identity
  sum(ro - jz) = sum(q) per slice, interior background density +1,
  brute-force per-particle loop on a tiny grid.
* ``move_estimate_wo_fields``/``move_smart`` - zero-field drift, analytic
  uniform-Ez kick, reflection cases by hand.
* ``step`` - quiet start (zero beam keeps everything at rest), mirror and
  x<->y symmetry of an axisymmetric beam, exact charge identity every
  slice, linear wake theory at small amplitude, and the on-axis wake-peak
  drift of the reference's diagnostic (lcode.py:992-1002): within 3 % over
  three plasma periods at two xi steps, peaks one plasma period apart.
  The composition (which field set feeds each push and trick term, what is
  returned) is pinned by mutation: six plausible mis-compositions, each
  built from these operators, drift by 13-58 % and fail that pin
  (tests/test_oracle_step.py).
* ``step_guided``/``predictor_fields`` - with the predictor's own fields as
  the guess the guided step equals ``step`` bit for bit at any threshold; a
  rejected guess gives ``step`` bit for bit; the residual of the perfect
  guess is the predictor-corrector gap computed directly; an accepted
  imperfect guess changes the result (tests/test_oracle_guided.py).
"""

import math
import os
import types

import numpy as np
import scipy.fft

ELECTRON_CHARGE = -1.0   # lcode.py:46
ELECTRON_MASS = 1.0      # lcode.py:47

_WORKERS = len(os.sched_getaffinity(0))


def set_fft_workers(n):
    """Thread count handed to scipy.fft (the CPU baseline reports it)."""
    global _WORKERS
    _WORKERS = int(n)


def fft_workers():
    return _WORKERS


class Arrays(types.SimpleNamespace):
    """Attribute bag of host arrays; mirrors ``GPUArrays`` (lcode.py:52-69)."""


# --------------------------------------------------------------------------
# Transforms: type-1 DST/DCT built from padded real FFTs
# --------------------------------------------------------------------------

def dst2d(a):
    """2D DST-I via an anti-symmetrically padded rFFT (lcode.py:124-141)."""
    assert a.shape[0] == a.shape[1]
    n = a.shape[0]
    p = np.zeros((2 * n + 2, 2 * n + 2))
    p[1:n + 1, 1:n + 1] = a
    p[1:n + 1, n + 2:] = -a[:, ::-1]
    p[n + 2:, 1:n + 1] = -a[::-1, :]
    p[n + 2:, n + 2:] = a[::-1, ::-1]
    return -scipy.fft.rfft2(p, workers=_WORKERS)[1:n + 1, 1:n + 1].real


def mix2d(a):
    """DST-I along axis 0 x DCT-I along axis 1 (lcode.py:190-212).

    Output shape is (M + 2, N); rows 0 and M + 1 are zero."""
    m, n = a.shape
    p = np.zeros((2 * m + 2, 2 * n - 2))
    p[1:m + 1, :n] = a
    p[m + 2:2 * m + 2, :n] = -a[::-1, :]
    p[1:m + 1, n - 1:2 * n - 2] = a[:, ::-1][:, :-1]
    p[m + 2:2 * m + 2, n - 1:2 * n - 2] = -a[::-1, ::-1][:, :-1]
    return -scipy.fft.rfft2(p, workers=_WORKERS)[:m + 2, :n].imag


def dct2d(a):
    """2D DCT-I via a symmetrically padded rFFT, sign-flipped
    (lcode.py:295-313)."""
    assert a.shape[0] == a.shape[1]
    n = a.shape[0]
    p = np.zeros((2 * n - 2, 2 * n - 2))
    p[:n, :n] = a
    p[n:, :n] = a[::-1, :][1:-1, :]
    p[:n, n:] = a[:, ::-1][:, 1:-1]
    p[n:, n:] = a[::-1, ::-1][1:-1, 1:-1]
    return -scipy.fft.rfft2(p, workers=_WORKERS)[:n, :n].real


# --------------------------------------------------------------------------
# Eigenvalue tables (the reference memoizes them with cp.memoize)
# --------------------------------------------------------------------------

def _lam(k, grid_steps, h):
    """Second-difference eigenvalues 4/h^2 sin^2(k pi / (2(N-1)))
    (lcode.py:152-154, :227-228, :326)."""
    return 4 / h**2 * np.sin(k * np.pi / (2 * (grid_steps - 1)))**2


def dirichlet_matrix(grid_steps, h):
    """lcode.py:144-157."""
    lam = _lam(np.arange(1, grid_steps - 1), grid_steps, h)
    mul = 1 / (lam[:, None] + lam[None, :])
    return mul / (2 * (grid_steps - 1))**2


def mixed_matrix(grid_steps, h, subtraction_trick):
    """lcode.py:215-231.  The trick enters as a boolean (:230)."""
    li = _lam(np.arange(1, grid_steps - 1), grid_steps, h)
    lj = _lam(np.arange(grid_steps), grid_steps, h)
    mul = 1 / (li[:, None] + lj[None, :] + (1 if subtraction_trick else 0))
    return mul / (2 * (grid_steps - 1))**2


def neumann_matrix(grid_steps, h):
    """lcode.py:316-330; the [0, 0] zero mode is set to 0 (:328-329)."""
    lam = _lam(np.arange(0, grid_steps), grid_steps, h)
    with np.errstate(divide='ignore'):
        mul = 1 / (lam[:, None] + lam[None, :])
    mul[0, 0] = 0
    return mul / (2 * (grid_steps - 1))**2


# --------------------------------------------------------------------------
# Field solvers
# --------------------------------------------------------------------------

def calculate_Ez(config, jx, jy):
    """Dirichlet Poisson solve for Ez (lcode.py:160-184).

    rhs_inner = -(djx/dx + djy/dy) over the interior (:166-168); note the code
    has this sign, which is what is restated."""
    h = config.grid_step_size
    djx_dx = jx[2:, 1:-1] - jx[:-2, 1:-1]
    djy_dy = jy[1:-1, 2:] - jy[1:-1, :-2]
    rhs_inner = -(djx_dx + djy_dy) / (h * 2)
    f = dst2d(rhs_inner)
    f *= dirichlet_matrix(config.grid_steps, h)
    return np.pad(dst2d(f), 1, 'constant', constant_values=0)


def dx_dy(arr, h):
    """Central differences with a zero perimeter (lcode.py:234-243)."""
    dx, dy = np.zeros_like(arr), np.zeros_like(arr)
    dx[1:-1, 1:-1] = arr[2:, 1:-1] - arr[:-2, 1:-1]
    dy[1:-1, 1:-1] = arr[1:-1, 2:] - arr[1:-1, :-2]
    return dx / (h * 2), dy / (h * 2)


def transverse_rhs(config, Ex_avg, Ey_avg, Bx_avg, By_avg,
                   beam_ro, ro, jx, jy, jz, jx_prev, jy_prev):
    """Right-hand sides of the four transverse equations
    (lcode.py:258-268)."""
    h = config.grid_step_size
    dro_dx, dro_dy = dx_dy(ro + beam_ro, h)
    djz_dx, djz_dy = dx_dy(jz + beam_ro, h)
    djx_dxi = (jx_prev - jx) / config.xi_step_size
    djy_dxi = (jy_prev - jy) / config.xi_step_size
    trick = config.field_solver_subtraction_trick
    Ex_rhs = -((dro_dx - djx_dxi) - Ex_avg * trick)
    Ey_rhs = -((dro_dy - djy_dxi) - Ey_avg * trick)
    Bx_rhs = +((djz_dy - djy_dxi) + Bx_avg * trick)
    By_rhs = -((djz_dx - djx_dxi) - By_avg * trick)
    return Ex_rhs, Ey_rhs, Bx_rhs, By_rhs


def mixed_solve(config, rhs):
    """mix2d(mix_mat * mix2d(rhs[1:-1, :])[1:-1, :]) (lcode.py:275-286)."""
    mat = mixed_matrix(config.grid_steps, config.grid_step_size,
                       config.field_solver_subtraction_trick)
    return mix2d(mat * mix2d(rhs[1:-1, :])[1:-1, :])


def calculate_Ex_Ey_Bx_By(config, Ex_avg, Ey_avg, Bx_avg, By_avg,
                          beam_ro, ro, jx, jy, jz, jx_prev, jy_prev):
    """Mixed-boundary Helmholtz solves (lcode.py:246-290).  Ey and Bx are
    solved as they are, Ex and By on the transposed RHS (:275-288)."""
    Ex_rhs, Ey_rhs, Bx_rhs, By_rhs = transverse_rhs(
        config, Ex_avg, Ey_avg, Bx_avg, By_avg,
        beam_ro, ro, jx, jy, jz, jx_prev, jy_prev)
    Ey = mixed_solve(config, Ey_rhs)
    Bx = mixed_solve(config, Bx_rhs)
    By = mixed_solve(config, By_rhs.T).T
    Ex = mixed_solve(config, Ex_rhs.T).T
    return Ex, Ey, Bx, By


def calculate_Bz(config, jx, jy):
    """Neumann Poisson solve for Bz, mean removed (lcode.py:333-364)."""
    h = config.grid_step_size
    djx_dy = np.pad(jx[1:-1, 2:] - jx[1:-1, :-2], 1, 'constant')
    djy_dx = np.pad(jy[2:, 1:-1] - jy[:-2, 1:-1], 1, 'constant')
    rhs = -(djx_dy - djy_dx) / (h * 2)
    f = dct2d(rhs)
    f *= neumann_matrix(config.grid_steps, h)
    Bz = dct2d(f)
    Bz -= Bz.mean()
    return Bz


# --------------------------------------------------------------------------
# Particles
# --------------------------------------------------------------------------

def move_estimate_wo_fields(config, m, x_init, y_init, prev_x_offt,
                            prev_y_offt, px, py, pz):
    """Field-free half-step estimate with coordinate-only reflection using
    >= / <= (lcode.py:369-391)."""
    x, y = x_init + prev_x_offt, y_init + prev_y_offt
    gamma_m = np.sqrt(m**2 + pz**2 + px**2 + py**2)
    x = x + px / (gamma_m - pz) * config.xi_step_size
    y = y + py / (gamma_m - pz) * config.xi_step_size
    r = config.reflect_boundary
    x = np.where(x >= +r, +2 * r - x, x)
    x = np.where(x <= -r, -2 * r - x, x)
    y = np.where(y >= +r, +2 * r - y, y)
    y = np.where(y <= -r, -2 * r - y, y)
    return x - x_init, y - y_init


def weights(x, y, grid_steps, h):
    """TSC-2D cell index and 9 weights (lcode.py:396-418).

    Returns i, j (int64 arrays) and the weights in the reference's order
    wMP, w0P, wPP, wM0, w00, wP0, wMM, w0M, wPM."""
    x_h, y_h = x / h + .5, y / h + .5
    fx, fy = np.floor(x_h), np.floor(y_h)
    i = (fx + grid_steps // 2).astype(np.int64)
    j = (fy + grid_steps // 2).astype(np.int64)
    x_loc, y_loc = x_h - fx - .5, y_h - fy - .5
    wx0, wy0 = .75 - x_loc**2, .75 - y_loc**2
    wxP, wyP = (.5 + x_loc)**2 / 2, (.5 + y_loc)**2 / 2
    wxM, wyM = (.5 - x_loc)**2 / 2, (.5 - y_loc)**2 / 2
    return (i, j,
            wxM * wyP, wx0 * wyP, wxP * wyP,
            wxM * wy0, wx0 * wy0, wxP * wy0,
            wxM * wyM, wx0 * wyM, wxP * wyM)


#: (di, dj) of the 9 stencil points in the reference's weight order
#: (lcode.py:427-429, :441-449).
STENCIL = ((-1, +1), (0, +1), (+1, +1),
           (-1, 0), (0, 0), (+1, 0),
           (-1, -1), (0, -1), (+1, -1))


def interp9(a, i, j, *w):
    """9-point gather, summed in the reference's order (lcode.py:421-430)."""
    acc = a[i - 1, j + 1] * w[0]
    for (di, dj), wk in zip(STENCIL[1:], w[1:]):
        acc = acc + a[i + di, j + dj] * wk
    return acc


def deposit9(out, i, j, val, *w):
    """9-point scatter-add of val * w_k (lcode.py:433-449), as bincount."""
    n0, n1 = out.shape
    idx = np.concatenate([((i + di) * n1 + (j + dj)).ravel()
                          for di, dj in STENCIL])
    vals = np.concatenate([(val * wk).ravel() for wk in w])
    out += np.bincount(idx, weights=vals, minlength=n0 * n1).reshape(n0, n1)


def make_coarse_plasma_grid(steps, step_size, coarseness=3):
    """lcode.py:454-465."""
    assert coarseness == int(coarseness)
    plasma_step = step_size * coarseness
    right_half = np.arange(steps // (coarseness * 2)) * plasma_step
    left_half = -right_half[:0:-1]
    grid = np.concatenate([left_half, right_half])
    assert np.array_equal(grid, -grid[::-1])
    return grid


def make_fine_plasma_grid(steps, step_size, fineness=2):
    """lcode.py:468-487; odd fineness puts particles on the axes."""
    assert fineness == int(fineness)
    plasma_step = step_size / fineness
    if fineness % 2:
        right_half = np.arange(steps // 2 * fineness) * plasma_step
        left_half = -right_half[:0:-1]
    else:
        right_half = (.5 + np.arange(steps // 2 * fineness)) * plasma_step
        left_half = -right_half[::-1]
    grid = np.concatenate([left_half, right_half])
    assert np.array_equal(grid, -grid[::-1])
    return grid


def virt_params_from_grids(coarse_grid, fine_grid, coarse_step):
    """Neighbour indices and bilinear influences of each fine position
    (lcode.py:526-546), for one axis (the same arrays serve x and y)."""
    nc = len(coarse_grid)
    indices = np.searchsorted(coarse_grid, fine_grid)
    indices_next = np.clip(indices, 0, nc - 1)
    indices_prev = np.clip(indices - 1, 0, nc - 1)
    influence_prev = (coarse_grid[indices_next] - fine_grid) / coarse_step
    influence_next = (fine_grid - coarse_grid[indices_prev]) / coarse_step
    influence_prev[indices_next == 0] = 0
    influence_next[indices_next == 0] = 1
    influence_next[indices_prev == nc - 1] = 0
    influence_prev[indices_prev == nc - 1] = 1
    return Arrays(influence_prev=influence_prev,
                  influence_next=influence_next,
                  indices_prev=indices_prev,
                  indices_next=indices_next,
                  fine_grid=fine_grid)


def make_plasma(steps, cell_size, coarseness=3, fineness=2):
    """Coarse particle arrays and coarse->fine interpolation parameters
    (lcode.py:490-573)."""
    coarse_grid = make_coarse_plasma_grid(steps, cell_size, coarseness)
    fine_grid = make_fine_plasma_grid(steps, cell_size, fineness)
    nc = len(coarse_grid)

    x_init = np.broadcast_to(coarse_grid[:, None], (nc, nc)).copy()
    y_init = np.broadcast_to(coarse_grid[None, :], (nc, nc)).copy()
    zeros = [np.zeros((nc, nc)) for _ in range(5)]
    m = np.ones((nc, nc)) * ELECTRON_MASS * coarseness**2
    q = np.ones((nc, nc)) * ELECTRON_CHARGE * coarseness**2
    virt_params = virt_params_from_grids(coarse_grid, fine_grid,
                                         cell_size * coarseness)
    x_offt, y_offt, px, py, pz = zeros
    return x_init, y_init, x_offt, y_offt, px, py, pz, m, q, virt_params


def coarse_to_fine(c_x_offt, c_y_offt, c_m, c_q, c_px, c_py, c_pz,
                   smallness, virt_params):
    """Bilinear coarse->fine interpolation for every fine particle at once
    (lcode.py:576-621).  Returns (Nf, Nf) arrays indexed [fi, fj]."""
    ip, inn = virt_params.influence_prev, virt_params.influence_next
    pi_, ni_ = virt_params.indices_prev, virt_params.indices_next
    A = ip[:, None] * ip[None, :]
    B = ip[:, None] * inn[None, :]
    C = inn[:, None] * ip[None, :]
    D = inn[:, None] * inn[None, :]
    P, Nn = pi_[:, None], ni_[:, None]
    Pj, Nj = pi_[None, :], ni_[None, :]

    def mix(c):
        return A * c[P, Pj] + B * c[P, Nj] + C * c[Nn, Pj] + D * c[Nn, Nj]

    fg = virt_params.fine_grid
    x = fg[:, None] + mix(c_x_offt)
    y = fg[None, :] + mix(c_y_offt)
    m = smallness * mix(c_m)
    q = smallness * mix(c_q)
    px = smallness * mix(c_px)
    py = smallness * mix(c_py)
    pz = smallness * mix(c_pz)
    return x, y, m, q, px, py, pz


def deposit(config, ro_initial, x_offt, y_offt, m, q, px, py, pz,
            virt_params):
    """Coarse->fine interpolation and TSC deposition of ro, jx, jy, jz;
    ro_initial is added last (lcode.py:626-689)."""
    n = config.grid_steps
    smallness = 1 / (config.plasma_coarseness * config.plasma_fineness)**2
    x, y, fm, fq, fpx, fpy, fpz = coarse_to_fine(
        x_offt, y_offt, m, q, px, py, pz, smallness, virt_params)
    gamma_m = np.sqrt(fm**2 + fpx**2 + fpy**2 + fpz**2)
    dro = fq / (1 - fpz / gamma_m)
    djx = fpx * (dro / gamma_m)
    djy = fpy * (dro / gamma_m)
    djz = fpz * (dro / gamma_m)
    i, j, *w = weights(x, y, n, config.grid_step_size)
    ro, jx, jy, jz = (np.zeros((n, n)) for _ in range(4))
    deposit9(ro, i, j, dro, *w)
    deposit9(jx, i, j, djx, *w)
    deposit9(jy, i, j, djy, *w)
    deposit9(jz, i, j, djz, *w)
    ro += ro_initial
    return ro, jx, jy, jz


def initial_deposition(config, x_offt, y_offt, px, py, pz, m, q,
                       virt_params):
    """Background ions = minus the initial electron density
    (lcode.py:692-699)."""
    ro_e, _, _, _ = deposit(config, 0, x_offt, y_offt, m, q, px, py, pz,
                            virt_params)
    return -ro_e


def move_smart(config, m, q, x_init, y_init, x_prev_offt, y_prev_offt,
               estimated_x_offt, estimated_y_offt, px_prev, py_prev, pz_prev,
               Ex_avg, Ey_avg, Ez_avg, Bx_avg, By_avg, Bz_avg):
    """Fused field interpolation + predictor-corrector pusher +
    reflection with momentum flip (lcode.py:704-823)."""
    dxi = config.xi_step_size
    r = config.reflect_boundary
    opx, opy, opz = px_prev, py_prev, pz_prev
    x_half = x_init + (x_prev_offt + estimated_x_offt) / 2
    y_half = y_init + (y_prev_offt + estimated_y_offt) / 2
    i, j, *w = weights(x_half, y_half, config.grid_steps,
                       config.grid_step_size)
    Ex = interp9(Ex_avg, i, j, *w)
    Ey = interp9(Ey_avg, i, j, *w)
    Ez = interp9(Ez_avg, i, j, *w)
    Bx = interp9(Bx_avg, i, j, *w)
    By = interp9(By_avg, i, j, *w)
    Bz = interp9(Bz_avg, i, j, *w)

    px, py, pz = opx, opy, opz
    for _ in range(2):  # exactly two half-step iterations (lcode.py:745-760)
        gamma_m = np.sqrt(m**2 + pz**2 + px**2 + py**2)
        vx, vy, vz = px / gamma_m, py / gamma_m, pz / gamma_m
        factor_1 = q * dxi / (1 - pz / gamma_m)
        dpx = factor_1 * (Ex + vy * Bz - vz * By)
        dpy = factor_1 * (Ey - vx * Bz + vz * Bx)
        dpz = factor_1 * (Ez + vx * By - vy * Bx)
        px, py, pz = opx + dpx / 2, opy + dpy / 2, opz + dpz / 2

    gamma_m = np.sqrt(m**2 + pz**2 + px**2 + py**2)
    x_offt = x_prev_offt + px / (gamma_m - pz) * dxi
    y_offt = y_prev_offt + py / (gamma_m - pz) * dxi
    px, py, pz = opx + dpx, opy + dpy, opz + dpz

    # Reflection, in the reference's order of the four tests (:774-789).
    x = x_init + x_offt
    y = y_init + y_offt
    for sign in (+1, -1):
        hit = (x > r) if sign > 0 else (x < -r)
        x = np.where(hit, sign * 2 * r - x, x)
        x_offt = np.where(hit, x - x_init, x_offt)
        px = np.where(hit, -px, px)
    for sign in (+1, -1):
        hit = (y > r) if sign > 0 else (y < -r)
        y = np.where(hit, sign * 2 * r - y, y)
        y_offt = np.where(hit, y - y_init, y_offt)
        py = np.where(hit, -py, py)
    return x_offt, y_offt, px, py, pz


# --------------------------------------------------------------------------
# The xi step and init
# --------------------------------------------------------------------------

def _solve_fields(config, prev, F_in, beam_ro, ro, jx, jy, jz):
    """Field half of a predictor or corrector pass (lcode.py:862-881 and
    :893-911).  F_in is the field set used in the Helmholtz trick term."""
    if config.field_solver_variant_A:
        ro_in, jz_in = (ro + prev.ro) / 2, (jz + prev.jz) / 2
    else:
        ro_in, jz_in = ro, jz
    Ex, Ey, Bx, By = calculate_Ex_Ey_Bx_By(
        config, F_in[0], F_in[1], F_in[3], F_in[4],
        beam_ro, ro_in, jx, jy, jz_in, prev.jx, prev.jy)
    if config.field_solver_variant_A:
        Ex, Ey = 2 * Ex - prev.Ex, 2 * Ey - prev.Ey
        Bx, By = 2 * Bx - prev.Bx, 2 * By - prev.By
    Ez = calculate_Ez(config, jx, jy)
    Bz = calculate_Bz(config, jx, jy)
    new = (Ex, Ey, Ez, Bx, By, Bz)
    olds = (prev.Ex, prev.Ey, prev.Ez, prev.Bx, prev.By, prev.Bz)
    avg = tuple((f + o) / 2 for f, o in zip(new, olds))
    return new, avg


def step(config, const, virt_params, prev, beam_ro):
    """One xi step: estimate, then predictor, corrector and final
    push/deposit passes (lcode.py:828-931).

    Pass 1 pushes with prev fields from the field-free estimate; its fields
    use prev fields in the trick term.  Pass 2 pushes with the pass-1
    averages from the pass-1 positions; its fields use the pass-1 averages in
    the trick term.  Pass 3 pushes with the pass-2 averages from the pass-2
    positions.  Returned: pass-2 fields, pass-3 particles and densities."""
    beam_ro = np.asarray(beam_ro, dtype=np.float64)
    prev_fields = (prev.Ex, prev.Ey, prev.Ez, prev.Bx, prev.By, prev.Bz)

    x_offt, y_offt = move_estimate_wo_fields(
        config, const.m, const.x_init, const.y_init,
        prev.x_offt, prev.y_offt, prev.px, prev.py, prev.pz)

    fields_for_push = prev_fields
    fields_for_trick = prev_fields
    for _ in range(2):
        x_offt, y_offt, px, py, pz = move_smart(
            config, const.m, const.q, const.x_init, const.y_init,
            prev.x_offt, prev.y_offt, x_offt, y_offt,
            prev.px, prev.py, prev.pz, *fields_for_push)
        ro, jx, jy, jz = deposit(config, const.ro_initial, x_offt, y_offt,
                                 const.m, const.q, px, py, pz, virt_params)
        fields, avg = _solve_fields(config, prev, fields_for_trick,
                                    beam_ro, ro, jx, jy, jz)
        fields_for_push = avg
        fields_for_trick = avg

    x_offt, y_offt, px, py, pz = move_smart(
        config, const.m, const.q, const.x_init, const.y_init,
        prev.x_offt, prev.y_offt, x_offt, y_offt,
        prev.px, prev.py, prev.pz, *fields_for_push)
    ro, jx, jy, jz = deposit(config, const.ro_initial, x_offt, y_offt,
                             const.m, const.q, px, py, pz, virt_params)

    Ex, Ey, Ez, Bx, By, Bz = (f.copy() for f in fields)
    return Arrays(x_offt=x_offt, y_offt=y_offt, px=px, py=py, pz=pz,
                  Ex=Ex, Ey=Ey, Ez=Ez, Bx=Bx, By=By, Bz=Bz,
                  ro=ro, jx=jx, jy=jy, jz=jz)


FIELD_NAMES = ("Ex", "Ey", "Ez", "Bx", "By", "Bz")


def step_guided(config, const, virt_params, prev, beam_ro, guess, threshold):
    """The xi step with its predictor pass (lcode.py:841-881) replaced by a
    field guess - the hook of the learned and extrapolated predictors.  Not
    in the reference; stated here from the reference's own passes:

    1. the predictor push with prev fields from the field-free estimate
       (lcode.py:843-854), as in step();
    2. the guess G (six fields) stands for the predictor's fields: the
       corrector pushes with (G + prev)/2 from the predictor positions,
       deposits and solves with (G + prev)/2 in the trick term
       (lcode.py:884-911);
    3. residual r = ||F_corr - G|| / ||F_corr||, the six fields summed;
    4. r > threshold (or not a number): the plain predictor and corrector
       passes run after all (the fields of step());
    5. the final push with the averages of the accepted pass and the final
       deposition (lcode.py:914-921).

    Returns (state, r, accepted).  With G equal to the plain predictor's
    fields the result is step()'s, bit for bit."""
    beam_ro = np.asarray(beam_ro, dtype=np.float64)
    prev_fields = tuple(getattr(prev, k) for k in FIELD_NAMES)
    guess = tuple(np.asarray(g, dtype=np.float64) for g in guess)
    x_e, y_e = move_estimate_wo_fields(
        config, const.m, const.x_init, const.y_init,
        prev.x_offt, prev.y_offt, prev.px, prev.py, prev.pz)
    p1 = move_smart(config, const.m, const.q, const.x_init, const.y_init,
                    prev.x_offt, prev.y_offt, x_e, y_e,
                    prev.px, prev.py, prev.pz, *prev_fields)

    def corrector(first_avg):
        p2 = move_smart(config, const.m, const.q, const.x_init,
                        const.y_init, prev.x_offt, prev.y_offt, p1[0], p1[1],
                        prev.px, prev.py, prev.pz, *first_avg)
        dens = deposit(config, const.ro_initial, p2[0], p2[1], const.m,
                       const.q, p2[2], p2[3], p2[4], virt_params)
        fields, avg = _solve_fields(config, prev, first_avg, beam_ro, *dens)
        return p2, fields, avg

    g_avg = tuple((g + o) / 2 for g, o in zip(guess, prev_fields))
    p2, fields, avg = corrector(g_avg)
    num = sum(float(np.sum((f - g) ** 2)) for f, g in zip(fields, guess))
    den = sum(float(np.sum(f ** 2)) for f in fields)
    r = math.sqrt(num / den) if den > 0 else (math.inf if num > 0 else 0.)
    accepted = r <= threshold
    if not accepted:
        dens = deposit(config, const.ro_initial, p1[0], p1[1], const.m,
                       const.q, p1[2], p1[3], p1[4], virt_params)
        _, avg1 = _solve_fields(config, prev, prev_fields, beam_ro, *dens)
        p2, fields, avg = corrector(avg1)
    x_offt, y_offt, px, py, pz = move_smart(
        config, const.m, const.q, const.x_init, const.y_init,
        prev.x_offt, prev.y_offt, p2[0], p2[1],
        prev.px, prev.py, prev.pz, *avg)
    ro, jx, jy, jz = deposit(config, const.ro_initial, x_offt, y_offt,
                             const.m, const.q, px, py, pz, virt_params)
    Ex, Ey, Ez, Bx, By, Bz = (f.copy() for f in fields)
    st = Arrays(x_offt=x_offt, y_offt=y_offt, px=px, py=py, pz=pz,
                Ex=Ex, Ey=Ey, Ez=Ez, Bx=Bx, By=By, Bz=Bz,
                ro=ro, jx=jx, jy=jy, jz=jz)
    return st, r, accepted


def predictor_fields(config, const, virt_params, prev, beam_ro):
    """The predictor pass's six fields (lcode.py:841-874): what a perfect
    guess of step_guided reproduces."""
    beam_ro = np.asarray(beam_ro, dtype=np.float64)
    prev_fields = tuple(getattr(prev, k) for k in FIELD_NAMES)
    x_e, y_e = move_estimate_wo_fields(
        config, const.m, const.x_init, const.y_init,
        prev.x_offt, prev.y_offt, prev.px, prev.py, prev.pz)
    p1 = move_smart(config, const.m, const.q, const.x_init, const.y_init,
                    prev.x_offt, prev.y_offt, x_e, y_e,
                    prev.px, prev.py, prev.pz, *prev_fields)
    dens = deposit(config, const.ro_initial, p1[0], p1[1], const.m,
                   const.q, p1[2], p1[3], p1[4], virt_params)
    fields, _ = _solve_fields(config, prev, prev_fields, beam_ro, *dens)
    return fields


def init(config):
    """Validate the config, write reflect_boundary into it, build const,
    virt_params and the zero state (lcode.py:936-975)."""
    assert config.grid_steps % 2 == 1
    assert config.reflect_padding_steps > config.plasma_coarseness + 1
    config.reflect_boundary = config.grid_step_size * (
        config.grid_steps / 2 - config.reflect_padding_steps)
    n = config.grid_steps
    grid = (np.arange(n) - n // 2) * config.grid_step_size
    xs, ys = grid[:, None], grid[None, :]
    (x_init, y_init, x_offt, y_offt, px, py, pz, m, q,
     virt_params) = make_plasma(n - config.plasma_padding_steps * 2,
                                config.grid_step_size,
                                coarseness=config.plasma_coarseness,
                                fineness=config.plasma_fineness)
    ro_initial = initial_deposition(config, x_offt, y_offt, px, py, pz,
                                    m, q, virt_params)
    const = Arrays(m=m, q=q, x_init=x_init, y_init=y_init,
                   ro_initial=ro_initial)
    z = lambda: np.zeros((n, n))  # noqa: E731
    state = Arrays(x_offt=x_offt, y_offt=y_offt, px=px, py=py, pz=pz,
                   Ex=z(), Ey=z(), Ez=z(), Bx=z(), By=z(), Bz=z(),
                   ro=z(), jx=z(), jy=z(), jz=z())
    return xs, ys, const, virt_params, state


# --------------------------------------------------------------------------
# Workloads (configs and beams)
# --------------------------------------------------------------------------

def make_config(grid_steps=641, grid_step_size=.025, xi_step_size=.005,
                xi_steps=1500, coarseness=3, fineness=2, trick=1,
                variant_A=True, reflect_padding_steps=5,
                plasma_padding_steps=10):
    """A config object with the attribute surface of config_example.py:1-17."""
    return types.SimpleNamespace(
        grid_steps=grid_steps, grid_step_size=grid_step_size,
        xi_step_size=xi_step_size, xi_steps=xi_steps,
        field_solver_subtraction_trick=trick,
        field_solver_variant_A=variant_A,
        reflect_padding_steps=reflect_padding_steps,
        plasma_padding_steps=plasma_padding_steps,
        plasma_coarseness=coarseness, plasma_fineness=fineness)


def beam(config, xi_i, xs, ys, amplitude=.05, sigma=1.0, shift_x=0.0,
         shift_y=0.0):
    """Cosine-Gaussian beam of config_example.py:21-28 (amplitude, sigma and
    shifts as parameters; the example shifts in y)."""
    xi = -xi_i * config.xi_step_size
    if xi < -2 * math.sqrt(2 * math.pi):
        return 0
    r = np.sqrt((xs - shift_x)**2 + (ys - shift_y)**2)
    return (amplitude * np.exp(-.5 * (r / sigma)**2) *
            (1 - math.cos(xi * math.sqrt(math.pi / 2))))
