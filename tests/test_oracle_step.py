"""Pins the oracle's xi step through properties the physics fixes: quiet
start, mirror/transpose symmetry of an axisymmetric beam, the exact charge
identity, linear wake theory at small amplitude, and the on-axis wake-peak
drift of the reference's run diagnostic (lcode.py:992-1002: behind the bunch
a small-amplitude wake keeps a constant peak, and its peaks are one plasma
period 2 pi apart).

The composition of the three passes (lcode.py:828-931: which field set
feeds each push and each Helmholtz trick term, what the step returns) is
pinned by a mutation test: a copy of the step built from the oracle's own
operators is first shown to equal oracle.step bit for bit, then each
plausible mis-composition of it must fail the wake-drift pin (CPU only)."""

import math

import numpy as np
import pytest
import scipy.signal
import scipy.sparse as sp
import scipy.sparse.linalg as spl


def _run(oracle, cfg, nsteps, amp, sigma=1.0, shift_x=0.0, on_step=None):
    xs, ys, const, vp, st = oracle.init(cfg)
    for i in range(nsteps):
        b = oracle.beam(cfg, i, xs, ys, amplitude=amp, sigma=sigma,
                        shift_x=shift_x)
        st = oracle.step(cfg, const, vp, st, b)
        if on_step:
            on_step(i, st)
    return xs, ys, const, vp, st


def test_quiet_start_stays_exactly_at_rest(oracle):
    cfg = oracle.make_config(grid_steps=33, grid_step_size=.1,
                             xi_step_size=.1, coarseness=2, fineness=2,
                             plasma_padding_steps=6)
    xs, ys, const, vp, st = oracle.init(cfg)
    for _ in range(3):
        st = oracle.step(cfg, const, vp, st, 0)
    for name in ('Ex', 'Ey', 'Ez', 'Bx', 'By', 'Bz', 'ro', 'jx', 'jy', 'jz',
                 'x_offt', 'y_offt', 'px', 'py', 'pz'):
        assert not np.any(getattr(st, name)), name


def test_axisymmetric_beam_symmetries_and_charge_identity(oracle):
    cfg = oracle.make_config(grid_steps=65, grid_step_size=.1,
                             xi_step_size=.1, coarseness=2, fineness=2)
    sums = []
    _run(oracle, cfg, 25, .05,
         on_step=lambda i, s: sums.append((s.ro - s.jz).sum()))
    xs, ys, const, vp, st = _run(oracle, cfg, 25, .05)
    assert np.abs(sums).max() < 1e-11
    ez = st.Ez
    scale = np.abs(ez).max()
    assert scale > 1e-3
    tol = 1e-10 * scale
    assert np.abs(ez - ez[::-1, :]).max() < tol
    assert np.abs(ez - ez[:, ::-1]).max() < tol
    assert np.abs(ez - ez.T).max() < tol
    # x <-> y exchange maps Ex to Ey and Bx to -By
    assert np.abs(st.Ex - st.Ey.T).max() < 1e-10 * np.abs(st.Ex).max()
    assert np.abs(st.Bx + st.By.T).max() < 1e-10 * np.abs(st.Bx).max()
    # mirror x -> -x flips Ex, keeps Ey
    assert np.abs(st.Ex + st.Ex[::-1, :]).max() < 1e-10 * np.abs(st.Ex).max()
    assert np.abs(st.Ey - st.Ey[::-1, :]).max() < 1e-10 * np.abs(st.Ey).max()
    # Bz is a pseudo-scalar: odd under each mirror and under x <-> y (the
    # square particle lattice allows a sin(4 theta)-like Bz)
    bz, tol_bz = st.Bz, 1e-10 * np.abs(st.Bx).max()
    assert np.abs(bz + bz[::-1, :]).max() < tol_bz
    assert np.abs(bz + bz.T).max() < tol_bz
    # the particles are mirror-symmetric too
    assert np.abs(st.px + st.px[::-1, :]).max() < 1e-12
    assert np.abs(st.pz - st.pz.T).max() < 1e-12


def _linear_wake_on_axis(xs, ys, h, s, amp, sigma=1.0):
    """Linear quasi-static theory: Ez(s) = -int_0^s cos(s - s') g0 f(s') ds',
    g = (1 - Lap_h)^-1 of the beam's transverse profile (Dirichlet box),
    f(s) = beam's longitudinal profile; solved with a sparse direct solve."""
    N = xs.shape[0]
    M = N - 2
    eye = sp.identity(M)
    d2 = sp.diags([np.full(M - 1, -1.), np.full(M, 2.), np.full(M - 1, -1.)],
                  [-1, 0, 1]) / h**2
    a = sp.identity(M * M) + sp.kron(d2, eye) + sp.kron(eye, d2)
    prof = np.exp(-.5 * (xs**2 + ys**2) / sigma**2)[1:-1, 1:-1]
    g0 = spl.spsolve(a.tocsc(), prof.ravel()).reshape(M, M)[M // 2, M // 2]
    fine = np.linspace(0, s[-1], 20001)
    f = np.where(fine <= 2 * math.sqrt(2 * math.pi),
                 amp * (1 - np.cos(fine * math.sqrt(math.pi / 2))), 0)
    out = []
    for si in s:
        sel = fine <= si
        out.append(-np.trapezoid(np.cos(si - fine[sel]) * g0 * f[sel],
                                 fine[sel]) if si > 0 else 0.)
    return np.array(out)


def test_linear_wake_theory_small_amplitude(oracle):
    """At amplitude 1e-3 the on-axis Ez follows linear theory to a few
    percent of its peak over more than one plasma period (the residue is
    the scheme's first-order xi error and the particle shape)."""
    cfg = oracle.make_config(grid_steps=129, grid_step_size=.05,
                             xi_step_size=.05, coarseness=2, fineness=2)
    n = 140
    ez = []
    xs, ys, *_ = _run(oracle, cfg, n, 1e-3,
                      on_step=lambda i, s: ez.append(s.Ez[64, 64]))
    lin = _linear_wake_on_axis(xs, ys, cfg.grid_step_size,
                               np.arange(n) * cfg.xi_step_size, 1e-3)
    assert np.abs(np.array(ez) - lin).max() < .05 * np.abs(lin).max()


# ---------------------------------------------------------------------------
# Wake-peak drift and the mutation test of the step composition
# ---------------------------------------------------------------------------

def _wake_peaks(oracle, step_fn, dxi, periods=3, amp=1e-3):
    """On-axis Ez(xi) over the bunch plus `periods` plasma periods at 65^2,
    h = 0.1; returns the xi of its local maxima and the drift of the last
    maximum from the first in percent (diags_peak_msg, lcode.py:992-1002)."""
    cfg = oracle.make_config(grid_steps=65, grid_step_size=.1,
                             xi_step_size=dxi, coarseness=2, fineness=2)
    xs, ys, const, vp, st = oracle.init(cfg)
    n = int((2 * math.sqrt(2 * math.pi) + periods * 2 * math.pi) / dxi)
    ez = np.zeros(n)
    for i in range(n):
        b = oracle.beam(cfg, i, xs, ys, amplitude=amp)
        st = step_fn(cfg, const, vp, st, b)
        ez[i] = st.Ez[32, 32]
    pk = scipy.signal.argrelmax(ez)[0]
    vals = ez[pk]
    return pk * dxi, 100 * (vals[-1] / vals[0] - 1)


MUTANTS = (
    "trick2_prev",         # pass 2's trick term fed prev fields (:895-898)
    "final_push_prev",     # final push with prev fields (:914-919)
    "return_avg",          # the step returns the averages (:926-929)
    "push2_unaveraged",    # pass 2 pushes with unaveraged fields (:884-889)
    "variantA_on_Ez_Bz",   # 2 F - F_prev applied to Ez and Bz (:869-874)
    "dj_this_pass",        # dj/dxi from this pass's currents (:260-261)
)


def _composed_step(oracle, mutation=None):
    """oracle.step written out from the oracle's operators, with one
    composition choice changed when `mutation` names it."""
    def solve(config, prev, F_in, beam_ro, ro, jx, jy, jz):
        va = config.field_solver_variant_A
        ro_in = (ro + prev.ro) / 2 if va else ro
        jz_in = (jz + prev.jz) / 2 if va else jz
        jxp, jyp = (jx, jy) if mutation == "dj_this_pass" else (prev.jx, prev.jy)
        Ex, Ey, Bx, By = oracle.calculate_Ex_Ey_Bx_By(
            config, F_in[0], F_in[1], F_in[3], F_in[4], beam_ro, ro_in, jx,
            jy, jz_in, jxp, jyp)
        if va:
            Ex, Ey = 2 * Ex - prev.Ex, 2 * Ey - prev.Ey
            Bx, By = 2 * Bx - prev.Bx, 2 * By - prev.By
        Ez = oracle.calculate_Ez(config, jx, jy)
        Bz = oracle.calculate_Bz(config, jx, jy)
        if va and mutation == "variantA_on_Ez_Bz":
            Ez, Bz = 2 * Ez - prev.Ez, 2 * Bz - prev.Bz
        new = (Ex, Ey, Ez, Bx, By, Bz)
        olds = (prev.Ex, prev.Ey, prev.Ez, prev.Bx, prev.By, prev.Bz)
        return new, tuple((f + g) / 2 for f, g in zip(new, olds))

    def step(config, const, vp, prev, beam_ro):
        beam_ro = np.asarray(beam_ro, dtype=np.float64)
        pf = (prev.Ex, prev.Ey, prev.Ez, prev.Bx, prev.By, prev.Bz)
        xo, yo = oracle.move_estimate_wo_fields(
            config, const.m, const.x_init, const.y_init, prev.x_offt,
            prev.y_offt, prev.px, prev.py, prev.pz)
        push_f = trick_f = pf

        def push(xo, yo, fields):
            return oracle.move_smart(
                config, const.m, const.q, const.x_init, const.y_init,
                prev.x_offt, prev.y_offt, xo, yo, prev.px, prev.py, prev.pz,
                *fields)

        for p in range(2):
            xo, yo, px, py, pz = push(xo, yo, push_f)
            ro, jx, jy, jz = oracle.deposit(config, const.ro_initial, xo, yo,
                                            const.m, const.q, px, py, pz, vp)
            if p == 1 and mutation == "trick2_prev":
                trick_f = pf
            fields, avg = solve(config, prev, trick_f, beam_ro, ro, jx, jy, jz)
            push_f = fields if (p == 0 and mutation == "push2_unaveraged") else avg
            trick_f = avg
        if mutation == "final_push_prev":
            push_f = pf
        xo, yo, px, py, pz = push(xo, yo, push_f)
        ro, jx, jy, jz = oracle.deposit(config, const.ro_initial, xo, yo,
                                        const.m, const.q, px, py, pz, vp)
        out = avg if mutation == "return_avg" else fields
        Ex, Ey, Ez, Bx, By, Bz = (f.copy() for f in out)
        return oracle.Arrays(x_offt=xo, y_offt=yo, px=px, py=py, pz=pz,
                             Ex=Ex, Ey=Ey, Ez=Ez, Bx=Bx, By=By, Bz=Bz,
                             ro=ro, jx=jx, jy=jy, jz=jz)
    return step


def test_composed_step_is_the_oracle_step(oracle):
    """The mutation harness starts from the oracle's own composition: with
    no mutation it reproduces oracle.step bit for bit (Variant A and B)."""
    for variant_A in (True, False):
        cfg = oracle.make_config(grid_steps=33, grid_step_size=.1,
                                 xi_step_size=.1, coarseness=2, fineness=2,
                                 plasma_padding_steps=6, variant_A=variant_A)
        xs, ys, const, vp, a = oracle.init(cfg)
        b = a
        mine = _composed_step(oracle)
        for i in range(12):
            beam = oracle.beam(cfg, i, xs, ys, amplitude=.05)
            a = oracle.step(cfg, const, vp, a, beam)
            b = mine(cfg, const, vp, b, beam)
        for name in ("Ex", "Ey", "Ez", "Bx", "By", "Bz", "ro", "jx", "jy",
                     "jz", "x_offt", "y_offt", "px", "py", "pz"):
            assert np.array_equal(getattr(a, name), getattr(b, name)), name
        assert np.abs(a.Ez).max() > 1e-4


@pytest.fixture(scope="module")
def wake_peaks_dxi_01(oracle):
    return _wake_peaks(oracle, oracle.step, .1)


def test_wake_peak_drift_small_and_converged(oracle, wake_peaks_dxi_01):
    """Behind the bunch the small-amplitude wake keeps its peak to within
    3 % over three plasma periods at dxi = 0.1 and 0.05 (measured: +0.05 %
    and +0.13 %), and successive peaks are one plasma period (2 pi) apart
    to within two slices."""
    for dxi, (xi_pk, drift) in ((.1, wake_peaks_dxi_01),
                                (.05, _wake_peaks(oracle, oracle.step, .05))):
        assert len(xi_pk) == 3, (dxi, xi_pk)
        assert abs(drift) < 3, (dxi, drift)
        assert np.all(np.abs(np.diff(xi_pk) - 2 * math.pi) <= 2 * dxi), (dxi, xi_pk)


@pytest.mark.parametrize("mutation", MUTANTS)
def test_step_mutants_fail_the_wake_pin(oracle, mutation):
    """Each plausible mis-composition of the step drifts by more than 3 %
    (measured at dxi = 0.1: 13 % to 58 %) or moves the peaks off the plasma
    period, so the pin above separates them from the oracle's step."""
    xi_pk, drift = _wake_peaks(oracle, _composed_step(oracle, mutation), .1)
    period_ok = len(xi_pk) == 3 and np.all(
        np.abs(np.diff(xi_pk) - 2 * math.pi) <= .2)
    assert abs(drift) > 3 or not period_ok, (mutation, drift, xi_pk)
