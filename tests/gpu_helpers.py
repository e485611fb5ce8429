"""Shared helpers of the GPU parity tests: seeded inputs, host<->device
conversion and error norms."""

import sys

import numpy as np
import torch


def say(*a):
    """Print and flush what names the case before each GPU step."""
    print(*a, flush=True)
    sys.stdout.flush()


def dev(a):
    return torch.as_tensor(np.ascontiguousarray(a), device="cuda")


def host(t):
    return t.detach().cpu().numpy()


def rel_l2(got, want, floor=0.0):
    got, want = np.asarray(got), np.asarray(want)
    den = max(np.linalg.norm(want), floor)
    if den == 0:
        return float(np.linalg.norm(got))
    return float(np.linalg.norm(got - want) / den)


def smooth_field(rng, n, amp=1.0, zero_rim=1):
    """A smooth random (n, n) field with a zero rim of zero_rim cells."""
    a = rng.standard_normal((n, n))
    k = np.array([.25, .5, .25])
    for _ in range(4):
        a = (np.roll(a, 1, 0) + 2 * a + np.roll(a, -1, 0)) / 4
        a = (np.roll(a, 1, 1) + 2 * a + np.roll(a, -1, 1)) / 4
    del k
    if zero_rim:
        a[:zero_rim] = a[-zero_rim:] = 0
        a[:, :zero_rim] = a[:, -zero_rim:] = 0
    return amp * a / np.abs(a).max()


def perturbed_state(oracle, cfg, seed=0, amp_x=None, amp_p=.05, amp_f=.05):
    """init() plus smooth offsets, momenta, fields and densities: a
    non-trivial `prev` for one-step comparisons."""
    xs, ys, const, vp, st = oracle.init(cfg)
    rng = np.random.default_rng(seed)
    nc = const.m.shape[0]
    n = cfg.grid_steps
    h = cfg.grid_step_size
    amp_x = 2 * h if amp_x is None else amp_x
    for name in ("x_offt", "y_offt"):
        setattr(st, name, smooth_field(rng, nc, amp_x, zero_rim=0))
    for name in ("px", "py", "pz"):
        setattr(st, name, smooth_field(rng, nc, amp_p, zero_rim=0))
    for name in ("Ex", "Ey", "Ez", "Bx", "By", "Bz", "jx", "jy", "jz"):
        setattr(st, name, smooth_field(rng, n, amp_f, zero_rim=2))
    st.ro = smooth_field(rng, n, amp_f, zero_rim=2)
    return xs, ys, const, vp, st


def to_device_state(lc, st):
    return lc.GPUArrays(**{k: dev(getattr(st, k)) for k in (
        "x_offt", "y_offt", "px", "py", "pz", "Ex", "Ey", "Ez", "Bx", "By",
        "Bz", "ro", "jx", "jy", "jz")})


STATE = ("x_offt", "y_offt", "px", "py", "pz", "Ex", "Ey", "Ez", "Bx", "By",
         "Bz", "ro", "jx", "jy", "jz")
