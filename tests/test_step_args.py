"""Argument checks of the Python step() that run before anything is
launched (CPU only: no device is touched)."""

import types

import numpy as np
import pytest
import torch

import lcode3d_amd as lc

STATE = ("x_offt", "y_offt", "px", "py", "pz", "Ex", "Ey", "Ez", "Bx", "By",
         "Bz", "ro", "jx", "jy", "jz")


def _arrays(nc=5, n=17, ro_dtype=torch.float64):
    const = types.SimpleNamespace(
        m=torch.ones(nc, nc, dtype=torch.float64),
        q=-torch.ones(nc, nc, dtype=torch.float64),
        x_init=torch.zeros(nc, nc, dtype=torch.float64),
        y_init=torch.zeros(nc, nc, dtype=torch.float64),
        ro_initial=torch.zeros(n, n, dtype=ro_dtype))
    prev = types.SimpleNamespace(**{
        k: torch.zeros((nc, nc) if k in STATE[:5] else (n, n),
                       dtype=torch.float64) for k in STATE})
    cfg = types.SimpleNamespace(grid_steps=n)
    return cfg, const, prev


def test_float32_ro_initial_is_rejected_before_launch():
    cfg, const, prev = _arrays(ro_dtype=torch.float32)
    with pytest.raises(ValueError, match="const.ro_initial"):
        lc.step(cfg, const, None, prev, 0)


def test_strided_ro_initial_is_rejected_before_launch():
    cfg, const, prev = _arrays()
    const.ro_initial = torch.zeros(17, 34, dtype=torch.float64)[:, ::2]
    with pytest.raises(ValueError, match="const.ro_initial"):
        lc.step(cfg, const, None, prev, 0)


def test_wrong_shape_prev_is_rejected_before_launch():
    cfg, const, prev = _arrays()
    prev.jx = torch.zeros(16, 16, dtype=torch.float64)
    with pytest.raises(ValueError, match="prev.jx"):
        lc.step(cfg, const, None, prev, 0)


def test_host_tensors_are_rejected_before_launch():
    cfg, const, prev = _arrays()
    with pytest.raises(ValueError, match="device"):
        lc.step(cfg, const, None, prev, 0)
    assert np.isfinite(0.0)
