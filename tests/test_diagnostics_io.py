"""Host-side parts of the run diagnostics and the state file format (no
GPU): the peak-drift report of lcode.py:992-1003 against hand-checked
histories and scipy.signal.argrelmax, and a save/load round trip."""

import json
import types

import numpy as np
import pytest
import scipy.signal
import torch

from lcode3d_amd import diagnostics, state_io


def _reference_peak_msg(hist):
    # lcode.py:992-1003 with the library call it makes
    a = np.array(hist)
    idx = scipy.signal.argrelmax(a)[0]
    if idx.size:
        v = a[idx]
        d = 100 * (v / v[0] - 1)
        return f'{v[-1]:0.4e} {d[-1]:+0.2f}%'
    return '...'


@pytest.mark.parametrize("hist", [
    [], [1.0], [1.0, 2.0], [0.0, 1.0, 0.0],
    [0.0, 1.0, 0.0, 1.1, 0.0],            # second peak +10 %
    [3.0, 2.0, 1.0],                      # maximum at the start: no peak
    [0.0, 1.0, 1.0, 0.0],                 # flat top: not strict
    list(np.sin(np.linspace(0, 40, 777)) * np.exp(-np.linspace(0, 1, 777))),
])
def test_peak_msg_matches_argrelmax(hist):
    assert diagnostics.diags_peak_msg(hist) == _reference_peak_msg(hist)


def test_peak_msg_worked_example():
    assert diagnostics.diags_peak_msg([0., 1., 0., 1.1, 0.]) == '1.1000e+00 +10.00%'


def test_peak_msg_accepts_device_style_history():
    h = torch.tensor([0., 2., 0., 1., 0.], dtype=torch.float64)
    assert diagnostics.diags_peak_msg(diagnostics._host(h)) == '1.0000e+00 -50.00%'


def test_state_round_trip(tmp_path):
    rng = np.random.default_rng(0)
    st = types.SimpleNamespace(**{k: torch.as_tensor(rng.standard_normal((7, 7)))
                                  for k in state_io.STATE_NAMES})
    const = types.SimpleNamespace(**{k: torch.as_tensor(rng.standard_normal((3, 3)))
                                     for k in state_io.CONST_NAMES})
    vp = types.SimpleNamespace(**{k: torch.as_tensor(rng.standard_normal(5))
                                  for k in state_io.VIRT_NAMES})
    vp.indices_prev = torch.arange(5, dtype=torch.int64)
    cfg = types.SimpleNamespace(grid_steps=7, grid_step_size=.5, xi_step_size=.1)
    p = tmp_path / "s.npz"
    state_io.save_state(p, st, cfg, xi_i=41, const=const, virt_params=vp)
    header, st2, const2, vp2 = state_io.load_state(p, device="cpu")
    assert header["xi_i"] == 41 and header["config"]["grid_steps"] == 7
    for k in state_io.STATE_NAMES:
        assert torch.equal(getattr(st2, k), getattr(st, k))
    for k in state_io.CONST_NAMES:
        assert torch.equal(getattr(const2, k), getattr(const, k))
    assert vp2.indices_prev.dtype == torch.int64
    # the file is plain arrays: loads without pickle
    with np.load(p, allow_pickle=False) as z:
        assert json.loads(str(z["header"]))["format"] == state_io.FORMAT


def test_state_without_const(tmp_path):
    st = types.SimpleNamespace(**{k: torch.zeros(3, 3) for k in state_io.STATE_NAMES})
    p = tmp_path / "s.npz"
    state_io.save_state(p, st)
    header, st2, const2, vp2 = state_io.load_state(p, device="cpu")
    assert const2 is None and vp2 is None and header["xi_i"] is None


def test_rejects_foreign_file(tmp_path):
    p = tmp_path / "x.npz"
    np.savez(p, header=np.array(json.dumps({"format": "other"})))
    with pytest.raises(ValueError):
        state_io.load_state(p, device="cpu")
