"""The reference's run diagnostics (lcode.py:978-1027) over device state:
the plasma-noise measure zn is computed on the GPU (lcode_diag_zn), the
on-axis Ez history stays on the device until a report needs it, and the
peak-drift report and the printed line keep the reference's formats."""

import ctypes
import os
import sys

import numpy as np
import torch

from . import _lib
from .lcode import _check_grid, _config_struct, _ptr, _stream, _workspace

#: running maximum of zn (the reference keeps it in a module global, lcode.py:980)
max_zn = 0.0


def zn(config, ro):
    """zn of one density grid as a 0-d device tensor, without synchronising:
    mean |ro - gaussian_filter(ro, 0.25 / grid_step_size)| / 4.23045376e-04
    (lcode.py:983-987)."""
    lib = _lib.load()
    (ro,) = _check_grid(config, ro)
    cs = _config_struct(config)
    out = torch.empty((), dtype=torch.float64, device=ro.device)
    _lib.check(lib.lcode_diag_zn(_workspace(cs), ctypes.byref(cs), _ptr(ro), _ptr(out),
                                 _stream()), "lcode_diag_zn")
    return out


def diags_ro_zn(config, ro):
    """lcode.py:981-989: update and return the running maximum of zn."""
    global max_zn
    max_zn = max(max_zn, zn(config, ro).item())
    return max_zn


def diags_peak_msg(Ez_00_history):
    """lcode.py:992-1003: the last local maximum of Ez(xi, 0, 0) and its
    deviation from the first one, or '...' before the first peak."""
    a = np.asarray(_host(Ez_00_history), dtype=np.float64)
    # strict local maxima, never at either end (scipy.signal.argrelmax)
    peaks = np.nonzero((a[1:-1] > a[:-2]) & (a[1:-1] > a[2:]))[0] + 1 \
        if a.size >= 3 else np.array([], dtype=int)
    if peaks.size:
        values = a[peaks]
        rel_deviations_perc = 100 * (values / values[0] - 1)
        return f'{values[-1]:0.4e} {rel_deviations_perc[-1]:+0.2f}%'
    return '...'


def diags_ro_slice(config, xi_i, xi, ro):
    """lcode.py:1006-1014: a PNG of ro every 1 / xi_step_size slices, in
    ./transverse (needs matplotlib)."""
    if xi_i % int(1 / config.xi_step_size):
        return
    import matplotlib
    matplotlib.use("Agg")
    import matplotlib.pyplot as plt
    if not os.path.isdir('transverse'):
        os.mkdir('transverse')
    fname = f'ro_{xi:+09.2f}.png' if xi else 'ro_-00000.00.png'
    ro = ro.cpu().numpy() if isinstance(ro, torch.Tensor) else np.asarray(ro)
    plt.imsave(os.path.join('transverse', fname), ro.T,
               origin='lower', vmin=-0.1, vmax=0.1, cmap='bwr')


def diagnostics(state, config, xi_i, Ez_00_history, write_png=True):
    """lcode.py:1017-1027 with `state` a GPUArrays (or GPUArraysView) and
    the history a list of floats or 0-d device tensors."""
    xi = -xi_i * config.xi_step_size
    hist = _host(Ez_00_history)
    Ez_00 = hist[-1]
    peak_report = diags_peak_msg(hist)
    ro = getattr(state, "_arrs", state).ro
    mz = diags_ro_zn(config, ro)
    if write_png:
        diags_ro_slice(config, xi_i, xi, ro)
    line = f'xi={xi:+.4f} {Ez_00:+.4e}|{peak_report}|zn={mz:.3f}'
    print(line)
    sys.stdout.flush()
    return line


def _host(history):
    if isinstance(history, torch.Tensor):
        return history.detach().cpu().numpy().tolist()
    if len(history) and isinstance(history[0], torch.Tensor):
        return torch.stack(list(history)).cpu().numpy().tolist()
    return list(history)
