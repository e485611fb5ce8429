"""The run driver of lcode.py:1032-1051 over the HIP step: `config` is the
reference's config module (config_example.py), imported as `config` when
not given.  Differences from the reference loop, none of which changes a
result: the on-axis Ez history is kept on the device and copied to the
host only when a report is printed (the reference copies all of Ez every
slice, lcode.py:1044-1045); zn runs on the GPU.

    python -m lcode3d_amd.main        # with config.py on sys.path
"""

import sys

import torch

from .diagnostics import diagnostics
from .lcode import init, step


def main(config=None, write_png=True):
    if config is None:
        import config   # noqa: the reference's convention (lcode.py:1033)
    torch.cuda.set_device(int(getattr(config, "gpu_index", 0)))
    xs, ys, const, virt_params, state = init(config)
    c = config.grid_steps // 2
    history = torch.empty(config.xi_steps, dtype=torch.float64,
                          device=state.Ez.device)
    for xi_i in range(config.xi_steps):
        beam_ro = config.beam(xi_i, xs, ys)
        state = step(config, const, virt_params, state, beam_ro)
        history[xi_i] = state.Ez[c, c]
        time_for_diags = xi_i % config.diagnostics_each_N_steps == 0
        last_step = xi_i == config.xi_steps - 1
        if time_for_diags or last_step:
            diagnostics(state, config, xi_i, history[:xi_i + 1], write_png)
    return state


if __name__ == "__main__":
    sys.path.insert(0, ".")
    main()
