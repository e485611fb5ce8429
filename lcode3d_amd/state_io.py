"""A state file format for the xi step: the 15 arrays of a state (the
reference's GPUArrays of lcode.py:926-929), optionally the constant inputs
of init (lcode.py:936-975), and the run position, in one NumPy .npz.  It
serves as checkpoint/resume (a run continues from slice xi_i + 1 with the
saved state and const) and as the exchange format for field/particle
datasets.  Files hold plain arrays and a JSON header: they load with
numpy's allow_pickle=False."""

import json

import numpy as np
import torch

STATE_NAMES = ("x_offt", "y_offt", "px", "py", "pz",
               "Ex", "Ey", "Ez", "Bx", "By", "Bz", "ro", "jx", "jy", "jz")
CONST_NAMES = ("m", "q", "x_init", "y_init", "ro_initial")
VIRT_NAMES = ("influence_prev", "influence_next", "indices_prev",
              "indices_next", "fine_grid")
FORMAT = "lcode3d-state-1"
_CONFIG_KEYS = ("grid_steps", "grid_step_size", "xi_step_size",
                "field_solver_subtraction_trick", "field_solver_variant_A",
                "reflect_padding_steps", "plasma_padding_steps",
                "plasma_coarseness", "plasma_fineness", "reflect_boundary")


def _np(a):
    return a.detach().cpu().numpy() if isinstance(a, torch.Tensor) else np.asarray(a)


def save_state(path, state, config=None, xi_i=None, const=None,
               virt_params=None):
    """Write `state` (and, if given, `const` and `virt_params`) to `path`."""
    arrays = {f"state.{k}": _np(getattr(state, k)) for k in STATE_NAMES}
    if const is not None:
        arrays.update({f"const.{k}": _np(getattr(const, k)) for k in CONST_NAMES})
    if virt_params is not None:
        arrays.update({f"virt.{k}": _np(getattr(virt_params, k)) for k in VIRT_NAMES})
    header = {"format": FORMAT, "xi_i": xi_i}
    if config is not None:
        header["config"] = {k: getattr(config, k) for k in _CONFIG_KEYS
                            if hasattr(config, k)}
    arrays["header"] = np.array(json.dumps(header))
    with open(path, "wb") as f:
        np.savez(f, **arrays)


class _Bag:
    pass


def load_state(path, device=None):
    """Read a file written by save_state.  Returns (header, state, const,
    virt_params); const and virt_params are None when the file has none.
    Arrays go to `device` (default: the current HIP device) as torch
    tensors; the state is a GPUArrays when the device is a HIP device."""
    with np.load(path, allow_pickle=False) as z:
        header = json.loads(str(z["header"]))
        if header.get("format") != FORMAT:
            raise ValueError(f"{path}: not an {FORMAT} file")
        data = {k: z[k] for k in z.files if k != "header"}
    if device is None:
        device = torch.device("cuda", torch.cuda.current_device())
    device = torch.device(device)

    def bag(prefix, names):
        if not all(f"{prefix}.{k}" in data for k in names):
            return None
        if device.type == "cuda":
            from .lcode import GPUArrays
            return GPUArrays(**{k: data[f"{prefix}.{k}"] for k in names})
        b = _Bag()
        for k in names:
            setattr(b, k, torch.as_tensor(data[f"{prefix}.{k}"], device=device))
        return b

    return header, bag("state", STATE_NAMES), bag("const", CONST_NAMES), \
        bag("virt", VIRT_NAMES)
