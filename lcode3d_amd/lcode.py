"""The reference's Python surface (lotov/lcode3d, lcode.py) over torch device
tensors, with every computation in the gfx950 kernels of
liblcode3d_hip.so.

Same names, argument meaning and ownership as the reference: every operator
allocates fresh outputs and never mutates its inputs, ``init`` writes
``config.reflect_boundary`` (lcode.py:947-949), ``step`` returns a new
``GPUArrays`` and leaves ``prev`` valid (lcode.py:926-931).  Arrays are
float64 torch tensors on the current CUDA (HIP) device; host NumPy arrays
are accepted wherever the reference accepted them and are copied over,
like ``cp.asarray``.

Differences from the reference, all on the error path: a particle whose
3x3 stencil would leave the grid is clamped into it and counted (see
``fault_count``) instead of reading or writing out of bounds.
"""

import ctypes
import math

import numpy as np
import torch

from . import _lib

ELECTRON_CHARGE = -1   # lcode.py:46
ELECTRON_MASS = 1      # lcode.py:47

_F64 = torch.float64


# ---------------------------------------------------------------------------
# Device arrays (lcode.py:52-119)
# ---------------------------------------------------------------------------

def _device():
    if not torch.cuda.is_available():
        raise RuntimeError("lcode3d_amd needs a HIP device (none visible)")
    return torch.device("cuda", torch.cuda.current_device())


def _as_device(array, dtype=None):
    """cp.asarray: host arrays are copied to the current device."""
    if isinstance(array, torch.Tensor):
        t = array
        if t.device.type != "cuda":
            t = t.to(_device())
    else:
        t = torch.as_tensor(np.asarray(array), device=_device())
    if dtype is not None:
        t = t.to(dtype)
    return t.contiguous()


class GPUArrays:
    """Attribute bag of device arrays (lcode.py:52-69)."""

    def __init__(self, **kwargs):
        for name, array in kwargs.items():
            setattr(self, name, _as_device(array))


class GPUArraysView:
    """Copying view of a GPUArrays (lcode.py:73-119): reading an attribute
    copies it to a host NumPy array, assigning copies the value back."""

    def __init__(self, gpu_arrays):
        super().__setattr__("_arrs", gpu_arrays)

    def __dir__(self):
        return list(set(super().__dir__() + dir(self._arrs)))

    def __getattr__(self, attrname):
        return getattr(self._arrs, attrname).cpu().numpy()

    def __setattr__(self, attrname, value):
        dst = getattr(self._arrs, attrname)
        dst.copy_(torch.as_tensor(np.asarray(value), dtype=dst.dtype))


# ---------------------------------------------------------------------------
# Workspaces (the reference memoizes its tables with cp.memoize)
# ---------------------------------------------------------------------------

_workspaces = {}


def _workspace(cfg_struct):
    key = (torch.cuda.current_device(), cfg_struct.grid_steps,
           cfg_struct.grid_step_size, cfg_struct.subtraction_trick != 0,
           cfg_struct.nc, cfg_struct.nf, cfg_struct.solver)
    ws = _workspaces.get(key)
    if ws is None:
        lib = _lib.load()
        handle = ctypes.c_void_p()
        _lib.check(lib.lcode_workspace_create(ctypes.byref(cfg_struct),
                                              ctypes.byref(handle)),
                   "lcode_workspace_create")
        ws = _workspaces[key] = handle
    return ws


def _config_struct(config, nc=0, nf=0):
    c = _lib.LcodeConfig()
    c.grid_steps = int(config.grid_steps)
    c.grid_step_size = float(config.grid_step_size)
    c.xi_step_size = float(config.xi_step_size)
    c.subtraction_trick = float(config.field_solver_subtraction_trick)
    c.variant_A = 1 if config.field_solver_variant_A else 0
    c.reflect_boundary = float(getattr(config, "reflect_boundary", math.nan))
    c.plasma_coarseness = int(config.plasma_coarseness)
    c.plasma_fineness = int(config.plasma_fineness)
    c.nc = int(nc)
    c.nf = int(nf)
    c.solver = _solver_of(config)
    return c


_SOLVERS = {
    ("fft", "fp64"): _lib.LCODE_SOLVER_FFT64,
    ("fft", "fp32"): _lib.LCODE_SOLVER_FFT32,
    ("fft", "mixed"): _lib.LCODE_SOLVER_MIXED,
    ("dense", "fp64"): _lib.LCODE_SOLVER_DENSE64,
    ("dense", "bf16x3"): _lib.LCODE_SOLVER_DENSE_BF16X3,
}


def _solver_of(config):
    """Field-solver back-end, from two attributes that extend the
    reference's config surface:
    config.field_solver_transform: 'fft' (default; mixed-radix FFT in LDS)
    or 'dense' (N x N sine/cosine matrices on MFMA);
    config.field_solver_precision: 'fp64' (default, the plain path), 'fp32'
    (FFT only: transforms, spectra and eigenvalue tables in FP32) or
    'bf16x3' (dense only: hi/lo BF16 splits, three MFMA products, FP32
    accumulation); in these three the RHS, Bz mean, Variant A and averages
    are FP64.  'mixed' (FFT only): the mixed-precision step - the ten grids
    of the state (float32 tensors from init), the deposition and the
    transforms in FP32; particles, the RHS arithmetic with its trick term,
    the Bz zero mode and mean and ro_initial stay FP64 (the standalone
    operators run as 'fp32' on float64 grids)."""
    p = getattr(config, "field_solver_precision", "fp64")
    t = getattr(config, "field_solver_transform", "dense" if p == "bf16x3" else "fft")
    try:
        return _SOLVERS[(t, p)]
    except KeyError:
        raise ValueError(
            f"unsupported field solver: transform {t!r} with precision {p!r} "
            f"(supported: {sorted(_SOLVERS)})") from None


def _stream():
    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)


def _ptr(t):
    return ctypes.c_void_p(t.data_ptr())


def _check_grid(config, *arrays, dtype=_F64):
    n = int(config.grid_steps)
    out = []
    for a in arrays:
        t = _as_device(a, dtype)
        if tuple(t.shape) != (n, n):
            raise ValueError(f"grid array of shape {tuple(t.shape)}, "
                             f"expected ({n}, {n})")
        out.append(t)
    return out


def _check_particles(*arrays):
    out = [_as_device(a, _F64) for a in arrays]
    shape = out[0].shape
    if len(shape) != 2 or shape[0] != shape[1]:
        raise ValueError(f"coarse particle arrays must be (nc, nc), got {tuple(shape)}")
    for t in out:
        if t.shape != shape:
            raise ValueError("coarse particle arrays differ in shape")
    return out


def _beam_arg(config, beam_ro):
    """Array beam -> (tensor, 0.0); scalar beam -> (None, value)."""
    if isinstance(beam_ro, (int, float, np.floating, np.integer)) or (
            np.ndim(beam_ro) == 0 and not isinstance(beam_ro, torch.Tensor)):
        return None, float(beam_ro)
    if isinstance(beam_ro, torch.Tensor) and beam_ro.dim() == 0:
        return None, float(beam_ro.item())
    (t,) = _check_grid(config, beam_ro)
    return t, 0.0


def fault_count(config=None, nc=None):
    """Stencil lookups clamped into the grid since the last call, summed over
    the workspaces of the current device (0 means no particle came within
    one cell of the grid edge).  Synchronises.  The arguments are accepted
    for call-site readability and ignored."""
    lib = _lib.load()
    dev = torch.cuda.current_device()
    total = 0
    for key, ws in _workspaces.items():
        if key[0] != dev:
            continue
        n = ctypes.c_int64()
        _lib.check(lib.lcode_workspace_fault_count(ws, _stream(),
                                                   ctypes.byref(n)),
                   "lcode_workspace_fault_count")
        total += n.value
    return total


def profile_begin(max_launches=8192, only=None):
    """Start bracketing every kernel launch of this device's workspaces
    with HIP events (instrumentation, not in the reference); with `only`
    (a kernel name such as "deposit_tile") just that kernel's launches."""
    lib = _lib.load()
    dev = torch.cuda.current_device()
    for key, ws in _workspaces.items():
        if key[0] == dev:
            _lib.check(lib.lcode_profile_begin(ws, int(max_launches)),
                       "lcode_profile_begin")
            _lib.check(lib.lcode_profile_only(ws, (only or "").encode()),
                       "lcode_profile_only")


def profile_pause(paused=True):
    """Stop (or resume) the timing begun by profile_begin; launches in
    between run untimed."""
    lib = _lib.load()
    dev = torch.cuda.current_device()
    for key, ws in _workspaces.items():
        if key[0] == dev:
            _lib.check(lib.lcode_profile_pause(ws, int(bool(paused))),
                       "lcode_profile_pause")


def profile_end():
    """Stop timing; returns {kernel name: (launches, total ms)}.
    Synchronises."""
    lib = _lib.load()
    dev = torch.cuda.current_device()
    out = {}
    buf = (_lib.LcodeKernelTime * 64)()
    n = ctypes.c_int32()
    for key, ws in _workspaces.items():
        if key[0] != dev:
            continue
        _lib.check(lib.lcode_profile_end(ws, _stream(), buf, 64,
                                         ctypes.byref(n)),
                   "lcode_profile_end")
        for k in range(n.value):
            name = buf[k].name.decode()
            calls, ms = out.get(name, (0, 0.0))
            out[name] = (calls + buf[k].calls, ms + buf[k].total_ms)
    return out


# ---------------------------------------------------------------------------
# Field solvers (lcode.py:160-364)
# ---------------------------------------------------------------------------

def calculate_Ez(config, jx, jy):
    """lcode.py:160-184."""
    lib = _lib.load()
    jx, jy = _check_grid(config, jx, jy)
    cs = _config_struct(config)
    Ez = torch.empty_like(jx)
    _lib.check(lib.lcode_calculate_Ez(_workspace(cs), ctypes.byref(cs),
                                      _ptr(jx), _ptr(jy), _ptr(Ez), _stream()),
               "lcode_calculate_Ez")
    return Ez


def calculate_Bz(config, jx, jy):
    """lcode.py:333-364."""
    lib = _lib.load()
    jx, jy = _check_grid(config, jx, jy)
    cs = _config_struct(config)
    Bz = torch.empty_like(jx)
    _lib.check(lib.lcode_calculate_Bz(_workspace(cs), ctypes.byref(cs),
                                      _ptr(jx), _ptr(jy), _ptr(Bz), _stream()),
               "lcode_calculate_Bz")
    return Bz


def calculate_Ex_Ey_Bx_By(config, Ex_avg, Ey_avg, Bx_avg, By_avg,
                          beam_ro, ro, jx, jy, jz, jx_prev, jy_prev):
    """lcode.py:246-290."""
    lib = _lib.load()
    grids = _check_grid(config, Ex_avg, Ey_avg, Bx_avg, By_avg, ro, jx, jy,
                        jz, jx_prev, jy_prev)
    beam, beam_scalar = _beam_arg(config, beam_ro)
    cs = _config_struct(config)
    outs = [torch.empty_like(grids[0]) for _ in range(4)]
    out_pp, keep = _lib.ptr_array([o.data_ptr() for o in outs])
    g = [_ptr(t) for t in grids]
    _lib.check(lib.lcode_calculate_Ex_Ey_Bx_By(
        _workspace(cs), ctypes.byref(cs), g[0], g[1], g[2], g[3],
        _ptr(beam) if beam is not None else None, beam_scalar,
        g[4], g[5], g[6], g[7], g[8], g[9], out_pp, _stream()),
        "lcode_calculate_Ex_Ey_Bx_By")
    return tuple(outs)


# ---------------------------------------------------------------------------
# Particles (lcode.py:369-391, :666-699, :796-823)
# ---------------------------------------------------------------------------

def move_estimate_wo_fields(config, m, x_init, y_init, prev_x_offt,
                            prev_y_offt, px, py, pz):
    """lcode.py:369-391."""
    lib = _lib.load()
    arrs = _check_particles(m, x_init, y_init, prev_x_offt, prev_y_offt,
                            px, py, pz)
    nc = arrs[0].shape[0]
    cs = _config_struct(config, nc)
    x_offt, y_offt = torch.empty_like(arrs[0]), torch.empty_like(arrs[0])
    _lib.check(lib.lcode_move_estimate_wo_fields(
        _workspace(cs), ctypes.byref(cs), *[_ptr(a) for a in arrs],
        _ptr(x_offt), _ptr(y_offt), _stream()),
        "lcode_move_estimate_wo_fields")
    return x_offt, y_offt


def move_smart(config, m, q, x_init, y_init, x_prev_offt, y_prev_offt,
               estimated_x_offt, estimated_y_offt, px_prev, py_prev, pz_prev,
               Ex_avg, Ey_avg, Ez_avg, Bx_avg, By_avg, Bz_avg):
    """lcode.py:796-823."""
    lib = _lib.load()
    parts = _check_particles(m, q, x_init, y_init, x_prev_offt, y_prev_offt,
                             estimated_x_offt, estimated_y_offt, px_prev,
                             py_prev, pz_prev)
    fields = _check_grid(config, Ex_avg, Ey_avg, Ez_avg, Bx_avg, By_avg,
                         Bz_avg)
    nc = parts[0].shape[0]
    cs = _config_struct(config, nc)
    outs = [torch.empty_like(parts[0]) for _ in range(5)]
    f_pp, kf = _lib.ptr_array([f.data_ptr() for f in fields])
    o_pp, ko = _lib.ptr_array([o.data_ptr() for o in outs])
    _lib.check(lib.lcode_move_smart(_workspace(cs), ctypes.byref(cs),
                                    *[_ptr(p) for p in parts], f_pp, o_pp,
                                    _stream()), "lcode_move_smart")
    return tuple(outs)


def _virt_struct(virt_params):
    vp = _lib.LcodeVirtParams()
    arrs = {}
    for name in ("influence_prev", "influence_next", "fine_grid"):
        arrs[name] = _as_device(getattr(virt_params, name), _F64)
    for name in ("indices_prev", "indices_next"):
        arrs[name] = _as_device(getattr(virt_params, name), torch.int64)
    nf = arrs["fine_grid"].numel()
    for name, t in arrs.items():
        if t.numel() != nf:
            raise ValueError(f"virt_params.{name} has {t.numel()} entries, "
                             f"expected {nf}")
        setattr(vp, name, t.data_ptr())
    return vp, nf, arrs


def deposit(config, ro_initial, x_offt, y_offt, m, q, px, py, pz,
            virt_params):
    """lcode.py:666-689.  ro_initial may be the scalar 0 (lcode.py:697)."""
    lib = _lib.load()
    parts = _check_particles(x_offt, y_offt, m, q, px, py, pz)
    vp, nf, keep = _virt_struct(virt_params)
    nc = parts[0].shape[0]
    if isinstance(ro_initial, torch.Tensor) or np.ndim(ro_initial) > 0:
        (ro0,) = _check_grid(config, ro_initial)
        ro0_ptr = _ptr(ro0)
    else:
        if float(ro_initial) != 0:
            raise ValueError("a scalar ro_initial must be 0")
        ro0_ptr = None
    cs = _config_struct(config, nc, nf)
    n = int(config.grid_steps)
    outs = [torch.empty((n, n), dtype=_F64, device=parts[0].device)
            for _ in range(4)]
    o_pp, ko = _lib.ptr_array([o.data_ptr() for o in outs])
    _lib.check(lib.lcode_deposit(_workspace(cs), ctypes.byref(cs), ro0_ptr,
                                 *[_ptr(p) for p in parts], ctypes.byref(vp),
                                 o_pp, _stream()), "lcode_deposit")
    return tuple(outs)


def initial_deposition(config, x_offt, y_offt, px, py, pz, m, q, virt_params):
    """lcode.py:692-699."""
    ro_electrons_initial, _, _, _ = deposit(config, 0, x_offt, y_offt, m, q,
                                            px, py, pz, virt_params)
    return -ro_electrons_initial


# ---------------------------------------------------------------------------
# The xi step (lcode.py:828-931)
# ---------------------------------------------------------------------------

FIELDS = ("Ex", "Ey", "Ez", "Bx", "By", "Bz")


#: Default acceptance threshold of a guess: the largest corrector residual
#: ||F_corr - guess|| / ||F_corr|| at which the guess stands in for the
#: predictor pass.  The reference's one-pass corrector leaves a residual of
#: about 0.1 dxi even for a perfect guess (3e-5 .. 6e-4 over a 641^2 run), so
#: a threshold of 1e-5 would reject every guess; 2e-3 accepts guesses on the
#: smooth slices and keeps Ez(xi, 0, 0) within 1e-5 of the plain step.
GUIDED_THRESHOLD = 2e-3


def is_mixed(config):
    """True when config selects the mixed-precision step (float32 grids)."""
    return getattr(config, "field_solver_precision", "fp64") == "mixed"


def grid_dtype(config):
    """dtype of the state's ten N x N grids: float32 in the mixed form."""
    return torch.float32 if is_mixed(config) else _F64


def _check_step_arrays(config, const, prev):
    """dtype, layout, shape and device of every array the fused step reads
    through a raw pointer (before anything is launched)."""
    nc = prev.x_offt.shape[0]
    n = int(config.grid_steps)
    coarse = ("x_offt", "y_offt", "px", "py", "pz")
    gdt = grid_dtype(config)
    checks = [(f"const.{k}", getattr(const, k), (nc, nc), _F64)
              for k in ("m", "q", "x_init", "y_init")]
    checks.append(("const.ro_initial", const.ro_initial, (n, n), _F64))
    checks += [(f"prev.{k}", getattr(prev, k),
                (nc, nc) if k in coarse else (n, n),
                _F64 if k in coarse else gdt) for k in _lib.STATE_FIELDS]
    for name, t, shape, dt in checks:
        if (not isinstance(t, torch.Tensor) or t.dtype != dt
                or not t.is_contiguous() or tuple(t.shape) != shape):
            raise ValueError(f"{name} must be a contiguous "
                             f"{str(dt).replace('torch.', '')} tensor "
                             f"of shape {shape}")
    for name, t, shape, _ in checks:
        if t.device.type != "cuda":
            raise ValueError(f"{name} must be a device (HIP) tensor")
    return nc


def check_nonfinite(ws, reset=True):
    """Raise if a kernel on workspace `ws` has met a non-finite particle
    state.  Reads a host-visible flag without synchronising, so a step
    whose kernels are still running reports at a later call."""
    flag = ctypes.c_int32()
    _lib.check(_lib.load().lcode_workspace_nonfinite(ws, 1 if reset else 0,
                                                     ctypes.byref(flag)),
               "lcode_workspace_nonfinite")
    if flag.value:
        raise FloatingPointError(
            "a particle offset or momentum became non-finite (or moved "
            "more than 1e6 cells) in an earlier step on this workspace; "
            "the grids deposited from it are NaN")


def step(config, const, virt_params, prev, beam_ro, guess=None,
         threshold=GUIDED_THRESHOLD):
    """One xi step; returns the new state (lcode.py:828-931).  beam_ro is
    an (N, N) array (host or device) or a scalar.

    Raises FloatingPointError when an earlier step on the same workspace
    met a non-finite particle state (checked without synchronising, so the
    error surfaces within a step or two of the cause).

    guess (not in the reference): the predictor hook of lcode.py:841-881.
    A sequence of (coefficient, fields) terms, fields being an object with
    Ex..Bz attributes (e.g. earlier states) or a 6-tuple of (N, N) device
    arrays; their linear combination replaces the predictor pass's
    deposition and field solve.  The corrector's residual against the guess
    is checked on the device and the plain predictor runs when it exceeds
    `threshold` (see guided_residual and GUIDED_THRESHOLD)."""
    nc = _check_step_arrays(config, const, prev)
    lib = _lib.load()
    beam, beam_scalar = _beam_arg(config, beam_ro)
    vp, nf, keep_vp = _virt_struct(virt_params)
    cs = _config_struct(config, nc, nf)
    ws = _workspace(cs)
    check_nonfinite(ws)
    c = _lib.LcodeConst(*(getattr(const, n).data_ptr()
                          for n in ("m", "q", "x_init", "y_init",
                                    "ro_initial")))
    p = _lib.LcodeState(*(getattr(prev, name).data_ptr()
                          for name in _lib.STATE_FIELDS))
    new = {name: torch.empty_like(getattr(prev, name))
           for name in _lib.STATE_FIELDS}
    o = _lib.LcodeState(*(new[name].data_ptr() for name in _lib.STATE_FIELDS))
    beam_ptr = _ptr(beam) if beam is not None else None
    mixed = is_mixed(config)
    if guess is None:
        fn = lib.lcode_step_mixed if mixed else lib.lcode_step
        _lib.check(fn(ws, ctypes.byref(cs), ctypes.byref(c), ctypes.byref(vp),
                      ctypes.byref(p), beam_ptr, beam_scalar,
                      ctypes.byref(o), _stream()),
                   "lcode_step_mixed" if mixed else "lcode_step")
    else:
        g, keep_g = _guess_struct(config, guess)
        fn = lib.lcode_step_guided_mixed if mixed else lib.lcode_step_guided
        _lib.check(fn(ws, ctypes.byref(cs), ctypes.byref(c), ctypes.byref(vp),
                      ctypes.byref(p), beam_ptr, beam_scalar, ctypes.byref(g),
                      float(threshold), ctypes.byref(o), _stream()),
                   "lcode_step_guided_mixed" if mixed else "lcode_step_guided")
    state = GPUArrays.__new__(GPUArrays)
    for name, t in new.items():
        setattr(state, name, t)
    return state


def predictor_fields(config, const, virt_params, prev, beam_ro):
    """The six fields of the plain step's predictor pass (lcode.py:841-874)
    from `prev` (FP64 state), composed from this module's operators: what a
    perfect guess of step(..., guess=...) is (the training target of the
    learned predictor)."""
    c = const
    x, y = move_estimate_wo_fields(config, c.m, c.x_init, c.y_init,
                                   prev.x_offt, prev.y_offt, prev.px, prev.py, prev.pz)
    x, y, px, py, pz = move_smart(config, c.m, c.q, c.x_init, c.y_init,
                                  prev.x_offt, prev.y_offt, x, y,
                                  prev.px, prev.py, prev.pz,
                                  *(getattr(prev, k) for k in FIELDS))
    ro, jx, jy, jz = deposit(config, c.ro_initial, x, y, c.m, c.q, px, py, pz,
                             virt_params)
    if config.field_solver_variant_A:
        ro_in, jz_in = (ro + prev.ro) / 2, (jz + prev.jz) / 2
    else:
        ro_in, jz_in = ro, jz
    Ex, Ey, Bx, By = calculate_Ex_Ey_Bx_By(config, prev.Ex, prev.Ey, prev.Bx, prev.By,
                                           beam_ro, ro_in, jx, jy, jz_in,
                                           prev.jx, prev.jy)
    if config.field_solver_variant_A:
        Ex, Ey = 2 * Ex - prev.Ex, 2 * Ey - prev.Ey
        Bx, By = 2 * Bx - prev.Bx, 2 * By - prev.By
    return (Ex, Ey, calculate_Ez(config, jx, jy), Bx, By, calculate_Bz(config, jx, jy))


class WorkspacePredictorFields:
    """A guess term that names the workspace's kept predictor-pass fields
    (lcode_workspace_predictor_fields): after a plain step with
    keep_predictor_fields on, its F1; after a guided step, its guess, or F1
    when the guess was rejected.  The six device pointers, in the grid type
    of the config's steps."""

    def __init__(self, config, nc, nf):
        self.ws = _workspace(_config_struct(config, nc, nf))
        ptrs = (ctypes.c_void_p * 6)()
        _lib.check(_lib.load().lcode_workspace_predictor_fields(self.ws, ptrs),
                   "lcode_workspace_predictor_fields")
        self.ptrs = tuple(int(p) for p in ptrs)

    def keep(self, on):
        _lib.check(_lib.load().lcode_workspace_keep_predictor_fields(self.ws, int(bool(on))),
                   "lcode_workspace_keep_predictor_fields")


def _guess_struct(config, guess):
    terms = list(guess)
    if not 1 <= len(terms) <= 4:
        raise ValueError("a guess has 1 to 4 (coefficient, fields) terms")
    g = _lib.LcodeGuess()
    keep = []
    for t, (coef, fields) in enumerate(terms):
        if isinstance(fields, WorkspacePredictorFields):
            g.coef[t] = float(coef)
            for f, ptr in enumerate(fields.ptrs):
                g.f[t][f] = ptr
            continue
        if not isinstance(fields, (tuple, list)):
            fields = tuple(getattr(fields, name) for name in FIELDS)
        arrs = _check_grid(config, *fields, dtype=grid_dtype(config))
        if len(arrs) != 6:
            raise ValueError("a guess term needs the six fields Ex..Bz")
        keep.extend(arrs)
        g.coef[t] = float(coef)
        for f, a in enumerate(arrs):
            g.f[t][f] = a.data_ptr()
    g.nterms = len(terms)
    return g, keep


def guided_residual(config, nc, nf):
    """||F_corr - guess|| / ||F_corr|| of the last guided step on this
    workspace (synchronises)."""
    lib = _lib.load()
    r = ctypes.c_double()
    _lib.check(lib.lcode_step_guided_residual(
        _workspace(_config_struct(config, nc, nf)), _stream(), ctypes.byref(r)),
        "lcode_step_guided_residual")
    return r.value


class FieldExtrapolator:
    """Predictor for the guided step: polynomial extrapolation in xi of the
    last slices' fields, quadratic 3 F(n-1) - 3 F(n-2) + F(n-3) once three
    slices are known (linear, then none, before that).

    alternating=True (for Variant A, lcode.py:869-871): the transverse
    fields leave each step as 2 F - F_prev, so an error in them returns
    with the opposite sign on the next slice and never decays; the
    quadratic extrapolation turns such an alternating part e (-1)^k into
    -7 e (-1)^n where the next slice carries +e (-1)^n.  Once four slices are known
    the guess is 2 F(n-1) - 2 F(n-3) + F(n-4): twice the quadratic
    extrapolation of the averages (F(k) + F(k-1)) / 2, minus F(n-1).  It
    is exact for fields quadratic in xi, and it carries an alternating
    part e (-1)^k on as e (-1)^n, as the step does."""

    COEFS = {1: None, 2: ((2., 0), (-1., 1)), 3: ((3., 0), (-3., 1), (1., 2))}
    ALTERNATING = ((2., 0), (-2., 2), (1., 3))

    def __init__(self, order=2, alternating=False):
        self.order = int(order)
        self.alternating = bool(alternating) and self.order == 2
        self.history = []            # newest first

    def push(self, state):
        self.history.insert(0, state)
        del self.history[self.order + (2 if self.alternating else 1):]

    def guess(self):
        if self.alternating and len(self.history) >= 4:
            return [(c, self.history[k]) for c, k in self.ALTERNATING]
        n = min(len(self.history), self.order + 1)
        if n < 2:
            return None
        return [(c, self.history[k]) for c, k in self.COEFS[n]]


class GuessBackoff:
    """When to offer a guess: after a rejected guess the next `wait` slices
    run the plain predictor; `wait` starts at `backoff`, doubles with every
    further rejection up to `max_backoff`, and goes back to 0 (guess every
    slice) after an accepted one.  backoff = 0: guess every slice."""

    def __init__(self, backoff=8, max_backoff=64):
        self.backoff = int(backoff)
        self.max_backoff = max(int(max_backoff), self.backoff)
        self.next_wait = self.backoff
        self.wait = 0

    def want_guess(self):
        """True when this slice should be guided; counts a skipped slice."""
        if self.wait > 0:
            self.wait -= 1
            return False
        return True

    def observe(self, accepted):
        """Outcome of the guided slice just run."""
        if accepted:
            self.next_wait = self.backoff
        else:
            self.wait = self.next_wait
            self.next_wait = min(2 * self.next_wait, self.max_backoff) if self.backoff else 0


def _takes_five(fn):
    """Whether fn accepts a fifth positional argument."""
    import inspect
    try:
        sig = inspect.signature(fn)
    except (TypeError, ValueError):
        return False
    n = 0
    for p in sig.parameters.values():
        if p.kind == p.VAR_POSITIONAL:
            return True
        if p.kind in (p.POSITIONAL_ONLY, p.POSITIONAL_OR_KEYWORD):
            n += 1
    return n >= 5


class GuidedStepper:
    """step() with a guess on the slices GuessBackoff selects: the guess is
    the quadratic extrapolation of the last three states' fields, or the
    fields a predictor(s1, s2, s3, beam) returns (states n-1, n-2, n-3;
    e.g. unet.UNetPredictor.guess).  The acceptance of a guess is
    decided on the device (the plain predictor runs there when the residual
    is above the threshold), so no slice waits for the host: each guided
    slice enqueues an asynchronous copy of its residual sums into a pinned
    ring and records an event, and the backoff policy observes outcomes as
    their copies complete (`lag` slices late at most, after which it
    waits; sync=True: each guided slice waits for its own outcome, so the
    policy's decisions do not depend on timing).  Counts guided, accepted
    and plain slices; `residuals` lists the observed residuals in slice
    order.

    extrapolation (without a predictor):
    'gap' (default) — a guess of the predictor pass's own fields F1:
      G(n) = E(n) - (F(n-1) - X(n-1)), the Variant-A-aware extrapolation E
      (FieldExtrapolator) of the corrector's fields minus the last slice's
      one-pass change, X(n-1) being that slice's F1 (kept by the plain step)
      or its guess (kept by the guided step); at 641^2 it is within
      1e-6..1e-5 of F1 along the whole run, where E is 1e-4..1e-3 off, so
      an accepted slice differs from the plain step by about that much and
      nearly every guess passes;
    'quadratic' — 3 F(n-1) - 3 F(n-2) + F(n-3), a guess of the corrector's
      fields (an accepted slice then differs from the plain step by the
      one-pass change, ~1e-3, in the transverse magnetic field);
    'alternating' — the Variant-A-aware form of it.
    A predictor is called as predictor(s1, s2, s3, beam), with slice n-4
    as a fifth argument once known when it takes one (UNetPredictor.guess
    does, for its alternating base).

    anchor: after this many guided slices in a row one plain slice runs
    (0: never).  The gap guess is built from the previous guess, so on
    guided slices alone its error e obeys e(n+1) = 2 e(n) - 2 e(n-2) +
    e(n-3) where the corrector passes it on unchanged (the transverse
    fields): a triple root at 1, polynomial growth; at 641^2 the magnetic
    field drifted 4e-3 from the plain step in 600 guided slices and 0.23 in
    1500.  A plain slice restarts the recursion from the true F1."""

    EXTRAPOLATIONS = ("gap", "quadratic", "alternating")

    def __init__(self, config, const, virt_params, threshold, predictor=None,
                 backoff=8, max_backoff=64, lag=8, sync=False, extrapolation="gap",
                 anchor=8):
        if extrapolation not in self.EXTRAPOLATIONS:
            raise ValueError(f"extrapolation must be one of {self.EXTRAPOLATIONS}")
        self.config, self.const, self.vp = config, const, virt_params
        self.threshold = float(threshold)
        self.predictor = predictor
        self.policy = GuessBackoff(backoff, max_backoff)
        variant_a = bool(getattr(config, "field_solver_variant_A", True))
        self.mode = extrapolation if predictor is None else "predictor"
        # four states of history (the alternating forms, and a predictor
        # that takes slice n-4 as its fifth argument)
        self._pred_four = predictor is not None and _takes_five(predictor)
        self.extrap = FieldExtrapolator(
            order=2, alternating=(variant_a and self.mode in ("gap", "alternating"))
            or self._pred_four)
        self._x = None               # kept predictor fields (gap mode)
        self._x_valid = False
        self.anchor = max(int(anchor), 0)
        self._streak = 0             # guided slices in a row
        self.guided = self.accepted = self.plain = 0
        self.residuals = []
        self.lag = max(int(lag), 1)
        self.sync = bool(sync)
        self._ring = torch.zeros((self.lag, 2), dtype=_F64, pin_memory=True)
        self._events = [None] * self.lag
        self._pending = []           # ring slots in slice order
        self._next = 0

    def push(self, state):
        """Record a state computed outside the stepper (history only)."""
        self.extrap.push(state)

    def _observe(self, slot):
        num, den = (float(v) for v in self._ring[slot])
        r = math.sqrt(num / den) if den > 0 else (math.inf if num > 0 else 0.)
        ok = r <= self.threshold
        self.residuals.append(r)
        self.accepted += int(ok)
        self.policy.observe(ok)

    def drain(self, wait=False):
        """Feed the policy every completed residual copy (all of them, in
        order, when wait is true)."""
        while self._pending:
            slot = self._pending[0]
            ev = self._events[slot]
            if wait:
                ev.synchronize()
            elif not ev.query():
                break
            self._pending.pop(0)
            self._observe(slot)

    def _gap_guess(self, h):
        """E(n) - F(n-1) + X(n-1) with E = 2 F(n-1) - 2 F(n-3) + F(n-4)
        (or the quadratic extrapolation before four slices are known)."""
        e = self.extrap.guess()
        if not self._x_valid:
            return e
        terms = {}
        for c, st in e:
            terms[id(st)] = [terms.get(id(st), [0.0, st])[0] + c, st]
        terms[id(h[0])][0] -= 1.0
        out = [(c, st) for c, st in terms.values() if c != 0.0]
        return out + [(1.0, self._x)]

    def step(self, prev, beam_ro):
        self.drain()
        h = self.extrap.history
        ready = len(h) >= 3
        if self.mode == "gap" and self._x is None:
            _, nf, _ = _virt_struct(self.vp)
            self._x = WorkspacePredictorFields(self.config, prev.x_offt.shape[0], nf)
        if self.anchor and self._streak >= self.anchor:
            ready = False            # an anchoring plain slice
        if ready and self.policy.want_guess():
            self._streak += 1
            if self.mode == "gap":
                guess = self._gap_guess(h)
            elif self.predictor is None:
                guess = self.extrap.guess()
            elif self._pred_four and len(h) >= 4:
                guess = [(1.0, self.predictor(h[0], h[1], h[2], beam_ro, h[3]))]
            else:
                guess = [(1.0, self.predictor(h[0], h[1], h[2], beam_ro))]
            st = step(self.config, self.const, self.vp, prev, beam_ro,
                      guess=guess, threshold=self.threshold)
            if len(self._pending) == self.lag:
                # every slot in flight: wait for the oldest
                old = self._pending.pop(0)
                self._events[old].synchronize()
                self._observe(old)
            slot = self._next
            self._next = (self._next + 1) % self.lag
            _, nf, _ = _virt_struct(self.vp)
            ws = _workspace(_config_struct(self.config, prev.x_offt.shape[0], nf))
            _lib.check(_lib.load().lcode_step_guided_residual_async(
                ws, ctypes.c_void_p(self._ring[slot].data_ptr()), _stream()),
                "lcode_step_guided_residual_async")
            ev = torch.cuda.Event()
            ev.record()
            self._events[slot] = ev
            self._pending.append(slot)
            self.guided += 1
            if self.sync:
                self.drain(wait=True)
        else:
            if self._x is not None:
                self._x.keep(True)
            try:
                st = step(self.config, self.const, self.vp, prev, beam_ro)
            finally:
                if self._x is not None:
                    self._x.keep(False)
            self.plain += 1
            self._streak = 0
        # after either kind of slice the workspace holds X(n) (gap mode)
        self._x_valid = self._x is not None
        self.extrap.push(st)
        return st


# ---------------------------------------------------------------------------
# Plasma set-up and init (lcode.py:454-573, :936-975); runs once, on the host
# like the reference's, then uploads
# ---------------------------------------------------------------------------

def make_coarse_plasma_grid(steps, step_size, coarseness=3):
    """lcode.py:454-465."""
    assert coarseness == int(coarseness)
    half = np.arange(steps // (coarseness * 2)) * (step_size * coarseness)
    grid = np.concatenate([-half[:0:-1], half])
    assert np.array_equal(grid, -grid[::-1])
    return grid


def make_fine_plasma_grid(steps, step_size, fineness=2):
    """lcode.py:468-487."""
    assert fineness == int(fineness)
    step = step_size / fineness
    if fineness % 2:
        half = np.arange(steps // 2 * fineness) * step
        grid = np.concatenate([-half[:0:-1], half])
    else:
        half = (.5 + np.arange(steps // 2 * fineness)) * step
        grid = np.concatenate([-half[::-1], half])
    assert np.array_equal(grid, -grid[::-1])
    return grid


def make_plasma(steps, cell_size, coarseness=3, fineness=2):
    """lcode.py:490-573 on the device (lcode_make_plasma): coarse particle
    arrays and the coarse->fine interpolation parameters (virt_params),
    bit for bit the reference's NumPy values."""
    assert coarseness == int(coarseness) and fineness == int(fineness)
    lib = _lib.load()
    nc, nf = ctypes.c_int32(), ctypes.c_int32()
    _lib.check(lib.lcode_plasma_sizes(int(steps), int(coarseness), int(fineness),
                                      ctypes.byref(nc), ctypes.byref(nf)),
               "lcode_plasma_sizes")
    nc, nf = nc.value, nf.value
    dev = _device()
    shape = (nc, nc)
    x_init, y_init, m, q = (torch.empty(shape, dtype=_F64, device=dev) for _ in range(4))
    ip, inn, fg = (torch.empty(nf, dtype=_F64, device=dev) for _ in range(3))
    prv, nxt = (torch.empty(nf, dtype=torch.int64, device=dev) for _ in range(2))
    _lib.check(lib.lcode_make_plasma(int(steps), float(cell_size), int(coarseness),
                                     int(fineness), *(_ptr(t) for t in (
                                         x_init, y_init, m, q, ip, inn, fg, prv, nxt)),
                                     _stream()), "lcode_make_plasma")
    virt_params = GPUArrays(influence_prev=ip, influence_next=inn,
                            indices_prev=prv, indices_next=nxt, fine_grid=fg)
    z = [torch.zeros(shape, dtype=_F64, device=dev) for _ in range(5)]
    return (x_init, y_init, z[0], z[1], z[2], z[3], z[4], m, q, virt_params)


def init(config):
    """lcode.py:936-975."""
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
    ro_initial = initial_deposition(config, x_offt, y_offt, px, py, pz, m, q,
                                    virt_params)
    const = GPUArrays(m=m, q=q, x_init=x_init, y_init=y_init,
                      ro_initial=ro_initial)
    dev = _device()

    def zeros():
        return torch.zeros((n, n), dtype=grid_dtype(config), device=dev)

    state = GPUArrays(x_offt=x_offt, y_offt=y_offt, px=px, py=py, pz=pz,
                      Ex=zeros(), Ey=zeros(), Ez=zeros(),
                      Bx=zeros(), By=zeros(), Bz=zeros(),
                      ro=zeros(), jx=zeros(), jy=zeros(), jz=zeros())
    return xs, ys, const, virt_params, state
