"""ctypes binding of liblcode3d_hip.so (the C ABI in include/lcode3d_hip.h).

The library is built in-tree by ``lcode3d_amd.build`` and links
``libamdhip64.so.7``; torch is imported first so that the library binds to
the HIP runtime torch already loaded (one runtime per process, so torch
streams and device pointers are valid on both sides).  There is no CPU
fallback: a missing library or GPU raises.
"""

import ctypes
import os

import torch  # noqa: F401  (must load the HIP runtime before the library)

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_NAME = "liblcode3d_hip.so"
#: LCODE3D_HIP_LIB selects another build of the same ABI (e.g. a debug one)
LIB_PATH = os.environ.get("LCODE3D_HIP_LIB") or os.path.join(HERE, LIB_NAME)

LCODE_OK = 0
LCODE_SOLVER_FFT64 = 0
LCODE_SOLVER_FFT32 = 1
LCODE_SOLVER_DENSE64 = 2
LCODE_SOLVER_DENSE_BF16X3 = 3
LCODE_SOLVER_MIXED = 4

_dp = ctypes.POINTER(ctypes.c_double)
_i64p = ctypes.POINTER(ctypes.c_int64)


class LcodeConfig(ctypes.Structure):
    _fields_ = [
        ("grid_steps", ctypes.c_int32),
        ("grid_step_size", ctypes.c_double),
        ("xi_step_size", ctypes.c_double),
        ("subtraction_trick", ctypes.c_double),
        ("variant_A", ctypes.c_int32),
        ("reflect_boundary", ctypes.c_double),
        ("plasma_coarseness", ctypes.c_int32),
        ("plasma_fineness", ctypes.c_int32),
        ("nc", ctypes.c_int32),
        ("nf", ctypes.c_int32),
        ("solver", ctypes.c_int32),
    ]


class LcodeConst(ctypes.Structure):
    _fields_ = [(n, ctypes.c_void_p)
                for n in ("m", "q", "x_init", "y_init", "ro_initial")]


class LcodeVirtParams(ctypes.Structure):
    _fields_ = [(n, ctypes.c_void_p)
                for n in ("influence_prev", "influence_next", "fine_grid",
                          "indices_prev", "indices_next")]


STATE_FIELDS = ("x_offt", "y_offt", "px", "py", "pz",
                "Ex", "Ey", "Ez", "Bx", "By", "Bz",
                "ro", "jx", "jy", "jz")


class LcodeState(ctypes.Structure):
    """lcode_state, and lcode_state_mixed (the same layout of pointers)."""
    _fields_ = [(n, ctypes.c_void_p) for n in STATE_FIELDS]


class LcodeGuess(ctypes.Structure):
    """lcode_guess, and lcode_guess_mixed (the same layout)."""
    _fields_ = [("f", (ctypes.c_void_p * 6) * 4),
                ("coef", ctypes.c_double * 4),
                ("nterms", ctypes.c_int32)]


class LcodeUnet(ctypes.Structure):
    _fields_ = [("S", ctypes.c_int32)] + [
        (n, ctypes.c_void_p) for n in ("params", "grad", "adam_m", "adam_v", "wpack",
                                       "act", "y", "partial", "loss_part", "loss")]


class LcodeUnetIo(ctypes.Structure):
    _fields_ = [("f1", ctypes.c_void_p * 6), ("f2", ctypes.c_void_p * 6),
                ("f3", ctypes.c_void_p * 6), ("f4", ctypes.c_void_p * 6),
                ("next", ctypes.c_void_p * 6),
                ("ro", ctypes.c_void_p), ("beam", ctypes.c_void_p),
                ("beam_scalar", ctypes.c_double), ("f32", ctypes.c_int32),
                ("grid_steps", ctypes.c_int32),
                ("scale", ctypes.c_float * 14), ("tscale", ctypes.c_float * 6),
                ("planes", ctypes.c_void_p), ("planes_f32", ctypes.c_void_p),
                ("targets", ctypes.c_void_p), ("guess", ctypes.c_void_p * 6)]


class LcodeKernelTime(ctypes.Structure):
    _fields_ = [("name", ctypes.c_char * 32), ("calls", ctypes.c_int64),
                ("total_ms", ctypes.c_double)]


#: every symbol the header declares, with (restype, argtypes)
_vp = ctypes.c_void_p
_SIGNATURES = {
    "lcode_last_error": (ctypes.c_char_p, []),
    "lcode_version": (ctypes.c_char_p, []),
    "lcode_workspace_create": (ctypes.c_int,
                               [ctypes.POINTER(LcodeConfig),
                                ctypes.POINTER(ctypes.c_void_p)]),
    "lcode_workspace_destroy": (ctypes.c_int, [_vp]),
    "lcode_profile_begin": (ctypes.c_int, [_vp, ctypes.c_int32]),
    "lcode_profile_only": (ctypes.c_int, [_vp, ctypes.c_char_p]),
    "lcode_workspace_keep_predictor_fields": (ctypes.c_int, [_vp, ctypes.c_int32]),
    "lcode_workspace_predictor_fields": (ctypes.c_int, [_vp, ctypes.c_void_p * 6]),
    "lcode_profile_pause": (ctypes.c_int, [_vp, ctypes.c_int32]),
    "lcode_profile_end": (ctypes.c_int,
                          [_vp, _vp, ctypes.POINTER(LcodeKernelTime),
                           ctypes.c_int32, ctypes.POINTER(ctypes.c_int32)]),
    "lcode_workspace_fault_count": (ctypes.c_int,
                                    [_vp, _vp, ctypes.POINTER(ctypes.c_int64)]),
    "lcode_workspace_nonfinite": (ctypes.c_int,
                                  [_vp, ctypes.c_int32,
                                   ctypes.POINTER(ctypes.c_int32)]),
    "lcode_move_estimate_wo_fields": (ctypes.c_int,
                                      [_vp, ctypes.POINTER(LcodeConfig)]
                                      + [_vp] * 10 + [_vp]),
    "lcode_move_smart": (ctypes.c_int,
                         [_vp, ctypes.POINTER(LcodeConfig)] + [_vp] * 11
                         + [ctypes.POINTER(_vp), ctypes.POINTER(_vp), _vp]),
    "lcode_deposit": (ctypes.c_int,
                      [_vp, ctypes.POINTER(LcodeConfig)] + [_vp] * 8
                      + [ctypes.POINTER(LcodeVirtParams),
                         ctypes.POINTER(_vp), _vp]),
    "lcode_plasma_sizes": (ctypes.c_int,
                           [ctypes.c_int32] * 3 + [ctypes.POINTER(ctypes.c_int32)] * 2),
    "lcode_make_plasma": (ctypes.c_int,
                          [ctypes.c_int32, ctypes.c_double, ctypes.c_int32, ctypes.c_int32]
                          + [_vp] * 9 + [_vp]),
    "lcode_deposit_noalloc": (ctypes.c_int,
                              [ctypes.POINTER(LcodeConfig)] + [_vp] * 8
                              + [ctypes.POINTER(LcodeVirtParams),
                                 ctypes.POINTER(_vp), _vp]),
    "lcode_calculate_Ex_Ey_Bx_By": (ctypes.c_int,
                                    [_vp, ctypes.POINTER(LcodeConfig)]
                                    + [_vp] * 5 + [ctypes.c_double]
                                    + [_vp] * 6
                                    + [ctypes.POINTER(_vp), _vp]),
    "lcode_calculate_Ez": (ctypes.c_int,
                           [_vp, ctypes.POINTER(LcodeConfig), _vp, _vp, _vp, _vp]),
    "lcode_calculate_Bz": (ctypes.c_int,
                           [_vp, ctypes.POINTER(LcodeConfig), _vp, _vp, _vp, _vp]),
    "lcode_step": (ctypes.c_int,
                   [_vp, ctypes.POINTER(LcodeConfig),
                    ctypes.POINTER(LcodeConst),
                    ctypes.POINTER(LcodeVirtParams),
                    ctypes.POINTER(LcodeState), _vp, ctypes.c_double,
                    ctypes.POINTER(LcodeState), _vp]),
    "lcode_step_guided": (ctypes.c_int,
                          [_vp, ctypes.POINTER(LcodeConfig),
                           ctypes.POINTER(LcodeConst),
                           ctypes.POINTER(LcodeVirtParams),
                           ctypes.POINTER(LcodeState), _vp, ctypes.c_double,
                           ctypes.POINTER(LcodeGuess), ctypes.c_double,
                           ctypes.POINTER(LcodeState), _vp]),
    "lcode_step_mixed": (ctypes.c_int,
                         [_vp, ctypes.POINTER(LcodeConfig),
                          ctypes.POINTER(LcodeConst),
                          ctypes.POINTER(LcodeVirtParams),
                          ctypes.POINTER(LcodeState), _vp, ctypes.c_double,
                          ctypes.POINTER(LcodeState), _vp]),
    "lcode_step_guided_mixed": (ctypes.c_int,
                                [_vp, ctypes.POINTER(LcodeConfig),
                                 ctypes.POINTER(LcodeConst),
                                 ctypes.POINTER(LcodeVirtParams),
                                 ctypes.POINTER(LcodeState), _vp, ctypes.c_double,
                                 ctypes.POINTER(LcodeGuess), ctypes.c_double,
                                 ctypes.POINTER(LcodeState), _vp]),
    "lcode_step_guided_residual": (ctypes.c_int,
                                   [_vp, _vp, ctypes.POINTER(ctypes.c_double)]),
    "lcode_step_guided_residual_async": (ctypes.c_int, [_vp, _vp, _vp]),
    "lcode_diag_zn": (ctypes.c_int,
                      [_vp, ctypes.POINTER(LcodeConfig), _vp, _vp, _vp]),
    "lcode_unet_sizes": (ctypes.c_int, [ctypes.c_int32, ctypes.POINTER(ctypes.c_int64)]),
    "lcode_unet_tensor_offset": (ctypes.c_int64, [ctypes.c_int32, ctypes.c_int32]),
    "lcode_unet_param_offset": (ctypes.c_int64, [ctypes.c_int32]),
    "lcode_unet_pack": (ctypes.c_int, [ctypes.POINTER(LcodeUnet), _vp]),
    "lcode_unet_forward": (ctypes.c_int, [ctypes.POINTER(LcodeUnet), _vp, _vp]),
    "lcode_unet_infer": (ctypes.c_int, [ctypes.POINTER(LcodeUnet), _vp, _vp]),
    "lcode_unet_backward": (ctypes.c_int, [ctypes.POINTER(LcodeUnet), _vp, _vp, _vp]),
    "lcode_unet_adam": (ctypes.c_int, [ctypes.POINTER(LcodeUnet), ctypes.c_int32,
                                       ctypes.c_float, _vp]),
    "lcode_unet_io": (ctypes.c_int, [ctypes.POINTER(LcodeUnet),
                                     ctypes.POINTER(LcodeUnetIo), ctypes.c_int32, _vp]),
}

SYMBOLS = tuple(_SIGNATURES)

_lib = None


def load():
    """Load the library once; raise if it was not built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(
            f"{LIB_PATH} is missing: build it with "
            "`python -c 'import lcode3d_amd.build as b; b.build()'`")
    lib = ctypes.CDLL(LIB_PATH)
    for name, (res, args) in _SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


class LcodeError(RuntimeError):
    pass


def check(rc, what):
    if rc != LCODE_OK:
        msg = load().lcode_last_error().decode(errors="replace")
        raise LcodeError(f"{what} failed ({rc}): {msg}")


def ptr_array(ptrs):
    arr = (ctypes.c_void_p * len(ptrs))(*ptrs)
    return ctypes.cast(arr, ctypes.POINTER(ctypes.c_void_p)), arr
