"""MI355X-native LCODE 3D xi step: the reference's ``init``/``step`` call
surface (lotov/lcode3d, lcode.py) over gfx950 HIP kernels behind the C ABI
of include/lcode3d_hip.h."""

from .lcode import (  # noqa: F401
    ELECTRON_CHARGE, ELECTRON_MASS, GPUArrays, GPUArraysView,
    calculate_Bz, calculate_Ex_Ey_Bx_By, calculate_Ez, deposit, fault_count,
    init, initial_deposition, make_coarse_plasma_grid, make_fine_plasma_grid,
    make_plasma, move_estimate_wo_fields, move_smart, profile_begin,
    profile_end, profile_pause, step, guided_residual, FieldExtrapolator, GuessBackoff, GuidedStepper,
    predictor_fields, is_mixed, grid_dtype,
)
