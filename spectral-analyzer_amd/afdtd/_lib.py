"""ctypes binding of libafdtd.so (the C ABI declared in include/afdtd.h).

The library is the only compute path: if it is missing or cannot be loaded
the import fails loudly; there is no CPU fallback.
"""
from __future__ import annotations

import ctypes as C
import os

_PKG_DIR = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
# AFDTD_LIB selects another build of the same library (e.g. a tuning variant).
LIB_PATH = os.environ.get("AFDTD_LIB") or os.path.join(_PKG_DIR, "libafdtd.so")

FDTD_AIR, FDTD_WALL, FDTD_PLATE, FDTD_PML = 0, 1, 2, 3
FDTD_OK, FDTD_ERR_NULL, FDTD_ERR_DIM, FDTD_ERR_UNSTABLE = 0, -1, -2, -3
FDTD_ERR_DEVICE, FDTD_ERR_ARG = -4, -5
FDTD_BOUNDARY_RIGID_PML, FDTD_BOUNDARY_IMPEDANCE = 0, 1
FDTD_STORE_F32, FDTD_STORE_BF16, FDTD_STORE_F16, FDTD_STORE_BFP16 = 0, 1, 2, 3
FDTD_STORE_F16S = 4
FDTD_KERNEL_AUTO, FDTD_KERNEL_GENERIC, FDTD_KERNEL_FUSED, FDTD_KERNEL_FUSED2 = 0, 1, 2, 3

# The 26 functions of the reference ABI (csrc/include/acoustic_fdtd.h:137-506).
REFERENCE_SYMBOLS = (
    "fdtd_create", "fdtd_destroy", "fdtd_set_bridge_sources", "fdtd_inject_bridge",
    "fdtd_inject_bridge_drive", "fdtd_inject_plate_force", "fdtd_set_face_fractions",
    "fdtd_sample_plate_weighted", "fdtd_precompute_pressure_samplers",
    "fdtd_precompute_velocity_samplers", "fdtd_sample_pressure_precomputed",
    "fdtd_sample_velocity_precomputed", "fdtd_step", "fdtd_get_step_count", "fdtd_get_dt",
    "fdtd_reset", "fdtd_get_pressure_field", "fdtd_get_plate_displacement",
    "fdtd_sample_plate_displacement_batch", "fdtd_sample_pressure", "fdtd_get_velocity_dims",
    "fdtd_get_velocity_x", "fdtd_get_velocity_y", "fdtd_get_velocity_z",
    "fdtd_sample_velocity", "fdtd_get_surface_emission",
)


class FdtdOptions(C.Structure):
    _fields_ = [
        ("struct_size", C.c_int), ("stencil_order", C.c_int), ("courant", C.c_float),
        ("boundary", C.c_int), ("refl", C.c_float * 6), ("rho", C.c_void_p),
        ("bulk", C.c_void_p), ("storage", C.c_int), ("kernel", C.c_int),
        ("tile_y", C.c_int), ("seg_x", C.c_int),
    ]


_lib = None

_vp, _i, _f = C.c_void_p, C.c_int, C.c_float

_SIGS = {
    "fdtd_create": (_vp, [_i, _i, _i, _f, _f, _f, _vp, _i, _vp, _f, _f, _f, _f, _i]),
    "fdtd_create_ex": (_vp, [_i, _i, _i, _f, _f, _f, _vp, _i, _vp, _f, _f, _f, _f, _i, _vp]),
    "fdtd_destroy": (None, [_vp]),
    "fdtd_set_bridge_sources": (_i, [_vp, _i, _vp, _vp]),
    "fdtd_inject_bridge": (_i, [_vp, _f, _f, _f]),
    "fdtd_inject_bridge_drive": (_i, [_vp, _vp, _f]),
    "fdtd_inject_plate_force": (_i, [_vp, _i, _vp, _vp, _f]),
    "fdtd_set_face_fractions": (_i, [_vp, _vp, _vp, _vp]),
    "fdtd_sample_plate_weighted": (_i, [_vp, _i, _vp, _vp, _vp]),
    "fdtd_precompute_pressure_samplers": (_i, [_vp, _i, _vp, _vp, _vp]),
    "fdtd_precompute_velocity_samplers": (_i, [_vp, _i, _vp] + [_vp] * 6),
    "fdtd_sample_pressure_precomputed": (_i, [_vp, _i, _vp, _vp, _vp]),
    "fdtd_sample_pressure_precomputed_noalloc": (_i, [_vp, _i, _vp, _vp, _vp]),
    "fdtd_sample_velocity_precomputed": (_i, [_vp, _i] + [_vp] * 9),
    "fdtd_step": (_i, [_vp, _i]),
    "fdtd_step_timed": (_i, [_vp, _i, _vp, _vp]),
    "fdtd_get_step_count": (_i, [_vp]),
    "fdtd_get_dt": (_f, [_vp]),
    "fdtd_get_kernel_family": (_i, [_vp]),
    "fdtd_get_sweep_steps": (_i, [_vp]),
    "fdtd_reset": (_i, [_vp]),
    "fdtd_get_pressure_field": (_i, [_vp, _vp, _i]),
    "fdtd_get_plate_displacement": (_i, [_vp, _vp, _i]),
    "fdtd_sample_plate_displacement_batch": (_i, [_vp, _i, _vp, _vp, _vp]),
    "fdtd_sample_pressure": (_i, [_vp, _i, _vp, _vp]),
    "fdtd_get_velocity_dims": (None, [_vp, _vp, _vp, _vp]),
    "fdtd_get_velocity_x": (_i, [_vp, _vp, _i]),
    "fdtd_get_velocity_y": (_i, [_vp, _vp, _i]),
    "fdtd_get_velocity_z": (_i, [_vp, _vp, _i]),
    "fdtd_sample_velocity": (_i, [_vp, _i, _vp, _vp, _vp, _vp]),
    "fdtd_get_surface_emission": (_i, [_vp, _i, _vp, _vp, _vp, _vp]),
    "fdtd_default_options": (None, [_vp]),
    "fdtd_set_point_sources": (_i, [_vp, _i, _vp, _vp, _i]),
    "fdtd_set_receivers": (_i, [_vp, _i, _vp]),
    "fdtd_record_receivers": (_i, [_vp, _i]),
    "fdtd_get_receiver_traces": (_i, [_vp, _vp, _i, _vp]),
    "fdtd_energy": (_i, [_vp, _vp]),
    "fdtd_set_state": (_i, [_vp, _vp, _vp, _vp, _vp]),
    "fdtd_last_error": (C.c_char_p, []),
    "fdtd_batch_create": (_vp, [_i, _vp, _f, _f, _f, _vp, _vp, _i, _i, _i, _i]),
    "fdtd_batch_destroy": (None, [_vp]),
    "fdtd_batch_cells": (_i, [_vp, _vp, _vp]),
    "fdtd_batch_autotune": (_i, [_vp, _i, _vp, _vp, _vp, _vp]),
    "fdtd_batch_set_sources": (_i, [_vp, _vp, _vp, _i]),
    "fdtd_batch_set_receivers": (_i, [_vp, _i, _vp]),
    "fdtd_batch_record": (_i, [_vp, _i]),
    "fdtd_batch_step": (_i, [_vp, _i, _i, _vp]),
    "fdtd_batch_get_traces": (_i, [_vp, _vp, _i, _vp]),
    "fdtd_batch_get_pressure": (_i, [_vp, _i, _vp, _i]),
    "fdtd_batch_get_step_count": (_i, [_vp]),
    "fdtd_batch_set_kernel": (_i, [_vp, _i]),
    "fdtd_batch_get_kernel_family": (_i, [_vp]),
    "fdtd_copy_calibration": (_i, [C.c_size_t, _i, _vp]),
    "fdtd_stream_calibration": (_i, [C.c_size_t, _i, _i, _i, _vp]),
    "fdtd_adjoint_create": (_vp, [_i, _i, _i, C.c_double, _i, _i]),
    "fdtd_adjoint_destroy": (None, [_vp]),
    "fdtd_adjoint_set_sources": (_i, [_vp, _i, _vp, _vp, _i]),
    "fdtd_adjoint_set_receivers": (_i, [_vp, _i, _vp, _vp]),
    "fdtd_adjoint_forward": (_i, [_vp, _vp, _vp, _vp, _i, _vp]),
    "fdtd_adjoint_backward": (_i, [_vp, _vp, _vp, _vp, _vp]),
    "fdtd_voxelise_body": (_i, [_vp, _i, _i, _i, _i] + [C.c_double] * 6 + [_i, _i, _vp, _vp]),
    "fdtd_voxelise_body_device": (_i, [_vp, _i, _i, _i, _i] + [C.c_double] * 6
                                  + [_i, _i, _vp, _vp, _vp, _vp]),
    "fdtd_cut_cell_fractions": (_i, [_vp, _vp, _vp, _i, _i, _i, _i] + [C.c_double] * 4
                                + [_i, _i, _vp, _vp, _vp]),
    "fdtd_cut_cell_fractions_device": (_i, [_vp, _vp, _vp, _i, _i, _i, _i] + [C.c_double] * 4
                                       + [_i, _i, _vp, _vp, _vp, _vp, _vp]),
    # instrument co-evolver (include/afdtd_coevolver.h)
    "coevolver_create": (_vp, [_i, _vp, _i, _vp, _i, _vp, _vp, _f, _i, _f]),
    "coevolver_destroy": (None, [_vp]),
    "coevolver_pluck_string": (_i, [_vp, _i, _f, _f, _i]),
    "coevolver_inject_string_force": (_i, [_vp, _i, _i, _vp, _f]),
    "coevolver_schedule_pluck": (_i, [_vp, _i, _i, _f, _f, _i]),
    "coevolver_clear_pluck_schedule": (None, [_vp]),
    "coevolver_step": (_i, [_vp, _i]),
    "coevolver_step_timed": (_i, [_vp, _i, _vp]),
    "coevolver_step_block_with_drive": (_i, [_vp, _vp, _i, _i, _f, _vp]),
    "coevolver_step_async": (_i, [_vp, _i]),
    "coevolver_get_progress": (None, [_vp, _vp]),
    "coevolver_wait": (_i, [_vp]),
    "coevolver_cancel": (_i, [_vp]),
    "coevolver_is_running": (_i, [_vp]),
    "coevolver_get_pickup_output": (_i, [_vp, _i, _vp, _i]),
    "coevolver_get_mic_output": (_i, [_vp, _i, _vp, _i]),
    "coevolver_get_pressure_field": (_i, [_vp, _vp, _i]),
    "coevolver_get_pressure_field_uniform": (_i, [_vp, _i, _i, _i, _vp, _vp, _vp, _i]),
    "coevolver_get_amr_cell_centers": (_i, [_vp, _vp, _i]),
    "coevolver_get_plate_displacement": (_i, [_vp, _vp, _i]),
    "coevolver_get_plate_dims": (_i, [_vp, _vp, _vp, _vp]),
    "coevolver_get_modal_amplitudes": (_i, [_vp, _i, _vp, _vp, _i]),
    "coevolver_sample_pressure": (_i, [_vp, _i, _vp, _vp]),
    "coevolver_get_string_velocity": (_i, [_vp, _i, _vp, _i]),
    "coevolver_get_string_displacement": (_i, [_vp, _i, _vp, _i]),
    "coevolver_get_string_position": (_i, [_vp, _i, _vp, _i]),
    "coevolver_get_n_strings": (_i, [_vp]),
    "coevolver_get_n_pickups": (_i, [_vp]),
    "coevolver_get_n_mics": (_i, [_vp]),
    "coevolver_get_dt_audio": (_f, [_vp]),
    "coevolver_get_dt_fdtd": (_f, [_vp]),
    "coevolver_set_damping_scale": (None, [_vp, _f]),
    "coevolver_reset": (_i, [_vp]),
    "coevolver_set_face_fractions": (_i, [_vp, _vp, _vp, _vp]),
    "coevolver_get_surface_emission": (_i, [_vp, _i, _vp, _vp, _vp, _vp]),
    "coevolver_sample_velocity": (_i, [_vp, _i, _vp, _vp, _vp, _vp]),
    "coevolver_get_fdtd": (_vp, [_vp]),
    "coevolver_get_substeps": (_i, [_vp]),
    "coevolver_get_string_substeps": (_i, [_vp, _i]),
    "coevolver_get_samples_written": (_i, [_vp]),
    "coevolver_get_string_state": (_i, [_vp, _i, _vp, _i]),
    "coevolver_get_neck_state": (_i, [_vp, _vp]),
    "coevolver_create_amr": (_vp, [_i, _vp, _i, _vp, _i, _vp, _vp, _f, _i]),
    "coevolver_create_amr_last_error": (C.c_char_p, []),
    "coevolver_amr_backend_error_code": (C.c_char_p, []),
    "coevolver_amr_backend_error_message": (C.c_char_p, []),
}

# The reference's uniform co-evolver ABI (csrc/include/acoustic_coevolver.h:306-897).
COEVOLVER_SYMBOLS = (
    "coevolver_create", "coevolver_destroy", "coevolver_pluck_string",
    "coevolver_inject_string_force", "coevolver_schedule_pluck",
    "coevolver_clear_pluck_schedule", "coevolver_step", "coevolver_step_block_with_drive",
    "coevolver_step_async", "coevolver_get_progress", "coevolver_wait", "coevolver_cancel",
    "coevolver_is_running", "coevolver_get_pickup_output", "coevolver_get_mic_output",
    "coevolver_get_pressure_field", "coevolver_get_pressure_field_uniform",
    "coevolver_get_amr_cell_centers", "coevolver_get_plate_displacement",
    "coevolver_get_plate_dims", "coevolver_get_modal_amplitudes", "coevolver_sample_pressure",
    "coevolver_get_string_velocity", "coevolver_get_string_displacement",
    "coevolver_get_string_position", "coevolver_get_n_strings", "coevolver_get_n_pickups",
    "coevolver_get_n_mics", "coevolver_get_dt_audio", "coevolver_get_dt_fdtd",
    "coevolver_set_damping_scale", "coevolver_reset", "coevolver_set_face_fractions",
    "coevolver_get_surface_emission", "coevolver_sample_velocity", "coevolver_create_amr",
    "coevolver_create_amr_last_error",
)


def _init_torch_first():
    """A process that also uses PyTorch holds two HIP runtimes: PyTorch's own
    (with its own HSA runtime) and the system ROCm one this library links.
    PyTorch finds no GPU if the system runtime opened the device first, while
    the other order works, so PyTorch's device context is created before the
    library is loaded whenever PyTorch is installed and sees a GPU."""
    try:
        import torch
    except ImportError:
        return
    if torch.cuda.is_available():
        torch.cuda.init()


def lib():
    """Load libafdtd.so once; raise if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(
                f"libafdtd.so not found at {LIB_PATH}: build it with "
                "`python -c 'import __graft_entry__ as g; g.build()'` (hipcc, gfx950)")
        _init_torch_first()
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in _SIGS.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def last_error() -> str:
    msg = lib().fdtd_last_error()
    return msg.decode() if msg else ""


def default_options() -> FdtdOptions:
    o = FdtdOptions()
    lib().fdtd_default_options(C.byref(o))
    return o
