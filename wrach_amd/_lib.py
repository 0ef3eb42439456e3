"""ctypes binding of libwrach_hip.so (include/wrach.h).

Loading fails loudly: there is no CPU fallback for the step.  Build the
library first (`python -m wrach_amd.build` or __graft_entry__.build()).
"""
from __future__ import annotations

import ctypes
import os

_PKG = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_PKG, "libwrach_hip.so")

WRACH_OK = 0
WRACH_E_INVALID = -1
WRACH_E_TOO_MANY_CELLS = -2
WRACH_E_CAPACITY = -3
WRACH_E_NONFINITE = -4
WRACH_E_DEVICE = -5
WRACH_E_BUFFER = -6
WRACH_E_STATE = -7
WRACH_FLAG_HOST_PACK = 0x1
WRACH_SOLVER_R = 0
WRACH_SOLVER_S = 1
WRACH_SOLVER_L = 2
WRACH_LEARN_FEATURES = 64
WRACH_LEARN_PARAMS = 25090
WRACH_LEARN_BUDGET = 4096
WRACH_NOALLOC_SCRATCH_BYTES = 32 << 20
WRACH_SOLVER_QUIET = 0x100
WRACH_MAX_PASSES = 64
WRACH_STENCIL_BUDGET = 4096
NUM_FRAME_KERNELS = 4
FRAME_KERNEL_NAMES = ("collide", "cells", "scan", "scatter")

# Every symbol include/wrach.h declares.
EXPORTS = (
    "wrach_config_default",
    "wrach_create",
    "wrach_create_batch",
    "wrach_create_worlds",
    "wrach_settings_world",
    "wrach_add_particles_to_world",
    "wrach_read_world",
    "wrach_worlds",
    "wrach_destroy",
    "wrach_add_particles",
    "wrach_step",
    "wrach_read",
    "wrach_capacity",
    "wrach_settings",
    "wrach_sync",
    "wrach_device_state",
    "wrach_canonical_device",
    "wrach_store_write_back",
    "wrach_load_packed",
    "wrach_get_tuning",
    "wrach_set_tuning",
    "wrach_autotune",
    "wrach_set_graph",
    "wrach_set_view_anchor",
    "wrach_set_solver",
    "wrach_solver_stats",
    "wrach_step_profiled",
    "wrach_prepare",
    "wrach_learn_set_params",
    "wrach_learn_get_params",
    "wrach_learn_features",
    "wrach_learn_forward",
    "wrach_learn_gradient",
    "wrach_learn_adam",
    "wrach_learn_sample",
    "wrach_learn_clear_samples",
    "wrach_learn_train",
    "wrach_learned_stats",
    "wrach_learn_profile",
    "wrach_frame_noalloc",
    "wrach_test_scan_wait",
    "wrach_check_pair_arithmetic",
    "wrach_last_error",
)


class WrachConfigC(ctypes.Structure):
    _fields_ = [
        ("width", ctypes.c_uint16),
        ("height", ctypes.c_uint16),
        ("cell_size", ctypes.c_uint16),
        ("boundaries_as_dimensions", ctypes.c_uint8),
        ("reserved0", ctypes.c_uint8),
        ("gravity", ctypes.c_float * 2),
        ("dt", ctypes.c_float),
        ("flags", ctypes.c_uint32),
        ("device", ctypes.c_int32),
    ]


class WorldSettingsC(ctypes.Structure):
    _fields_ = [
        ("view_dimensions", ctypes.c_float * 2),
        ("view_anchor", ctypes.c_float * 2),
        ("grid_dimensions", ctypes.c_uint32 * 2),
        ("cell_size", ctypes.c_uint32),
        ("particles_in_frame_count", ctypes.c_uint32),
    ]


class LearnedReportC(ctypes.Structure):
    _fields_ = [
        ("start_overlap_sum", ctypes.c_uint64),
        ("start_pairs", ctypes.c_uint64),
        ("warm_overlap_sum", ctypes.c_uint64),
        ("warm_pairs", ctypes.c_uint64),
        ("fell_back", ctypes.c_uint32),
        ("reserved", ctypes.c_uint32),
    ]


class SolverReportC(ctypes.Structure):
    _fields_ = [
        ("max_overlap_before", ctypes.c_float),
        ("max_overlap_after", ctypes.c_float),
        ("passes", ctypes.c_uint32),
        ("after_measured", ctypes.c_uint32),
        ("overlap_sum_before", ctypes.c_uint64),
        ("pairs_before", ctypes.c_uint64),
        ("overlap_sum_after", ctypes.c_uint64),
        ("pairs_after", ctypes.c_uint64),
    ]


class TuningC(ctypes.Structure):
    _fields_ = [
        ("scan_threads", ctypes.c_uint32),
        ("scan_items", ctypes.c_uint32),
        ("scatter_per_lane", ctypes.c_uint32),
        ("cells_grid", ctypes.c_uint32),
        ("frame_us", ctypes.c_float),
    ]


class DeviceViewC(ctypes.Structure):
    _fields_ = [
        ("indices", ctypes.c_void_p),
        ("positions", ctypes.c_void_p),
        ("velocities", ctypes.c_void_p),
        ("n_particles", ctypes.c_uint32),
        ("n_indices", ctypes.c_uint32),
        ("stream", ctypes.c_void_p),
    ]


assert ctypes.sizeof(WorldSettingsC) == 32


class WrachError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"wrach error {code}: {msg}")
        self.code = code


_lib = None


def load(path: str = LIB_PATH) -> ctypes.CDLL:
    """Load the HIP library (once).  Raises if it is missing."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise RuntimeError(
            f"{path} is missing: build it with `python -m wrach_amd.build` "
            "(the particle step has no CPU fallback)")
    L = ctypes.CDLL(path)
    vp, f32p, u32p = ctypes.c_void_p, ctypes.POINTER(ctypes.c_float), ctypes.POINTER(ctypes.c_uint32)
    L.wrach_config_default.argtypes = [ctypes.POINTER(WrachConfigC)]
    L.wrach_config_default.restype = None
    L.wrach_create.argtypes = [ctypes.POINTER(WrachConfigC), ctypes.POINTER(vp)]
    L.wrach_destroy.argtypes = [vp]
    L.wrach_destroy.restype = None
    L.wrach_add_particles.argtypes = [vp, f32p, f32p, ctypes.c_uint32]
    L.wrach_step.argtypes = [vp, ctypes.c_uint32]
    L.wrach_read.argtypes = [vp, f32p, f32p, ctypes.c_size_t, u32p, ctypes.c_size_t, u32p]
    L.wrach_capacity.argtypes = [vp, u32p, u32p]
    L.wrach_settings.argtypes = [vp, ctypes.POINTER(WorldSettingsC)]
    L.wrach_sync.argtypes = [vp]
    L.wrach_device_state.argtypes = [vp, ctypes.POINTER(DeviceViewC)]
    L.wrach_canonical_device.argtypes = [vp, ctypes.POINTER(DeviceViewC)]
    L.wrach_create_batch.argtypes = [ctypes.POINTER(WrachConfigC), ctypes.c_uint32, ctypes.POINTER(vp)]
    L.wrach_create_worlds.argtypes = [ctypes.POINTER(WrachConfigC), ctypes.c_uint32, ctypes.POINTER(vp)]
    L.wrach_settings_world.argtypes = [vp, ctypes.c_uint32, ctypes.POINTER(WorldSettingsC)]
    L.wrach_add_particles_to_world.argtypes = [vp, ctypes.c_uint32, f32p, f32p, ctypes.c_uint32]
    L.wrach_read_world.argtypes = [vp, ctypes.c_uint32, f32p, f32p, ctypes.c_size_t, u32p,
                                   ctypes.c_size_t, u32p]
    L.wrach_worlds.argtypes = [vp, u32p]
    L.wrach_store_write_back.argtypes = [vp]
    L.wrach_load_packed.argtypes = [vp, u32p, ctypes.c_size_t, f32p, f32p, ctypes.c_uint32]
    L.wrach_get_tuning.argtypes = [vp, ctypes.POINTER(TuningC)]
    L.wrach_set_tuning.argtypes = [vp, ctypes.POINTER(TuningC)]
    L.wrach_autotune.argtypes = [vp, ctypes.c_uint32, ctypes.POINTER(TuningC)]
    L.wrach_set_graph.argtypes = [vp, ctypes.c_int]
    L.wrach_set_solver.argtypes = [vp, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_float]
    L.wrach_set_view_anchor.argtypes = [vp, ctypes.c_float, ctypes.c_float]
    L.wrach_solver_stats.argtypes = [vp, ctypes.POINTER(SolverReportC)]
    L.wrach_step_profiled.argtypes = [vp, ctypes.c_uint32, ctypes.POINTER(ctypes.c_double)]
    L.wrach_prepare.argtypes = [vp, ctypes.c_uint32]
    u16p = ctypes.POINTER(ctypes.c_uint16)
    u64p = ctypes.POINTER(ctypes.c_uint64)
    L.wrach_learn_set_params.argtypes = [vp, f32p, ctypes.c_size_t]
    L.wrach_learn_get_params.argtypes = [vp, f32p, ctypes.c_size_t]
    L.wrach_learn_features.argtypes = [vp, u16p, ctypes.c_size_t, u64p]
    L.wrach_learn_forward.argtypes = [vp, u16p, ctypes.c_uint32, f32p]
    L.wrach_learn_gradient.argtypes = [vp, u16p, f32p, ctypes.c_uint32, f32p, f32p]
    L.wrach_learn_adam.argtypes = [vp, f32p, ctypes.c_float]
    L.wrach_learn_sample.argtypes = [vp, ctypes.c_uint32, u64p]
    L.wrach_learn_clear_samples.argtypes = [vp]
    L.wrach_learn_train.argtypes = [vp, u32p, ctypes.c_size_t, ctypes.c_uint32, ctypes.c_float, f32p]
    L.wrach_learned_stats.argtypes = [vp, ctypes.POINTER(LearnedReportC)]
    L.wrach_learn_profile.argtypes = [vp, ctypes.c_uint32, ctypes.POINTER(ctypes.c_double)]
    L.wrach_frame_noalloc.argtypes = [ctypes.POINTER(WorldSettingsC), f32p, ctypes.c_float, vp, vp, vp,
                                      ctypes.c_uint32, vp]
    L.wrach_test_scan_wait.argtypes = [vp, ctypes.c_uint32, ctypes.c_uint32]
    L.wrach_check_pair_arithmetic.argtypes = [ctypes.c_int, u64p]
    L.wrach_last_error.argtypes = []
    L.wrach_last_error.restype = ctypes.c_char_p
    for name in EXPORTS:
        if name not in ("wrach_config_default", "wrach_destroy", "wrach_last_error"):
            getattr(L, name).restype = ctypes.c_int
    _lib = L
    return L


def check(rc: int) -> None:
    if rc != WRACH_OK:
        msg = load().wrach_last_error().decode(errors="replace")
        raise WrachError(rc, msg)
