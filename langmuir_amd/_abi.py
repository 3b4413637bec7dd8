"""ctypes declarations of the C-ABI in include/langmuir_gpu.h.

The shared library ``liblangmuir_gpu.so`` is built in-tree next to this file by
``langmuir_amd.build``.  Loading fails loudly when it is missing: there is no CPU
fallback anywhere in this package.
"""
from __future__ import annotations

import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liblangmuir_gpu.so")

I32 = C.c_int32
U32 = C.c_uint32
U64 = C.c_uint64
I64 = C.c_int64
F64 = C.c_double
VP = C.c_void_p
PI32 = C.POINTER(C.c_int32)
PF64 = C.POINTER(C.c_double)

LMG_OK = 0
LMG_ERR_ARG = -1
LMG_ERR_NO_DEVICE = -2
LMG_ERR_HIP = -3
LMG_ERR_CAPACITY = -4
LMG_ERR_STATE = -5
LMG_ERR_DEVICE = -6

LMG_GAUSS = 1
LMG_IMAGE_LEFT = 2
LMG_NO_CUTOFF = 4
LMG_MAP_PAIRS = 8
LMG_MAP_CULLED = 16

FLUX_NAMES = ("eSL", "eSR", "hSL", "hSR", "xS", "eDL", "eDR", "hDL", "hDR", "xD")


class LmgParams(C.Structure):
    _fields_ = [
        ("grid_x", I32), ("grid_y", I32), ("grid_z", I32), ("cutoff", I32),
        ("hopping_range", I32), ("defects_charge", I32), ("device_id", I32), ("reserved", I32),
        ("prefactor", F64), ("inverse_kt", F64), ("gaussian_sigma", F64),
        ("exciton_binding", F64),
    ]


class LmgSimConfig(C.Structure):
    _fields_ = [
        ("simulation_type", I32), ("coulomb_carriers", I32), ("source_metropolis", I32),
        ("source_coulomb", I32), ("balance_charges", I32), ("recombination_range", I32),
        ("max_electrons", I32), ("max_holes", I32), ("e_drain_slot", I32 * 2),
        ("h_drain_slot", I32 * 2), ("voltage_left", F64), ("voltage_right", F64),
        ("rate", F64 * 10), ("coupling", F64 * 27),
    ]


class LmgSimState(C.Structure):
    _fields_ = [
        ("n_electrons", I32), ("n_holes", I32), ("n_defects", I32), ("reserved", I32),
        ("e_site", PI32), ("e_lifetime", PI32), ("e_pathlength", PI32),
        ("h_site", PI32), ("h_lifetime", PI32), ("h_pathlength", PI32),
        ("defect_site", PI32), ("potential", PF64),
        ("rng_window", U32 * 624), ("rng_consumed", U64),
        ("flux_attempts", U64 * 10), ("flux_successes", U64 * 10), ("current_step", I64),
    ]


class LmgError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"langmuir_gpu error {code}: {msg}")
        self.code = code


_lib = None


def lib():
    """Load liblangmuir_gpu.so (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(
            f"{LIB_PATH} is missing: build it with `python -c 'import langmuir_amd.build as b; b.build()'`"
            " (there is no CPU fallback)")
    L = C.CDLL(LIB_PATH)
    L.lmg_last_error.restype = C.c_char_p
    L.lmg_last_error.argtypes = [VP]
    L.lmg_device_count.restype = I32
    L.lmg_create.argtypes = [C.POINTER(LmgParams), C.POINTER(VP)]
    L.lmg_destroy.argtypes = [VP]
    L.lmg_destroy.restype = None
    L.lmg_sync.argtypes = [VP]
    L.lmg_stream.argtypes = [VP]
    L.lmg_stream.restype = VP
    L.lmg_pair_potential.argtypes = [VP, VP, VP, I32, VP, I32, I32, VP]
    L.lmg_pair_potential_dev.argtypes = [VP, VP, VP, I32, VP, I32, I32, VP]
    L.lmg_pair_potential_noalloc_dev.argtypes = [VP, VP, VP, I32, VP, I32, I32, VP]
    L.lmg_potential_map.argtypes = [VP, VP, VP, I32, I32, VP]
    L.lmg_potential_map_dev.argtypes = [VP, VP, VP, I32, I32, VP]
    L.lmg_sim_load.argtypes = [VP, C.POINTER(LmgSimConfig), C.POINTER(LmgSimState)]
    L.lmg_sim_store.argtypes = [VP, C.POINTER(LmgSimState)]
    L.lmg_sim_step.argtypes = [VP, I32]
    L.lmg_sim_set_sync.argtypes = [VP, I32]
    L.lmg_sim_set_de_path.argtypes = [VP, I32]
    L.lmg_sim_set_graphs.argtypes = [VP, I32]
    L.lmg_batch_set_graphs.argtypes = [VP, I32]
    L.lmg_sim_get_tuning.argtypes = [VP, C.POINTER(I32), C.POINTER(C.c_double)]
    L.lmg_sim_set_de_chunks.argtypes = [VP, I32]
    L.lmg_batch_get_tuning.argtypes = [VP, C.POINTER(I32), C.POINTER(C.c_double)]
    L.lmg_sim_get_de_path.argtypes = [VP, C.POINTER(I32)]
    L.lmg_sim_delta_e.argtypes = [VP, VP, VP, VP, VP]
    L.lmg_sim_device_error.argtypes = [VP, PI32]
    L.lmg_sim_step_timed.argtypes = [VP, I32, PF64]
    L.lmg_batch_create.argtypes = [I32, VP, VP, VP, I32, C.POINTER(VP)]
    L.lmg_batch_create_with_path.argtypes = [I32, VP, VP, VP, I32, I32, C.POINTER(VP)]
    L.lmg_batch_get_de_paths.argtypes = [VP, C.POINTER(I32)]
    L.lmg_batch_tuning_report.argtypes = [VP]
    L.lmg_batch_tuning_report.restype = C.c_char_p
    L.lmg_batch_destroy.argtypes = [VP]
    L.lmg_batch_destroy.restype = None
    L.lmg_batch_last_error.argtypes = [VP]
    L.lmg_batch_last_error.restype = C.c_char_p
    L.lmg_batch_size.argtypes = [VP]
    L.lmg_batch_size.restype = I32
    L.lmg_batch_step.argtypes = [VP, I32]
    L.lmg_batch_step_timed.argtypes = [VP, I32, PF64]
    L.lmg_batch_step_phases.argtypes = [VP, I32, PF64]
    L.lmg_batch_delta_e.argtypes = [VP, I32, VP, VP, VP, VP]
    L.lmg_batch_sync.argtypes = [VP]
    L.lmg_batch_store.argtypes = [VP, I32, C.POINTER(LmgSimState)]
    L.lmg_get_params.argtypes = [VP, C.POINTER(LmgParams)]
    L.lmg_surrogate_create.argtypes = [VP, I32, I32, C.POINTER(VP)]
    L.lmg_surrogate_destroy.argtypes = [VP]
    L.lmg_surrogate_destroy.restype = None
    L.lmg_surrogate_last_error.argtypes = [VP]
    L.lmg_surrogate_last_error.restype = C.c_char_p
    L.lmg_surrogate_shape.argtypes = [VP, PI32, C.POINTER(I64)]
    L.lmg_surrogate_get_weights.argtypes = [VP, VP]
    L.lmg_surrogate_set_weights.argtypes = [VP, VP]
    L.lmg_surrogate_occupancy_dev.argtypes = [VP, VP, VP, I32, VP]
    L.lmg_surrogate_infer_dev.argtypes = [VP, VP, I32, VP]
    L.lmg_surrogate_gradient_dev.argtypes = [VP, VP, VP, I32, VP, PF64]
    L.lmg_surrogate_train_dev.argtypes = [VP, VP, VP, I32, I32, I32, F64, U64, VP]
    L.lmg_surrogate_map_dev.argtypes = [VP, VP, VP, I32, VP, F64, I32, U64, PF64, PI32]
    L.lmg_surrogate_last_gate.argtypes = [VP, PF64, PI32]
    L.lmg_surrogate_map_batch_dev.argtypes = [VP, VP, VP, PI32, I32, VP, F64, I32, U64, PF64, PI32]
    L.lmg_surrogate_last_gates.argtypes = [VP, I32, PF64, PI32]
    L.lmg_surrogate_set_graphs.argtypes = [VP, I32]
    L.lmg_potential_map_if_dev.argtypes = [VP, VP, VP, I32, I32, VP, VP]
    _lib = L
    return L


EXPORTED = (
    "lmg_create", "lmg_destroy", "lmg_last_error", "lmg_device_count", "lmg_sync", "lmg_stream",
    "lmg_pair_potential", "lmg_pair_potential_dev", "lmg_pair_potential_noalloc_dev", "lmg_potential_map", "lmg_potential_map_dev",
    "lmg_sim_load", "lmg_sim_store", "lmg_sim_step", "lmg_sim_set_sync", "lmg_sim_set_de_path",
    "lmg_sim_set_graphs", "lmg_batch_set_graphs", "lmg_sim_get_tuning", "lmg_batch_get_tuning",
    "lmg_sim_set_de_chunks",
    "lmg_sim_get_de_path", "lmg_sim_delta_e",
    "lmg_sim_device_error", "lmg_sim_step_timed",
    "lmg_batch_create", "lmg_batch_create_with_path", "lmg_batch_get_de_paths", "lmg_batch_tuning_report",
    "lmg_batch_destroy", "lmg_batch_last_error", "lmg_batch_size",
    "lmg_batch_step", "lmg_batch_step_timed", "lmg_batch_step_phases", "lmg_batch_delta_e", "lmg_batch_sync", "lmg_batch_store",
    "lmg_get_params",
    "lmg_surrogate_create", "lmg_surrogate_destroy", "lmg_surrogate_last_error", "lmg_surrogate_shape",
    "lmg_surrogate_get_weights", "lmg_surrogate_set_weights", "lmg_surrogate_occupancy_dev",
    "lmg_surrogate_infer_dev", "lmg_surrogate_gradient_dev", "lmg_surrogate_train_dev",
    "lmg_surrogate_map_dev", "lmg_surrogate_last_gate", "lmg_surrogate_set_graphs", "lmg_potential_map_if_dev",
    "lmg_surrogate_map_batch_dev", "lmg_surrogate_last_gates",
)
