"""MI355X-native implementation of Langmuir's data-parallel hot path.

The device work lives in ``liblangmuir_gpu.so`` (HIP for gfx950, C-ABI in
``include/langmuir_gpu.h``); this package is the host-side mirror of the
reference's interfaces over that C-ABI (see :mod:`langmuir_amd.host`).
"""
from .host import (Batch, Context, GpuHelper, Simulation, SimulationParameters, background_potential,
                   coupling_table, device_count, drain_slots, flux_rates, make_config,
                   mt19937_seed_window)
from ._abi import LMG_GAUSS, LMG_IMAGE_LEFT, LMG_MAP_CULLED, LMG_MAP_PAIRS, LMG_NO_CUTOFF, LmgError
from .device import DeviceBuffer
from .surrogate import SnapshotSet, Surrogate

__all__ = [
    "Batch", "Context", "GpuHelper", "Simulation", "SimulationParameters", "background_potential",
    "coupling_table", "device_count", "drain_slots", "flux_rates", "make_config",
    "mt19937_seed_window", "LMG_GAUSS", "LMG_IMAGE_LEFT", "LMG_NO_CUTOFF", "LMG_MAP_PAIRS", "LMG_MAP_CULLED", "LmgError",
    "DeviceBuffer", "SnapshotSet", "Surrogate",
]
