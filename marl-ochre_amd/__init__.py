"""marl-ochre_amd: the MI355X (gfx950) per-tick field-and-cell path of MARL.

The compute lives in libmarl_hip.so (hand-written HIP kernels behind the C ABI
of include/marl_hip.h); this package is the thin Python host side over it.
The directory name is not a Python identifier, so callers load it with
importlib under the name `marl_ochre_amd` (see __graft_entry__.load_package).
"""
from ._abi import (CELL_DTYPE, RULESET_DTYPE, BufferMap, DeviceRequest, GpuError, InvalidInput, NoAdapter,
                   load_library)
from .config import SimulationConfig
from .engine import GpuFieldDiffuser, MarlEnsemble, MarlGpu
from . import output

__all__ = [
    "CELL_DTYPE", "RULESET_DTYPE", "GpuError", "NoAdapter", "DeviceRequest", "BufferMap", "InvalidInput",
    "load_library", "SimulationConfig", "GpuFieldDiffuser", "MarlGpu", "MarlEnsemble", "output",
]
