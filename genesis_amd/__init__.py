"""genesis_amd — MI355X-native force-and-integration path of GENESIS.

The compute path is libgenesis_gpu.so (hand-written HIP for gfx950) behind
the C ABI in include/genesis_gpu.h.  This package is the host-side mirror:
parameters (``params``), seeded workloads (``inputs``), the ctypes binding
(``_lib``), the resident-universe handle (``universe``) and the System
interface mirror (``systems``).
"""
from .params import Physics, GsParams, SYMBOLS, SPECIES_INDEX, fs_to_time  # noqa: F401
from .inputs import State, Workload  # noqa: F401

__all__ = ["Physics", "GsParams", "SYMBOLS", "SPECIES_INDEX", "fs_to_time", "State", "Workload",
           "GpuUniverse", "load_library"]


def load_library():
    from . import _lib
    return _lib.load()


def __getattr__(name):
    if name == "GpuUniverse":
        from .universe import GpuUniverse
        return GpuUniverse
    raise AttributeError(name)
