"""ctypes binding of libgenesis_gpu.so (the C ABI of include/genesis_gpu.h).

There is no CPU fallback: if the shared library is missing or cannot be
loaded, every entry point raises ``GpuUnavailable``.
"""
from __future__ import annotations

import ctypes
import os
import re
from typing import List

from .params import GsParams

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libgenesis_gpu.so")
HEADER = os.path.join(os.path.dirname(HERE), "include", "genesis_gpu.h")

GS_OK = 0
GS_E_INVALID = -1
GS_E_HIP = -2
GS_E_STATE = -3
GS_E_OVERFLOW = -4
GS_E_NONFINITE = -5


class GpuUnavailable(RuntimeError):
    pass


class GsError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"genesis_gpu error {code}: {msg}")
        self.code = code


class Observables(ctypes.Structure):
    _fields_ = [("kinetic", ctypes.c_double), ("potential", ctypes.c_double), ("bond_energy", ctypes.c_double),
                ("temperature", ctypes.c_double), ("lambda_", ctypes.c_double), ("energy_total", ctypes.c_double)]


class ClusterStats(ctypes.Structure):
    _fields_ = [("monomers", ctypes.c_int64), ("aggregates", ctypes.c_int64), ("largest", ctypes.c_int64),
                ("mean_size", ctypes.c_double), ("bound_pairs", ctypes.c_int64)]


class Timing(ctypes.Structure):
    _fields_ = [("steps", ctypes.c_int64), ("step_ms", ctypes.c_double), ("rebuild_ms", ctypes.c_double),
                ("pair_ms", ctypes.c_double), ("kick_ms", ctypes.c_double), ("pairs_in_range", ctypes.c_int64)]


class Tuning(ctypes.Structure):
    _fields_ = [("pair_mode", ctypes.c_int32), ("lanes_per_atom", ctypes.c_int32), ("skin", ctypes.c_double),
                ("sort_every", ctypes.c_int64), ("list_builds", ctypes.c_int64),
                ("list_replays", ctypes.c_int64), ("ms_per_tick", ctypes.c_double)]


class BondStats(ctypes.Structure):
    _fields_ = [("tick", ctypes.c_int64), ("bonded_pairs", ctypes.c_int64), ("bonded_entities", ctypes.c_int64),
                ("mean_coordination", ctypes.c_double), ("bonds_formed", ctypes.c_int64),
                ("lifetime_sum_ticks", ctypes.c_double), ("species_matrix", ctypes.c_int64 * 100)]


class GrSpec(ctypes.Structure):
    _fields_ = [("nbins", ctypes.c_int32), ("pad", ctypes.c_int32), ("r_max", ctypes.c_double),
                ("sigma_b", ctypes.c_double)]


_D = ctypes.POINTER(ctypes.c_double)
_U8 = ctypes.POINTER(ctypes.c_uint8)
_U32 = ctypes.POINTER(ctypes.c_uint32)
_U64 = ctypes.POINTER(ctypes.c_uint64)
_I32 = ctypes.POINTER(ctypes.c_int32)
_H = ctypes.c_void_p

_SIGNATURES = {
    "gs_create": (ctypes.c_int, [ctypes.POINTER(GsParams), ctypes.c_int64, ctypes.POINTER(_H)]),
    "gs_destroy": (ctypes.c_int, [_H]),
    "gs_last_error": (ctypes.c_char_p, []),
    "gs_device_info": (ctypes.c_int, [ctypes.POINTER(ctypes.c_int), ctypes.c_char_p, ctypes.c_int]),
    "gs_upload_state": (ctypes.c_int, [_H] + [_D] * 9 + [_U8, _D]),
    "gs_download_state": (ctypes.c_int, [_H] + [_D] * 9),
    "gs_forces": (ctypes.c_int, [_H, _D, _D]),
    "gs_download_force_extras": (ctypes.c_int, [_H, _D, _D]),
    "gs_download_cells": (ctypes.c_int, [_H, _U32, _U32]),
    "gs_half_kick": (ctypes.c_int, [_H]),
    "gs_drift": (ctypes.c_int, [_H]),
    "gs_wrap": (ctypes.c_int, [_H]),
    "gs_thermostat": (ctypes.c_int, [_H, _D]),
    "gs_kinetic": (ctypes.c_int, [_H, _D]),
    "gs_speed_stats": (ctypes.c_int, [_H, _D, _D]),
    "gs_step": (ctypes.c_int, [_H, ctypes.c_int64, ctypes.POINTER(Observables)]),
    "gs_radial_distribution": (ctypes.c_int, [_H, ctypes.c_double, ctypes.c_int32, _U64, _D, _D, _D]),
    "gs_radial_distribution_between": (ctypes.c_int, [_H, ctypes.c_int32, ctypes.c_int32, ctypes.c_double,
                                                      ctypes.c_int32, _U64, _D, _D, _D]),
    "gs_radial_distribution_full": (ctypes.c_int, [_H, ctypes.c_int32, ctypes.c_int32, ctypes.c_double,
                                                   ctypes.c_int32, _U64, _D, _D, _D]),
    "gs_get_tuning": (ctypes.c_int, [_H, ctypes.POINTER(Tuning)]),
    "gs_set_tuning": (ctypes.c_int, [_H, ctypes.POINTER(Tuning)]),
    "gs_set_bound_search": (ctypes.c_int, [_H, ctypes.c_int32]),
    "gs_autotune": (ctypes.c_int, [_H, ctypes.c_int32, ctypes.POINTER(Tuning)]),
    "gs_hvp": (ctypes.c_int, [_H, _D, _U8, _D, _D, _D, ctypes.c_double, _D, _D, _D, _D]),
    "gs_adjoint_forward": (ctypes.c_int, [_H, _D, _U8, ctypes.c_int64, ctypes.c_int64, ctypes.c_int64,
                                          ctypes.POINTER(GrSpec), _D, _D, _D]),
    "gs_adjoint_backward": (ctypes.c_int, [_H, _D]),
    "gs_batch_create": (ctypes.c_int, [ctypes.c_int32, ctypes.POINTER(GsParams), ctypes.POINTER(ctypes.c_int64),
                                       ctypes.POINTER(_H)]),
    "gs_batch_destroy": (ctypes.c_int, [_H]),
    "gs_batch_upload": (ctypes.c_int, [_H, ctypes.c_int32] + [_D] * 9 + [_U8, _D]),
    "gs_batch_download": (ctypes.c_int, [_H, ctypes.c_int32] + [_D] * 9),
    "gs_batch_step": (ctypes.c_int, [_H, ctypes.c_int64, ctypes.c_int32, ctypes.POINTER(Observables)]),
    "gs_batch_autotune": (ctypes.c_int, [_H, ctypes.c_int32, _D, _D]),
    "gs_batch_get_tuning": (ctypes.c_int, [_H, ctypes.POINTER(Tuning)]),
    "gs_batch_set_tuning": (ctypes.c_int, [_H, ctypes.POINTER(Tuning)]),
    "gs_timing_enable": (ctypes.c_int, [_H, ctypes.c_int, ctypes.c_int]),
    "gs_timing_read": (ctypes.c_int, [_H, ctypes.POINTER(Timing)]),
    "gs_clusters": (ctypes.c_int, [_H, ctypes.c_double, _U32, ctypes.POINTER(ClusterStats)]),
    "gs_bond_observe": (ctypes.c_int, [_H, ctypes.c_int64]),
    "gs_bond_stats_read": (ctypes.c_int, [_H, ctypes.POINTER(BondStats)]),
    "gs_download_bonds": (ctypes.c_int, [_H, _I32, _I32, ctypes.c_int64, ctypes.POINTER(ctypes.c_int64)]),
    "gs_upload_bonds": (ctypes.c_int, [_H, _I32, _I32]),
    "gs_bound_pairs": (ctypes.c_int, [_H, _U32, _U32, ctypes.c_int64, ctypes.POINTER(ctypes.c_int64)]),
    "gs_bond_components": (ctypes.c_int, [_H, _U32, ctypes.POINTER(ctypes.c_int64)]),
    "gs_bond_binding": (ctypes.c_int, [_H, _D]),
    "gs_velocity_histogram": (ctypes.c_int, [_H, ctypes.c_double, ctypes.c_int32, _U64,
                                             ctypes.POINTER(ctypes.c_uint64), ctypes.POINTER(ctypes.c_uint64), _D]),
    "gs_bond_reset": (ctypes.c_int, [_H]),
    "gs_snapshot_append": (ctypes.c_int, [_H, ctypes.c_char_p]),
    "gs_snapshot_load": (ctypes.c_int, [_H, ctypes.c_char_p, ctypes.c_int64]),
    "gs_set_tick": (ctypes.c_int, [_H, ctypes.c_int64]),
    "gs_get_tick": (ctypes.c_int, [_H, ctypes.POINTER(ctypes.c_int64)]),
    "gs_radial_distribution_noalloc": (ctypes.c_int, [_H, ctypes.c_double, ctypes.c_int32, ctypes.c_void_p, _D, _D]),
}

_lib = None


def header_symbols() -> List[str]:
    """Every function name declared in include/genesis_gpu.h."""
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(gs_[a-z0-9_]+)\s*\(", text)))


def load(path: str = LIB_PATH):
    """Load the library (once).  Raises GpuUnavailable when it is absent."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise GpuUnavailable(f"{path} not built: run __graft_entry__.build() (hipcc, gfx950)")
    try:
        lib = ctypes.CDLL(path)
    except OSError as e:
        raise GpuUnavailable(f"cannot load {path}: {e}") from e
    for name, (res, args) in _SIGNATURES.items():
        if hasattr(lib, name):
            fn = getattr(lib, name)
            fn.restype = res
            fn.argtypes = args
    _lib = lib
    return lib


def check(code: int) -> None:
    if code != GS_OK:
        msg = _lib.gs_last_error().decode() if _lib is not None else "?"
        raise GsError(code, msg)
