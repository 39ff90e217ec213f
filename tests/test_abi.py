"""The C-ABI library loads and exports every symbol the header declares
(no compute calls: this runs without a GPU)."""
import ctypes
import os
import subprocess

import pytest

from genesis_amd import _lib
from genesis_amd.params import GsParams, Physics


def test_library_exports_every_header_symbol():
    lib = _lib.load()
    syms = _lib.header_symbols()
    assert len(syms) >= 15
    missing = [s for s in syms if not hasattr(lib, s)]
    assert not missing, missing


def test_binding_covers_every_header_symbol():
    assert set(_lib.header_symbols()) <= set(_lib._SIGNATURES), set(_lib.header_symbols()) - set(_lib._SIGNATURES)


def test_params_struct_layout_matches_header():
    # 4 (box+dt) + 40 (tables) + 10 scalars doubles + 8 int32
    assert ctypes.sizeof(GsParams) == (4 + 40 + 10) * 8 + 8 * 4
    p = Physics().to_c()
    assert p.lj_cutoff_factor == 2.5 and p.enable_boundaries == 1


def test_library_is_built_for_gfx950():
    # The offload bundle of the library built here names its target.
    blob = open(_lib.LIB_PATH, "rb").read()
    assert b"amdgcn-amd-amdhsa--gfx950" in blob or b"gfx950" in blob


def test_missing_library_fails_loudly(tmp_path):
    saved = _lib._lib
    _lib._lib = None
    try:
        with pytest.raises(_lib.GpuUnavailable):
            _lib.load(str(tmp_path / "nope.so"))
    finally:
        _lib._lib = saved
