"""librtg.so loads without a GPU and exports every entry point its headers
declare; calls that need a device fail cleanly (no compute calls here)."""
from __future__ import annotations

import os
import re

import raytrace_amd
from raytrace_amd import _capi

from conftest import ROOT


def declared(header):
    text = open(os.path.join(ROOT, "include", header)).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"^\s*(?:const\s+)?[\w\s\*]+?\b(rtg_\w+)\s*\(", text, flags=re.M)))


def test_every_declared_symbol_is_exported():
    lib = _capi.load()
    names = declared("rtg.h") + declared("rtg_host.h")
    assert len(names) >= 30
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    # and the Python binding covers exactly the declared set
    assert sorted(_capi.SIGNATURES) == sorted(names)


def test_abi_version_and_error_path():
    lib = _capi.load()
    assert lib.rtg_abi_version() == 1
    assert lib.rtg_scene_create(None, None) == _capi.RTG_E_INVALID
    assert "NULL" in _capi.last_error()
    assert lib.rtg_render(None, None, None, None, None, None) == _capi.RTG_E_INVALID


def test_tuning_values_are_validated_before_any_device_call():
    lib = _capi.load()
    ok = _capi.rtg_tuning(tile_w=16, grab=2, grid_pct=75)
    assert lib.rtg_set_tuning(None) == _capi.RTG_E_INVALID
    for bad in (dict(tile_w=6, grab=1, grid_pct=100), dict(tile_w=64, grab=1, grid_pct=100),
                dict(tile_w=8, grab=3, grid_pct=100), dict(tile_w=8, grab=1, grid_pct=24),
                dict(tile_w=8, grab=1, grid_pct=101)):
        t = _capi.rtg_tuning(**bad)
        assert lib.rtg_set_tuning(t) == _capi.RTG_E_INVALID, bad
    # a valid tuning gets as far as the device check (no rtg_init here)
    assert lib.rtg_set_tuning(ok) == _capi.RTG_E_STATE
    assert lib.rtg_autotune(None, 0, 3, None, None, 0, None) == _capi.RTG_E_INVALID


def test_package_imports_without_gpu():
    assert hasattr(raytrace_amd, "DeviceScene")
