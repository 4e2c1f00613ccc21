"""The C-ABI library loads and exports every entry point its header declares
(no device calls: runs without a GPU)."""

from __future__ import annotations

import ctypes
import re
from pathlib import Path

import pytest

from jaxrans_amd import device

ROOT = Path(__file__).resolve().parents[1]


def declared() -> list[str]:
    text = (ROOT / "include" / "jrhip.h").read_text()
    return sorted(set(re.findall(r"^\s*(?:int|const char\*)\s+(jr_\w+)\s*\(", text, re.M)))


def test_header_declares_the_binding_exports():
    assert set(declared()) == set(device.EXPORTS)


def test_library_exports_every_declared_symbol():
    if not device.LIB_PATH.is_file():
        pytest.skip("libjrhip.so not built (run __graft_entry__.build())")
    lib = ctypes.CDLL(str(device.LIB_PATH))
    missing = [name for name in declared() if not hasattr(lib, name)]
    assert not missing


def test_struct_layouts_match_the_header():
    # jr_grid_desc: two int32 then pointers; jr_numerics: doubles and int32s in declaration order
    assert device.GridDesc.area.offset == 8
    assert device.Numerics.flux.offset == 16
    assert device.Numerics.max_relative_update.offset == ctypes.sizeof(device.Numerics) - 8
