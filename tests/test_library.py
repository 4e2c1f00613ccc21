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


def test_struct_layouts_match_the_header(tmp_path):
    """Every field of the ctypes mirrors sits at the offset the C compiler
    gives it in include/jrhip.h, and the structs have the same size."""
    import shutil
    import subprocess
    if shutil.which("gcc") is None:
        pytest.skip("gcc not available")
    structs = {"jr_grid_desc": device.GridDesc, "jr_grid_desc3": device.GridDesc3, "jr_numerics": device.Numerics}
    lines = ['#include <stdio.h>', '#include <stddef.h>', f'#include "{ROOT / "include" / "jrhip.h"}"', "int main(void) {"]
    for cname, py in structs.items():
        lines.append(f'printf("{cname} sizeof %zu\\n", sizeof({cname}));')
        for field, *_ in py._fields_:
            lines.append(f'printf("{cname} {field} %zu\\n", offsetof({cname}, {field}));')
    lines.append("return 0; }")
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines))
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-o", str(exe), str(src)], check=True)
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout.split("\n")
    got = {(a, b): int(c) for a, b, c in (line.split() for line in out if line)}
    for cname, py in structs.items():
        assert got[(cname, "sizeof")] == ctypes.sizeof(py), cname
        for field, *_ in py._fields_:
            assert got[(cname, field)] == getattr(py, field).offset, (cname, field)
