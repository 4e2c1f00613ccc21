"""Result files under the reference's names and columns (``Result.write``,
``src/jaxrans/simulation.py:135-160``; writers ``src/jaxrans/io/output.py:80-180``):

* ``<name>_metadata.json``: ``config`` (flow, numerics, implicit), ``git``,
  ``environment`` and ``extra`` with the reference's flat keys (``cl``,
  ``cd``, ``cd_friction``, ``cm``, ``bodies``, ``exit_reason``,
  ``orders_dropped``, ``iterations``, ``seconds``, ``cl_spread``, ``timing``,
  ``grid_quality``, ``cells``, ``grid``, ``case_setup``);
* ``<name>_residual.dat`` (iteration, residual) and ``<name>_lift.dat``
  (iteration, CL at each block end);
* ``<name>_surface.dat``: the wall distributions, one row per wall face
  (``post.surface``; the loads come from the device);
* ``<name>_fields.npz``: per block ``U_k`` and its nodes ``x_node_k``,
  ``y_node_k`` [, ``z_node_k``], and ``n_blocks``;
* 3D: ``<name>_block<k>.vts`` per block, a VTK XML structured grid (raw
  appended little-endian binary, UInt64 headers) with density, velocity,
  pressure and Mach number per cell.
"""

from __future__ import annotations

import json
import platform
import subprocess
import time
from dataclasses import asdict, is_dataclass
from pathlib import Path

import numpy as np


def _jsonable(v):
    if is_dataclass(v):
        return _jsonable(asdict(v))
    if isinstance(v, dict):
        return {str(k): _jsonable(x) for k, x in v.items()}
    if isinstance(v, (list, tuple)):
        return [_jsonable(x) for x in v]
    if isinstance(v, Path):
        return str(v)
    if isinstance(v, np.ndarray):
        return v.tolist()
    if isinstance(v, np.generic):
        return v.item()
    if isinstance(v, float) and not np.isfinite(v):
        return None if np.isnan(v) else ("inf" if v > 0 else "-inf")
    return v


def provenance() -> dict:
    """The commit this package was run from (``unknown`` outside a git
    checkout) and the host environment."""
    root = Path(__file__).resolve().parents[2]
    try:
        commit = subprocess.run(["git", "-C", str(root), "rev-parse", "HEAD"], capture_output=True, text=True,
                                timeout=10, check=True).stdout.strip()
        dirty = bool(subprocess.run(["git", "-C", str(root), "status", "--porcelain"], capture_output=True,
                                    text=True, timeout=10, check=True).stdout.strip())
        git = {"commit": commit, "dirty": dirty}
    except (subprocess.SubprocessError, OSError):
        git = {"commit": "unknown", "dirty": False}
    env = {"python": platform.python_version(), "platform": platform.platform(), "numpy": np.__version__,
           "timestamp": time.strftime("%Y-%m-%dT%H:%M:%S%z"), "device": "HIP (libjrhip.so)"}
    return {"git": git, "environment": env}


def primitive_cells(U: np.ndarray, gamma: float) -> dict:
    """Density, velocity ``(3, ...)``, pressure and Mach number of a 3D
    block's conservative state ``(nvar, ...)``."""
    rho = U[0]
    vel = U[1:4] / rho
    p = (gamma - 1.0) * (U[4] - 0.5 * rho * np.sum(vel * vel, axis=0))
    mach = np.linalg.norm(vel, axis=0) / np.sqrt(gamma * p / rho)
    return {"density": rho, "velocity": vel, "pressure": p, "mach": mach}


def write_vts(path: Path, nodes: np.ndarray, cells: dict) -> Path:
    """One structured block for ParaView: nodes ``(3, ni+1, nj+1, nk+1)``,
    cell fields ``(ni, nj, nk)`` or vectors ``(3, ni, nj, nk)``; VTK orders
    points and cells with ``i`` fastest."""
    n = nodes.shape[1:]
    extent = f"0 {n[0] - 1} 0 {n[1] - 1} 0 {n[2] - 1}"
    chunks, tags, offset = [], [], 0

    def put(name, arr, comps):
        nonlocal offset
        raw = np.ascontiguousarray(arr, dtype="<f8").tobytes()
        chunks.append(np.array([len(raw)], "<u8").tobytes() + raw)
        tags.append(f'<DataArray type="Float64" Name="{name}" NumberOfComponents="{comps}" '
                    f'format="appended" offset="{offset}"/>')
        offset += 8 + len(raw)

    def i_fastest(a):
        if a.ndim == 3:
            return np.transpose(a, (2, 1, 0)).ravel()
        return np.stack([np.transpose(c, (2, 1, 0)).ravel() for c in a], axis=1).ravel()

    put("Points", i_fastest(nodes), 3)
    for name, a in cells.items():
        a = np.asarray(a, np.float64)
        put(name, i_fastest(a), 3 if a.ndim == 4 else 1)
    head = ('<?xml version="1.0"?>\n<VTKFile type="StructuredGrid" version="1.0" byte_order="LittleEndian" '
            f'header_type="UInt64"><StructuredGrid WholeExtent="{extent}"><Piece Extent="{extent}">'
            f"<Points>{tags[0]}</Points><CellData>{''.join(tags[1:])}</CellData>"
            '</Piece></StructuredGrid><AppendedData encoding="raw">_')
    with open(path, "wb") as f:
        f.write(head.encode())
        for c in chunks:
            f.write(c)
        f.write(b"</AppendedData></VTKFile>\n")
    return path


def _block_nodes(block) -> list[np.ndarray]:
    if block.dim == 3:
        return [np.asarray(c) for c in block.nodes]
    return [np.asarray(block.x_node), np.asarray(block.y_node)]


def write_result(result, directory: str | Path | None = None, extra: dict | None = None) -> Path:
    """Write ``result``'s files to ``directory`` (default ``<case output>/<case
    name>``); returns it.  The wall distributions need the device (their
    loads are computed there)."""
    from .mesh.quality import quality
    case = result.case
    directory = Path(directory) if directory is not None else Path(case.output) / case.name
    directory.mkdir(parents=True, exist_ok=True)
    name = case.name
    f = result.forces
    spec = case.grid_spec
    payload = {"config": {"flow": _jsonable(case.flow), "numerics": _jsonable(case.numerics),
                          "implicit": _jsonable(case.implicit)},
               **provenance(),
               "extra": _jsonable({"case_setup": case.raw, "grid": spec.describe() if spec is not None else None,
                                   "cells": result.cells, "grid_quality": quality(result.grid.mesh).record(),
                                   "bodies": [dict(zip(("cl", "cd", "cm"), b)) for b in f.bodies],
                                   "cl": f.cl, "cd": f.cd, "cd_friction": f.cd_friction, "cm": f.cm,
                                   "exit_reason": result.history.reason,
                                   "orders_dropped": result.history.orders_dropped,
                                   "iterations": result.history.n_iters, "seconds": result.seconds,
                                   "cl_spread": result.cl_spread, "timing": result.timing.record(),
                                   **({"grid_sequencing": list(result.levels)} if result.levels else {}),
                                   **(extra or {})})}
    (directory / f"{name}_metadata.json").write_text(json.dumps(payload, indent=2, sort_keys=True))
    residual = np.asarray(result.history.residual, np.float64).reshape(-1)
    np.savetxt(directory / f"{name}_residual.dat", np.column_stack([np.arange(1, residual.size + 1), residual]),
               header="iteration residual_l2", fmt=["%d", "%.12e"])
    cl = np.asarray(result.cl_history, np.float64).reshape(-1)
    np.savetxt(directory / f"{name}_lift.dat", np.column_stack([case.block_iters * np.arange(1, cl.size + 1), cl]),
               header="iteration cl", fmt=["%d", "%.12e"])
    surf = result.surface()
    columns = {"x": "x_over_c", "y": "y_over_c", "z": "z_over_c"}
    np.savetxt(directory / f"{name}_surface.dat", np.column_stack([np.asarray(v) for v in surf.values()]),
               header=" ".join(columns.get(k, k) for k in surf), fmt="%.12e")
    grid = result.grid
    blocks = grid.unpack(np.asarray(result.U))
    arrays = {"n_blocks": np.asarray(len(grid.mesh))}
    for k, (U, block) in enumerate(zip(blocks, grid.mesh)):
        arrays[f"U_{k}"] = U
        for axis, c in zip("xyz", _block_nodes(block)):
            arrays[f"{axis}_node_{k}"] = c
    np.savez_compressed(directory / f"{name}_fields.npz", **arrays)
    if grid.dim == 3:
        for k, (U, block) in enumerate(zip(blocks, grid.mesh)):
            write_vts(directory / f"{name}_block{k}.vts", np.asarray(block.nodes),
                      primitive_cells(U, case.flow.gamma))
    return directory
