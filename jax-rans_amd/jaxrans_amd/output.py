"""Result files (``src/jaxrans/io/output.py:54-135``, ``Result.write`` at
``src/jaxrans/simulation.py:135-160``): the resolved configs as JSON, the
residual and lift histories as two-column text, and the per-block fields with
their nodes as a compressed ``.npz``, under the reference's file names.  The
wall distributions (``_surface.dat``) and the grid-quality record are not
written by this implementation."""

from __future__ import annotations

import json
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
    return v


def write_result(result, directory: str | Path | None = None, extra: dict | None = None) -> Path:
    """Write ``result``'s metadata, residual and lift histories and fields to
    ``directory`` (default ``<case output>/<case name>``); returns it."""
    case = result.case
    directory = Path(directory) if directory is not None else Path(case.output) / case.name
    directory.mkdir(parents=True, exist_ok=True)
    name = case.name
    payload = {"config": {"flow": _jsonable(case.flow), "numerics": _jsonable(case.numerics),
                          "implicit": _jsonable(case.implicit)},
               "extra": _jsonable({"case_setup": case.raw, "cells": result.cells,
                                   "forces": result.forces, "exit_reason": result.history.reason,
                                   "iterations": result.history.n_iters, "seconds": result.seconds,
                                   **({"grid_sequencing": list(result.levels)} if result.levels else {}),
                                   **(extra or {})})}
    (directory / f"{name}_metadata.json").write_text(json.dumps(payload, indent=2, sort_keys=True))
    residual = np.asarray(result.history.residual, np.float64).reshape(-1)
    np.savetxt(directory / f"{name}_residual.dat", np.column_stack([np.arange(1, residual.size + 1), residual]),
               header="iteration residual_l2", fmt=["%d", "%.12e"])
    cl = np.asarray(result.cl_history, np.float64).reshape(-1)
    np.savetxt(directory / f"{name}_lift.dat", np.column_stack([case.block_iters * np.arange(1, cl.size + 1), cl]),
               header="iteration cl", fmt=["%d", "%.12e"])
    grid = result.grid
    if hasattr(grid, "unpack") and hasattr(grid, "mesh") and grid.mesh and grid.mesh[0].dim == 2:
        arrays = {"n_blocks": np.asarray(len(grid.mesh))}
        for k, (U, block) in enumerate(zip(grid.unpack(result.U), grid.mesh)):
            arrays[f"U_{k}"] = U
            arrays[f"x_node_{k}"] = np.asarray(block.x_node)
            arrays[f"y_node_{k}"] = np.asarray(block.y_node)
    else:   # 3D: the whole canvas (cells in the canvas layout of the run)
        arrays = {"n_blocks": np.asarray(0), "U_canvas": np.asarray(result.U)}
    np.savez_compressed(directory / f"{name}_fields.npz", **arrays)
    return directory
