"""Writes the oracle's converged forces of a case (FP64, to the case's own
residual drop ``convergence.tol`` from its peak), the reference point of the
device run() parity tests:

    python tests/golden/make_converged.py                        # the small Euler case
    python tests/golden/make_converged.py cases/naca0012_o_euler  # the 128 x 64 Euler O-grid

Output: ``tests/golden/<case name>_converged.json``.  Run from the
repository root.
"""

from __future__ import annotations

import json
import sys
from pathlib import Path

ROOT = Path(__file__).resolve().parents[2]
sys.path[:0] = [str(ROOT / "jax-rans_amd"), str(ROOT)]

from jaxrans_amd import load_case  # noqa: E402
from oracle import residual as OR  # noqa: E402
from oracle import solver as OS  # noqa: E402


def main(case_dir: str = "tests/golden/cases/euler_o48") -> None:
    case = load_case(ROOT / case_dir)
    atlas = case.grid()
    setup = OR.setup_for(atlas, case.flow, case.numerics)
    s, hist, f = OS.run_oracle(setup, case.implicit, tol=case.numerics.tol, max_iters=case.numerics.max_iters,
                               block_iters=case.block_iters)
    out = {"case": case_dir, "cells": atlas.cells, "iterations": int(hist.size),
           "residual_first": float(hist[0]), "residual_peak": float(hist.max()), "residual_last": float(hist[-1]),
           "cl": float(f["cl"]), "cd": float(f["cd"]), "cm": float(f["cm"])}
    path = ROOT / "tests/golden" / f"{case.name}_converged.json"
    path.write_text(json.dumps(out, indent=2) + "\n")
    print(json.dumps(out, indent=2))


if __name__ == "__main__":
    main(*sys.argv[1:])
