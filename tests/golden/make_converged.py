"""Writes tests/golden/euler_o48_converged.json: the oracle's converged forces
on the small Euler case tests/golden/cases/euler_o48 (FP64, to a 1e-10 drop of
the residual from its peak), the reference point of the device run() parity
test.  Run from the repository root:  python tests/golden/make_converged.py
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


def main() -> None:
    case = load_case(ROOT / "tests/golden/cases/euler_o48")
    atlas = case.grid()
    setup = OR.setup_for(atlas, case.flow, case.numerics)
    s, hist, f = OS.run_oracle(setup, case.implicit, tol=case.numerics.tol, max_iters=case.numerics.max_iters,
                               block_iters=case.block_iters)
    out = {"case": "tests/golden/cases/euler_o48", "cells": atlas.cells, "iterations": int(hist.size),
           "residual_first": float(hist[0]), "residual_peak": float(hist.max()), "residual_last": float(hist[-1]),
           "cl": float(f["cl"]), "cd": float(f["cd"]), "cm": float(f["cm"])}
    path = ROOT / "tests/golden/euler_o48_converged.json"
    path.write_text(json.dumps(out, indent=2) + "\n")
    print(json.dumps(out, indent=2))


if __name__ == "__main__":
    main()
