"""Grid sequencing through run() (``src/jaxrans/simulation.py:230-240``): the
coarse grid first, its solution injected into the fine start; the fine grid
converges to the same forces as a march from the freestream."""

from __future__ import annotations

from pathlib import Path

import pytest

pytestmark = pytest.mark.gpu

CASE = Path(__file__).resolve().parent / "golden" / "cases" / "euler_o48"


def test_sequenced_run_converges_to_the_unsequenced_forces():
    from jaxrans_amd import load_case, run
    plain = run(load_case(CASE))
    seen = []
    seq = run(load_case(CASE, ("numerical.convergence.grid_sequencing=1",)),
              on_level=lambda level, grid: seen.append((level, grid.cells)))
    print(f"[gpu] sequencing: levels {seq.levels}; fine {seq.history.summary()} vs plain "
          f"{plain.history.summary()}", flush=True)
    assert [lvl for lvl, _ in seen] == [1, 0] and seen[0][1] * 4 == seen[1][1]
    assert len(seq.levels) == 1 and seq.levels[0]["exit_reason"] == "converged"
    assert seq.history.reason == "converged" and plain.history.reason == "converged"
    for out in ("cl", "cd", "cm"):
        assert abs(getattr(seq.forces, out) - getattr(plain.forces, out)) < 1e-8, out
