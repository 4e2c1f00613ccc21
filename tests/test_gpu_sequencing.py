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


def test_result_write_files(tmp_path):
    """Result.write: the reference's file names (io/output.py:80-135) with
    the run's own numbers in them."""
    import json

    import numpy as np

    from jaxrans_amd import load_case, run
    result = run(load_case(CASE, ("numerical.convergence.max_iters=40",)))
    out = result.write(tmp_path / "out")
    name = result.case.name
    meta = json.loads((out / f"{name}_metadata.json").read_text())
    assert meta["config"]["flow"]["mach"] == result.case.flow.mach and meta["extra"]["cells"] == result.cells
    res = np.loadtxt(out / f"{name}_residual.dat")
    assert res.shape == (40, 2) and np.allclose(res[:, 1], result.history.residual, rtol=1e-11, atol=0)
    lift = np.loadtxt(out / f"{name}_lift.dat", ndmin=2)
    assert np.allclose(lift[:, 1], result.cl_history, rtol=1e-11, atol=0)
    with np.load(out / f"{name}_fields.npz") as f:
        assert int(f["n_blocks"]) == 1
        assert np.array_equal(f["U_0"], result.grid.unpack(result.U)[0])
        assert f["x_node_0"].shape == (f["U_0"].shape[1] + 1, f["U_0"].shape[2] + 1)


def test_sequenced_3d_run_converges_to_the_unsequenced_forces():
    """The M6 file case (SA, a pole and an O seam) with one coarse level:
    the coarse 3D grid runs first and its injected solution starts the fine
    run, which converges to the forces of the march from the freestream."""
    from jaxrans_amd import load_case, run
    m6 = CASE.parent / "onera_m6_sa"
    over = ("numerical.convergence.forces_tol=null", "numerical.convergence.tol=1e-9",
            "numerical.convergence.max_iters=3000")
    plain = run(load_case(m6, over))
    seq = run(load_case(m6, over + ("numerical.convergence.grid_sequencing=1",)))
    print(f"[gpu] 3D sequencing: levels {seq.levels}; fine {seq.history.summary()} vs plain "
          f"{plain.history.summary()}", flush=True)
    assert len(seq.levels) == 1 and seq.levels[0]["cells"] * 8 == seq.cells
    assert seq.history.reason == "converged" and plain.history.reason == "converged"
    for out in ("cl", "cd", "cm"):
        assert abs(getattr(seq.forces, out) - getattr(plain.forces, out)) < 1e-7, out
