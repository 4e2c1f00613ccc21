"""Result.write on the device (``src/jaxrans/simulation.py:135-160``,
``io/output.py:80-180``): the reference's files and metadata keys, and wall
distributions whose per-face loads integrate back to the device's force
coefficients (to the 13 digits the files hold; the loads are the force kernel's own
per-face terms, reordered and normalised on the host)."""

from __future__ import annotations

import json
import re
from pathlib import Path

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

CASES = Path(__file__).resolve().parent / "golden" / "cases"
FLAT = ("cl", "cd", "cd_friction", "cm", "bodies", "exit_reason", "orders_dropped", "iterations", "seconds",
        "cl_spread", "timing", "grid_quality", "cells", "grid", "case_setup")


def _meta_checks(out, result):
    name = result.case.name
    meta = json.loads((out / f"{name}_metadata.json").read_text())
    extra = meta["extra"]
    assert set(FLAT) <= set(extra), set(FLAT) - set(extra)
    assert extra["cl"] == result.forces.cl and extra["cd"] == result.forces.cd and extra["cm"] == result.forces.cm
    assert extra["iterations"] == result.history.n_iters and extra["cells"] == result.cells
    assert extra["grid_quality"]["cells"] == result.cells and extra["grid_quality"]["min_area"] > 0
    assert "commit" in meta["git"] and "python" in meta["environment"]
    return meta


def test_write_2d_euler_surface_integrates_to_the_forces(tmp_path):
    from jaxrans_amd import load_case, run
    from jaxrans_amd.post.surface import face_vectors_2d, loop_order
    print("[gpu] output 2D Euler", flush=True)
    result = run(load_case(CASES / "euler_o48", ("numerical.convergence.max_iters=40",)))
    out = result.write(tmp_path / "out")
    name = result.case.name
    _meta_checks(out, result)
    surf = np.loadtxt(out / f"{name}_surface.dat")
    header = (out / f"{name}_surface.dat").read_text().splitlines()[0]
    assert header.split()[1:] == ["x_over_c", "y_over_c", "body", "cp"]
    at = result.grid
    assert surf.shape == (at.walls.n, 4)
    sx, sy = face_vectors_2d(at, loop_order(at.walls))
    a = np.deg2rad(result.case.flow.alpha_deg)
    cp, x, y = surf[:, 3], surf[:, 0], surf[:, 1]
    fx, fy = -np.sum(cp * sx), -np.sum(cp * sy)             # per dynamic pressure and unit chord
    # the file holds 13 significant digits
    assert abs((fy * np.cos(a) - fx * np.sin(a)) - result.forces.cl) < 1e-10
    assert abs((fx * np.cos(a) + fy * np.sin(a)) - result.forces.cd) < 1e-10
    cm = np.sum(y * (-cp * sx) - (x - 0.25) * (-cp * sy))
    assert abs(cm - result.forces.cm) < 1e-11
    with np.load(out / f"{name}_fields.npz") as f:
        assert int(f["n_blocks"]) == 1 and np.array_equal(f["U_0"], at.unpack(result.U)[0])


def test_write_2d_laminar_friction(tmp_path):
    """Laminar O-grid, Re 500: the per-face viscous loads sum to the device's
    friction drag, and the skin friction is positive over most of the wall
    (attached flow at 1.25 deg)."""
    from jaxrans_amd import load_case, run
    print("[gpu] output 2D laminar", flush=True)
    d = tmp_path / "laminar"
    d.mkdir()
    doc = json.loads((CASES / "euler_o48" / "case.json").read_text())
    doc["general"]["case_name"] = "laminar_o48"
    doc["boundary_conditions"]["wall"]["type"] = "no_slip_wall"
    doc["freestream"]["reynolds"] = 500.0
    (d / "case.json").write_text(json.dumps(doc))
    (d / "numerical.json").write_text((CASES / "euler_o48" / "numerical.json").read_text())
    case = load_case(d, ("numerical.convergence.max_iters=60",))
    result = run(case)
    from jaxrans_amd.simulation import make_context, freestream_of
    ctx = make_context(case, result.grid)
    q, mu = freestream_of(case.flow)
    ctx.set_freestream(q, mu)
    ctx.set_state(result.U[None])
    loads = ctx.wall_loads()[0]
    ctx.close()
    a = np.deg2rad(case.flow.alpha_deg)
    qd = 0.5 * case.flow.mach ** 2
    cdf = (np.cos(a) * loads[1].sum() + np.sin(a) * loads[2].sum()) / qd
    assert abs(cdf - result.forces.cd_friction) < 1e-12 * max(1.0, abs(cdf))
    out = result.write(tmp_path / "lam")
    surf = np.loadtxt(out / f"{case.name}_surface.dat")
    assert surf.shape[1] == 5 and np.mean(surf[:, 4] > 0) > 0.8


def _read_vts(path):
    raw = path.read_bytes()
    head, data = raw.split(b'<AppendedData encoding="raw">_', 1)
    names = re.findall(rb'Name="(\w+)"', head)
    return head.decode(), [n.decode() for n in names], len(data)


def test_write_3d_files_and_surface(tmp_path):
    from jaxrans_amd import load_case, run
    print("[gpu] output 3D M6", flush=True)
    case = load_case(CASES / "onera_m6_sa", ("numerical.convergence.max_iters=8",
                                             "numerical.convergence.block_iters=4"))
    result = run(case)
    out = result.write(tmp_path / "m6")
    name = case.name
    _meta_checks(out, result)
    at = result.grid
    with np.load(out / f"{name}_fields.npz") as f:
        nb = int(f["n_blocks"])
        assert nb == len(at.mesh)
        for k, (U, blk) in enumerate(zip(at.unpack(result.U), at.mesh)):
            assert np.array_equal(f[f"U_{k}"], U)
            assert f[f"z_node_{k}"].shape == tuple(n + 1 for n in blk.shape)
    for k, blk in enumerate(at.mesh):
        head, names, size = _read_vts(out / f"{name}_block{k}.vts")
        ni, nj, nk = blk.shape
        assert f'WholeExtent="0 {ni} 0 {nj} 0 {nk}"' in head
        assert names == ["Points", "density", "velocity", "pressure", "mach"]
        n_pts, n_cells = (ni + 1) * (nj + 1) * (nk + 1), ni * nj * nk
        assert size == 5 * 8 + 8 * (3 * n_pts + 6 * n_cells) + len(b"</AppendedData></VTKFile>\n")
    surf = np.loadtxt(out / f"{name}_surface.dat")
    cols = (out / f"{name}_surface.dat").read_text().splitlines()[0].split()[1:]
    assert cols == ["block", "i", "k", "x_over_c", "y_over_c", "z_over_c", "cp", "cf_x", "cf_y", "cf_z"]
    assert surf.shape[0] == at.walls.n
    from jaxrans_amd.post.surface import w_face_vectors
    S = w_face_vectors(at) * at.walls.into_fluid
    spec = case.grid_spec
    a = np.deg2rad(case.flow.alpha_deg)
    fp = -np.sum(surf[:, 6] * S, axis=1) / spec.reference_area
    area = np.linalg.norm(S, axis=0)
    ff = np.sum(surf[:, 7:10].T * area, axis=1) / spec.reference_area
    total = fp + ff
    assert abs((total[2] * np.cos(a) - total[0] * np.sin(a)) - result.forces.cl) < 1e-10
    assert abs((total[0] * np.cos(a) + total[2] * np.sin(a)) - result.forces.cd) < 1e-10
