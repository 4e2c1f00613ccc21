"""Result-file helpers on the CPU: the grid-quality record on grids whose
figures are known in closed form, the VTK structured-grid writer read back
byte by byte, and the wall loop order of the surface distributions."""

from __future__ import annotations

import re
from types import SimpleNamespace

import numpy as np
import pytest

from conftest import o_atlas
from jaxrans_amd.mesh.block import Grid, rectangle_sides
from jaxrans_amd.mesh.block3d import Grid3D, box_sides
from jaxrans_amd.mesh.quality import quality
from jaxrans_amd.output import primitive_cells, write_vts
from jaxrans_amd.post.surface import loop_order

KINDS = {"wall": "no_slip_wall", "far": "far_field", "sym": "symmetry"}


def _channel(ni=12, nj=8, dx=0.5, dy0=0.01, ratio=1.0):
    """Cartesian cells of width dx, heights dy0 * ratio^j, wall on j_min."""
    h = dy0 * ratio ** np.arange(nj)
    y = np.concatenate([[0.0], np.cumsum(h)])
    x = np.arange(ni + 1) * dx
    X, Y = np.meshgrid(x, y, indexing="ij")
    sides = rectangle_sides(ni, nj, j_min="wall", j_max="far", i_min="sym", i_max="sym").with_conditions(KINDS)
    return Grid(X, Y, sides), h


def test_quality_of_a_cartesian_channel_is_known_in_closed_form():
    g, h = _channel(ratio=1.2)
    q = quality((g,))
    assert q.cells == 12 * 8
    assert q.min_area == pytest.approx(0.5 * 0.01, rel=1e-12)
    assert q.min_angle_deg == pytest.approx(90.0, abs=1e-9)
    assert q.max_growth == pytest.approx(1.2, rel=1e-12)
    assert q.max_aspect == pytest.approx(0.5 / 0.01, rel=1e-12)
    assert q.wall_spacing == pytest.approx((0.01, 0.01), rel=1e-12)
    assert q.wall_angle_deg == pytest.approx(0.0, abs=1e-9)
    # far-field nodes (every 7th) to wall face midpoints (half a cell off in x)
    assert q.far_field == pytest.approx(np.hypot(0.25, h.sum()), rel=1e-12)
    assert not q.folded and set(q.record()) == {"cells", "min_area", "min_angle_deg", "max_growth", "max_aspect",
                                                "wall_spacing", "wall_angle_deg", "far_field"}


def test_quality_sees_skew_at_the_wall():
    """Shearing every node by x += s y tilts the off-wall lines by atan(s)."""
    g, _ = _channel()
    s = 0.3
    sheared = Grid(g.x_node + s * g.y_node, g.y_node, g.sides)
    q = quality((sheared,))
    assert q.wall_angle_deg == pytest.approx(np.degrees(np.arctan(s)), rel=1e-9)
    assert q.min_angle_deg == pytest.approx(90.0 - np.degrees(np.arctan(s)), rel=1e-9)


def test_quality_3d_of_a_box():
    ni, nj, nk = 4, 5, 3
    x, y, z = np.meshgrid(np.arange(ni + 1) * 0.2, np.arange(nj + 1) * 0.05, np.arange(nk + 1) * 0.4,
                          indexing="ij")
    sides = box_sides((ni, nj, nk), j_min="wall", j_max="far", i_min="sym", i_max="sym", k_min="sym",
                      k_max="sym").with_conditions(KINDS)
    q = quality((Grid3D(np.stack([x, y, z]), sides),))
    assert q.cells == ni * nj * nk
    assert q.min_area == pytest.approx(0.2 * 0.05 * 0.4, rel=1e-12)
    assert q.min_angle_deg == pytest.approx(90.0, abs=1e-9)
    assert q.max_growth == pytest.approx(1.0, rel=1e-12)
    assert q.max_aspect == pytest.approx(0.4 / 0.05, rel=1e-12)
    assert q.wall_spacing == pytest.approx((0.05, 0.05), rel=1e-12)
    assert q.wall_angle_deg == pytest.approx(0.0, abs=1e-9)
    assert q.far_field == pytest.approx(np.sqrt(0.1 ** 2 + (nj * 0.05) ** 2 + 0.2 ** 2), rel=1e-12)


def _read_vts(path):
    raw = path.read_bytes()
    head, data = raw.split(b'<AppendedData encoding="raw">_', 1)
    arrays = {}
    for m in re.finditer(rb'Name="(\w+)" NumberOfComponents="(\d)" format="appended" offset="(\d+)"', head):
        off = int(m.group(3))
        n = int(np.frombuffer(data[off:off + 8], "<u8")[0])
        arrays[m.group(1).decode()] = np.frombuffer(data[off + 8:off + 8 + n], "<f8")
    return head.decode(), arrays


def test_vts_writer_round_trip(tmp_path):
    rng = np.random.default_rng(4)
    nodes = rng.standard_normal((3, 3, 4, 5))
    gamma = 1.4
    rho = 1.0 + 0.1 * rng.random((2, 3, 4))
    vel = 0.3 * rng.standard_normal((3, 2, 3, 4))
    p = 1.0 / gamma + 0.05 * rng.random((2, 3, 4))
    U = np.concatenate([rho[None], rho * vel, (p / (gamma - 1) + 0.5 * rho * np.sum(vel * vel, 0))[None]])
    cells = primitive_cells(U, gamma)
    assert np.allclose(cells["pressure"], p, rtol=1e-12) and np.allclose(cells["velocity"], vel, rtol=1e-12)
    head, arrays = _read_vts(write_vts(tmp_path / "b.vts", nodes, cells))
    assert 'WholeExtent="0 2 0 3 0 4"' in head
    pts = arrays["Points"].reshape(-1, 3)                     # i fastest
    assert np.array_equal(pts[1], nodes[:, 1, 0, 0]) and np.array_equal(pts[3], nodes[:, 0, 1, 0])
    assert np.array_equal(pts[3 * 4], nodes[:, 0, 0, 1])
    assert np.array_equal(arrays["density"][1], rho[1, 0, 0]) and np.array_equal(arrays["density"][2], rho[0, 1, 0])
    v = arrays["velocity"].reshape(-1, 3)
    assert np.array_equal(v[2 * 3], vel[:, 0, 0, 1])
    mach = np.linalg.norm(vel, axis=0) / np.sqrt(gamma * p / rho)
    assert np.allclose(arrays["mach"], np.transpose(mach, (2, 1, 0)).ravel(), rtol=1e-12)


def test_wall_loop_order_follows_the_contour_from_the_trailing_edge():
    at = o_atlas()
    w = at.walls
    order = loop_order(w)
    assert sorted(order.tolist()) == list(range(w.n))
    assert w.a[order[0], 0] == w.a[:, 0].max()
    assert np.allclose(w.b[order[:-1]], w.a[order[1:]], rtol=0, atol=1e-12)
    # a shuffled wall list gives the same loop
    perm = np.random.default_rng(0).permutation(w.n)
    shuffled = SimpleNamespace(a=w.a[perm], b=w.b[perm], body=w.body[perm], bodies=w.bodies)
    assert np.array_equal(perm[loop_order(shuffled)], order)
