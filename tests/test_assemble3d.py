"""3D grid files and their assembly: the Plot3D reader's variants, interfaces
found in every relative orientation, poles, patch rules, and the assembled
grids against hand-built ones through the oracle's residual.

The ONERA M6 grid (``tests/golden/cases/onera_m6_sa/om6_strct_16x44x24.ufmt``) is the data file
of the reference's 3D example case (``examples/onera_m6_sa``); its expected
assembly (a pole on ``i_min`` of period 24 turned by 12, ``k_min`` seamed to
``k_max``) is the one ``tests/unit/test_assemble3d.py:147-155`` of the
reference asserts.
"""

from __future__ import annotations

from pathlib import Path

import numpy as np
import pytest

from grids3d import cell_map, cut, reindex, right_handed_orientations, warped_box, with_kinds
from jaxrans_amd.core.config import FlowConditions, NumericsConfig
from jaxrans_amd.mesh.assemble3d import PatchRule, assemble_3d, grid_report, is_3d, load_grid_3d
from jaxrans_amd.mesh.atlas3d import build_atlas_3d
from jaxrans_amd.mesh.block3d import SIDES_3D, Grid3D, InterfaceRect, box_sides
from jaxrans_amd.mesh.plot3d import read_plot3d_3d, write_plot3d_3d
from jaxrans_amd.mesh.wing import wing_blocks, wing_patch_rules
from oracle import residual3d as O3

GOLDEN = Path(__file__).resolve().parent / "golden"
M6_CASE = GOLDEN / "cases" / "onera_m6_sa"
M6 = M6_CASE / "om6_strct_16x44x24.ufmt"
M6_RULES = (PatchRule("symmetry", y_max=1e-9),
            PatchRule("wall", x_min=-1.0, x_max=3.0, y_max=3.0, z_min=-1.0, z_max=1.0),
            PatchRule("farfield"))


def _record(payload: bytes, order="<") -> bytes:
    n = np.array([len(payload)], order + "i4").tobytes()
    return n + payload + n


@pytest.mark.parametrize("variant", ["le-f8-multi", "be-f8-multi", "le-f4-single", "le-f8-iblank",
                                     "text-multi", "text-single-iblank"])
def test_plot3d_variants_read_alike(tmp_path, variant):
    blocks = [warped_box((3, 4, 2), 0.1, 1), warped_box((2, 2, 5), 0.1, 2) + 7.0]
    if "single" in variant:
        blocks = blocks[:1]
    path = tmp_path / "g.x"
    if variant == "le-f8-multi":
        write_plot3d_3d(path, blocks)
    elif variant.startswith(("le", "be")):
        order, prec = ("<" if variant.startswith("le") else ">"), ("f4" if "f4" in variant else "f8")
        parts = [] if "single" in variant else [_record(np.array([len(blocks)], order + "i4").tobytes(), order)]
        parts.append(_record(np.array([b.shape[1:] for b in blocks], order + "i4").tobytes(), order))
        for b in blocks:
            xyz = np.concatenate([b[c].T.ravel() for c in range(3)]).astype(order + prec).tobytes()
            if "iblank" in variant:
                xyz += np.ones(b[0].size, order + "i4").tobytes()
            parts.append(_record(xyz, order))
        path.write_bytes(b"".join(parts))
    else:
        lines = [] if "single" in variant else [str(len(blocks))]
        lines += [" ".join(str(n) for n in b.shape[1:]) for b in blocks]
        for b in blocks:
            lines += [" ".join(f"{v:.17e}" for v in b[c].T.ravel()) for c in range(3)]
            if "iblank" in variant:
                lines.append(" ".join("1" for _ in range(b[0].size)))
        path.write_text("\n".join(lines) + "\n")
    got = read_plot3d_3d(path)
    assert len(got) == len(blocks)
    for a, b in zip(got, blocks):
        assert a.shape == b.shape
        tol = 1e-6 if "f4" in variant else 0.0
        np.testing.assert_allclose(a, b, rtol=tol, atol=tol * 10)


def test_not_a_grid_is_refused(tmp_path):
    path = tmp_path / "junk.x"
    path.write_text("hello grid\n")
    with pytest.raises(ValueError, match="not a 3D Plot3D grid"):
        read_plot3d_3d(path)
    assert not is_3d(path)


def test_the_m6_grid_assembles_with_a_pole_and_a_seam():
    (grid,) = load_grid_3d(M6, rules=M6_RULES)
    assert grid.shape == (16, 44, 24)
    (pole,) = grid.sides.i_min
    assert pole.side == "i_min" and pole.block == 0 and pole.period == (0, 24) and pole.origin == (0, 12)
    assert grid.sides.k_min[0].side == "k_max" and grid.sides.k_max[0].side == "k_min"
    report = grid_report((grid,))
    assert report["patch_faces"] == {"wall": 16 * 24, "farfield": 16 * 24, "symmetry": 44 * 24}
    assert report["interfaces"] == 3


@pytest.mark.parametrize("turbulent", [False, True], ids=["euler", "sa"])
def test_freestream_is_preserved_on_the_m6_grid(turbulent):
    """Every patch far field or symmetry: a uniform freestream must have a zero
    residual, through the pole, the seam and every composed corner ghost."""
    mesh = with_kinds(load_grid_3d(M6, rules=M6_RULES),
                      {"wall": "far_field", "farfield": "far_field", "symmetry": "symmetry"})
    atlas = build_atlas_3d(mesh)
    flow = FlowConditions(mach=0.84, alpha_deg=3.06, reynolds=14.6e6 if turbulent else None,
                          turbulence="sa_noft2" if turbulent else None,
                          sutherland_ratio=198.6 / 540.0 if turbulent else None)
    st = O3.setup_for(atlas, flow, NumericsConfig(flux="hllc", limiter="van_albada", solver="implicit"))
    U = O3.U0_of(st)
    scale = np.abs(U[:, atlas.interior]).max() * max(np.abs(a).max() for a in atlas.faces)
    with np.errstate(divide="ignore", invalid="ignore"):
        for first_order in (True, False):
            R = O3.residual(U, st, first_order)
            assert np.abs(R[:, atlas.interior]).max() < 1e-14 * scale


def test_the_wing_reads_back_from_plot3d_as_generated(tmp_path):
    kw = dict(cells_airfoil=32, cells_wake=8, cells_normal=12, cells_span=4, blocks=3, wall_spacing=1e-3,
              far_field=15.0, section_points=81)
    made = wing_blocks(**kw)
    path = write_plot3d_3d(tmp_path / "wing.x", [g.nodes for g in made])
    assert is_3d(path)
    mesh = load_grid_3d(path, rules=wing_patch_rules())
    assert len(mesh) == len(made)
    for a, b in zip(made, mesh):
        np.testing.assert_array_equal(a.nodes, b.nodes)
        for side in SIDES_3D:
            assert sorted(a.sides.side(side), key=repr) == sorted(b.sides.side(side), key=repr), side


def _mutual(mesh):
    seen = 0
    for b, grid in enumerate(mesh):
        for side in SIDES_3D:
            for r in grid.sides.side(side):
                if not isinstance(r, InterfaceRect) or r.period != (0, 0):
                    continue
                other = mesh[r.block]
                for a0 in range(r.start[0], r.stop[0]):
                    for a1 in range(r.start[1], r.stop[1]):
                        p = r.partner((a0, a1))
                        back = [q for q in other.sides.side(r.side) if isinstance(q, InterfaceRect)
                                and all(q.start[m] <= p[m] < q.stop[m] for m in range(2))]
                        assert len(back) == 1 and back[0].block == b and back[0].side == side
                        assert back[0].partner(p) == (a0, a1)
                        seen += 1
    return seen


@pytest.mark.parametrize("seed", [1, 2])
def test_blocks_in_any_orientation_give_the_single_block_residual(seed):
    """A warped box cut into eight blocks, each re-indexed by a different
    right-handed permutation and reversal: the assembled interfaces are
    mutual, and the oracle's second-order Euler residual in every cell equals
    that of the uncut box."""
    shape = (4, 6, 5)
    nodes = warped_box(shape, amplitude=0.15, seed=seed)
    rng = np.random.default_rng(seed)
    orients = right_handed_orientations()
    picks = rng.choice(len(orients), 8, replace=False)
    pieces = cut(nodes, ((2,), (3,), (2,)))
    raw = [reindex(b, *orients[k]) for (b, _), k in zip(pieces, picks)]
    mesh = assemble_3d(raw, (PatchRule("outer"),))
    assert _mutual(mesh) == 2 * (6 * 5 + 4 * 5 + 4 * 6)
    assert len({r.axes for g in mesh for s in SIDES_3D for r in g.sides.side(s)
                if isinstance(r, InterfaceRect)}) == 2
    kinds = {"outer": "far_field"}
    whole = build_atlas_3d(with_kinds((Grid3D(nodes, box_sides(shape, **dict.fromkeys(SIDES_3D, "outer"))),),
                                      kinds))
    split = build_atlas_3d(with_kinds(mesh, kinds))
    flow = FlowConditions(mach=0.6, alpha_deg=3.0)
    num = NumericsConfig(flux="roe", limiter="van_albada_smooth", solver="implicit")
    st_w, st_s = O3.setup_for(whole, flow, num), O3.setup_for(split, flow, num)
    field = np.random.default_rng(5).uniform(-1.0, 1.0, (5,) + shape)
    U_whole = O3.U0_of(st_w) * (1.0 + 0.02 * whole.pack([field]))
    blocks = []
    for (piece, origin), k, blk in zip(pieces, picks, mesh):
        old_shape = tuple(n - 1 for n in piece.shape[1:])
        new = cell_map(old_shape, *orients[k])
        out = np.empty((5,) + blk.shape)
        for c in np.ndindex(old_shape):
            out[(slice(None),) + new(c)] = field[(slice(None),) + tuple(o + x for o, x in zip(origin, c))]
        blocks.append(out)
    U_split = O3.U0_of(st_s) * (1.0 + 0.02 * split.pack(blocks))
    for first_order in (True, False):
        R_w = whole.unpack(O3.residual(U_whole, st_w, first_order))[0]
        R_s = split.unpack(O3.residual(U_split, st_s, first_order))
        scale = np.abs(R_w).max()
        for (piece, origin), k, r in zip(pieces, picks, R_s):
            old_shape = tuple(n - 1 for n in piece.shape[1:])
            new = cell_map(old_shape, *orients[k])
            for c in np.ndindex(old_shape):
                want = R_w[(slice(None),) + tuple(o + x for o, x in zip(origin, c))]
                got = r[(slice(None),) + new(c)]
                np.testing.assert_allclose(got, want, rtol=0, atol=1e-12 * scale)


def test_a_left_handed_block_is_turned_round():
    nodes = warped_box((3, 3, 4), 0.1, 3)
    (grid,) = assemble_3d([np.ascontiguousarray(nodes[..., ::-1])], (PatchRule("outer"),))
    np.testing.assert_array_equal(grid.nodes, nodes)
    assert np.all(grid.volume > 0.0)


def test_a_block_of_mixed_volumes_is_refused():
    nodes = warped_box((3, 3, 4), 0.0, 0)
    nodes[2, :, :, 2] = nodes[2, :, :, 0] - 0.5      # folds the second layer of cells
    with pytest.raises(ValueError, match="non-positive volume"):
        assemble_3d([nodes], (PatchRule("outer"),))


def test_an_unnamed_face_names_its_block_and_side():
    with pytest.raises(ValueError, match="block 0: faces of j_min"):
        assemble_3d([warped_box((2, 2, 2))], (PatchRule("top", z_min=1e9),))


def test_a_face_against_two_blocks_is_cut_in_two():
    nodes = warped_box((4, 4, 3), amplitude=0.1, seed=4)
    bottom = nodes[:, :, :3]
    left, right = nodes[:, :3, 2:], nodes[:, 2:, 2:]
    mesh = assemble_3d([bottom, left, reindex(right, (2, 0, 1), (False, False, False))], (PatchRule("outer"),))
    assert [r.block for r in mesh[0].sides.j_max] == [1, 2]
    assert _mutual(mesh) == 2 * (4 * 3 + 2 * 3)


def test_the_m6_case_loads_as_a_3d_atlas():
    from jaxrans_amd.case.loader import load_case

    case = load_case(M6_CASE)
    assert case.grid_spec.grid == "file" and case.grid_spec.reference_area == pytest.approx(1.15315084119)
    atlas = case.grid()
    assert atlas.dim == 3 and atlas.cells == 16 * 44 * 24
    kinds = {r.patch: r.kind for b in atlas.mesh for _, r in b.sides.boundaries()}
    assert kinds == {"wall": "no_slip_wall", "farfield": "far_field", "symmetry": "symmetry"}
    assert case.flow.sutherland_ratio == pytest.approx(198.6 / 540.0)


def test_a_3d_case_without_a_reference_area_is_refused(tmp_path):
    import json
    import shutil

    from jaxrans_amd.case.loader import CaseError, load_case

    for name in ("case.json", "numerical.json", M6.name):
        shutil.copy(M6_CASE / name, tmp_path / name)
    doc = json.loads((tmp_path / "case.json").read_text())
    del doc["domain"]["reference_area"]
    (tmp_path / "case.json").write_text(json.dumps(doc))
    with pytest.raises(CaseError, match="reference_area"):
        load_case(tmp_path).grid()
    doc["domain"]["reference_area"] = 1.0
    doc["domain"]["patches"] = [{"name": "symmetry", "y_max": 1e-9, "w_max": 1.0}]
    (tmp_path / "case.json").write_text(json.dumps(doc))
    with pytest.raises(CaseError, match="w_max"):
        load_case(tmp_path).grid()
    doc["domain"]["patches"] = [{"name": "symmetry", "y_max": 1e-9}]
    (tmp_path / "case.json").write_text(json.dumps(doc))
    with pytest.raises(CaseError, match="match no patch rule"):
        load_case(tmp_path).grid()
