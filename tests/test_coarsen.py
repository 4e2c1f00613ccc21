"""Grid sequencing's host side (``src/jaxrans/mesh/coarsen.py:58-83``):
every other node, segments halved, injection; odd counts refused."""

from __future__ import annotations

from pathlib import Path

import numpy as np
import pytest

from jaxrans_amd import CaseError, load_case
from jaxrans_amd.mesh.coarsen import prolong

CASE = Path(__file__).resolve().parent / "golden" / "cases" / "euler_o48"


def test_coarse_grid_is_every_other_node_and_keeps_its_patches():
    from jaxrans_amd.mesh.coarsen import coarsen
    case = load_case(CASE)
    fine, coarse = case.grid(), case.grid(coarsen=1)
    assert coarse.cells * 4 == fine.cells
    assert sorted(set(k for st in coarse.stages for k in st.kind.tolist())) == \
        sorted(set(k for st in fine.stages for k in st.kind.tolist()))
    mesh = case.grid_spec.build(case.directory)
    half = coarsen(mesh)
    for b, h in zip(mesh, half):
        assert np.array_equal(h.x_node, b.x_node[::2, ::2]) and np.array_equal(h.y_node, b.y_node[::2, ::2])
        for name in ("j_min", "j_max", "i_min", "i_max"):
            for sf, sc in zip(b.sides.side(name), h.sides.side(name)):
                assert (sc.start, sc.stop) == (sf.start // 2, sf.stop // 2)
                assert type(sc) is type(sf)


def test_injection_gives_each_child_its_parents_value():
    case = load_case(CASE)
    fine, coarse = case.grid(), case.grid(coarsen=1)
    rng = np.random.default_rng(0)
    U = rng.standard_normal((4,) + coarse.shape)
    start = np.full((4,) + fine.shape, 7.0)
    out = prolong(U, coarse, fine, start)
    (fi, fj), (ci, cj) = fine.windows[0], coarse.windows[0]
    parent = U[:, ci, cj]
    for di in (0, 1):
        for dj in (0, 1):
            assert np.array_equal(out[:, fi, fj][:, di::2, dj::2], parent)
    outside = np.ones(fine.shape, bool)
    outside[fi, fj] = False
    assert np.all(out[:, outside] == 7.0)


def test_odd_cell_counts_are_refused():
    case = load_case(CASE, ("case.domain.cells_normal=25",))
    with pytest.raises(CaseError, match="odd"):
        case.grid(coarsen=1)


M6 = Path(__file__).resolve().parent / "golden" / "cases" / "onera_m6_sa"


def test_3d_coarse_grid_of_the_m6_file_case():
    """The reference's M6 grid (an O seam and a pole): every other node of
    each block, every rectangle halved, the same condition kinds, and the
    coarse atlas assembles (its interfaces and pole still match)."""
    from jaxrans_amd.mesh.block3d import SIDES_3D
    from jaxrans_amd.mesh.coarsen import coarsen
    case = load_case(M6)
    fine, coarse = case.grid(), case.grid(coarsen=1)
    assert coarse.cells * 8 == fine.cells
    assert sorted(set(coarse.chains.kind.tolist())) == sorted(set(fine.chains.kind.tolist()))
    mesh = case.grid_spec.build(case.directory)
    for b, h in zip(mesh, coarsen(mesh)):
        assert np.array_equal(h.nodes, b.nodes[:, ::2, ::2, ::2])
        for name in SIDES_3D:
            for rf, rc in zip(b.sides.side(name), h.sides.side(name)):
                assert rc.start == tuple(x // 2 for x in rf.start) and rc.stop == tuple(x // 2 for x in rf.stop)


def test_3d_injection_gives_each_child_its_parents_value():
    case = load_case(M6)
    fine, coarse = case.grid(), case.grid(coarsen=1)
    rng = np.random.default_rng(1)
    U = rng.standard_normal((6,) + coarse.shape)
    out = prolong(U, coarse, fine, np.zeros((6,) + fine.shape))
    for wf, wc in zip(fine.windows, coarse.windows):
        parent = U[(Ellipsis,) + tuple(wc)]
        child = out[(Ellipsis,) + tuple(wf)]
        for d in np.ndindex(2, 2, 2):
            assert np.array_equal(child[:, d[0]::2, d[1]::2, d[2]::2], parent)
