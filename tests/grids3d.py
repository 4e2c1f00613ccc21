"""Small 3D grids for the tests: a 2D mesh extruded along the span, and a
small version of the four-block wing."""

from __future__ import annotations

import numpy as np

from jaxrans_amd.mesh.block import Boundary, Interface
from jaxrans_amd.mesh.block3d import BoundaryRect, Grid3D, InterfaceRect, box_sides
from jaxrans_amd.mesh.atlas3d import build_atlas_3d
from jaxrans_amd.mesh.wing import wing_blocks

KINDS = {"wall": "no_slip_wall", "farfield": "far_field", "symmetry": "symmetry"}


def extrude(mesh2d, nk: int, dy: float, k_sides=("symmetry", "symmetry")):
    """Each 2D block (nodes ``x, y``) as a 3D block: ``x = x``, ``z = y``,
    ``y = -k dy`` (right-handed), its 2D sides carried over every ``k``, the
    ``k`` sides the given patches.  The 3D residual of a state uniform in
    ``k`` with ``v = 0`` is then ``dy`` times the 2D one."""
    out = []
    for blk in mesh2d:
        ni, nj = blk.ni, blk.nj
        nodes = np.empty((3, ni + 1, nj + 1, nk + 1))
        for k in range(nk + 1):
            nodes[0, :, :, k] = blk.x_node
            nodes[1, :, :, k] = -k * dy
            nodes[2, :, :, k] = blk.y_node
        sides = {}
        for name in ("j_min", "j_max", "i_min", "i_max"):
            rects = []
            for seg in blk.sides.side(name):
                if isinstance(seg, Boundary):
                    rects.append(BoundaryRect(seg.patch, (seg.start, 0), (seg.stop, nk), seg.kind))
                else:
                    first = seg.partner(seg.start)
                    rects.append(InterfaceRect((seg.start, 0), (seg.stop, nk), seg.block, seg.side, (first, 0),
                                               flip=(seg.reversed, False)))
            sides[name] = tuple(rects)
        sides["k_min"], sides["k_max"] = k_sides
        out.append(Grid3D(nodes, box_sides((ni, nj, nk), **sides)))
    return out


def with_kinds(mesh, kinds=KINDS):
    return tuple(b.with_sides(b.sides.with_conditions(kinds)) for b in mesh)


def extruded_atlas(atlas2d, nk=3, dy=0.1, kinds=None):
    kinds = dict(kinds or {})
    kinds.setdefault("symmetry", "symmetry")
    mesh = extrude(atlas2d.mesh, nk, dy)
    for b in mesh:
        for _, r in b.sides.boundaries():
            if r.kind is not None:
                kinds.setdefault(r.patch, r.kind)
    return build_atlas_3d(with_kinds(mesh, kinds))


def small_wing(cells_airfoil=24, cells_wake=6, cells_normal=8, cells_span=4, blocks=4, wall="no_slip_wall"):
    mesh = wing_blocks(cells_airfoil=cells_airfoil, cells_wake=cells_wake, cells_normal=cells_normal,
                       cells_span=cells_span, blocks=blocks, far_field=15.0, wall_spacing=2e-3,
                       section_points=101)
    return build_atlas_3d(with_kinds(mesh, {**KINDS, "wall": wall}))
