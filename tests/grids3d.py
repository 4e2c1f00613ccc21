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


def warped_box(cells, amplitude=0.1, seed=0):
    """Nodes ``(3, ni+1, nj+1, nk+1)`` of a unit-spaced box with every
    interior node moved by a smooth field (boundary nodes stay on the
    box's planes), right-handed."""
    ni, nj, nk = cells
    i, j, k = np.meshgrid(np.arange(ni + 1.0), np.arange(nj + 1.0), np.arange(nk + 1.0), indexing="ij")
    rng = np.random.default_rng(seed)
    phase = rng.uniform(0.0, 2.0 * np.pi, 3)
    bump = np.sin(np.pi * i / ni) * np.sin(np.pi * j / nj) * np.sin(np.pi * k / nk)
    x = i + amplitude * bump * np.sin(0.7 * j + 0.5 * k + phase[0])
    y = j + amplitude * bump * np.sin(0.6 * k + 0.8 * i + phase[1])
    z = k + amplitude * bump * np.sin(0.9 * i + 0.4 * j + phase[2])
    return np.stack([x, y, z])


def reindex(nodes, perm, flips):
    """The same block with its index directions permuted (new axis ``a`` is old
    axis ``perm[a]``) and then reversed where ``flips``; a permutation of odd
    parity is paired with an odd number of reversals so the block stays
    right-handed."""
    out = np.transpose(nodes, (0, *(1 + p for p in perm)))
    for axis, flip in enumerate(flips):
        if flip:
            out = np.flip(out, axis=1 + axis)
    return np.ascontiguousarray(out)


def cell_map(shape, perm, flips):
    """Index of old cell ``c`` in the reindexed block: ``new[a] = c[perm[a]]``,
    reversed where ``flips``."""
    n_new = tuple(shape[p] for p in perm)

    def new(c):
        return tuple(n_new[a] - 1 - c[perm[a]] if flips[a] else c[perm[a]] for a in range(3))
    return new


def cut(nodes, cuts):
    """Blocks of ``nodes`` cut at the cell planes ``cuts`` (one tuple per
    index direction), with each block's first cell in the whole box."""
    edges = [(0, *c, nodes.shape[1 + a] - 1) for a, c in enumerate(cuts)]
    out = []
    for i0, i1 in zip(edges[0], edges[0][1:]):
        for j0, j1 in zip(edges[1], edges[1][1:]):
            for k0, k1 in zip(edges[2], edges[2][1:]):
                out.append((nodes[:, i0:i1 + 1, j0:j1 + 1, k0:k1 + 1], (i0, j0, k0)))
    return out


def right_handed_orientations():
    """Every permutation with reversals that keeps a block right-handed."""
    import itertools
    out = []
    for perm in itertools.permutations(range(3)):
        odd = sum(1 for a in range(3) for b in range(a + 1, 3) if perm[a] > perm[b]) % 2
        for flips in itertools.product((False, True), repeat=3):
            if sum(flips) % 2 == odd:
                out.append((perm, flips))
    return out
