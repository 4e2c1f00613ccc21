"""3D node blocks (as a grid file holds them) -> a mesh of ``Grid3D`` blocks
with their interfaces found and their boundary patches named.

The same assembly as ``src/jaxrans/mesh/assemble3d.py:1-19`` (steps
``:227-276``):

1. a block whose cell volumes are all negative has ``k`` reversed (mixed
   signs are an error);
2. the nodes of every block side are merged into shared points where they
   lie within ``pair_tol`` times the local scale (the smaller of the
   shortest face edge and the cell width across the side, over the faces
   round the node); two side faces whose four corners are the same four
   points form an interface, in whichever of the eight relative
   orientations of two quadrilaterals maps one onto the other -- between
   blocks or within one (an O seam, a wake cut); the matched faces of a side
   are cut into ``InterfaceRect`` regions of one affine index map each;
3. a side whose faces are all degenerate (fewer than three distinct
   points) is a pole: an interface with itself half a turn round the running
   direction that is an O seam;
4. every other face takes the name of the first patch rule (a box in
   ``x, y, z``) that holds its midpoint; a side whose faces get several
   labels is tiled by rectangles of one label each.

Plot3D files carry no boundary declarations, so every name comes from the
rules.  CGNS files need h5py, which this build does not use.
"""

from __future__ import annotations

from dataclasses import dataclass
from pathlib import Path

import numpy as np
from scipy.spatial import cKDTree

from .block3d import SIDES_3D, BoundaryRect, Grid3D, InterfaceRect, Sides3D, folded_3d, metrics3d, \
    normal_axis, running_axes
from .plot3d import read_plot3d_3d

FORMATS = ("auto", "plot3d", "cgns")

# A face's corners in its running indices, in order round it.
CORNERS = np.array([(0, 0), (1, 0), (1, 1), (0, 1)])

# For each relative orientation (four rotations, then four reflections): the
# partner's corner that lies on each of my corners.
TURNS = np.array([[(r + s * c) % 4 for c in range(4)] for s in (1, -1) for r in range(4)])


@dataclass(frozen=True)
class PatchRule:
    """Faces whose midpoints lie in the box take patch ``name``; a missing
    bound is unbounded, so a rule with no bounds takes every face still
    unnamed (``src/jaxrans/mesh/assemble.py:37-52``)."""

    name: str
    x_min: float = -np.inf
    x_max: float = np.inf
    y_min: float = -np.inf
    y_max: float = np.inf
    z_min: float = -np.inf
    z_max: float = np.inf

    def holds(self, p: np.ndarray) -> np.ndarray:
        x, y, z = p[..., 0], p[..., 1], p[..., 2]
        return ((x >= self.x_min) & (x <= self.x_max) & (y >= self.y_min) & (y <= self.y_max)
                & (z >= self.z_min) & (z <= self.z_max))


def is_3d(path: str | Path, format: str = "auto") -> bool:
    """A Plot3D file of 3D blocks, none of them a two-node-thick extrusion."""
    if format == "cgns":
        return True
    try:
        blocks = read_plot3d_3d(path)
    except ValueError:
        return False
    return all(min(b.shape[1:]) > 2 for b in blocks)


def load_grid_3d(path: str | Path, format: str = "auto", coordinate_scale: float = 1.0,
                 rules: tuple = (), pair_tol: float = 0.05) -> tuple:
    """Read a 3D Plot3D grid file and assemble it (``assemble3d.py:54-66``)."""
    if format not in FORMATS:
        raise ValueError(f"unknown grid format {format!r}; choose from {FORMATS}")
    path = Path(path)
    if format == "cgns" or path.read_bytes()[:8] == b"\x89HDF\r\n\x1a\n":
        raise ValueError(f"{path.name}: CGNS (HDF5) grids need h5py, which this build does not read; "
                         f"convert the grid to Plot3D")
    return assemble_3d([coordinate_scale * b for b in read_plot3d_3d(path)], rules, pair_tol)


def right_handed(nodes: np.ndarray, label: str = "") -> np.ndarray:
    _, volume, _ = metrics3d(nodes)
    if np.all(volume > 0.0):
        return nodes
    if np.all(volume < 0.0):
        return np.ascontiguousarray(nodes[..., ::-1])
    raise ValueError(f"block {label}has {int(np.sum(volume <= 0.0))} cells of non-positive volume among "
                     f"positive ones")


def side_nodes(nodes: np.ndarray, side: str, depth: int = 0) -> np.ndarray:
    """Node coordinates ``(na+1, nb+1, 3)`` of the plane ``depth`` planes in
    from ``side``, over its two running axes in ascending order."""
    axis = normal_axis(side)
    index = depth if side.endswith("min") else nodes.shape[1 + axis] - 1 - depth
    return np.moveaxis(np.take(nodes, index, axis=1 + axis), 0, -1)


def face_corners(plane: np.ndarray) -> np.ndarray:
    """The four corners ``(na, nb, 4, ...)`` of each face of a node plane,
    in ``CORNERS`` order (round the face)."""
    return np.stack([plane[a:plane.shape[0] - 1 + a, b:plane.shape[1] - 1 + b] for a, b in CORNERS], axis=2)


class SideFaces:
    """Every side face of every block in one flat table.

    Matching is topological: the nodes of all block sides are merged into
    shared points (union of nodes closer than ``pair_tol`` times the local
    scale -- the shortest non-degenerate edge or the cell width across the
    side, whichever is smaller), each face becomes the four point ids of its
    corners, and two faces coincide when they have the same four points.  A
    face with fewer than three distinct points is degenerate (a pole side is
    all degenerate faces).  Columns: ``block``, ``side`` (index into
    ``SIDES_3D``), running ``cell (n, 2)``, ``corners (n, 4, 3)``, point ids
    ``ids (n, 4)``, ``degenerate``, ``mid``."""

    def __init__(self, blocks, pair_tol: float):
        per_face: dict[str, list] = {k: [] for k in ("block", "side", "cell", "corners", "node")}
        node_xyz, node_scale = [], []
        offset = 0
        for b, xyz in enumerate(blocks):
            for s, side in enumerate(SIDES_3D):
                plane = side_nodes(xyz, side)
                quads = face_corners(plane)                                  # (na, nb, 4, 3)
                na, nb = quads.shape[:2]
                edges = np.linalg.norm(np.roll(quads, -1, axis=2) - quads, axis=-1)
                shortest = np.where(edges > 0.0, edges, np.inf).min(axis=-1)
                width = np.linalg.norm(quads.mean(axis=2) - face_corners(side_nodes(xyz, side, 1)).mean(axis=2),
                                       axis=-1)
                scale = np.minimum(shortest, width)                           # (na, nb)
                # a node's scale: the smallest of the faces round it
                ns = np.full(plane.shape[:2], np.inf)
                for a, c in CORNERS:
                    view = ns[a:a + na, c:c + nb]
                    np.minimum(view, scale, out=view)
                ids = np.arange(plane.shape[0] * plane.shape[1]).reshape(plane.shape[:2]) + offset
                offset += ids.size
                node_xyz.append(plane.reshape(-1, 3))
                node_scale.append(ns.ravel())
                ia, ib = np.meshgrid(np.arange(na), np.arange(nb), indexing="ij")
                per_face["block"].append(np.full(na * nb, b))
                per_face["side"].append(np.full(na * nb, s))
                per_face["cell"].append(np.stack([ia.ravel(), ib.ravel()], axis=1))
                per_face["corners"].append(quads.reshape(-1, 4, 3))
                per_face["node"].append(face_corners(ids).reshape(-1, 4))
        for key, parts in per_face.items():
            setattr(self, key, np.concatenate(parts))
        self.mid = self.corners.mean(axis=1)
        points = self._merge(np.concatenate(node_xyz), pair_tol * np.concatenate(node_scale))
        self.ids = points[self.node]
        distinct = np.array([len(set(row)) for row in self.ids.tolist()])
        self.degenerate = distinct < 3

    @staticmethod
    def _merge(xyz: np.ndarray, tol: np.ndarray) -> np.ndarray:
        """Point id per node: nodes closer than the smaller of their two
        tolerances share one (union-find over the close pairs)."""
        parent = np.arange(len(xyz))
        radius = np.where(np.isfinite(tol), tol, 0.0)
        near = cKDTree(xyz).query_ball_point(xyz, r=radius)
        first = np.repeat(np.arange(len(xyz)), [len(c) for c in near])
        second = np.fromiter((o for c in near for o in c), dtype=np.int64, count=len(first))
        keep = first < second
        pairs = np.stack([first[keep], second[keep]], axis=1)
        if len(pairs):
            gap = np.linalg.norm(xyz[pairs[:, 0]] - xyz[pairs[:, 1]], axis=-1)
            pairs = pairs[gap <= np.minimum(radius[pairs[:, 0]], radius[pairs[:, 1]])]

        def root(i):
            while parent[i] != i:
                parent[i] = parent[parent[i]]
                i = parent[i]
            return i

        for i, j in pairs.tolist():
            ri, rj = root(i), root(j)
            if ri != rj:
                parent[max(ri, rj)] = min(ri, rj)
        return np.array([root(i) for i in range(len(xyz))])

    def of(self, block: int, side: str) -> np.ndarray:
        return np.flatnonzero((self.block == block) & (self.side == SIDES_3D.index(side)))


def match_faces(faces: SideFaces):
    """Partner of each face (-1: none) -- the other non-degenerate face with
    the same four points -- and the orientation (row of ``TURNS``) that takes
    my corner ``c`` onto the partner's corner ``TURNS[t][c]``."""
    n = len(faces.ids)
    partner = np.full(n, -1)
    orient = np.zeros(n, dtype=np.int64)
    groups: dict[tuple, list] = {}
    for f in np.flatnonzero(~faces.degenerate):
        groups.setdefault(tuple(sorted(faces.ids[f].tolist())), []).append(int(f))
    for members in groups.values():
        if len(members) == 1:
            continue
        if len(members) > 2:
            raise ValueError("a side face coincides with more than one other face")
        f, g = members
        for me, other in ((f, g), (g, f)):
            hits = np.flatnonzero(np.all(faces.ids[other][TURNS] == faces.ids[me][None], axis=1))
            if hits.size == 0:
                raise ValueError("two side faces share their corner points in no quadrilateral orientation")
            partner[me], orient[me] = other, int(hits[0])
    return partner, orient


def interface_label(faces: SideFaces, row: int, partner: int, turn: int) -> tuple:
    """The affine index map from face ``row`` to its partner as a label that is
    constant over one ``InterfaceRect``: ``(block, side, axes, flip, c)`` with
    partner running cell ``p[m] = c[m] + s_m a[axes[m]]``."""
    sigma = TURNS[turn]
    origin = CORNERS[sigma[0]]
    axes, flip = [0, 0], [False, False]
    for mine, corner in ((0, 1), (1, 3)):
        step = CORNERS[sigma[corner]] - origin
        m = int(np.flatnonzero(step)[0])
        axes[m], flip[m] = mine, bool(step[m] < 0)
    a, p = faces.cell[row], faces.cell[partner]
    c = tuple(int(p[m] + a[axes[m]]) if flip[m] else int(p[m] - a[axes[m]]) for m in range(2))
    return int(faces.block[partner]), SIDES_3D[faces.side[partner]], tuple(axes), tuple(flip), c


def tile(labels: np.ndarray) -> list:
    """Rectangles ``(start, stop, label)`` of one label each that tile
    ``labels`` exactly: from the first uncovered cell in row-major order,
    grow along the second index, then along the first while the whole run
    keeps the label."""
    na, nb = labels.shape
    done = np.zeros(labels.shape, dtype=bool)
    out = []
    for a in range(na):
        for b in range(nb):
            if done[a, b]:
                continue
            lab = labels[a, b]
            b1 = b + 1
            while b1 < nb and not done[a, b1] and labels[a, b1] == lab:
                b1 += 1
            a1 = a + 1
            while a1 < na and not done[a1, b:b1].any() and all(labels[a1, x] == lab for x in range(b, b1)):
                a1 += 1
            done[a:a1, b:b1] = True
            out.append(((a, b), (a1, b1), lab))
    return out


def _rect(start, stop, label):
    if isinstance(label, str):
        return BoundaryRect(label, start, stop)
    block, side, axes, flip, c = label
    origin = tuple(c[m] - start[axes[m]] if flip[m] else c[m] + start[axes[m]] for m in range(2))
    return InterfaceRect(start, stop, block, side, origin, axes, flip)


def pole(block: int, side: str, shape, sides: dict) -> tuple:
    """The collapsed ``side`` as an interface with itself half a turn round
    its periodic running direction (``assemble3d.py:205-220``)."""
    run = running_axes(side)
    extent = tuple(shape[x] for x in run)
    seams = [m for m, axis in enumerate(run)
             if any(isinstance(r, InterfaceRect) and r.block == block and normal_axis(r.side) == axis
                    for s, rects in sides.items() if normal_axis(s) == axis for r in rects)]
    if len(seams) != 1 or extent[seams[0]] % 2:
        raise ValueError(f"block {block}: side {side} is collapsed but not round one periodic direction of "
                         f"even extent")
    m = seams[0]
    origin, period = [0, 0], [0, 0]
    origin[m], period[m] = extent[m] // 2, extent[m]
    return (InterfaceRect((0, 0), extent, block, side, tuple(origin), period=tuple(period)),)


def assemble_3d(blocks, rules: tuple = (), pair_tol: float = 0.05) -> tuple:
    """Node blocks ``(3, ni+1, nj+1, nk+1)`` -> ``Grid3D`` blocks with
    interfaces and named patches (conditions still unassigned)."""
    nodes = [right_handed(np.asarray(b, np.float64), f"{k} ") for k, b in enumerate(blocks)]
    faces = SideFaces(nodes, pair_tol)
    partner, turn = match_faces(faces)
    names = np.full(len(faces.block), None, dtype=object)
    for rule in rules:
        take = (names == None) & (partner < 0) & rule.holds(faces.mid)  # noqa: E711
        names[take] = rule.name
    mesh = []
    for b, xyz in enumerate(nodes):
        shape = tuple(n - 1 for n in xyz.shape[1:])
        sides, poles = {}, []
        for side in SIDES_3D:
            rows = faces.of(b, side)
            if np.all(faces.degenerate[rows]):
                poles.append(side)
                continue
            labels = np.empty(len(rows), dtype=object)
            for k, r in enumerate(rows):
                if partner[r] >= 0:
                    labels[k] = interface_label(faces, r, partner[r], turn[r])
                elif names[r] is not None:
                    labels[k] = names[r]
                else:
                    raise ValueError(f"block {b}: faces of {side} match no patch rule and no other face; "
                                     f"add a rule without bounds (`domain.patches`)")
            extent = tuple(shape[x] for x in running_axes(side))
            sides[side] = tuple(_rect(s, e, lab) for s, e, lab in tile(labels.reshape(extent)))
        for side in poles:
            sides[side] = pole(b, side, shape, sides)
        grid = Grid3D(xyz, Sides3D(shape, *(sides[s] for s in SIDES_3D)))
        if folded_3d(grid):
            raise ValueError(f"block {b} has cells of non-positive volume")
        mesh.append(grid)
    return tuple(mesh)


def grid_report(mesh) -> dict:
    """Blocks, cells, interface rectangles and faces per patch."""
    counts: dict[str, int] = {}
    interfaces = 0
    for grid in mesh:
        for side in SIDES_3D:
            for r in grid.sides.side(side):
                if isinstance(r, InterfaceRect):
                    interfaces += 1
                else:
                    counts[r.patch] = counts.get(r.patch, 0) + (r.stop[0] - r.start[0]) * (r.stop[1] - r.start[1])
    return {"blocks": len(mesh), "cells": sum(int(np.prod(g.shape)) for g in mesh), "interfaces": interfaces,
            "patch_faces": counts}
