"""Grid quality record of a mesh, for the result metadata
(``quality(mesh).record()``, ``src/jaxrans/mesh/quality.py:24-117``; its keys
are the reference's).

One formulation for 2D and 3D blocks: every cell's mean edge along each
index (the mean of its 2^(d-1) parallel edges; in 2D that is the segment
between the midpoints of the two opposite faces) gives its width along that
index, the smallest angle between two index directions, the aspect ratio
(largest over smallest width) and the growth ratio to its neighbour along
the index.  At walls: the first cell's height (distance from a wall node to
the next node off the wall) and how far the wall cell's off-wall mean edge
deviates from the wall face's normal.  The far field's distance from the
walls is the smallest distance from a far-field node to a wall face
midpoint (every 7th node, at most 2000 wall faces).
"""

from __future__ import annotations

from dataclasses import asdict, dataclass

import numpy as np

from .block import WALL_KINDS


@dataclass(frozen=True)
class Quality:
    cells: int
    min_area: float              # the smallest cell area (2D) or volume (3D)
    min_angle_deg: float
    max_growth: float
    max_aspect: float
    wall_spacing: tuple          # (smallest, largest) first-cell height
    wall_angle_deg: float        # largest deviation from orthogonal at a wall
    far_field: float

    @property
    def folded(self) -> bool:
        return not self.min_area > 0.0

    def record(self) -> dict:
        return asdict(self)


def _nodes(block) -> np.ndarray:
    """Node coordinates ``(d, n_1 + 1, ..., n_d + 1)``."""
    if block.dim == 3:
        return np.asarray(block.nodes, np.float64)
    return np.stack([np.asarray(block.x_node, np.float64), np.asarray(block.y_node, np.float64)])


def mean_edges(nodes: np.ndarray) -> list[np.ndarray]:
    """Per index direction a, each cell's mean edge along a, ``(d, *cells)``."""
    d = nodes.shape[0]
    out = []
    for a in range(d):
        e = np.diff(nodes, axis=1 + a)
        for b in range(d):
            if b != a:
                n = e.shape[1 + b]
                e = 0.5 * (np.take(e, range(n - 1), axis=1 + b) + np.take(e, range(1, n), axis=1 + b))
        out.append(e)
    return out


def _angle(u: np.ndarray, v: np.ndarray) -> np.ndarray:
    """Angle in degrees between vectors ``(d, ...)`` taken as lines (0..90),
    NaN where one of them has zero length (a pole)."""
    nu, nv = np.linalg.norm(u, axis=0), np.linalg.norm(v, axis=0)
    with np.errstate(invalid="ignore", divide="ignore"):
        c = np.abs(np.sum(u * v, axis=0)) / (nu * nv)
    return np.degrees(np.arccos(np.clip(c, 0.0, 1.0)))


def _side_index(side: str, n: int) -> int:
    return 0 if side.endswith("min") else n


def _wall_faces(block, side, span):
    """(first-layer node heights, wall-cell off-wall mean edges, wall face
    normals) of the wall cells on ``side`` over ``span`` (running slices)."""
    nodes = _nodes(block)
    d = nodes.shape[0]
    axis = "ijk".index(side[0])
    w = _side_index(side, nodes.shape[1 + axis] - 1)
    off = w + (1 if w == 0 else -1)
    plane = np.take(nodes, w, axis=1 + axis)
    inner = np.take(nodes, off, axis=1 + axis)
    node_span = tuple(slice(s.start, s.stop + 1) for s in span)
    h = (inner - plane)[(slice(None),) + node_span]
    # off-wall mean edge of each wall cell: the mean over its 2^(d-1) wall nodes
    hc = h
    for b in range(d - 1):
        n = hc.shape[1 + b]
        hc = 0.5 * (np.take(hc, range(n - 1), axis=1 + b) + np.take(hc, range(1, n), axis=1 + b))
    # wall face normals: from the face's own corner nodes
    p = plane[(slice(None),) + node_span]
    if d == 2:
        t = p[:, 1:] - p[:, :-1]
        normal = np.stack([t[1], -t[0]])
    else:
        a = p[:, 1:, 1:] - p[:, :-1, :-1]
        b_ = p[:, :-1, 1:] - p[:, 1:, :-1]
        normal = np.cross(a, b_, axis=0)
    return np.linalg.norm(h, axis=0).ravel(), hc, normal


def _spans(rect_or_seg, d: int):
    if d == 2:
        return (slice(rect_or_seg.start, rect_or_seg.stop),)
    return tuple(slice(a, b) for a, b in zip(rect_or_seg.start, rect_or_seg.stop))


def quality(mesh) -> Quality:
    """The quality of a 2D or 3D mesh (walls: patches whose kind is a wall)."""
    mesh = tuple(mesh)
    d = mesh[0].dim
    sizes, angles, growth, aspect = [], [], [1.0], [1.0]
    heights, wall_angles, wall_mids, far_nodes = [], [0.0], [], []
    for block in mesh:
        nodes = _nodes(block)
        edges = mean_edges(nodes)
        widths = [np.linalg.norm(e, axis=0) for e in edges]
        sizes.append(float(np.min(block.area if d == 2 else block.volume)))
        for a in range(d):
            for b in range(a + 1, d):
                ang = _angle(edges[a], edges[b])
                if np.isfinite(ang).any():
                    angles.append(float(np.nanmin(ang)))
        for a, w in enumerate(widths):
            n = w.shape[a]
            if n > 1:
                lo, hi = np.take(w, range(n - 1), axis=a), np.take(w, range(1, n), axis=a)
                ok = (lo > 0) & (hi > 0)
                if ok.any():
                    growth.append(float(np.max(np.maximum(lo / hi, hi / lo)[ok])))
        stack = np.stack(widths)
        ok = stack.min(axis=0) > 0
        if ok.any():
            aspect.append(float(np.max(stack.max(axis=0)[ok] / stack.min(axis=0)[ok])))
        for side, seg in block.sides.boundaries():
            span = _spans(seg, d)
            if seg.kind in WALL_KINDS:
                h, hc, normal = _wall_faces(block, side, span)
                heights.append(h)
                ang = _angle(hc, normal)
                if np.isfinite(ang).any():
                    wall_angles.append(float(np.nanmax(ang)))
                axis = "ijk".index(side[0])
                plane = np.take(nodes, _side_index(side, nodes.shape[1 + axis] - 1), axis=1 + axis)
                p = plane[(slice(None),) + tuple(slice(s.start, s.stop + 1) for s in span)]
                for b in range(d - 1):
                    n = p.shape[1 + b]
                    p = 0.5 * (np.take(p, range(n - 1), axis=1 + b) + np.take(p, range(1, n), axis=1 + b))
                wall_mids.append(p.reshape(d, -1).T)
            elif seg.kind == "far_field":
                axis = "ijk".index(side[0])
                plane = np.take(nodes, _side_index(side, nodes.shape[1 + axis] - 1), axis=1 + axis)
                far_nodes.append(plane[(slice(None),) + tuple(slice(s.start, s.stop + 1) for s in span)]
                                 .reshape(d, -1).T[::7])
    far = float("inf")
    if wall_mids and far_nodes:
        mids = np.concatenate(wall_mids)
        mids = mids[:: max(1, len(mids) // 2000)]
        for pts in far_nodes:
            dist = np.linalg.norm(pts[:, None, :] - mids[None], axis=-1)
            far = min(far, float(dist.min()))
    h = np.concatenate(heights) if heights else np.array([np.nan])
    cells = int(sum(int(np.prod(b.shape)) for b in mesh))
    return Quality(cells, float(min(sizes)), float(min(angles)) if angles else 90.0, float(max(growth)),
                   float(max(aspect)), (float(np.min(h)), float(np.max(h))), float(max(wall_angles)), far)
