"""Raw node blocks -> a mesh: interfaces found, boundary patches named.

As ``src/jaxrans/mesh/assemble.py:210-299``: coincident side faces (both end
nodes within a tolerance) become interfaces -- a seam or a wake cut within one
block, or a joint between blocks; every other side face is a boundary,
named by the orientation of the boundary loop it lies on, walked with the
fluid on the left: counter-clockwise (the outer loop) ``farfield``,
clockwise (a body) ``wall``.
"""

from __future__ import annotations

import numpy as np
from scipy.spatial import cKDTree

from .block import SIDES, BlockSides, Boundary, Grid, Interface

Mesh = tuple


def _side_nodes(x, y, side):
    idx = {"j_min": (slice(None), 0), "j_max": (slice(None), -1),
           "i_min": (0, slice(None)), "i_max": (-1, slice(None))}[side]
    return x[idx], y[idx]


def _right_handed(x, y):
    d1x, d1y = x[1:, 1:] - x[:-1, :-1], y[1:, 1:] - y[:-1, :-1]
    d2x, d2y = x[:-1, 1:] - x[1:, :-1], y[:-1, 1:] - y[1:, :-1]
    area = 0.5 * (d1x * d2y - d1y * d2x)
    if np.all(area > 0.0):
        return x, y
    if not np.all(area < 0.0):
        raise ValueError(f"block of {x.shape} nodes folds over: cell areas of both signs")
    return x[::-1].copy(), y[::-1].copy()


def assemble(raw: list[tuple[np.ndarray, np.ndarray]], pair_tol: float = 0.05) -> Mesh:
    blocks = [_right_handed(np.asarray(x, np.float64), np.asarray(y, np.float64)) for x, y in raw]
    rec = []            # (block, side index, cell, a(2), b(2), width)
    for k, (x, y) in enumerate(blocks):
        xc = 0.25 * (x[:-1, :-1] + x[1:, :-1] + x[1:, 1:] + x[:-1, 1:])
        yc = 0.25 * (y[:-1, :-1] + y[1:, :-1] + y[1:, 1:] + y[:-1, 1:])
        for s, side in enumerate(SIDES):
            sx, sy = _side_nodes(x, y, side)
            cx, cy = _side_nodes(np.pad(xc, 0), np.pad(yc, 0), side)
            for n in range(sx.size - 1):
                a, b = np.array([sx[n], sy[n]]), np.array([sx[n + 1], sy[n + 1]])
                e, w = b - a, np.array([cx[n], cy[n]]) - a
                width = 2.0 * abs(e[0] * w[1] - e[1] * w[0]) / max(np.linalg.norm(e), 1e-300)
                rec.append((k, s, n, a, b, width))
    nf = len(rec)
    A = np.array([r[3] for r in rec])
    B = np.array([r[4] for r in rec])
    mid = 0.5 * (A + B)
    length = np.hypot(*(B - A).T)
    tol = pair_tol * np.minimum(length, np.array([r[5] for r in rec]))
    partner = np.full(nf, -1)
    reverse = np.zeros(nf, dtype=bool)
    tree = cKDTree(mid)
    for f in range(nf):
        for g in tree.query_ball_point(mid[f], tol[f]):
            if g == f:
                continue
            same = np.hypot(*(A[f] - A[g])) <= tol[f] and np.hypot(*(B[f] - B[g])) <= tol[f]
            flip = np.hypot(*(A[f] - B[g])) <= tol[f] and np.hypot(*(B[f] - A[g])) <= tol[f]
            if same or flip:
                if partner[f] >= 0:
                    raise ValueError("a face coincides with more than one other face")
                partner[f], reverse[f] = g, flip
    names = _loop_names(rec, partner < 0)
    mesh = []
    for k, (x, y) in enumerate(blocks):
        spec = {}
        for s, side in enumerate(SIDES):
            rows = [f for f in range(nf) if rec[f][0] == k and rec[f][1] == s]
            rows.sort(key=lambda f: rec[f][2])
            targets = []
            for f in rows:
                if partner[f] < 0:
                    targets.append(names[f])
                else:
                    g = partner[f]
                    targets.append((rec[g][0], SIDES[rec[g][1]], rec[g][2], bool(reverse[f])))
            spec[side] = _segments(targets)
        mesh.append(Grid(x, y, BlockSides(x.shape[0] - 1, x.shape[1] - 1, **spec)))
    return tuple(mesh)


def _segments(targets: list) -> tuple:
    segs, k = [], 0
    while k < len(targets):
        t, stop = targets[k], k + 1
        if isinstance(t, str):
            while stop < len(targets) and targets[stop] == t:
                stop += 1
            segs.append(Boundary(t, k, stop))
        else:
            blk, side, cell, rev = t
            step = -1 if rev else 1
            while (stop < len(targets) and not isinstance(targets[stop], str)
                   and targets[stop][:2] == (blk, side) and targets[stop][3] == rev
                   and targets[stop][2] == targets[stop - 1][2] + step):
                stop += 1
            segs.append(Interface(k, stop, blk, side, min(cell, targets[stop - 1][2]), rev))
        k = stop
    return tuple(segs)


def _loop_names(rec, boundary: np.ndarray) -> list:
    """``farfield``/``wall`` per boundary face by the orientation of its loop."""
    names = [None] * len(rec)
    rows = np.flatnonzero(boundary)
    if not rows.size:
        return names
    fwd = np.array([SIDES[rec[f][1]] in ("j_min", "i_max") for f in rows])
    a = np.array([rec[f][3] if w else rec[f][4] for f, w in zip(rows, fwd)])
    b = np.array([rec[f][4] if w else rec[f][3] for f, w in zip(rows, fwd)])
    tol = 1e-3 * np.hypot(*(b - a).T).min()
    key = lambda p: tuple(np.round(p / tol).astype(np.int64))  # noqa: E731
    start_of = {key(p): n for n, p in enumerate(a)}
    nxt = np.array([start_of.get(key(p), -1) for p in b])
    seen = np.zeros(len(rows), dtype=bool)
    for s in range(len(rows)):
        if seen[s]:
            continue
        loop, f = [], s
        while f >= 0 and not seen[f]:
            seen[f] = True
            loop.append(f)
            f = nxt[f]
        closed = f == s
        signed = 0.5 * np.sum(a[loop, 0] * b[loop, 1] - b[loop, 0] * a[loop, 1])
        for f in loop:
            names[rows[f]] = "farfield" if closed and signed > 0.0 else "wall"
    return names
