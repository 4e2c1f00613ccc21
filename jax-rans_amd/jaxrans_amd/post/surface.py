"""Wall distributions of a result (``surface``, ``src/jaxrans/post/forces.py:203-246``).

The per-face loads come from the device (``Context.wall_loads``: the wall
cell's gauge pressure and the viscous force on the body, the same values the
force kernels integrate); this module only orders and normalises them.

2D, per wall face, body by body (bodies numbered by leading-edge ``x``), each
body's faces in loop order (fluid on the left) starting at its trailing
edge: ``x``, ``y`` (face midpoint), ``body``, ``cp`` and, viscous, ``cf``
(the shear along the downstream-pointing tangent over ``q_inf``, negative
where the flow is separated).  3D, per wall face in the atlas's wall order:
``block``, ``i``, ``k`` (the face's running indices on its side), ``x``,
``y``, ``z``, ``cp`` and, viscous, ``cf_x``, ``cf_y``, ``cf_z``.
"""

from __future__ import annotations

import numpy as np


def dynamic_pressure(q_inf: np.ndarray, dim: int) -> float:
    """``rho |V|^2 / 2`` of freestream primitives ``(rho, V (dim), p[, nu])``."""
    v = np.asarray(q_inf[1:1 + dim], np.float64)
    return 0.5 * float(q_inf[0]) * float(v @ v)


def loop_order(walls) -> np.ndarray:
    """Wall face indices body by body, each body's loop followed from end node
    to the next face's start node, starting at the face whose start node lies
    furthest downstream (largest ``x``: the trailing edge)."""
    tol = 1e-6 * max(float(np.hypot(*(walls.b - walls.a).T).min()), 1e-300)
    key = lambda p: (round(p[0] / tol), round(p[1] / tol))  # noqa: E731
    starting = {key(p): n for n, p in enumerate(walls.a)}
    out = []
    for body in range(walls.bodies):
        faces = np.flatnonzero(walls.body == body)
        first = faces[np.argmax(walls.a[faces, 0])]
        seen, f = set(), first
        while f is not None and f not in seen and walls.body[f] == body:
            seen.add(f)
            out.append(f)
            f = starting.get(key(walls.b[f]))
        out.extend(sorted(set(faces.tolist()) - seen))     # an open body: the rest in atlas order
    return np.asarray(out, np.int64)


def face_vectors_2d(atlas, idx: np.ndarray) -> np.ndarray:
    """Area vectors ``(2, n)`` of wall faces ``idx``, into the fluid."""
    w = atlas.walls
    out = np.empty((2, idx.size))
    for fam, arr in ((0, atlas.face_i.reshape(2, -1)), (1, atlas.face_j.reshape(2, -1))):
        sel = w.family[idx] == fam
        out[:, sel] = arr[:, w.face[idx][sel]]
    return out * w.into_fluid[idx]


def surface_2d(loads: np.ndarray, atlas, q_inf: np.ndarray, viscous: bool) -> dict:
    """Wall distributions of one case from its loads ``(3, n_walls)``."""
    w = atlas.walls
    order = loop_order(w)
    q = dynamic_pressure(q_inf, 2)
    mid = w.mid[order]
    out = {"x": mid[:, 0], "y": mid[:, 1], "body": w.body[order].astype(float), "cp": loads[0, order] / q}
    if viscous:
        sx, sy = face_vectors_2d(atlas, order)
        length = np.hypot(sx, sy)
        tx, ty = sy / length, -sx / length                 # along the wall, fluid on the left
        sign = np.where(tx >= 0.0, 1.0, -1.0)              # downstream-pointing
        out["cf"] = sign * (loads[1, order] * tx + loads[2, order] * ty) / length / q
    return out


def surface_3d(loads: np.ndarray, atlas, q_inf: np.ndarray, viscous: bool, blocks: np.ndarray,
               running: np.ndarray) -> dict:
    """Wall distributions of one 3D case from its loads ``(4, n_walls)``;
    ``blocks`` and ``running (n_walls, 2)`` give each face's block and its
    running indices on its side."""
    w = atlas.walls
    q = dynamic_pressure(q_inf, 3)
    mid = w.mid
    out = {"block": blocks.astype(float), "i": running[:, 0].astype(float), "k": running[:, 1].astype(float),
           "x": mid[:, 0], "y": mid[:, 1], "z": mid[:, 2], "cp": loads[0] / q}
    if viscous:
        area = np.linalg.norm(w_face_vectors(atlas), axis=0)
        for c, name in enumerate("xyz"):
            out[f"cf_{name}"] = loads[1 + c] / area / q
    return out


def w_face_vectors(atlas) -> np.ndarray:
    """Area vectors ``(3, n_walls)`` of the 3D wall faces."""
    w = atlas.walls
    out = np.empty((3, w.n))
    for fam in range(3):
        sel = w.family == fam
        out[:, sel] = atlas.faces[fam].reshape(3, -1)[:, w.face[sel]]
    return out


def wall_running_indices(atlas) -> tuple[np.ndarray, np.ndarray]:
    """Block and running indices on its side of every 3D wall face, in the
    atlas's wall order (block by block, wall side by side, the side's first
    running index outermost, as ``mesh.atlas3d`` lists them)."""
    blocks, running = [], []
    for b, blk in enumerate(atlas.mesh):
        for _, rect in blk.sides.walls():
            for x in range(rect.start[0], rect.stop[0]):
                for y in range(rect.start[1], rect.stop[1]):
                    blocks.append(b)
                    running.append((x, y))
    return np.asarray(blocks, np.int64), np.asarray(running, np.int64).reshape(-1, 2)


def surface(ctx, case_index: int = 0) -> dict:
    """Wall distributions of case ``case_index`` of a context holding a state
    and its freestream (the device computes the loads)."""
    loads = ctx.wall_loads()[case_index]
    q_inf = ctx.freestream[case_index]
    if ctx.dim == 3:
        blocks, running = wall_running_indices(ctx.atlas)
        return surface_3d(loads, ctx.atlas, q_inf, ctx.viscous, blocks, running)
    return surface_2d(loads, ctx.atlas, q_inf, ctx.viscous)


def swept_surface(ctx, atlas2, sweep_deg: float) -> dict:
    """Wall distributions of a swept case (the reference's swept columns,
    ``src/jaxrans/post/forces.py:203-226``) from a context on its extruded
    canvas: per wall face of the 2D section (the faces of plane ``k = 0``),
    ``x``, ``x_streamwise`` (``x / cos(sweep)``), ``y``, ``body``, ``cp``
    and, viscous, ``cf`` (in the section's plane, along the downstream
    tangent) and ``cf_z`` (along the sweep line), all over the whole
    freestream's dynamic pressure."""
    loads = ctx.wall_loads()[0]
    at3 = ctx.atlas
    w3, w2 = at3.walls, atlas2.walls
    first = w3.mid[:, 2] < 0.5 * (at3.mesh[0].nodes[2, 0, 0, 1] + at3.mesh[0].nodes[2, 0, 0, 0]) + 1e-12
    order = loop_order(w2)
    mid2 = w2.mid[order]
    # the plane-0 face over each 2D wall face: same (x, y) midpoint
    from scipy.spatial import cKDTree
    idx3 = np.flatnonzero(first)
    _, near = cKDTree(w3.mid[idx3, :2]).query(mid2)
    pick = idx3[near]
    q = dynamic_pressure(ctx.freestream[0], 3)
    out = {"x": mid2[:, 0], "x_streamwise": mid2[:, 0] / np.cos(np.deg2rad(sweep_deg)), "y": mid2[:, 1],
           "body": w2.body[order].astype(float), "cp": loads[0, pick] / q}
    if ctx.viscous:
        S = w_face_vectors(at3)[:, pick] * w3.into_fluid[pick]
        length = np.hypot(S[0], S[1])
        tx, ty = S[1] / length, -S[0] / length
        sign = np.where(tx >= 0.0, 1.0, -1.0)
        out["cf"] = sign * (loads[1, pick] * tx + loads[2, pick] * ty) / length / q
        out["cf_z"] = loads[3, pick] / length / q
    return out
