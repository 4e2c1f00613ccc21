"""The canvas of a 3D mesh: every block with its ghost ring packed into one
array, and the static plans the device kernels consume.

Semantics follow ``src/jaxrans/mesh/atlas3d.py:1-24``, ``119-287``:

* tiles (block + ``g`` ghost layers) packed in the plane of two canvas axes,
  as deep as the deepest along the third, the depth axis chosen to make the
  canvas smallest; flat index ``(i NJ + j) NK + k``;
* geometry: each tile's nodes extended by linearly extrapolated layers, the
  block's own metrics inside; gap cells are inert unit cells;
* the ghost fill as three gather stages per pass, each block's side pairs in
  its ``fill_order`` (its line axis first, then the other two ascending),
  each side extended over the ghosts of the stages before it; the passes
  repeat until the last one reads filled cells only;
* ghost centres: reflections through physical faces, copies across
  interfaces (one-layer plans);
* line chains along each block's ``line_axis`` (normal of most of its wall
  faces, else of its far-field faces), one per pair of the other two
  indices, ending at the block's ends; their links are flat indices into the
  three face families concatenated;
* the wall faces (bilinear quads) and the distance of every cell centre to
  the nearest one.

On top of the reference's plans: ``chains`` composes the staged fill per
final ghost cell -- the interior cell its value comes from and the boundary
tasks applied on the way, in order -- the form the device fill evaluates.
"""

from __future__ import annotations

import os
from dataclasses import dataclass

import numpy as np

from .atlas import KIND_CODES, KIND_COPY, Lines, Stage, _pack
from .block import WALL_KINDS
from .block3d import SIDES_3D, Grid3D, InterfaceRect, metrics3d, normal_axis, running_axes

MAX_PASSES = 4
MAX_OPS = 4


def cell_on(side: str, a, layer: int, shape, inward: bool) -> tuple[int, int, int]:
    """Block cell ``layer`` cells inward from (``inward=False``: beyond)
    ``side`` at running position ``a``."""
    n = normal_axis(side)
    out = [0, 0, 0]
    for m, axis in enumerate(running_axes(side)):
        out[axis] = a[m]
    low = side.endswith("min")
    if inward:
        out[n] = layer if low else shape[n] - 1 - layer
    else:
        out[n] = -1 - layer if low else shape[n] + layer
    return tuple(out)


def _extend_nodes(nodes: np.ndarray, g: int) -> np.ndarray:
    for axis in range(1, 4):
        a = np.moveaxis(nodes, axis, 0)
        lo = [a[0] - (k + 1) * (a[1] - a[0]) for k in range(g)][::-1]
        hi = [a[-1] + (k + 1) * (a[-1] - a[-2]) for k in range(g)]
        nodes = np.moveaxis(np.concatenate([np.stack(lo), a, np.stack(hi)]), 0, axis)
    return nodes


def _side_geometry(block: Grid3D, side: str):
    """Outward unit normals and node-mean midpoints ``(3, na, nb)`` of a side's faces."""
    n = normal_axis(side)
    index = 0 if side.endswith("min") else block.shape[n]
    S = np.take(block.faces[n], index, axis=1 + n)
    nodes = np.take(block.nodes, index, axis=1 + n)
    length = np.linalg.norm(S, axis=0)
    sign = -1.0 if side.endswith("min") else 1.0
    normal = np.where(length > 0.0, sign * S / np.where(length > 0.0, length, 1.0), 0.0)
    mid = 0.25 * (nodes[:, :-1, :-1] + nodes[:, 1:, :-1] + nodes[:, :-1, 1:] + nodes[:, 1:, 1:])
    return normal, mid


def _pack_3d(tiles):
    best = None
    for depth in range(3):
        a, b = (x for x in range(3) if x != depth)
        flat, (na, nb) = _pack([(t[a], t[b]) for t in tiles])
        shape = [0, 0, 0]
        shape[a], shape[b], shape[depth] = na, nb, max(t[depth] for t in tiles)
        if best is None or np.prod(shape) < np.prod(best[1]):
            origins = []
            for x, y in flat:
                o = [0, 0, 0]
                o[a], o[b] = x, y
                origins.append(tuple(o))
            best = (origins, tuple(shape))
    return best


def line_axis(block: Grid3D) -> int:
    """The normal axis of most of the block's wall faces, else of its
    far-field faces, else ``j`` (``src/jaxrans/mesh/atlas3d.py:98-111``)."""
    for kinds in (WALL_KINDS, ("far_field",)):
        count = [0, 0, 0]
        for side, rect in block.sides.boundaries():
            if rect.kind in kinds:
                count[normal_axis(side)] += (rect.stop[0] - rect.start[0]) * (rect.stop[1] - rect.start[1])
        if max(count):
            return int(np.argmax(count))
    return 1


def fill_order(block: Grid3D) -> tuple[int, int, int]:
    first = line_axis(block)
    return (first, *(a for a in range(3) if a != first))


@dataclass(frozen=True)
class Walls3D:
    """Wall faces: interior ``cell``, ``family`` (normal axis), flat ``face``
    index in that family's array, ``into_fluid`` sign (+1 on a min side),
    ``body`` and ``corners (n, 4, 3)`` in running order ``a b c d``."""

    cell: np.ndarray
    family: np.ndarray
    face: np.ndarray
    into_fluid: np.ndarray
    body: np.ndarray
    corners: np.ndarray

    @property
    def n(self) -> int:
        return self.cell.size

    @property
    def mid(self) -> np.ndarray:
        return self.corners.mean(axis=1)

    @property
    def bodies(self) -> int:
        return int(self.body.max()) + 1 if self.body.size else 0


@dataclass(frozen=True)
class FillChains:
    """The composed ghost fill: final ghost ``dest[t]`` takes the primitives
    of interior cell ``root[t]`` through the boundary tasks ``ops[t, :n_ops[t]]``
    (indices into the task table ``kind``, ``normal``, ``midpoint``) in order."""

    dest: np.ndarray
    root: np.ndarray
    n_ops: np.ndarray
    ops: np.ndarray           # (n, MAX_OPS), -1 unused
    kind: np.ndarray          # (n_tasks,)
    normal: np.ndarray        # (3, n_tasks)
    midpoint: np.ndarray      # (3, n_tasks)


@dataclass(frozen=True)
class Atlas3D:
    mesh: tuple
    g: int
    origins: tuple
    windows: tuple
    shape: tuple[int, int, int]
    interior: np.ndarray        # (NI, NJ, NK) bool
    volume: np.ndarray          # (NI, NJ, NK)
    faces: tuple                # (3, NI+1, NJ, NK), (3, NI, NJ+1, NK), (3, NI, NJ, NK+1)
    centres: np.ndarray         # (3, NI+2, NJ+2, NK+2): edge-padded by one, layer-1 ghosts reflected
    stages: tuple               # the staged fill, passes x 3 stages, in order
    chains: FillChains
    neighbours: np.ndarray      # (n_canvas,)
    lines: Lines                # links: flat indices into the concatenated face families
    walls: Walls3D
    distance: np.ndarray        # (NI, NJ, NK)

    dim = 3

    @property
    def ni(self) -> int:
        return self.shape[0]

    @property
    def nj(self) -> int:
        return self.shape[1]

    @property
    def nk(self) -> int:
        return self.shape[2]

    @property
    def n_canvas(self) -> int:
        return int(np.prod(self.shape))

    @property
    def cells(self) -> int:
        return int(self.interior.sum())

    @property
    def face_offsets(self) -> np.ndarray:
        """Start of each family in the concatenated face numbering."""
        return np.cumsum([0] + [int(np.prod(f.shape[1:])) for f in self.faces])

    def pack(self, blocks, fill=0.0) -> np.ndarray:
        lead = blocks[0].shape[:blocks[0].ndim - 3]
        out = np.empty(lead + self.shape, dtype=np.float64)
        out[...] = np.asarray(fill).reshape(np.shape(fill) + (1, 1, 1)) if np.ndim(fill) else fill
        for blk, w in zip(blocks, self.windows):
            out[(Ellipsis,) + w] = blk
        return out

    def unpack(self, canvas: np.ndarray) -> tuple:
        return tuple(canvas[(Ellipsis,) + w] for w in self.windows)


def build_atlas_3d(mesh, g: int = 2) -> Atlas3D:
    """The atlas of a 3D mesh whose patches have conditions assigned."""
    mesh = tuple(mesh)
    unset = sorted({r.patch for b in mesh for _, r in b.sides.boundaries() if r.kind is None})
    if unset:
        raise ValueError(f"patch(es) {unset} have no boundary condition; use assign_conditions")
    origins, canvas = _pack_3d([tuple(n + 2 * g for n in b.shape) for b in mesh])
    NI, NJ, NK = canvas
    n_canvas = NI * NJ * NK

    def pos(b, i, j, k):
        oi, oj, ok = origins[b]
        return ((oi + g + i) * NJ + (oj + g + j)) * NK + (ok + g + k)

    windows = tuple(tuple(slice(o + g, o + g + n) for o, n in zip(origins[b], blk.shape)) for b, blk in enumerate(mesh))
    interior = np.zeros(canvas, dtype=bool)
    for w in windows:
        interior[w] = True

    lattice = np.stack(np.meshgrid(*(np.arange(n, dtype=np.float64) for n in canvas), indexing="ij"))
    centres, volume = 1e3 + 10.0 * lattice, np.ones(canvas)
    faces = [np.zeros((3,) + tuple(n + (a == x) for x, n in enumerate(canvas))) for a in range(3)]
    for blk, origin, w in zip(mesh, origins, windows):
        tc, tv, tf = metrics3d(_extend_nodes(np.asarray(blk.nodes, np.float64), g))
        tile = tuple(slice(o, o + n) for o, n in zip(origin, tv.shape))
        centres[(slice(None),) + tile] = tc
        volume[tile] = np.where(np.abs(tv) > 1e-300, tv, 1.0)
        volume[w] = blk.volume
        for a, f in enumerate(tf):
            faces[a][(slice(None),) + tuple(slice(o, o + n) for o, n in zip(origin, f.shape[1:]))] = f

    interior_flat = interior.ravel()
    plans = {layers: [_stage(mesh, layers, s, pos) for s in range(3)] for layers in (g, 1)}
    passes = {layers: _passes(p, interior_flat) for layers, p in plans.items()}
    flat = centres.reshape(3, -1)
    for stage in plans[1] * passes[1]:
        flat = _reflect(stage, flat)
    centres = flat.reshape((3,) + canvas)
    padded = np.stack([np.pad(c, 1, mode="edge") for c in centres])
    stages = tuple(plans[g] * passes[g])
    walls = _walls(mesh, origins, pos, g, canvas)
    distance = np.ones(canvas)
    for w, d in zip(windows, wall_distance_3d(mesh, walls)):
        distance[w] = d
    return Atlas3D(mesh, g, tuple(origins), windows, canvas, interior, volume, tuple(faces), padded, stages,
                   _compose(stages, interior_flat), _neighbour_map(mesh, pos, n_canvas),
                   _lines(mesh, origins, g, canvas), walls, distance)


def _stage(mesh, g: int, stage: int, pos) -> Stage:
    """Tasks of fill stage ``stage``: every block's side pair of that place
    in its fill order, each side extended ``g`` cells along running axes that
    an earlier stage filled (``src/jaxrans/mesh/atlas3d.py:160-205``)."""
    dest, src, kind, normal, midpoint = [], [], [], [], []
    for b, blk in enumerate(mesh):
        shape = blk.shape
        order = fill_order(blk)
        handled = set(order[:stage])
        name = "ijk"[order[stage]]
        for side in (f"{name}_min", f"{name}_max"):
            nrm, mid = _side_geometry(blk, side)
            axes = running_axes(side)
            n = tuple(shape[a] for a in axes)
            for rect in blk.sides.side(side):
                ranges = []
                for m, axis in enumerate(axes):
                    ext = g if axis in handled else 0
                    lo = rect.start[m] - (ext if rect.start[m] == 0 else 0)
                    hi = rect.stop[m] + (ext if rect.stop[m] == n[m] else 0)
                    ranges.append(range(lo, hi))
                for x in ranges[0]:
                    for y in ranges[1]:
                        a = (x, y)
                        f = (min(max(x, 0), n[0] - 1), min(max(y, 0), n[1] - 1))
                        for layer in range(g):
                            dest.append(pos(b, *cell_on(side, a, layer, shape, inward=False)))
                            if isinstance(rect, InterfaceRect):
                                partner = mesh[rect.block].shape
                                p = list(rect.partner(a))
                                for m, axis in enumerate(running_axes(rect.side)):
                                    corner = g if axes[rect.axes[m]] in handled else 0
                                    p[m] = min(max(p[m], -corner), partner[axis] - 1 + corner)
                                src.append(pos(rect.block, *cell_on(rect.side, p, layer, partner, inward=True)))
                                kind.append(KIND_COPY)
                            else:
                                read = layer if rect.kind in (*WALL_KINDS, "symmetry") else 0
                                src.append(pos(b, *cell_on(side, a, read, shape, inward=True)))
                                kind.append(KIND_CODES[rect.kind])
                            normal.append(nrm[:, f[0], f[1]])
                            midpoint.append(mid[:, f[0], f[1]])
    return Stage(np.array(dest, np.int64), np.array(src, np.int64), np.array(kind, np.int32),
                 np.array(normal, np.float64).reshape(-1, 3).T.copy(),
                 np.array(midpoint, np.float64).reshape(-1, 3).T.copy())


def _passes(stages, interior: np.ndarray) -> int:
    """Passes of the stages after which the last one reads filled cells only."""
    filled = interior.copy()
    for n in range(1, MAX_PASSES + 1):
        clean = True
        for st in stages:
            ok = filled[st.src]
            clean &= bool(np.all(ok))
            nxt = filled.copy()
            nxt[st.dest] = ok
            filled = nxt
        if clean:
            return n
    raise ValueError(f"the ghost fill does not settle in {MAX_PASSES} passes: an interface map is inconsistent")


def _reflect(stage: Stage, centres: np.ndarray) -> np.ndarray:
    out = centres.copy()
    x = centres[:, stage.src]
    bc = stage.kind != KIND_COPY
    d = np.sum((x - stage.midpoint) * stage.normal, axis=0)
    out[:, stage.dest] = np.where(bc[None], x - 2.0 * d * stage.normal, x)
    return out


def _compose(stages, interior: np.ndarray) -> FillChains:
    """Run the staged fill symbolically: every canvas cell holds (root cell,
    boundary tasks applied); a stage gathers from the state before it."""
    n = interior.size
    root = np.where(interior, np.arange(n), -1)
    ops = np.full((n, MAX_OPS), -1, np.int64)
    n_ops = np.zeros(n, np.int64)
    written = np.zeros(n, bool)
    task_kind, task_normal, task_mid = [], [], []
    offset = 0
    for st in stages:
        r, o, c = root[st.src], ops[st.src].copy(), n_ops[st.src].copy()
        bc = np.flatnonzero(st.kind != KIND_COPY)
        ids = offset + np.arange(st.dest.size)
        if bc.size:
            if np.any(c[bc] >= MAX_OPS):
                raise ValueError(f"a ghost composes more than {MAX_OPS} boundary tasks")
            o[bc, c[bc]] = ids[bc]
            c[bc] += 1
        new_root, new_ops, new_n = root.copy(), ops.copy(), n_ops.copy()
        new_root[st.dest], new_ops[st.dest], new_n[st.dest] = r, o, c
        root, ops, n_ops = new_root, new_ops, new_n
        written[st.dest] = True
        task_kind.append(st.kind)
        task_normal.append(st.normal)
        task_mid.append(st.midpoint)
        offset += st.dest.size
    dest = np.flatnonzero(written & ~interior)
    if np.any(root[dest] < 0):
        raise ValueError("a ghost cell's fill does not reach an interior cell")
    return FillChains(dest, root[dest], n_ops[dest], ops[dest], np.concatenate(task_kind),
                      np.concatenate(task_normal, axis=1), np.concatenate(task_mid, axis=1))


def _neighbour_map(mesh, pos, n_canvas: int) -> np.ndarray:
    """Across every block face, the cell on the other side, or the cell itself."""
    index = np.arange(n_canvas)
    for b, blk in enumerate(mesh):
        for side in SIDES_3D:
            for rect in blk.sides.side(side):
                for x in range(rect.start[0], rect.stop[0]):
                    for y in range(rect.start[1], rect.stop[1]):
                        ghost = pos(b, *cell_on(side, (x, y), 0, blk.shape, inward=False))
                        if isinstance(rect, InterfaceRect):
                            partner = mesh[rect.block].shape
                            index[ghost] = pos(rect.block, *cell_on(rect.side, rect.partner((x, y)), 0, partner,
                                                                   inward=True))
                        else:
                            index[ghost] = pos(b, *cell_on(side, (x, y), 0, blk.shape, inward=True))
    return index


def _lines(mesh, origins, g: int, canvas) -> Lines:
    """One chain per block and pair of the other indices, along the block's
    line axis; ragged (``Lines``), links into the concatenated face families."""
    dims = [tuple(n + (a == x) for x, n in enumerate(canvas)) for a in range(3)]
    offsets = np.cumsum([0] + [int(np.prod(d)) for d in dims])
    cells_all, faces_all, starts = [], [], [0]
    for blk, origin in zip(mesh, origins):
        axis = line_axis(blk)
        idx = np.stack(np.meshgrid(*(o + g + np.arange(n) for o, n in zip(origin, blk.shape)), indexing="ij"))
        idx = np.moveaxis(idx, 1 + axis, -1).reshape(3, -1, blk.shape[axis])
        above = idx.copy()
        above[axis] += 1
        cells = np.ravel_multi_index(tuple(idx), canvas)                   # (chains, L)
        faces = offsets[axis] + np.ravel_multi_index(tuple(above), dims[axis])
        for c, f in zip(cells, faces):
            cells_all.append(c)
            faces_all.append(f)
            starts.append(starts[-1] + c.size)
    cells = np.concatenate(cells_all).astype(np.int64)
    up_face = np.concatenate(faces_all).astype(np.int64)
    start = np.array(starts, np.int64)
    up_sign = np.ones(cells.size, np.int64)
    up_sign[start[1:] - 1] = 0
    up_face = np.where(up_sign != 0, up_face, 0)
    return Lines(start, cells, up_face, up_sign, up_face.copy(), -up_sign)


def _walls(mesh, origins, pos, g: int, canvas) -> Walls3D:
    cell, family, face, sign, corners = [], [], [], [], []
    for b, blk in enumerate(mesh):
        nodes = blk.nodes
        for side, rect in blk.sides.walls():
            n = normal_axis(side)
            plane = np.take(nodes, 0 if side.endswith("min") else blk.shape[n], axis=1 + n)
            dims = list(canvas)
            dims[n] += 1
            for x in range(rect.start[0], rect.stop[0]):
                for y in range(rect.start[1], rect.stop[1]):
                    c = cell_on(side, (x, y), 0, blk.shape, inward=True)
                    cell.append(pos(b, *c))
                    f = [o + g + v for o, v in zip(origins[b], c)]
                    f[n] += side.endswith("max")
                    family.append(n)
                    face.append((f[0] * dims[1] + f[1]) * dims[2] + f[2])
                    sign.append(1.0 if side.endswith("min") else -1.0)
                    corners.append(np.stack([plane[:, x, y], plane[:, x + 1, y], plane[:, x + 1, y + 1],
                                             plane[:, x, y + 1]]))
    if not cell:
        z = np.zeros(0, np.int64)
        return Walls3D(z, z, z, np.zeros(0), z, np.zeros((0, 4, 3)))
    return Walls3D(np.array(cell, np.int64), np.array(family, np.int64), np.array(face, np.int64), np.array(sign),
                   np.zeros(len(cell), np.int64), np.stack(corners))


def _segment_distance(p, s, e):
    edge = e - s
    t = np.clip(np.sum((p - s) * edge, -1) / np.maximum(np.sum(edge * edge, -1), 1e-300), 0.0, 1.0)
    r = p - s - t[..., None] * edge
    return np.sqrt(np.sum(r * r, -1))


def patch_distance(p: np.ndarray, quads: np.ndarray, iterations: int = 12) -> np.ndarray:
    """Distance from points ``p (..., 3)`` to bilinear patches ``(..., 4, 3)``
    (corners ``a b c d``): Newton from the centre for the interior stationary
    point of ``|x(u, v) - p|^2``, kept if inside the unit square, and the four
    straight edges (``src/jaxrans/mesh/wall3d.py:112-137``)."""
    a, b, c, d = (quads[..., k, :] for k in range(4))
    eu, ev, w = b - a, d - a, a - b + c - d
    u = np.full(p.shape[:-1], 0.5)
    v = np.full(p.shape[:-1], 0.5)
    for _ in range(iterations):
        x_u = eu + v[..., None] * w
        x_v = ev + u[..., None] * w
        r = a + u[..., None] * eu + v[..., None] * ev + (u * v)[..., None] * w - p
        g_u, g_v = np.sum(r * x_u, -1), np.sum(r * x_v, -1)
        h_uu, h_vv = np.sum(x_u * x_u, -1), np.sum(x_v * x_v, -1)
        h_uv = np.sum(x_u * x_v, -1) + np.sum(r * w, -1)
        det = h_uu * h_vv - h_uv * h_uv
        ok = np.abs(det) > 1e-300
        safe = np.where(ok, det, 1.0)
        u = np.where(ok, u - (h_vv * g_u - h_uv * g_v) / safe, u)
        v = np.where(ok, v - (h_uu * g_v - h_uv * g_u) / safe, v)
        u, v = np.clip(u, -0.5, 1.5), np.clip(v, -0.5, 1.5)
    r = a + u[..., None] * eu + v[..., None] * ev + (u * v)[..., None] * w - p
    inside = (u >= 0.0) & (u <= 1.0) & (v >= 0.0) & (v <= 1.0)
    best = np.where(inside, np.sqrt(np.sum(r * r, -1)), np.inf)
    for s_, e_ in ((a, b), (b, c), (c, d), (d, a)):
        best = np.minimum(best, _segment_distance(p, s_, e_))
    return best


def _nearest_patches(args):
    pts, mids, corners, k = args
    from scipy.spatial import cKDTree
    _, idx = cKDTree(mids).query(pts, k=k)
    return patch_distance(pts[:, None, :], corners[idx.reshape(len(pts), k)]).min(axis=1)


def wall_distance_3d(mesh, walls: Walls3D, candidates: int = 32, chunk: int = 8192,
                     workers: int | None = None) -> list[np.ndarray]:
    """Distance from every cell centre to the nearest wall quad: the exact
    patch distance over the ``candidates`` quads whose corner-mean is
    nearest (a k-d tree); the reference ranks candidates by the distance to
    each quad's two triangles instead (``wall3d.py:140-158``).  Large grids
    spread the chunks over a process pool (``workers``: default the CPU
    count, at most 16)."""
    if walls.n == 0:
        return [np.full(b.shape, np.inf) for b in mesh]
    k = min(candidates, walls.n)
    pts = np.concatenate([np.moveaxis(b.centres, 0, -1).reshape(-1, 3) for b in mesh])
    jobs = [(pts[s:s + chunk], walls.mid, walls.corners, k) for s in range(0, len(pts), chunk)]
    n = min(16, os.cpu_count() or 1) if workers is None else workers
    if n > 1 and len(jobs) > 4:
        import multiprocessing as mp
        with mp.get_context("fork").Pool(n) as pool:
            parts = pool.map(_nearest_patches, jobs, chunksize=1)
    else:
        parts = [_nearest_patches(j) for j in jobs]
    d = np.concatenate(parts)
    out, s = [], 0
    for b in mesh:
        n_b = int(np.prod(b.shape))
        out.append(d[s:s + n_b].reshape(b.shape))
        s += n_b
    return out
