"""The canvas: every block of a 2D mesh with its ghost ring packed into one
array, and the static plans the device kernels consume.

Layout and plans follow ``src/jaxrans/mesh/atlas.py:1-19``, ``189-228``:

* block ``b`` cell ``(i, j)`` (ghosts ``-g <= i < ni + g``) at canvas cell
  ``(oi_b + g + i, oj_b + g + j)``; canvas cells are interior, ghost or gap;
* the ghost fill in two gather stages -- ``j`` sides, then ``i`` sides
  extended ``g`` cells over the corners (``atlas.py:331-367``);
* the face-neighbour map behind face averages (``atlas.py:163-183``, ``381-395``);
* the line chains along ``j``, continued through a block's own ``j``-side
  cut (``atlas.py:398-470``);
* the wall faces (``atlas.py:473-489``) and the exact wall distance
  (``src/jaxrans/mesh/mesh.py:163-192``).

Everything here is NumPy, built once on the host, and handed to the device
library as flat integer/double tables.
"""

from __future__ import annotations

from dataclasses import dataclass

import numpy as np

from .block import OUTWARD, SIDES, WALL_KINDS, Grid, Interface, metrics2d

# Ghost-fill task kinds, shared with the device library.
KIND_COPY = 0
KIND_CODES = {"slip_wall": 1, "no_slip_wall": 2, "symmetry": 3, "far_field": 4,
              "inflow_total": 5, "outflow_static": 6}


def interior_cell(side: str, a: int, layer: int, ni: int, nj: int) -> tuple[int, int]:
    return {"j_min": (a, layer), "j_max": (a, nj - 1 - layer),
            "i_min": (layer, a), "i_max": (ni - 1 - layer, a)}[side]


def ghost_cell(side: str, a: int, layer: int, ni: int, nj: int) -> tuple[int, int]:
    return {"j_min": (a, -1 - layer), "j_max": (a, nj + layer),
            "i_min": (-1 - layer, a), "i_max": (ni + layer, a)}[side]


@dataclass(frozen=True)
class Stage:
    """One ghost-fill stage as a task list: ``dest`` canvas cell <- ``src``
    canvas cell, copied (``kind`` 0) or through the boundary condition
    ``kind`` with the face's outward unit ``normal`` and ``midpoint``."""

    dest: np.ndarray
    src: np.ndarray
    kind: np.ndarray
    normal: np.ndarray      # (2, n)
    midpoint: np.ndarray    # (2, n)


@dataclass(frozen=True)
class Lines:
    """Ragged chains: chain ``c`` occupies slots ``start[c]:start[c+1]``.

    ``cells[s]`` is the canvas cell of slot ``s``; between slots ``s`` and
    ``s+1`` of one chain, slot ``s`` couples across j-face ``up_face[s]``
    (flat index into the j-face array) with outward sign ``up_sign[s]``, and
    slot ``s+1`` across ``down_face[s]`` with ``down_sign[s]``; the last slot
    of a chain has sign 0."""

    start: np.ndarray
    cells: np.ndarray
    up_face: np.ndarray
    up_sign: np.ndarray
    down_face: np.ndarray
    down_sign: np.ndarray

    @property
    def n_lines(self) -> int:
        return self.start.size - 1

    @property
    def max_length(self) -> int:
        return int(np.diff(self.start).max())


@dataclass(frozen=True)
class Walls:
    """Wall faces: interior ``cell``, ``family`` (0 i, 1 j), flat ``face``
    index in that family's array, ``into_fluid`` sign, ``body`` index, and
    end nodes ``a``, ``b`` (fluid on the left)."""

    cell: np.ndarray
    family: np.ndarray
    face: np.ndarray
    into_fluid: np.ndarray
    body: np.ndarray
    a: np.ndarray
    b: np.ndarray

    @property
    def n(self) -> int:
        return self.cell.size

    @property
    def mid(self) -> np.ndarray:
        return 0.5 * (self.a + self.b)

    @property
    def bodies(self) -> int:
        return int(self.body.max()) + 1 if self.body.size else 0


@dataclass(frozen=True)
class Atlas:
    mesh: tuple
    g: int
    origins: tuple
    windows: tuple
    shape: tuple[int, int]
    interior: np.ndarray        # (NI, NJ) bool
    area: np.ndarray            # (NI, NJ)
    face_i: np.ndarray          # (2, NI+1, NJ)
    face_j: np.ndarray          # (2, NI, NJ+1)
    centres: np.ndarray         # (2, NI+2, NJ+2): edge-padded by one, layer-1 ghosts reflected
    stages: tuple
    neighbours: np.ndarray      # (NI*NJ,) int64
    lines: Lines
    walls: Walls
    distance: np.ndarray        # (NI, NJ) wall distance (1 outside blocks, inf without walls)

    dim = 2

    @property
    def ni(self) -> int:
        return self.shape[0]

    @property
    def nj(self) -> int:
        return self.shape[1]

    @property
    def n_canvas(self) -> int:
        return self.shape[0] * self.shape[1]

    @property
    def cells(self) -> int:
        return int(self.interior.sum())

    def pack(self, blocks, fill=0.0) -> np.ndarray:
        lead = blocks[0].shape[:blocks[0].ndim - 2]
        out = np.empty(lead + self.shape, dtype=np.float64)
        out[...] = np.asarray(fill).reshape(np.shape(fill) + (1, 1)) if np.ndim(fill) else fill
        for blk, (wi, wj) in zip(blocks, self.windows):
            out[..., wi, wj] = blk
        return out

    def unpack(self, canvas: np.ndarray) -> tuple:
        return tuple(canvas[..., wi, wj] for wi, wj in self.windows)


def _pack(shapes):
    """Tiles, tallest first, stacked in columns as tall as the tallest tile."""
    height = max(h for _, h in shapes)
    order = sorted(range(len(shapes)), key=lambda k: (-shapes[k][1], -shapes[k][0]))
    origins = [None] * len(shapes)
    cols: list[list[int]] = []
    for k in order:
        w, h = shapes[k]
        for col in cols:
            if w <= col[1] and col[2] + h <= height:
                origins[k] = (col[0], col[2])
                col[2] += h
                break
        else:
            x = cols[-1][0] + cols[-1][1] if cols else 0
            cols.append([x, w, h])
            origins[k] = (x, 0)
    return origins, (cols[-1][0] + cols[-1][1], height)


def _extend(a: np.ndarray, g: int) -> np.ndarray:
    """Nodes extended by ``g`` linearly extrapolated layers on every side."""
    for axis in (0, 1):
        a = np.moveaxis(a, axis, 0)
        lo = [a[0] - (k + 1) * (a[1] - a[0]) for k in range(g)][::-1]
        hi = [a[-1] + (k + 1) * (a[-1] - a[-2]) for k in range(g)]
        a = np.moveaxis(np.concatenate([np.stack(lo), a, np.stack(hi)]), 0, axis)
    return a


def _side_geometry(block: Grid, side: str):
    """Outward unit normals and midpoints of a side's faces."""
    if side.startswith("j"):
        j = 0 if side == "j_min" else block.nj
        sx, sy = block.sx_j[:, j], block.sy_j[:, j]
        x, y = block.x_node[:, j], block.y_node[:, j]
    else:
        i = 0 if side == "i_min" else block.ni
        sx, sy = block.sx_i[i, :], block.sy_i[i, :]
        x, y = block.x_node[i, :], block.y_node[i, :]
    length = np.hypot(sx, sy)
    sign = OUTWARD[side]
    return sign * sx / length, sign * sy / length, 0.5 * (x[:-1] + x[1:]), 0.5 * (y[:-1] + y[1:])


def build_atlas(mesh, g: int = 2) -> Atlas:
    mesh = (mesh,) if isinstance(mesh, Grid) else tuple(mesh)
    unset = sorted({seg.patch for b in mesh for _, seg in b.sides.boundaries() if seg.kind is None})
    if unset:
        raise ValueError(f"patch(es) {unset} have no boundary condition; use assign_conditions")
    shapes = [(b.ni + 2 * g, b.nj + 2 * g) for b in mesh]
    origins, (NI, NJ) = _pack(shapes)

    def pos(b, i, j):
        oi, oj = origins[b]
        return (oi + g + i) * NJ + (oj + g + j)

    interior = np.zeros((NI, NJ), dtype=bool)
    windows = []
    for b, blk in enumerate(mesh):
        oi, oj = origins[b]
        w = (slice(oi + g, oi + g + blk.ni), slice(oj + g, oj + g + blk.nj))
        interior[w] = True
        windows.append(w)

    # Geometry: real faces keep their block's exact vectors; ghosts get the
    # geometry of linearly extrapolated nodes; gaps are inert.
    area = np.ones((NI, NJ))
    li, lj = np.meshgrid(np.arange(NI), np.arange(NJ), indexing="ij")
    xc, yc = 1e3 + 10.0 * li, 1e3 + 10.0 * lj
    fi = np.zeros((2, NI + 1, NJ))
    fj = np.zeros((2, NI, NJ + 1))
    for b, blk in enumerate(mesh):
        oi, oj = origins[b]
        ti, tj = shapes[b]
        txc, tyc, tarea, (tsxi, tsyi), (tsxj, tsyj) = metrics2d(_extend(blk.x_node, g), _extend(blk.y_node, g))
        area[oi:oi + ti, oj:oj + tj] = np.where(np.abs(tarea) > 1e-300, tarea, 1.0)
        xc[oi:oi + ti, oj:oj + tj], yc[oi:oi + ti, oj:oj + tj] = txc, tyc
        fi[0, oi:oi + ti + 1, oj:oj + tj], fi[1, oi:oi + ti + 1, oj:oj + tj] = tsxi, tsyi
        fj[0, oi:oi + ti, oj:oj + tj + 1], fj[1, oi:oi + ti, oj:oj + tj + 1] = tsxj, tsyj
        wi, wj = windows[b]
        area[wi, wj] = blk.area
        fi[0, wi.start:wi.stop + 1, wj], fi[1, wi.start:wi.stop + 1, wj] = blk.sx_i, blk.sy_i
        fj[0, wi, wj.start:wj.stop + 1], fj[1, wi, wj.start:wj.stop + 1] = blk.sx_j, blk.sy_j
        xc[wi, wj], yc[wi, wj] = blk.xc, blk.yc

    n_canvas = NI * NJ
    centres = np.stack([xc.ravel(), yc.ravel()])
    for stage in (_stage(mesh, 1, ("j_min", "j_max"), pos, False), _stage(mesh, 1, ("i_min", "i_max"), pos, True)):
        centres = _reflect(stage, centres)
    centres = centres.reshape(2, NI, NJ)
    padded_centres = np.stack([np.pad(c, 1, mode="edge") for c in centres])

    stages = (_stage(mesh, g, ("j_min", "j_max"), pos, False), _stage(mesh, g, ("i_min", "i_max"), pos, True))
    walls = _walls(mesh, pos, origins, g, NI, NJ)
    distance = np.ones((NI, NJ))
    for (wi, wj), d in zip(windows, wall_distance(mesh, walls)):
        distance[wi, wj] = d
    return Atlas(tuple(mesh), g, tuple(origins), tuple(windows), (NI, NJ), interior, area, fi, fj,
                 padded_centres, stages, _neighbour_map(mesh, pos, n_canvas), _lines(mesh, pos, origins, g, NI, NJ),
                 walls, distance)


def _stage(mesh, g: int, sides, pos, pad: bool) -> Stage:
    """Tasks of one fill stage.  On the i sides (``pad``) each side extends
    ``g`` cells into the j ghost rows of stage 1, so the corners take the
    neighbour's values, or the condition applied to the corner column."""
    dest, src, kind, normal, midpoint = [], [], [], [], []
    for b, blk in enumerate(mesh):
        ni, nj = blk.ni, blk.nj
        for side in sides:
            segs = blk.sides.side(side)
            n = blk.sides.length(side)
            ext = g if pad else 0
            nx, ny, mx, my = _side_geometry(blk, side)
            for k, seg in enumerate(segs):
                lo = seg.start - (ext if k == 0 else 0)
                hi = seg.stop + (ext if k == len(segs) - 1 else 0)
                for a in range(lo, hi):
                    face = min(max(a, 0), n - 1)
                    for layer in range(g):
                        dest.append(pos(b, *ghost_cell(side, a, layer, ni, nj)))
                        if isinstance(seg, Interface):
                            partner = mesh[seg.block]
                            span = partner.sides.length(seg.side)
                            corner = g if (pad and seg.side.startswith("i")) else 0
                            p = min(max(seg.partner(a), -corner), span - 1 + corner)
                            src.append(pos(seg.block, *interior_cell(seg.side, p, layer, partner.ni, partner.nj)))
                            kind.append(KIND_COPY)
                        else:
                            read = layer if seg.kind in (*WALL_KINDS, "symmetry") else 0
                            src.append(pos(b, *interior_cell(side, a, read, ni, nj)))
                            kind.append(KIND_CODES[seg.kind])
                        normal.append((nx[face], ny[face]))
                        midpoint.append((mx[face], my[face]))
    return Stage(np.array(dest, np.int64), np.array(src, np.int64), np.array(kind, np.int32),
                 np.array(normal, np.float64).reshape(-1, 2).T.copy(),
                 np.array(midpoint, np.float64).reshape(-1, 2).T.copy())


def _reflect(stage: Stage, centres: np.ndarray) -> np.ndarray:
    """A centre stage: copies across interfaces, reflections through physical faces."""
    out = centres.copy()
    x = centres[:, stage.src]
    bc = stage.kind != KIND_COPY
    d = np.sum((x - stage.midpoint) * stage.normal, axis=0)
    val = np.where(bc[None], x - 2.0 * d * stage.normal, x)
    out[:, stage.dest] = val
    return out


def _neighbour_map(mesh, pos, n_canvas: int) -> np.ndarray:
    index = np.arange(n_canvas)
    for b, blk in enumerate(mesh):
        for side in SIDES:
            for seg in blk.sides.side(side):
                for a in range(seg.start, seg.stop):
                    ghost = pos(b, *ghost_cell(side, a, 0, blk.ni, blk.nj))
                    if isinstance(seg, Interface):
                        partner = mesh[seg.block]
                        index[ghost] = pos(seg.block, *interior_cell(seg.side, seg.partner(a), 0, partner.ni, partner.nj))
                    else:
                        index[ghost] = pos(b, *interior_cell(side, a, 0, blk.ni, blk.nj))
    return index


def _lines(mesh, pos, origins, g, NI, NJ) -> Lines:
    """Chains of cells linked along j; a block's own j-side cut continues them."""

    def jface(b, i, j):
        oi, oj = origins[b]
        return (oi + g + i) * (NJ + 1) + (oj + g + j)

    links: dict = {}

    def link(cell, other, face, sign):
        links.setdefault(cell, []).append((other, face, sign))

    for b, blk in enumerate(mesh):
        ni, nj = blk.ni, blk.nj
        for i in range(ni):
            for j in range(nj - 1):
                link((b, i, j), (b, i, j + 1), jface(b, i, j + 1), 1)
                link((b, i, j + 1), (b, i, j), jface(b, i, j + 1), -1)
        for side, j, fj, sign in (("j_min", 0, 0, -1), ("j_max", nj - 1, nj, 1)):
            for seg in blk.sides.side(side):
                if not isinstance(seg, Interface) or not seg.side.startswith("j") or seg.block != b:
                    continue
                partner = mesh[seg.block]
                for a in range(seg.start, seg.stop):
                    pj = 0 if seg.side == "j_min" else partner.nj - 1
                    link((b, a, j), (seg.block, seg.partner(a), pj), jface(b, a, fj), sign)

    def nbrs(cell):
        return [r[0] for r in links.get(cell, [])]

    seen: set = set()
    chains = []
    cells = [(b, i, j) for b, blk in enumerate(mesh) for i in range(blk.ni) for j in range(blk.nj)]
    for start in [c for c in cells if len(nbrs(c)) < 2] + cells:
        if start in seen:
            continue
        chain, prev, cell = [], None, start
        while cell is not None and cell not in seen:
            seen.add(cell)
            chain.append(cell)
            nxt = [c for c in nbrs(cell) if c != prev and c not in seen]
            prev, cell = cell, (nxt[0] if nxt else None)
        chains.append(chain)

    total = sum(len(c) for c in chains)
    start = np.zeros(len(chains) + 1, np.int64)
    out = {k: np.zeros(total, np.int64) for k in ("cells", "up_face", "up_sign", "down_face", "down_sign")}
    s = 0
    for n, chain in enumerate(chains):
        start[n] = s
        for k, cell in enumerate(chain):
            out["cells"][s] = pos(*cell)
            if k + 1 < len(chain):
                nxt = chain[k + 1]
                for other, face, sign in links[cell]:
                    if other == nxt:
                        out["up_face"][s], out["up_sign"][s] = face, sign
                for other, face, sign in links[nxt]:
                    if other == cell:
                        out["down_face"][s], out["down_sign"][s] = face, sign
            s += 1
    start[-1] = s
    return Lines(start, out["cells"], out["up_face"], out["up_sign"], out["down_face"], out["down_sign"])


def _walls(mesh, pos, origins, g, NI, NJ) -> Walls:
    rows = []
    for b, blk in enumerate(mesh):
        for side, seg in blk.sides.walls():
            if side.startswith("j"):
                j = 0 if side == "j_min" else blk.nj
                xs, ys = blk.x_node[seg.start:seg.stop + 1, j], blk.y_node[seg.start:seg.stop + 1, j]
            else:
                i = 0 if side == "i_min" else blk.ni
                xs, ys = blk.x_node[i, seg.start:seg.stop + 1], blk.y_node[i, seg.start:seg.stop + 1]
            forward = side in ("j_min", "i_max")
            for n in range(seg.stop - seg.start):
                a_ = seg.start + n
                i, j = interior_cell(side, a_, 0, blk.ni, blk.nj)
                ci, cj = origins[b][0] + g + i, origins[b][1] + g + j
                if side.startswith("j"):
                    fam, face = 1, ci * (NJ + 1) + cj + (side == "j_max")
                else:
                    fam, face = 0, (ci + (side == "i_max")) * NJ + cj
                p, q = (xs[n], ys[n]), (xs[n + 1], ys[n + 1])
                a, bb = (p, q) if forward else (q, p)
                rows.append((pos(b, i, j), fam, face, 1.0 if side in ("j_min", "i_min") else -1.0, *a, *bb))
    if not rows:
        z = np.zeros(0, np.int64)
        return Walls(z, z, z, np.zeros(0), z, np.zeros((0, 2)), np.zeros((0, 2)))
    r = np.array(rows)
    a, bb = r[:, 4:6], r[:, 6:8]
    body = _bodies(a, bb)
    return Walls(r[:, 0].astype(np.int64), r[:, 1].astype(np.int64), r[:, 2].astype(np.int64), r[:, 3].copy(),
                 body, a, bb)


def _bodies(a: np.ndarray, b: np.ndarray) -> np.ndarray:
    """Group wall faces into bodies by following end node -> start node."""
    tol = 1e-3 * np.hypot(*(b - a).T).min()
    key = lambda p: tuple(np.round(p / tol).astype(np.int64))  # noqa: E731
    start_of = {key(p): n for n, p in enumerate(a)}
    nxt = np.array([start_of.get(key(p), -1) for p in b])
    body = np.full(len(a), -1)
    loops = []
    for s in range(len(a)):
        if body[s] >= 0:
            continue
        f, loop = s, []
        while f >= 0 and body[f] < 0:
            body[f] = len(loops)
            loop.append(f)
            f = nxt[f]
        loops.append(loop)
    order = np.argsort([a[loop, 0].min() for loop in loops])
    out = np.empty(len(a), np.int64)
    for new, old in enumerate(order):
        out[loops[old]] = new
    return out


def quarter_chord(walls: Walls) -> tuple[float, float]:
    """The point-vortex centre (``Atlas.quarter_chord``,
    ``src/jaxrans/mesh/atlas.py:492-507``): on the body of longest chord, a
    quarter of the way from its leading edge (the wall node farthest from
    the trailing edge) to its trailing edge (its wall node of largest ``x``,
    the first in wall order on a tie); ``(0.25, 0)`` without walls."""
    if walls.n == 0:
        return 0.25, 0.0
    chords = []
    for body in range(walls.bodies):
        pts = walls.a[walls.body == body]
        te = pts[int(np.argmax(pts[:, 0]))]
        le = pts[int(np.argmax(np.hypot(pts[:, 0] - te[0], pts[:, 1] - te[1])))]
        chords.append((float(np.hypot(te[0] - le[0], te[1] - le[1])), te, le))
    _, te, le = max(chords, key=lambda c: c[0])
    return float(le[0] + 0.25 * (te[0] - le[0])), float(le[1] + 0.25 * (te[1] - le[1]))


def wall_distance(mesh, walls: Walls, chunk: int = 4096) -> list[np.ndarray]:
    """Exact distance from each block's cell centres to the nearest wall face
    (clamped projection onto every segment ``a -> b``)."""
    if walls.n == 0:
        return [np.full(blk.shape, np.inf) for blk in mesh]
    ax, ay = walls.a.T
    ex, ey = (walls.b - walls.a).T
    len2 = ex * ex + ey * ey
    out = []
    for blk in mesh:
        px, py = blk.xc.ravel(), blk.yc.ravel()
        d = np.empty_like(px)
        for s in range(0, px.size, chunk):
            qx = px[s:s + chunk, None] - ax[None]
            qy = py[s:s + chunk, None] - ay[None]
            t = np.clip((qx * ex + qy * ey) / len2, 0.0, 1.0)
            d[s:s + chunk] = np.sqrt(np.min((qx - t * ex) ** 2 + (qy - t * ey) ** 2, axis=1))
        out.append(d.reshape(blk.shape))
    return out


def patches(mesh) -> list[str]:
    names: list[str] = []
    for blk in mesh:
        for _, seg in blk.sides.boundaries():
            if seg.patch not in names:
                names.append(seg.patch)
    return names


def assign_conditions(mesh, kinds: dict[str, str]) -> tuple:
    missing = [p for p in patches(mesh) if p not in kinds]
    if missing:
        raise ValueError(f"no boundary condition for patch(es) {missing}; the grid's patches are {patches(mesh)}")
    return tuple(blk.with_sides(blk.sides.with_conditions(kinds)) for blk in mesh)
