"""3D structured blocks: node geometry, metrics, and what lies beyond each side.

Conventions are the reference's (``src/jaxrans/mesh/block3d.py:1-24``,
``src/jaxrans/mesh/block.py:225-268``):

* nodes ``(3, ni+1, nj+1, nk+1)``, cells ``(ni, nj, nk)``, right-handed;
* the face of family ``a`` spanned by its corners ``pa pb pc pd`` (in the
  cyclic order of the two running indices after ``a``) has the vector area
  ``S = (pc - pa) x (pd - pb) / 2``, pointing toward increasing ``a``; the six
  vectors of a cell then sum to zero exactly and a collapsed face has
  ``S = 0``;
* volumes by the divergence theorem ``V = (1/3) sum_f S_f . x_f`` (outward
  ``S``, ``x_f`` the face's node mean) -- the trilinear hexahedron's volume;
  centres the mean of the eight nodes;
* sides ``i_min`` ... ``k_max``, each tiled by rectangles over its two running
  indices (ascending axes): ``BoundaryRect`` patches and ``InterfaceRect``
  maps onto a side of a block (this one or another).
"""

from __future__ import annotations

from dataclasses import dataclass, replace

import numpy as np

from .block import CONDITION_KINDS, WALL_KINDS

SIDES_3D = ("j_min", "j_max", "i_min", "i_max", "k_min", "k_max")
AXIS = {"i": 0, "j": 1, "k": 2}


def normal_axis(side: str) -> int:
    return AXIS[side[0]]


def running_axes(side: str) -> tuple[int, int]:
    n = normal_axis(side)
    return tuple(a for a in range(3) if a != n)


@dataclass(frozen=True)
class BoundaryRect:
    """Running cells ``[start, stop)`` of a side on patch ``patch``."""

    patch: str
    start: tuple[int, int]
    stop: tuple[int, int]
    kind: str | None = None

    def __post_init__(self) -> None:
        if self.kind is not None and self.kind not in CONDITION_KINDS:
            raise ValueError(f"unknown boundary condition {self.kind!r}; known: {CONDITION_KINDS}")


@dataclass(frozen=True)
class InterfaceRect:
    """Running cells ``[start, stop)`` facing side ``side`` of block ``block``:
    running position ``a`` meets the partner's ``p[m] = origin[m] + s_m
    (a[axes[m]] - start[axes[m]])`` (``s_m = -1`` where ``flip[m]``; modulo
    ``period[m]`` when positive).  Ghost layer ``l`` is the partner's
    ``l``-th layer inward (``src/jaxrans/mesh/block3d.py:9-17``)."""

    start: tuple[int, int]
    stop: tuple[int, int]
    block: int
    side: str
    origin: tuple[int, int]
    axes: tuple[int, int] = (0, 1)
    flip: tuple[bool, bool] = (False, False)
    period: tuple[int, int] = (0, 0)

    def __post_init__(self) -> None:
        if self.side not in SIDES_3D:
            raise ValueError(f"unknown side {self.side!r}; known: {SIDES_3D}")

    def partner(self, a) -> tuple[int, int]:
        out = []
        for m in range(2):
            d = a[self.axes[m]] - self.start[self.axes[m]]
            p = self.origin[m] + (-d if self.flip[m] else d)
            out.append(p % self.period[m] if self.period[m] else p)
        return tuple(out)


@dataclass(frozen=True)
class Sides3D:
    """Rectangles tiling each of a block's six sides exactly once."""

    shape: tuple[int, int, int]
    j_min: tuple
    j_max: tuple
    i_min: tuple
    i_max: tuple
    k_min: tuple
    k_max: tuple

    def __post_init__(self) -> None:
        for side in SIDES_3D:
            n = self.extent(side)
            count = np.zeros(n, dtype=int)
            for r in self.side(side):
                if not all(0 <= r.start[m] < r.stop[m] <= n[m] for m in range(2)):
                    raise ValueError(f"rectangle {r} lies outside {side} of extent {n}")
                count[r.start[0]:r.stop[0], r.start[1]:r.stop[1]] += 1
            if np.any(count != 1):
                raise ValueError(f"the rectangles of {side} do not tile it exactly once")

    @property
    def ni(self) -> int:
        return self.shape[0]

    @property
    def nj(self) -> int:
        return self.shape[1]

    @property
    def nk(self) -> int:
        return self.shape[2]

    def side(self, name: str) -> tuple:
        return getattr(self, name)

    def extent(self, name: str) -> tuple[int, int]:
        a, b = running_axes(name)
        return self.shape[a], self.shape[b]

    def boundaries(self) -> list:
        return [(s, r) for s in SIDES_3D for r in self.side(s) if isinstance(r, BoundaryRect)]

    def walls(self) -> list:
        return [(s, r) for s, r in self.boundaries() if r.kind in WALL_KINDS]

    def with_conditions(self, kinds: dict) -> "Sides3D":
        def assign(rects):
            return tuple(replace(r, kind=kinds[r.patch]) if isinstance(r, BoundaryRect) else r for r in rects)
        return replace(self, **{s: assign(self.side(s)) for s in SIDES_3D})


def box_sides(shape, **sides) -> Sides3D:
    """Each side one patch name (a string) or the given rectangles."""
    def side(name):
        v = sides[name]
        if isinstance(v, str):
            return (BoundaryRect(v, (0, 0), tuple(int(shape[a]) for a in running_axes(name))),)
        return tuple(v)
    return Sides3D(tuple(int(n) for n in shape), *(side(n) for n in SIDES_3D))


def face_vectors(nodes: np.ndarray):
    """Per family ``a``: the face area vectors ``(3, ...)`` (shape of the cells
    with one more along ``a``) and the faces' node means."""
    out = []
    for axis in range(3):
        a1, a2 = (axis + 1) % 3, (axis + 2) % 3
        n = np.moveaxis(nodes, (1 + axis, 1 + a1, 1 + a2), (1, 2, 3))
        pa, pb, pc, pd = n[:, :, :-1, :-1], n[:, :, 1:, :-1], n[:, :, 1:, 1:], n[:, :, :-1, 1:]
        S = 0.5 * np.cross(pc - pa, pd - pb, axis=0)
        mid = 0.25 * (pa + pb + pc + pd)
        out.append((np.moveaxis(S, (1, 2, 3), (1 + axis, 1 + a1, 1 + a2)),
                    np.moveaxis(mid, (1, 2, 3), (1 + axis, 1 + a1, 1 + a2))))
    return out


def metrics3d(nodes: np.ndarray):
    """``(centres (3, ...), volumes, (S_i, S_j, S_k))`` of nodes ``(3, ni+1, nj+1, nk+1)``."""
    faces = face_vectors(nodes)
    volume = 0.0
    for axis, (S, mid) in enumerate(faces):
        flux = np.sum(S * mid, axis=0)
        hi = np.take(flux, np.arange(1, flux.shape[axis]), axis=axis)
        lo = np.take(flux, np.arange(0, flux.shape[axis] - 1), axis=axis)
        volume = volume + (hi - lo)
    volume = volume / 3.0
    centres = 0.0
    for di in (0, 1):
        for dj in (0, 1):
            for dk in (0, 1):
                centres = centres + nodes[:, di:di + nodes.shape[1] - 1, dj:dj + nodes.shape[2] - 1,
                                          dk:dk + nodes.shape[3] - 1]
    return centres / 8.0, volume, tuple(S for S, _ in faces)


@dataclass(frozen=True)
class Grid3D:
    """One 3D block: nodes, centres, volumes, face vectors, and its sides."""

    nodes: np.ndarray
    sides: Sides3D

    def __post_init__(self) -> None:
        expected = (3,) + tuple(n + 1 for n in self.sides.shape)
        if self.nodes.shape != expected:
            raise ValueError(f"node array must be {expected} for these sides, got {self.nodes.shape}")
        object.__setattr__(self, "_m", metrics3d(np.asarray(self.nodes, np.float64)))

    dim = 3

    @property
    def shape(self) -> tuple[int, int, int]:
        return self.sides.shape

    @property
    def ni(self) -> int:
        return self.shape[0]

    @property
    def nj(self) -> int:
        return self.shape[1]

    @property
    def nk(self) -> int:
        return self.shape[2]

    centres = property(lambda self: self._m[0])
    volume = property(lambda self: self._m[1])
    faces = property(lambda self: self._m[2])

    def with_sides(self, sides: Sides3D) -> "Grid3D":
        return Grid3D(self.nodes, sides)


def normal_closure_3d(grid: Grid3D) -> np.ndarray:
    """Sum of the outward area vectors round every cell, ``(3, ni, nj, nk)``."""
    total = 0.0
    for a, S in enumerate(grid.faces):
        total = total + (np.take(S, np.arange(1, S.shape[1 + a]), axis=1 + a)
                         - np.take(S, np.arange(S.shape[1 + a] - 1), axis=1 + a))
    return total


def folded_3d(grid: Grid3D) -> bool:
    return bool(np.any(grid.volume <= 0.0))
