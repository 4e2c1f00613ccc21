"""Structured blocks: node geometry, metrics, and what lies beyond each side.

Conventions are the reference's (``src/jaxrans/mesh/block.py:1-14``, ``225-252``):
nodes ``(ni+1, nj+1)``, cells ``(ni, nj)``; face area vectors point toward
increasing index and carry the face length -- i-faces ``(dy, -dx)`` of the edge
``(i, j) -> (i, j+1)``, j-faces ``(-dy, dx)`` of the edge ``(i, j) -> (i+1, j)``;
areas are half the cross product of the cell diagonals, centres the mean of
the four nodes.  A side is tiled by ``Boundary`` patches and ``Interface``
segments (a seam, a wake cut, or another block).
"""

from __future__ import annotations

from dataclasses import dataclass, replace

import numpy as np

SIDES = ("j_min", "j_max", "i_min", "i_max")
OUTWARD = {"j_min": -1, "j_max": 1, "i_min": -1, "i_max": 1}
CONDITION_KINDS = ("slip_wall", "no_slip_wall", "symmetry", "far_field", "inflow_total", "outflow_static")
WALL_KINDS = ("slip_wall", "no_slip_wall")


@dataclass(frozen=True)
class Boundary:
    """Cells ``[start, stop)`` of a side on patch ``patch``; ``kind`` its condition."""

    patch: str
    start: int
    stop: int
    kind: str | None = None

    def __post_init__(self) -> None:
        if self.kind is not None and self.kind not in CONDITION_KINDS:
            raise ValueError(f"unknown boundary condition {self.kind!r}; known: {CONDITION_KINDS}")


@dataclass(frozen=True)
class Interface:
    """Cells ``[start, stop)`` facing cells ``[partner_start, partner_start + n)``
    of side ``side`` of block ``block``, in order or reversed."""

    start: int
    stop: int
    block: int
    side: str
    partner_start: int
    reversed: bool

    def __post_init__(self) -> None:
        if self.side not in SIDES:
            raise ValueError(f"unknown side {self.side!r}; known: {SIDES}")

    def partner(self, a: int) -> int:
        """The partner's running index facing cell ``a`` of this segment."""
        off = a - self.start
        return self.partner_start + (self.stop - self.start - 1 - off if self.reversed else off)


@dataclass(frozen=True)
class BlockSides:
    ni: int
    nj: int
    j_min: tuple
    j_max: tuple
    i_min: tuple
    i_max: tuple

    def __post_init__(self) -> None:
        for side in SIDES:
            n = self.length(side)
            segs = self.side(side)
            stops = [0] + [s.stop for s in segs]
            starts = [s.start for s in segs] + [n]
            if stops != starts or any(s.stop <= s.start for s in segs):
                raise ValueError(f"segments of {side} do not tile [0, {n}) in order: {segs}")

    def side(self, name: str) -> tuple:
        return getattr(self, name)

    def length(self, name: str) -> int:
        return self.ni if name.startswith("j") else self.nj

    def boundaries(self) -> list[tuple[str, Boundary]]:
        return [(s, seg) for s in SIDES for seg in self.side(s) if isinstance(seg, Boundary)]

    def walls(self) -> list[tuple[str, Boundary]]:
        return [(s, seg) for s, seg in self.boundaries() if seg.kind in WALL_KINDS]

    def with_conditions(self, kinds: dict[str, str]) -> "BlockSides":
        def assign(segs):
            return tuple(replace(s, kind=kinds[s.patch]) if isinstance(s, Boundary) else s for s in segs)
        return replace(self, **{s: assign(self.side(s)) for s in SIDES})


def rectangle_sides(ni: int, nj: int, **segments) -> BlockSides:
    """Each side one named patch (a string) or the given segments."""
    def side(name):
        v = segments[name]
        return (Boundary(v, 0, ni if name.startswith("j") else nj),) if isinstance(v, str) else tuple(v)
    return BlockSides(ni, nj, *(side(n) for n in SIDES))


def o_grid_sides(ni: int, nj: int) -> BlockSides:
    """O-grid: ``wall`` on j_min, ``farfield`` on j_max, seam i_min <-> i_max."""
    if ni < 4 or ni % 2 or nj < 1:
        raise ValueError(f"an O-grid needs ni even and >= 4, nj >= 1; got {ni} x {nj}")
    return rectangle_sides(ni, nj, j_min="wall", j_max="farfield",
                           i_min=(Interface(0, nj, 0, "i_max", 0, False),),
                           i_max=(Interface(0, nj, 0, "i_min", 0, False),))


def metrics2d(x: np.ndarray, y: np.ndarray):
    """``(xc, yc, area, (sx_i, sy_i), (sx_j, sy_j))`` of 2D nodes ``(ni+1, nj+1)``."""
    x1, y1, x2, y2 = x[:-1, :-1], y[:-1, :-1], x[1:, :-1], y[1:, :-1]
    x3, y3, x4, y4 = x[1:, 1:], y[1:, 1:], x[:-1, 1:], y[:-1, 1:]
    area = 0.5 * ((x3 - x1) * (y4 - y2) - (y3 - y1) * (x4 - x2))
    xc = 0.25 * (x1 + x2 + x3 + x4)
    yc = 0.25 * (y1 + y2 + y3 + y4)
    ex_i, ey_i = x[:, 1:] - x[:, :-1], y[:, 1:] - y[:, :-1]
    ex_j, ey_j = x[1:, :] - x[:-1, :], y[1:, :] - y[:-1, :]
    return xc, yc, area, (ey_i, -ex_i), (-ey_j, ex_j)


@dataclass(frozen=True)
class Grid:
    """One 2D block: nodes, centres, areas, face vectors, and its sides."""

    x_node: np.ndarray
    y_node: np.ndarray
    sides: BlockSides

    def __post_init__(self) -> None:
        expected = (self.sides.ni + 1, self.sides.nj + 1)
        if self.x_node.shape != expected or self.y_node.shape != expected:
            raise ValueError(f"node arrays must be {expected} for these sides, got {self.x_node.shape}")
        xc, yc, area, si, sj = metrics2d(self.x_node, self.y_node)
        object.__setattr__(self, "_m", (xc, yc, area, si, sj))

    dim = 2

    @property
    def ni(self) -> int:
        return self.sides.ni

    @property
    def nj(self) -> int:
        return self.sides.nj

    @property
    def shape(self) -> tuple[int, int]:
        return (self.ni, self.nj)

    xc = property(lambda self: self._m[0])
    yc = property(lambda self: self._m[1])
    area = property(lambda self: self._m[2])
    sx_i = property(lambda self: self._m[3][0])
    sy_i = property(lambda self: self._m[3][1])
    sx_j = property(lambda self: self._m[4][0])
    sy_j = property(lambda self: self._m[4][1])

    def with_sides(self, sides: BlockSides) -> "Grid":
        return Grid(self.x_node, self.y_node, sides)


def normal_closure(grid: Grid) -> np.ndarray:
    """Sum of the outward area vectors of every cell, ``(2, ni, nj)``: zero up to round-off."""
    return np.stack([grid.sx_i[1:] - grid.sx_i[:-1] + grid.sx_j[:, 1:] - grid.sx_j[:, :-1],
                     grid.sy_i[1:] - grid.sy_i[:-1] + grid.sy_j[:, 1:] - grid.sy_j[:, :-1]])


def folded(grid: Grid) -> bool:
    return bool(np.any(grid.area <= 0.0))
