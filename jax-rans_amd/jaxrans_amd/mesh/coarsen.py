"""Grid sequencing (``src/jaxrans/mesh/coarsen.py:36-83``): a mesh coarsened
by keeping every other node in every index direction, and a coarse canvas
solution carried to the finer canvas by injection.

A coarse cell is the union of its 2^d children, so every index that names a
cell boundary -- block cell counts, the limits of boundary patches and
interfaces, partner offsets, periods -- must be even and is halved; an
interface whose partner runs backwards names its partner range by the far end,
which must then be odd (the last child of a coarse cell).  Injection (each
child takes its parent's value) is conservative.
"""

from __future__ import annotations

from dataclasses import replace

import numpy as np

from .block import SIDES, BlockSides, Boundary, Grid, Interface
from .block3d import SIDES_3D, BoundaryRect, Grid3D, InterfaceRect, Sides3D


class _Halver:
    """Halves the cell indices of one block, naming the first odd one it meets."""

    def __init__(self, block: int):
        self.block = block

    def __call__(self, value, what: str, last_child: bool = False):
        if isinstance(value, tuple):
            return tuple(self(v, what, last_child) for v in value)
        value = int(value)
        want_odd = last_child
        if (value % 2 == 1) != want_odd:
            parity = "even" if want_odd else "odd"
            raise ValueError(f"block {self.block}: cannot coarsen the grid: {what} is {value}, {parity} (every "
                             "block's cell counts and every patch and interface limit must be even)")
        return value // 2


def _coarse_2d(grid: Grid, half: _Halver) -> Grid:
    s = grid.sides

    def segment(seg, side):
        out = replace(seg, start=half(seg.start, f"a {side} limit"), stop=half(seg.stop, f"a {side} limit"))
        if isinstance(seg, Interface):
            out = replace(out, partner_start=half(seg.partner_start, f"a {side} interface offset"))
        return out

    sides = BlockSides(ni=half(s.ni, "a cell count"), nj=half(s.nj, "a cell count"),
                       **{name: tuple(segment(seg, name) for seg in s.side(name)) for name in SIDES})
    return Grid(np.ascontiguousarray(grid.x_node[::2, ::2]), np.ascontiguousarray(grid.y_node[::2, ::2]), sides)


def _coarse_3d(grid: Grid3D, half: _Halver) -> Grid3D:
    s = grid.sides

    def rect(r, side):
        out = replace(r, start=half(r.start, f"a {side} limit"), stop=half(r.stop, f"a {side} limit"))
        if isinstance(r, InterfaceRect):
            # origin: the partner cell facing ``start`` -- the near end of an
            # ascending partner run, the far end (an odd last child) of a flipped one
            origin = tuple(half(o, f"a {side} interface offset", last_child=f) for o, f in zip(r.origin, r.flip))
            out = replace(out, origin=origin, period=half(r.period, f"a {side} interface period"))
        return out

    shape = half(tuple(s.shape), "a cell count")
    sides = Sides3D(shape, *(tuple(rect(r, name) for r in s.side(name)) for name in SIDES_3D))
    return Grid3D(np.ascontiguousarray(grid.nodes[:, ::2, ::2, ::2]), sides)


def coarsen(mesh) -> tuple:
    """Every block with every other node (2D or 3D blocks)."""
    out = []
    for b, block in enumerate(mesh):
        half = _Halver(b)
        out.append(_coarse_2d(block, half) if block.dim == 2 else _coarse_3d(block, half))
    return tuple(out)


def prolong(U: np.ndarray, coarse, fine, start: np.ndarray) -> np.ndarray:
    """The coarse canvas solution ``U`` on the fine canvas: each fine cell
    takes its parent's value; cells outside every block keep ``start``'s."""
    canvas = np.array(start, dtype=np.float64, copy=True)
    dim = fine.dim
    for block, window in zip(coarse.unpack(U), fine.windows):
        for axis in range(-dim, 0):
            block = np.repeat(block, 2, axis=axis)
        canvas[(Ellipsis,) + tuple(window)] = block
    return canvas

