"""Grid sequencing for 2D meshes (``src/jaxrans/mesh/coarsen.py:58-83``): a
mesh coarsened by taking every other node, and a coarse canvas solution
carried to the fine canvas by injection.

Every boundary and interface segment keeps its patch, condition, partner and
orientation with its cell indices halved, so a coarse cell is the union of
its four children and injection (each child takes its parent's value) is
conservative.
"""

from __future__ import annotations

from dataclasses import replace

import numpy as np

from .block import SIDES, BlockSides, Boundary, Grid


def _half(n: int, what: str) -> int:
    if n % 2:
        raise ValueError(f"cannot coarsen the grid: {what} is {n}, odd (every block's cell counts and every "
                         "patch and interface limit must be even)")
    return n // 2


def _segment(seg, side: str):
    if isinstance(seg, Boundary):
        return replace(seg, start=_half(seg.start, f"a {side} limit"), stop=_half(seg.stop, f"a {side} limit"))
    return replace(seg, start=_half(seg.start, f"a {side} limit"), stop=_half(seg.stop, f"a {side} limit"),
                   partner_start=_half(seg.partner_start, f"a {side} interface offset"))


def coarsen(mesh) -> tuple:
    """Every block with every other node (2D blocks)."""
    out = []
    for b, block in enumerate(mesh):
        if block.dim != 2:
            raise ValueError("grid sequencing is built for 2D meshes only")
        s = block.sides
        try:
            sides = BlockSides(ni=_half(s.ni, "a cell count"), nj=_half(s.nj, "a cell count"),
                               **{name: tuple(_segment(seg, name) for seg in s.side(name)) for name in SIDES})
        except ValueError as error:
            raise ValueError(f"block {b}: {error}") from None
        out.append(Grid(np.ascontiguousarray(block.x_node[::2, ::2]), np.ascontiguousarray(block.y_node[::2, ::2]),
                        sides))
    return tuple(out)


def prolong(U: np.ndarray, coarse, fine, start: np.ndarray) -> np.ndarray:
    """The coarse canvas solution ``U`` on the fine canvas: each fine cell
    takes its parent's value; cells outside every block keep ``start``'s."""
    canvas = np.array(start, dtype=np.float64, copy=True)
    for block, (wi, wj) in zip(coarse.unpack(U), fine.windows):
        canvas[..., wi, wj] = np.repeat(np.repeat(block, 2, axis=-2), 2, axis=-1)
    return canvas
