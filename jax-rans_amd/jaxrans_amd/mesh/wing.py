"""A synthetic swept, tapered wing between two symmetry planes: four
point-matched C-H blocks along the span.

The reference reads 3D grids from files and generates none
(``docs/wing-3d.md:1-40``); this generator makes the multi-block wing grid
of the 3D benchmark case directly:

* section: the NACA 4-digit thickness law (``geometry/airfoil.py:naca4``,
  trailing edge closed so that a C-grid applies), thickness ratio ``t``;
* planform: root chord 1 at ``y = 0``, taper ``taper``, leading-edge sweep
  ``sweep_deg``, semi-span from the aspect ratio ``AR = b^2 / S`` of the full
  trapezoid (``b = AR (1 + taper) / 2``);
* one 2D C-grid (``mesh/generate.py:c_grid``) of the unit-chord section in
  the ``x-z`` plane; station ``y`` places it at the local leading edge and
  chord near the wall and blends to the mid-span chord's placement toward
  the far field, so the outer boundary is one prism along ``y``;
* the span split into ``blocks`` blocks of ``nk`` cells each, joined
  face to face; ``k_min`` of the first block (root) and ``k_max`` of the last
  (tip) are symmetry planes; each block's wake cut is an interface with
  itself (``i`` reversed), its section part the wall, everything else far
  field.

Index directions: ``i`` round the section from the far end of the upper
wake (upper surface first), ``j`` away from the wall, ``k`` along ``+y``;
right-handed.
"""

from __future__ import annotations

import numpy as np

from ..geometry.airfoil import naca4
from .block3d import BoundaryRect, Grid3D, InterfaceRect, box_sides
from .generate import c_grid


def wing_planform(aspect_ratio: float, taper: float):
    """``(semi_span, area of the half wing)`` for root chord 1."""
    span = aspect_ratio * (1.0 + taper) / 2.0
    semi = 0.5 * span
    return semi, semi * (1.0 + taper) / 2.0


def wing_blocks(cells_airfoil: int = 128, cells_wake: int = 32, cells_normal: int = 49, cells_span: int = 29,
                blocks: int = 4, thickness: float = 0.10, taper: float = 0.55, sweep_deg: float = 28.0,
                aspect_ratio: float = 7.0, wall_spacing: float = 1e-6, far_field: float = 50.0,
                section_points: int = 201) -> list[Grid3D]:
    """The wing grid as ``blocks`` ``Grid3D`` of ``(cells_airfoil + 2
    cells_wake) x cells_normal x cells_span`` cells (patches ``wall``,
    ``farfield``, ``symmetry``)."""
    foil = naca4(f"00{int(round(100 * thickness)):02d}", n=section_points, sharp=True)
    X, Z = c_grid(foil, cells_airfoil, cells_wake, cells_normal, wall_spacing, far_field)
    X, Z = X[::-1], Z[::-1]                       # upper surface first: right-handed with k along +y
    ni, nj = X.shape[0] - 1, X.shape[1] - 1
    semi, _ = wing_planform(aspect_ratio, taper)
    nk_total = blocks * cells_span
    y = np.linspace(0.0, semi, nk_total + 1)
    chord = 1.0 - (1.0 - taper) * y / semi
    x_le = y * np.tan(np.deg2rad(sweep_deg))
    mid = 0.5 * semi
    chord_mid, x_mid = 1.0 - (1.0 - taper) * 0.5, mid * np.tan(np.deg2rad(sweep_deg))
    # blend weight: 1 at the wall, 0 at the far field, along j (by distance from the section)
    reach = np.hypot(X - X[:, :1], Z - Z[:, :1])
    w = np.clip(1.0 - reach / reach[:, -1:], 0.0, 1.0) ** 2
    nodes = np.empty((3, ni + 1, nj + 1, nk_total + 1))
    for k in range(nk_total + 1):
        near_x, near_z = x_le[k] + chord[k] * X, chord[k] * Z
        far_x, far_z = x_mid + chord_mid * X, chord_mid * Z
        nodes[0, :, :, k] = w * near_x + (1.0 - w) * far_x
        nodes[1, :, :, k] = y[k]
        nodes[2, :, :, k] = w * near_z + (1.0 - w) * far_z
    out = []
    wake, air = cells_wake, cells_airfoil
    for b in range(blocks):
        nk = cells_span
        j_min = (InterfaceRect((0, 0), (wake, nk), b, "j_min", (ni - 1, 0), flip=(True, False)),
                 BoundaryRect("wall", (wake, 0), (wake + air, nk)),
                 InterfaceRect((wake + air, 0), (ni, nk), b, "j_min", (wake - 1, 0), flip=(True, False)))
        k_min = "symmetry" if b == 0 else (InterfaceRect((0, 0), (ni, nj), b - 1, "k_max", (0, 0)),)
        k_max = "symmetry" if b == blocks - 1 else (InterfaceRect((0, 0), (ni, nj), b + 1, "k_min", (0, 0)),)
        sides = box_sides((ni, nj, nk), j_min=j_min, j_max="farfield", i_min="farfield", i_max="farfield",
                          k_min=k_min, k_max=k_max)
        out.append(Grid3D(nodes[:, :, :, b * nk:(b + 1) * nk + 1].copy(), sides))
    return out


def wing_reference(aspect_ratio: float = 7.0, taper: float = 0.55) -> dict:
    """Force reference values of the half wing: area, mean aerodynamic chord
    as the moment length, moments about the root quarter chord."""
    semi, area = wing_planform(aspect_ratio, taper)
    mac = 2.0 / 3.0 * (1.0 + taper + taper ** 2) / (1.0 + taper)
    return {"area": area, "length": mac, "centre": (0.25, 0.0, 0.0), "semi_span": semi}


def wing_patch_rules(aspect_ratio: float = 7.0, taper: float = 0.55) -> tuple:
    """Box rules naming the wing grid's boundary faces when it is read back
    from a grid file: the root and tip planes ``symmetry``, the section's
    surface ``wall`` (every unmatched face within a chord of it), the rest
    ``farfield``."""
    from .assemble3d import PatchRule

    semi, _ = wing_planform(aspect_ratio, taper)
    reach = semi * np.tan(np.deg2rad(40.0)) + 2.0
    return (PatchRule("symmetry", y_max=1e-9), PatchRule("symmetry", y_min=semi - 1e-9),
            PatchRule("wall", x_min=-1.0, x_max=reach, z_min=-1.0, z_max=1.0), PatchRule("farfield"))


def write_wing_grid(path, **kw):
    """The wing grid's blocks as a multi-block Plot3D file."""
    from .plot3d import write_plot3d_3d

    return write_plot3d_3d(path, [g.nodes for g in wing_blocks(**kw)])


if __name__ == "__main__":
    import argparse

    ap = argparse.ArgumentParser(description="write the synthetic wing grid as multi-block Plot3D")
    ap.add_argument("path")
    ap.add_argument("--cells-airfoil", type=int, default=128)
    ap.add_argument("--cells-wake", type=int, default=32)
    ap.add_argument("--cells-normal", type=int, default=49)
    ap.add_argument("--cells-span", type=int, default=29)
    ap.add_argument("--blocks", type=int, default=4)
    ap.add_argument("--wall-spacing", type=float, default=1e-6)
    ap.add_argument("--far-field", type=float, default=50.0)
    a = ap.parse_args()
    out = write_wing_grid(a.path, cells_airfoil=a.cells_airfoil, cells_wake=a.cells_wake,
                          cells_normal=a.cells_normal, cells_span=a.cells_span, blocks=a.blocks,
                          wall_spacing=a.wall_spacing, far_field=a.far_field)
    print(f"wrote {out} ({a.blocks} blocks of {a.cells_airfoil + 2 * a.cells_wake} x {a.cells_normal} x "
          f"{a.cells_span} cells)")
