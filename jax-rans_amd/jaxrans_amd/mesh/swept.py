"""The infinite swept wing (2.5D) on the 3D device path.

The reference solves a swept case on its 2D grid with three velocity
components (``lift_atlas``, ``src/jaxrans/mesh/atlas.py:231-247``): every
spanwise derivative is zero, ``w`` is carried.  Its own verification
(``tests/verification/test_swept.py:86-115``) shows that the same operator is
the 3D operator on the 2D grid extruded in ``z`` with periodic end planes and
a spanwise-uniform state: the 3D residual of every plane is the 2.5D residual
times the cell depth.  That is how this implementation runs a swept case: the
section extruded ``SPAN_CELLS`` cells deep with periodic ``k`` sides, on the
3D kernels (three velocity components, 5x5 flow blocks, the 3D diamond,
vorticity and QCR), so that no second set of 2D kernels with a third
velocity component is needed.

``SPAN_CELLS`` equals the ghost depth: each of a side's two ghost layers is
then a real cell of the same column (layer ``l`` of ``k_min`` is cell
``nk - 1 - l``), and a spanwise-uniform state stays uniform bit for bit
(every plane sees the same stencil).  The 3D atlas gives an interface ghost
its partner's cell centre; across the periodic ends the ghost's centre is
the partner's moved by the span, which ``lift`` sets.
"""

from __future__ import annotations

from dataclasses import replace

import numpy as np

from .atlas3d import Atlas3D, build_atlas_3d
from .block import Boundary
from .block3d import BoundaryRect, Grid3D, InterfaceRect, box_sides

SPAN_CELLS = 2
SPAN_DEPTH = 1.0          # cell depth in z; the residual of a plane is the 2.5D one times it


def extrude_periodic(mesh2d, nk: int = SPAN_CELLS, dz: float = SPAN_DEPTH) -> tuple:
    """Each 2D block ``(x, y)`` as a 3D block ``(x, y, k dz)``, ``k = 0..nk``
    (right-handed for a counter-clockwise 2D block); its 2D sides carried
    over every ``k``, its ``k`` sides periodic with each other."""
    out = []
    for b, blk in enumerate(mesh2d):
        ni, nj = blk.ni, blk.nj
        nodes = np.empty((3, ni + 1, nj + 1, nk + 1))
        nodes[0] = np.asarray(blk.x_node, np.float64)[..., None]
        nodes[1] = np.asarray(blk.y_node, np.float64)[..., None]
        nodes[2] = (np.arange(nk + 1) * dz)[None, None, :]
        sides = {}
        for name in ("j_min", "j_max", "i_min", "i_max"):
            rects = []
            for seg in blk.sides.side(name):
                if isinstance(seg, Boundary):
                    rects.append(BoundaryRect(seg.patch, (seg.start, 0), (seg.stop, nk), seg.kind))
                else:
                    rects.append(InterfaceRect((seg.start, 0), (seg.stop, nk), seg.block, seg.side,
                                               (seg.partner(seg.start), 0), flip=(seg.reversed, False)))
            sides[name] = tuple(rects)
        sides["k_min"] = (InterfaceRect((0, 0), (ni, nj), b, "k_max", (0, 0)),)
        sides["k_max"] = (InterfaceRect((0, 0), (ni, nj), b, "k_min", (0, 0)),)
        out.append(Grid3D(nodes, box_sides((ni, nj, nk), **sides)))
    return tuple(out)


def lift(mesh2d, g: int = 2, nk: int = SPAN_CELLS, dz: float = SPAN_DEPTH) -> Atlas3D:
    """The 3D atlas of a 2D mesh (patches with conditions assigned) for a
    swept case, every ghost centre at its true position: the ``k`` ghosts'
    (and the edge padding's) ``z`` at their periodic images."""
    if nk != g:
        raise ValueError(f"the periodic span must be {g} cells deep (the ghost depth), got {nk}")
    atlas = build_atlas_3d(extrude_periodic(mesh2d, nk, dz), g)
    centres = atlas.centres.copy()
    NK = atlas.shape[2]
    for w in atlas.windows:
        k0 = w[2].start - g                      # the block's tile starts g cells before its first cell
        lo = max(0, k0 + 1)
        hi = min(NK + 2, k0 + 1 + nk + 2 * g)
        if k0 == 0:
            lo = 0                               # the edge padding below the canvas
        if k0 + nk + 2 * g == NK:
            hi = NK + 2                          # and above it
        kp = np.arange(lo, hi)
        cols = (slice(w[0].start - g, w[0].stop + g + 2), slice(w[1].start - g, w[1].stop + g + 2))
        centres[2][cols[0], cols[1], lo:hi] = ((kp - 1 - w[2].start) + 0.5)[None, None, :] * dz
    return replace(atlas, centres=centres)


def swept_freestream(mach: float, alpha_deg: float, sweep_deg: float, gamma: float) -> np.ndarray:
    """``(rho, u, v, w, p)``: ``M (cos a cos L, sin a, cos a sin L)`` in the
    frame of the section normal to the sweep line (``x`` chordwise, ``y`` up,
    ``z`` along the sweep line; ``freestream_vector``,
    ``src/jaxrans/core/state.py:123-139``)."""
    a, s = np.deg2rad(alpha_deg), np.deg2rad(sweep_deg)
    u = mach * np.cos(a)
    return np.array([1.0, u * np.cos(s), mach * np.sin(a), u * np.sin(s), 1.0 / gamma])


def planes_to_section(U3: np.ndarray, atlas3: Atlas3D, atlas2) -> np.ndarray:
    """The solution of plane ``k = 0`` of every block on the 2D canvas,
    ``(nvar, NI, NJ)``: the reference's 2.5D state layout."""
    blocks = [b[..., 0] for b in atlas3.unpack(U3)]
    return atlas2.pack(blocks, fill=np.asarray(U3)[(Ellipsis, 0, 0, 0)])


def section_to_planes(U2: np.ndarray, atlas2, atlas3: Atlas3D, fill: np.ndarray) -> np.ndarray:
    """A 2.5D state ``(nvar, NI, NJ)`` on the extruded canvas, the same in
    every plane; cells outside every block take ``fill``'s."""
    canvas = np.array(fill, dtype=np.float64, copy=True)
    for blk2, w3, blk3 in zip(atlas2.unpack(U2), atlas3.windows, atlas3.mesh):
        canvas[(Ellipsis,) + tuple(w3)] = np.repeat(blk2[..., None], blk3.shape[2], axis=-1)
    return canvas


def _coefficients(loads, atlas3: Atlas3D, q_inf, sweep_deg: float, length: float, centre, viscous: bool) -> dict:
    """``swept_coefficients`` on arrays of any numeric type: real, or complex
    for the complex-step derivative (every operation is analytic in the loads
    and the freestream)."""
    w = atlas3.walls
    span = float(atlas3.mesh[0].nodes[2, 0, 0, -1] - atlas3.mesh[0].nodes[2, 0, 0, 0])
    S = np.empty((3, w.n))
    for fam in range(3):
        sel = w.family == fam
        S[:, sel] = atlas3.faces[fam].reshape(3, -1)[:, w.face[sel]]
    S = S * w.into_fluid / span                          # per unit length along the sweep line
    pressure = -loads[0] * S
    shear = (loads[1:4] if viscous else np.zeros((3, w.n))) / span
    load = pressure + shear
    c, s = float(np.cos(np.deg2rad(sweep_deg))), float(np.sin(np.deg2rad(sweep_deg)))

    def wing(f):
        return np.stack([f[0] + (s / c) * f[2], f[1] / c])

    side = (c * load[2] - s * load[0]) / c
    load_w, pressure_w, shear_w = wing(load), wing(pressure), wing(shear)
    vel = np.array([c * q_inf[1] + s * q_inf[3], q_inf[2]])
    speed = np.sqrt(q_inf[1] ** 2 + q_inf[2] ** 2 + q_inf[3] ** 2)
    drag = vel / speed
    lift_dir = np.array([-drag[1], drag[0]])
    q_area = 0.5 * q_inf[0] * speed ** 2 * length
    mid = w.mid[:, :2] / np.array([c, 1.0])
    r = mid - np.asarray(centre[:2], np.float64)
    moment = r[:, 1] * pressure_w[0] - r[:, 0] * pressure_w[1]
    return {"cl": lift_dir @ load_w.sum(axis=1) / q_area, "cd": drag @ load_w.sum(axis=1) / q_area,
            "cm": np.sum(moment / length) / q_area, "cd_friction": drag @ shear_w.sum(axis=1) / q_area,
            "cy": side.sum() / q_area, "force": np.concatenate([load_w.sum(axis=1), [side.sum()]])}


def swept_coefficients(loads: np.ndarray, atlas3: Atlas3D, q_inf: np.ndarray, sweep_deg: float, length: float,
                       centre, viscous: bool) -> dict:
    """Force coefficients of a swept case from the device's per-wall-face
    loads of the extruded canvas (``(4, n_walls)``: gauge pressure, viscous
    force), as ``force_coefficients`` forms them for a swept flow
    (``src/jaxrans/post/forces.py:11-21``, ``136-160``): the loads per unit
    length along the sweep line taken to the wing frame per unit span,
    ``F_X = F_x + (s / c) F_z``, ``F_Y = F_y / c``, ``F_S = (c F_z - s F_x) /
    c``; lift and drag in the ``X-Y`` plane per ``q_inf`` (the whole
    freestream's) times ``length``; the pressure moment about ``centre`` at
    streamwise positions ``x / c``; ``cy`` the spanwise force."""
    out = _coefficients(np.asarray(loads, np.float64), atlas3, np.asarray(q_inf, np.float64), sweep_deg, length,
                        centre, viscous)
    return {k: (v if k == "force" else float(v)) for k, v in out.items()}


def swept_coefficients_tangent(loads, dloads, atlas3: Atlas3D, q_inf, dq_inf, sweep_deg: float, length: float,
                               centre, viscous: bool) -> dict:
    """The derivative of ``swept_coefficients`` along ``(dloads, dq_inf)``,
    exact to round-off by the complex step (Squire and Trapp, SIAM Review 40
    (1998) 110-112): Im f(x + i h dx) / h with h = 1e-30, no subtraction."""
    h = 1e-30
    z = _coefficients(np.asarray(loads, np.float64) + 1j * h * np.asarray(dloads, np.float64), atlas3,
                      np.asarray(q_inf, np.float64) + 1j * h * np.asarray(dq_inf, np.float64), sweep_deg, length,
                      centre, viscous)
    return {k: np.imag(v) / h if k == "force" else float(np.imag(v) / h) for k, v in z.items()}
