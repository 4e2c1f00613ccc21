"""Body-fitted O- and C-grids by hyperbolic marching.

The wall line is distributed with Vinokur's two-sided stretching (J. Comput.
Phys. 50 (1983) 215-234), clustered at the leading and trailing edges, and
marched outward one layer at a time by solving the linearized orthogonality
and cell-area equations

    x_xi x_eta + y_xi y_eta = 0,      x_xi y_eta - y_xi x_eta = dV

(Steger and Chaussee, SIAM J. Sci. Stat. Comput. 1 (1980) 431-437) with
implicit second-difference smoothing of the step (Chan and Steger, Appl. Math.
Comput. 51 (1992) 181-205).  Same inputs and outputs as the reference's
generator (``src/jaxrans/mesh/generate.py:238-461``); the arithmetic is this
module's own, so grids agree in shape and quality, not bitwise.
"""

from __future__ import annotations

import numpy as np
from scipy.optimize import brentq
from scipy.sparse import csc_matrix
from scipy.sparse.linalg import spsolve

from ..geometry.airfoil import Airfoil

LEADING_EDGE_SPACING = 2e-3
TRAILING_EDGE_SPACING = 2e-3


def vinokur(n: int, ds0: float, ds1: float) -> np.ndarray:
    """``n + 1`` points on [0, 1] whose end intervals approximate ``ds0`` and
    ``ds1`` (fractions of the length): Vinokur's tanh (or tan) two-sided map."""
    t = np.linspace(0.0, 1.0, n + 1)
    s0, s1 = 1.0 / (n * ds0), 1.0 / (n * ds1)
    a, b = np.sqrt(s0 / s1), np.sqrt(s0 * s1)
    if abs(b - 1.0) < 1e-6:
        u = t
    elif b > 1.0:
        delta = brentq(lambda d: np.sinh(d) / d - b, 1e-12, 50.0)
        u = 0.5 * (1.0 + np.tanh(delta * (t - 0.5)) / np.tanh(0.5 * delta))
    else:
        delta = brentq(lambda d: np.sin(d) / d - b, 1e-12, np.pi - 1e-12)
        u = 0.5 * (1.0 + np.tan(delta * (t - 0.5)) / np.tan(0.5 * delta))
    s = u / (a + (1.0 - a) * u)
    s[0], s[-1] = 0.0, 1.0
    return s


def geometric(n: int, first: float, total: float) -> np.ndarray:
    """``n + 1`` stations on [0, total], steps growing by a constant ratio from ``first``."""
    if first * n >= total:
        raise ValueError(f"a first step of {first:g} over {n} steps already exceeds {total:g}")
    ratio = brentq(lambda r: first * (r**n - 1.0) / (r - 1.0) - total, 1.0 + 1e-12, 10.0)
    return np.concatenate([[0.0], first * np.cumsum(ratio ** np.arange(n))])


def geometric_total(n: int, first: float, ratio: float) -> float:
    """Length of ``n`` steps from ``first`` growing by ``ratio``: what a
    stretching ratio means as a far-field distance."""
    return first * (ratio**n - 1.0) / (ratio - 1.0)


def surface_arcs(airfoil: Airfoil, n: int, le: float, te: float) -> np.ndarray:
    """``n + 1`` arc lengths trailing edge to trailing edge, ``n/2`` cells per side."""
    if n % 2:
        raise ValueError(f"cells round the airfoil must be even, got {n}")
    s_le, total, c = float(airfoil.s[airfoil.leading_edge]), airfoil.length, airfoil.chord
    lower = s_le * vinokur(n // 2, te * c / s_le, le * c / s_le)
    upper = s_le + (total - s_le) * vinokur(n // 2, le * c / (total - s_le), te * c / (total - s_le))
    return np.concatenate([lower, upper[1:]])


def surface(airfoil: Airfoil, n: int, le: float, te: float):
    x, y = airfoil.points(surface_arcs(airfoil, n, le, te))
    if airfoil.sharp:
        x[-1], y[-1] = x[0], y[0]
    return x, y


def wake_line(airfoil: Airfoil, x_te: float, y_te: float, n: int, first: float, length: float):
    """``n + 1`` nodes from the trailing edge along the bisector of the two
    surface tangents there, geometric from ``first`` to ``length``."""
    fx, fy = airfoil.spline()
    total = airfoil.length
    t_lower = np.array([fx(0.0, 1), fy(0.0, 1)])
    t_upper = -np.array([fx(total, 1), fy(total, 1)])
    d = -(t_lower / np.linalg.norm(t_lower) + t_upper / np.linalg.norm(t_upper))
    d /= np.linalg.norm(d)
    s = geometric(n, first, length)
    return x_te + s * d[0], y_te + s * d[1], d


def _d1(a: np.ndarray, periodic: bool) -> np.ndarray:
    if periodic:
        return 0.5 * (np.roll(a, -1, axis=0) - np.roll(a, 1, axis=0))
    out = np.empty_like(a)
    out[1:-1] = 0.5 * (a[2:] - a[:-2])
    out[0], out[-1] = a[1] - a[0], a[-1] - a[-2]
    return out


def _d2(a: np.ndarray, periodic: bool) -> np.ndarray:
    if periodic:
        return np.roll(a, -1, axis=0) - 2.0 * a + np.roll(a, 1, axis=0)
    out = np.zeros_like(a)
    out[1:-1] = a[2:] - 2.0 * a[1:-1] + a[:-2]
    return out


def _smooth_areas(area: np.ndarray, periodic: bool) -> np.ndarray:
    for _ in range(2):
        if periodic:
            area = 0.5 * area + 0.25 * (np.roll(area, 1) + np.roll(area, -1))
        else:
            area = np.concatenate([[area[0]], 0.5 * area[1:-1] + 0.25 * (area[:-2] + area[2:]), [area[-1]]])
    return area


def march(line: np.ndarray, steps: np.ndarray, periodic: bool, outflow: np.ndarray | None = None,
          smoothing: float = 3.0) -> np.ndarray:
    """Layers ``(n, nsteps + 1, 2)`` marched from ``line`` ``(n, 2)``.

    Per layer the unknowns are the new positions; with ``t = r_xi`` (central
    differences of the current layer) and the target area ``dV = ds |t|``, the
    linearized equations give ``r_new + C (r_new[i+1] - r_new[i-1]) / 2 =
    r + g`` with ``C = B^-1 A``, ``g = B^-1 (0, 2 dV)``, ``B = [[x_xi, y_xi],
    [-y_xi, x_xi]]``, ``A = [[x_eta, y_eta], [y_eta, -x_eta]]`` at the
    orthogonal guess ``r_eta = dV/|t|^2 (-y_xi, x_xi)``; the step is smoothed
    implicitly by ``-eps d2(step)``.  Open lines keep their two ends on the
    lines through them along ``outflow``, level with their inner neighbours.
    """
    n = len(line)
    steps = np.asarray(steps, dtype=np.float64)
    layers = np.empty((n, steps.size + 1, 2))
    layers[:, 0] = line
    r = line.astype(np.float64).copy()
    eps = float(smoothing)
    for k, ds in enumerate(steps):
        t = _d1(r, periodic)
        xx, yx = t[:, 0], t[:, 1]
        L2 = xx * xx + yx * yx
        area = _smooth_areas(ds * np.sqrt(L2), periodic)
        xe, ye = -area / L2 * yx, area / L2 * xx
        inv = 1.0 / L2
        c00 = inv * (xx * xe - yx * ye)
        c01 = inv * (xx * ye + yx * xe)
        c10 = inv * (yx * xe + xx * ye)
        c11 = inv * (yx * ye - xx * xe)
        g = np.stack([-yx, xx], axis=1) * (2.0 * area * inv)[:, None]
        rhs = r + g - eps * _d2(r, periodic)
        rows, cols, vals = [], [], []

        def put(i, j, blk):
            for a in range(2):
                for b in range(2):
                    if blk[a][b] != 0.0:
                        rows.append(2 * i + a)
                        cols.append(2 * j + b)
                        vals.append(blk[a][b])

        inner = range(n) if periodic else range(1, n - 1)
        for i in inner:
            half = ((0.5 * c00[i], 0.5 * c01[i]), (0.5 * c10[i], 0.5 * c11[i]))
            put(i, (i - 1) % n, ((-half[0][0] - eps, -half[0][1]), (-half[1][0], -half[1][1] - eps)))
            put(i, i, ((1.0 + 2.0 * eps, 0.0), (0.0, 1.0 + 2.0 * eps)))
            put(i, (i + 1) % n, ((half[0][0] - eps, half[0][1]), (half[1][0], half[1][1] - eps)))
        if not periodic:
            w = outflow / np.linalg.norm(outflow)
            m = np.array([-w[1], w[0]])
            for end, inner_node in ((0, 1), (n - 1, n - 2)):
                put(end, end, ((w[0], w[1]), (m[0], m[1])))
                put(end, inner_node, ((0.0, 0.0), (-m[0], -m[1])))
                rhs[end] = [w @ r[end], 0.0]
        A = csc_matrix((vals, (rows, cols)), shape=(2 * n, 2 * n))
        r = spsolve(A, rhs.ravel()).reshape(n, 2)
        layers[:, k + 1] = r
    return layers


def _reach(wall: np.ndarray, outer: np.ndarray) -> float:
    d = np.hypot(outer[:, None, 0] - wall[None, ::4, 0], outer[:, None, 1] - wall[None, ::4, 1])
    return float(d.min())


def march_to(line, wall, cells_normal, first, far_field, iterations=4, **kwargs):
    """March with geometric steps, rescaling the total so the outer layer
    ends ``far_field`` from the wall (within 1 %)."""
    total = far_field
    for _ in range(iterations):
        layers = march(line, np.diff(geometric(cells_normal, first, total)), **kwargs)
        reach = _reach(wall, layers[:, -1])
        if abs(reach - far_field) <= 0.01 * far_field:
            break
        total *= far_field / reach
    return layers


def o_grid(airfoil: Airfoil, cells_airfoil: int, cells_normal: int, wall_spacing: float,
           far_field: float = 100.0, le: float = LEADING_EDGE_SPACING, te: float = TRAILING_EDGE_SPACING,
           smoothing: float = 3.0):
    """O-grid nodes ``(ni+1, nj+1)`` each: i round the body from the trailing
    edge, lower surface first, the seam node line repeated; an open trailing
    edge adds ``max(2, cells_airfoil // 32)`` cells across its base."""
    c = airfoil.chord
    x, y = surface(airfoil, cells_airfoil, le, te)
    if airfoil.sharp:
        x, y = x[:-1], y[:-1]
    else:
        n_base = max(2, cells_airfoil // 32)
        t = np.linspace(0.0, 1.0, n_base + 1)[1:-1]
        x = np.concatenate([x, x[-1] + t * (x[0] - x[-1])])
        y = np.concatenate([y, y[-1] + t * (y[0] - y[-1])])
    line = np.stack([x, y], axis=1)
    layers = march_to(line, line, cells_normal, wall_spacing * c, far_field * c, periodic=True,
                      smoothing=smoothing)
    layers = np.concatenate([layers, layers[:1]], axis=0)
    return layers[..., 0], layers[..., 1]


def c_grid(airfoil: Airfoil, cells_airfoil: int, cells_wake: int, cells_normal: int, wall_spacing: float,
           far_field: float = 100.0, le: float = LEADING_EDGE_SPACING, te: float = TRAILING_EDGE_SPACING,
           smoothing: float = 3.0):
    """C-grid nodes: i from the far end of the lower wake, round the airfoil
    (lower surface first), out along the upper wake; the wake lines coincide."""
    if not airfoil.sharp:
        raise ValueError(f"airfoil {airfoil.name!r} has a blunt trailing edge "
                         f"(gap {airfoil.trailing_edge_gap:.3g}); use an O-grid, or close it "
                         f"(case key domain.trailing_edge = \"close\")")
    c = airfoil.chord
    x, y = surface(airfoil, cells_airfoil, le, te)
    first = np.hypot(x[1] - x[0], y[1] - y[0])
    wx, wy, direction = wake_line(airfoil, x[0], y[0], cells_wake, first, far_field * c)
    line = np.stack([np.concatenate([wx[::-1], x[1:-1], wx]), np.concatenate([wy[::-1], y[1:-1], wy])], axis=1)
    layers = march_to(line, np.stack([x, y], axis=1), cells_normal, wall_spacing * c, far_field * c,
                      periodic=False, outflow=direction, smoothing=smoothing)
    return layers[..., 0], layers[..., 1]
