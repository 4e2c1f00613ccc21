"""Airfoil contours: NACA 4-digit law, Karman-Trefftz sections, coordinate files.

A contour is stored clockwise from the trailing edge: lower surface, leading
edge, upper surface (the orientation of ``src/jaxrans/geometry/airfoil.py:25-106``),
and is splined against arc length for redistribution by the grid generator.
"""

from __future__ import annotations

from dataclasses import dataclass
from pathlib import Path

import numpy as np
from scipy.interpolate import CubicSpline


def _arc_length(z: np.ndarray) -> np.ndarray:
    """Cumulative chord-segment length along the complex points ``z``."""
    return np.concatenate([[0.0], np.cumsum(np.abs(np.diff(z)))])


@dataclass(frozen=True)
class Airfoil:
    """A closed-or-open section contour ``(x, y)``, clockwise from the trailing
    edge (lower surface first).  Derived quantities are computed from the
    complex points ``z = x + i y``: the trailing-edge midpoint, the leading
    edge (the point farthest from it), the chord (their distance) and the
    arc length the grid generator parametrises by."""

    name: str
    x: np.ndarray
    y: np.ndarray

    def __post_init__(self) -> None:
        n = self.x.size
        if n < 5:
            raise ValueError(f"airfoil {self.name!r}: at least 5 points are needed, got {n}")
        z = self.x + 1j * self.y
        if not np.all(np.abs(np.diff(z)) > 0.0):
            raise ValueError(f"airfoil {self.name!r}: repeated consecutive points")
        te = 0.5 * (z[0] + z[-1])
        le = int(np.argmax(np.abs(z - te)))
        object.__setattr__(self, "_geo", {"z": z, "te": te, "le": le, "s": _arc_length(z)})

    @property
    def s(self) -> np.ndarray:
        return self._geo["s"]

    @property
    def length(self) -> float:
        return float(self._geo["s"][-1])

    @property
    def trailing_edge_gap(self) -> float:
        z = self._geo["z"]
        return float(abs(z[-1] - z[0]))

    @property
    def leading_edge(self) -> int:
        """Index of the contour point farthest from the trailing-edge midpoint."""
        return self._geo["le"]

    @property
    def chord(self) -> float:
        g = self._geo
        return float(abs(g["z"][g["le"]] - g["te"]))

    @property
    def sharp(self) -> bool:
        """A trailing-edge gap below a millionth of the chord counts as closed."""
        return self.trailing_edge_gap <= 1e-6 * self.chord

    def spline(self) -> tuple[CubicSpline, CubicSpline]:
        """Cubic splines ``x(s)``, ``y(s)`` against arc length; a sharp
        contour's last point is pinned onto its first."""
        x, y = self.x.copy(), self.y.copy()
        if self.sharp:
            x[-1], y[-1] = x[0], y[0]
        return CubicSpline(self.s, x), CubicSpline(self.s, y)

    def points(self, s: np.ndarray) -> tuple[np.ndarray, np.ndarray]:
        fx, fy = self.spline()
        return fx(s), fy(s)

    def unit_chord(self) -> "Airfoil":
        """Translated so that the leading edge sits at the origin, scaled to chord 1."""
        g = self._geo
        z = (g["z"] - g["z"][g["le"]]) / self.chord
        return Airfoil(self.name, z.real.copy(), z.imag.copy())

    def closed(self) -> "Airfoil":
        """An open trailing edge closed at its midpoint: each surface is
        sheared towards the midpoint by a displacement that grows linearly with
        the distance along the leading edge -> trailing-edge-midpoint line,
        from nothing at the leading edge to the whole half gap at its end point."""
        if self.sharp:
            return self
        g = self._geo
        z, te, le = g["z"], g["te"], g["le"]
        axis = (te - z[le]) / abs(te - z[le])
        along = ((z - z[le]) * np.conj(axis)).real          # projection on the chord line
        n = z.size
        ends = np.where(np.arange(n) < le, 0, n - 1)           # lower surface -> point 0, upper -> point n-1
        weight = np.clip(along / along[ends], 0.0, 1.0)
        closed = z + weight * (te - z[ends])
        closed[0] = closed[-1] = te
        return Airfoil(self.name, closed.real.copy(), closed.imag.copy())


def from_points(x, y, name: str = "airfoil") -> Airfoil:
    """A contour given in any direction: repeated consecutive points dropped,
    then reversed if it runs counter-clockwise (positive shoelace area), so
    the result runs clockwise from its first point."""
    z = np.asarray(x, dtype=np.float64) + 1j * np.asarray(y, dtype=np.float64)
    z = z[np.concatenate([[True], np.abs(np.diff(z)) > 0.0])]
    area = 0.5 * float(np.imag(np.sum(np.conj(z) * np.roll(z, -1))))
    if area > 0.0:
        z = z[::-1]
    return Airfoil(name, z.real.copy(), z.imag.copy())


def thickness(xc: np.ndarray, t: float, a4: float) -> np.ndarray:
    """NACA 4-digit half thickness ``5 t (0.2969 sqrt x - 0.1260 x - 0.3516 x^2
    + 0.2843 x^3 + a4 x^4)``; ``a4 = -0.1015`` standard (open trailing edge),
    ``-0.1036`` closed."""
    return 5.0 * t * (0.2969 * np.sqrt(xc) - 0.1260 * xc - 0.3516 * xc**2 + 0.2843 * xc**3 + a4 * xc**4)


def naca4(code: str, n: int = 201, sharp: bool = True) -> Airfoil:
    """NACA 4-digit section with ``n`` cosine-spaced stations per surface."""
    if len(code) != 4 or not code.isdigit():
        raise ValueError(f"a NACA 4-digit code is four digits, got {code!r}")
    m, p, t = int(code[0]) / 100.0, int(code[1]) / 10.0, int(code[2:]) / 100.0
    xc = 0.5 * (1.0 - np.cos(np.linspace(0.0, np.pi, n)))
    yt = thickness(xc, t, -0.1036 if sharp else -0.1015)
    if m > 0.0 and p > 0.0:
        fore = xc < p
        yc = np.where(fore, m / p**2 * (2 * p * xc - xc**2),
                      m / (1 - p) ** 2 * (1 - 2 * p + 2 * p * xc - xc**2))
        slope = np.where(fore, 2 * m / p**2 * (p - xc), 2 * m / (1 - p) ** 2 * (p - xc))
    else:
        yc, slope = np.zeros_like(xc), np.zeros_like(xc)
    th = np.arctan(slope)
    xu, yu = xc - yt * np.sin(th), yc + yt * np.cos(th)
    xl, yl = xc + yt * np.sin(th), yc - yt * np.cos(th)
    x = np.concatenate([xl[::-1], xu[1:]])
    y = np.concatenate([yl[::-1], yu[1:]])
    if sharp:
        y[0] = y[-1] = 0.5 * (y[0] + y[-1])
    return from_points(x, y, f"NACA {code}")


def karman_trefftz(te_angle_deg: float, centre: tuple[float, float], n: int = 257) -> Airfoil:
    """Karman-Trefftz section: the circle through ``zeta = 1`` centred at
    ``centre`` mapped by ``z = k [(zeta+1)^k + (zeta-1)^k] / [(zeta+1)^k - (zeta-1)^k]``,
    ``k = 2 - tau/pi`` (tau the trailing-edge angle), sampled at ``n``
    points, scaled to unit chord.  The point ``zeta = 1`` maps to the sharp
    trailing edge."""
    k = 2.0 - np.deg2rad(te_angle_deg) / np.pi
    c = complex(*centre)
    radius = abs(1.0 - c)
    start = np.angle(1.0 - c)
    theta = start + 2.0 * np.pi * np.linspace(0.0, 1.0, n)
    zeta = c + radius * np.exp(1j * theta)
    # ((zeta - 1)/(zeta + 1))^k is single-valued on the circle: the ratio is a
    # negative real number only on (-1, 1), inside it.
    w = ((zeta - 1.0) / (zeta + 1.0)) ** k
    z = k * (1.0 + w) / (1.0 - w)
    z[0] = z[-1] = k  # zeta = 1: the sharp trailing edge
    foil = from_points(z.real, z.imag, f"Karman-Trefftz {te_angle_deg:g} deg {centre}")
    return foil.unit_chord()


def read_airfoil(path: str | Path) -> Airfoil:
    """Selig or plain two-column coordinate file (a name line is allowed);
    Lednicer (a first row of point counts) as well."""
    path = Path(path)
    rows, name = [], path.stem
    for line in path.read_text().splitlines():
        parts = line.replace(",", " ").split()
        try:
            values = [float(p) for p in parts]
        except ValueError:
            if not rows and parts:
                name = line.strip()
            continue
        if len(values) == 2:
            rows.append(values)
    data = np.array(rows)
    if len(data) < 5:
        raise ValueError(f"{path.name}: fewer than 5 coordinate pairs")
    if data[0, 0] > 1.5 and data[0, 1] > 1.5:
        nu, nl = int(data[0, 0]), int(data[0, 1])
        upper, lower = data[1:1 + nu], data[1 + nu:1 + nu + nl]
        contour = np.concatenate([upper[::-1], lower[1:]])
        return from_points(contour[:, 0], contour[:, 1], name)
    return from_points(data[:, 0], data[:, 1], name)


def write_selig(airfoil: Airfoil, path: str | Path) -> Path:
    """Write ``airfoil`` as a Selig file (name line, then x y per line)."""
    path = Path(path)
    lines = [airfoil.name] + [f"{x:.16e} {y:.16e}" for x, y in zip(airfoil.x, airfoil.y)]
    path.write_text("\n".join(lines) + "\n")
    return path


def airfoil_source(spec: str, base: Path | None = None) -> Airfoil:
    """``naca####`` (closed TE) / ``naca####_blunt`` (open TE), else a file."""
    lowered = spec.lower()
    if lowered.startswith("naca") and lowered[4:8].isdigit() and len(lowered) in (8, 14):
        return naca4(lowered[4:8], sharp=not lowered.endswith("_blunt"))
    for root in ([base] if base else []) + [Path.cwd()]:
        path = Path(spec) if Path(spec).is_absolute() else root / spec
        if path.is_file():
            return read_airfoil(path)
    raise FileNotFoundError(f"airfoil {spec!r}: not a NACA 4-digit name, nor a file in {base} or {Path.cwd()}")
