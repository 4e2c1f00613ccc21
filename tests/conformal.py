"""Closed-form conformal O-grids round the NACA 0012 thickness law, for tests.

The construction is Vassberg and Jameson's, "In Pursuit of Grid Convergence for
Two-Dimensional Euler Solutions", J. Aircraft 47 (2010) 1152-1166, sections
III-IV, which the reference uses for its own test and validation grids
(``tests/karman_trefftz.py``).  It must produce the same nodes as the
reference's helper for the published end-to-end values (CL 0.33567, CD 0.025711
at M 0.8, alpha 1.25 on 128 x 128, ``docs/dev/notes.md:82-86``) to apply, so
the steps are the published ones; the code is this repository's.

Section: ``y = +-0.6 (0.2969 sqrt(x) - 0.126 x - 0.3516 x^2 + 0.2843 x^3 -
0.1015 x^4)`` with the chord extended to the thickness law's next root
``x_te`` (reference chord stays 1).  Map: Karman-Trefftz,
``((zeta - 1)/(zeta + 1))^P = (z - z1)/(z - z2)`` with ``P = 2 - tau/pi``
(``tau`` the included trailing-edge angle), ``z1 = x_te``, ``z2 = rho_le / 2``.
Stations equally spaced in angle on the contour's image (a near circle);
radial levels ``R_1 exp(2 pi j / n)`` (square cells, far field ~148 chords);
linear blend from the near circle at the wall to a true circle outside;
mapped back and made mirror-symmetric about ``y = 0`` bit for bit.
"""

from __future__ import annotations

import numpy as np

COEFFS = (0.2969, -0.126, -0.3516, 0.2843, -0.1015)
SCALE = 0.6                       # t / 0.2, t = 0.12
RHO_LE = 0.01586736               # 1.1019 t^2, as V&J publish it


def half_thickness(x):
    x = np.asarray(x, np.float64)
    a = COEFFS
    return SCALE * (a[0] * np.sqrt(x) + a[1] * x + a[2] * x**2 + a[3] * x**3 + a[4] * x**4)


def _slope(x):
    a = COEFFS
    return SCALE * (0.5 * a[0] / np.sqrt(x) + a[1] + 2 * a[2] * x + 3 * a[3] * x**2 + 4 * a[4] * x**3)


def _root_past_one() -> float:
    """The thickness law's root past x = 1, bisected to the last bit."""
    lo, hi = 1.0, 1.05                    # positive at 1, negative at 1.05
    while True:
        mid = 0.5 * (lo + hi)
        if mid in (lo, hi):
            return lo if abs(half_thickness(lo)) <= abs(half_thickness(hi)) else hi
        lo, hi = (mid, hi) if half_thickness(mid) > 0.0 else (lo, mid)


X_TE = _root_past_one()
TAU = 2.0 * np.arctan(abs(_slope(X_TE)))
POWER = 2.0 - TAU / np.pi
Z1, Z2 = complex(X_TE, 0.0), complex(0.5 * RHO_LE, 0.0)


def to_physical(zeta):
    """zeta -> z, in whichever of the two equivalent forms keeps the factor's
    modulus at most one (so both singular points map exactly)."""
    zeta = np.asarray(zeta, np.complex128)
    a, b = zeta - 1.0, zeta + 1.0
    swap = np.abs(a) > np.abs(b)
    w = np.power(np.where(swap, b, a) / np.where(swap, a, b), POWER)
    return np.where(swap, (Z1 * w - Z2) / (w - 1.0), (Z1 - w * Z2) / (1.0 - w))


def to_circle(z):
    """z -> zeta, principal branch."""
    v = np.power((np.asarray(z, np.complex128) - Z1) / (np.asarray(z, np.complex128) - Z2), 1.0 / POWER)
    return (1.0 + v) / (1.0 - v)


def _lower(t):
    """Lower-surface point at parameter t in [0, pi] (trailing edge at 0, nose
    at pi): x = x_te cos^2(t/2)."""
    x = X_TE * np.cos(0.5 * t) ** 2
    return x - 1j * half_thickness(x)


def contour(n: int) -> np.ndarray:
    """``n`` stations, trailing edge -> lower -> nose -> upper, equally spaced
    in angle about the centre of the contour's image (the wrap not repeated)."""
    if n < 4 or n % 2:
        raise ValueError(f"n must be even and >= 4, got {n}")
    centre = 0.5 * (1.0 + to_circle(np.array([0j]))[0].real)
    half = n // 2
    want = -2.0 * np.pi * np.arange(1, half) / n
    lo, hi = np.zeros(want.size), np.full(want.size, np.pi)
    for _ in range(64):
        mid = 0.5 * (lo + hi)
        ang = np.angle(to_circle(_lower(mid)) - centre)
        lo, hi = np.where(ang > want, mid, lo), np.where(ang > want, hi, mid)
    lower = _lower(0.5 * (lo + hi))
    z = np.empty(n, np.complex128)
    z[0], z[half] = complex(X_TE, 0.0), 0j
    z[1:half] = lower
    z[half + 1:] = np.conj(lower[::-1])
    return z


def o_grid_nodes(n: int) -> tuple[np.ndarray, np.ndarray]:
    """Nodes ``(n + 1, n + 1)`` of the conformal ``n x n`` O-grid: ``i`` round
    the body (node ``n`` repeats node 0), ``j`` from the wall outward."""
    z0 = contour(n)
    zeta = to_circle(z0)
    centre = 0.5 * (zeta[0] + zeta[n // 2])
    r_wall, ang = np.abs(zeta - centre), np.angle(zeta - centre)
    r1 = np.abs(np.roll(zeta, -1) - zeta).sum() / (2.0 * np.pi)
    levels = r1 * np.exp(2.0 * np.pi * np.arange(n + 1) / n)
    r_out = levels[-1]
    r = (r_wall[:, None] * (r_out - levels[None]) + r_out * (levels[None] - r1)) / (r_out - r1)
    z = to_physical(centre + r * np.exp(1j * ang[:, None]))
    z = np.concatenate([z, z[:1]])
    mirror = np.concatenate([(n - np.arange(n)) % n, [0]])
    return 0.5 * (z.real + z.real[mirror]), 0.5 * (z.imag - z.imag[mirror])


_ATLAS: dict = {}


def conformal_atlas(n: int, wall: str = "slip_wall"):
    """The atlas of the conformal ``n x n`` O-grid (cached)."""
    key = (n, wall)
    if key not in _ATLAS:
        from jaxrans_amd.mesh.assemble import assemble
        from jaxrans_amd.mesh.atlas import assign_conditions, build_atlas
        X, Y = o_grid_nodes(n)
        mesh = assign_conditions(assemble([(X, Y)]), {"wall": wall, "farfield": "far_field"})
        _ATLAS[key] = build_atlas(mesh)
    return _ATLAS[key]
