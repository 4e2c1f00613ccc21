"""State layout and the freestream (host side).

Rows, as the reference (``src/jaxrans/core/state.py:1-21``): 0 density,
1..d momentum (velocity), d+1 energy (pressure); with Spalart-Allmaras one
more row, ``rho nu~ / nu_inf`` (conservative) or ``nu~ / nu_inf`` (primitive).
Every field on the device is variable-major on the canvas, ``(nvar, NI, NJ)``.
"""

from __future__ import annotations

import numpy as np

from .config import FlowConditions


def freestream_vector(mach, alpha_rad, gamma: float, dim: int = 2) -> np.ndarray:
    """``(rho, V, p)`` of the freestream, ``V = M (cos a, sin a)`` in 2D and
    ``M (cos a, 0, sin a)`` in 3D (``src/jaxrans/core/state.py:176-192``)."""
    u, w = mach * np.cos(alpha_rad), mach * np.sin(alpha_rad)
    velocity = [u, w] if dim == 2 else [u, 0.0 * u, w]
    return np.array([1.0, *velocity, 1.0 / gamma], dtype=np.float64)


def freestream_primitive(flow: FlowConditions, dim: int = 2) -> np.ndarray:
    if flow.swept:
        raise NotImplementedError("swept (2.5D) flow is not supported by this implementation")
    return freestream_vector(flow.mach, flow.alpha_rad, flow.gamma, dim)


def q_infinity(flow: FlowConditions, dim: int = 2) -> np.ndarray:
    """Freestream primitives, with ``nu~/nu_inf`` appended for SA."""
    q = freestream_primitive(flow, dim)
    if flow.turbulence is not None:
        q = np.concatenate([q, [flow.nu_tilde_ratio]])
    return q


def conservative(q: np.ndarray, gamma: float, dim: int = 2) -> np.ndarray:
    """Conservative vector of a primitive one ``(nvar, ...)``."""
    rho, vel, p = q[0], q[1:1 + dim], q[1 + dim]
    energy = p / (gamma - 1.0) + 0.5 * rho * np.sum(vel * vel, axis=0)
    out = [rho[None] if np.ndim(rho) else np.atleast_1d(rho), rho * vel,
           energy[None] if np.ndim(energy) else np.atleast_1d(energy)]
    if q.shape[0] > dim + 2:
        out.append(rho * q[dim + 2:])
    return np.concatenate(out, axis=0)


def uniform_canvas(q: np.ndarray, shape: tuple[int, ...], gamma: float) -> np.ndarray:
    """The primitive state ``q`` on every canvas cell, conservative ``(nvar, *shape)``."""
    u = conservative(np.asarray(q, dtype=np.float64), gamma, len(shape))
    return np.broadcast_to(u.reshape((-1,) + (1,) * len(shape)), (u.size,) + tuple(shape)).copy()
