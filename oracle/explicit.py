"""CPU restatement of the explicit multistage Runge-Kutta step with local time
steps (test infrastructure: only tests, the smoke check and the CPU baseline
use it; never the product path).

``local_time_step`` follows ``src/jaxrans/numerics/residual.py:142-166``:
``dt = CFL V / sum_faces lambda_f`` per canvas cell with ``lambda = |v . S| +
a |S|`` at the cell's own state on each of its faces (lower i, upper i, lower
j, upper j, summed in that order), plus the diffusive limit ``4 (gamma / Pr)
(mu / rho) (S_i^2 + S_j^2) / V`` when viscous, ``S_a`` the mean area of the
two faces normal to ``a`` and ``mu`` laminar plus the SA eddy viscosity.
``local_time_step3`` is the same on a 3D canvas (three face families, lower
and upper per family, ``|S| = sqrt(S . S)``, ``S_i^2 + S_j^2 + S_k^2``).
``rk_step`` follows ``src/jaxrans/solvers/explicit.py:48-71``:
``U^(k) = U^n - alpha_k (dt / V) R(U^(k-1))`` with ``dt`` frozen at ``U^n``;
the reported norm is the first stage's.  Pinned by
``tests/test_oracle_explicit.py`` (closed form on a uniform state, the Euler
scheme as one explicit step, the stage recursion by hand).
"""

from __future__ import annotations

import numpy as np

from . import physics as P
from .residual import Setup, norm, residual, rows

RK_SCHEMES = {"euler": (1.0,), "rk4_jameson": (0.25, 1.0 / 3.0, 0.5, 1.0),
              "rk5_jameson": (0.25, 1.0 / 6.0, 0.375, 0.5, 1.0)}   # explicit.py:23-27


def local_time_step(U, setup: Setup, cfl: float):
    at, gamma = setup.atlas, setup.gamma
    q = np.stack(P.primitive(rows(U), gamma))
    a = np.sqrt(gamma * q[3] / q[0])
    total = None
    for s in (at.face_i, at.face_j):
        lo, hi = (s[:, :-1, :], s[:, 1:, :]) if s is at.face_i else (s[:, :, :-1], s[:, :, 1:])
        lam = [np.abs(q[1] * f[0] + q[2] * f[1]) + a * P.hypot_jnp(f[0], f[1]) for f in (lo, hi)]
        total = lam[0] + lam[1] if total is None else total + lam[0] + lam[1]
    if setup.viscous:
        mean = [0.5 * (P.hypot_jnp(lo[0], lo[1]) + P.hypot_jnp(hi[0], hi[1])) for lo, hi in
                ((at.face_i[:, :-1, :], at.face_i[:, 1:, :]), (at.face_j[:, :, :-1], at.face_j[:, :, 1:]))]
        squares = mean[0] ** 2 + mean[1] ** 2
        mu = P.viscosity(gamma * q[3] / q[0], setup.mu_inf, setup.sutherland)
        if setup.sa:
            mu = mu + P.eddy_viscosity(q[0], q[4] * setup.mu_inf, mu / q[0])
        total = total + 4.0 * (gamma / P.PR * mu / q[0]) * squares / at.area
    return cfl * at.area / np.maximum(total, 1e-300)


def _lower_upper(s, a):
    n = s.shape[1 + a]
    return np.take(s, range(n - 1), axis=1 + a), np.take(s, range(1, n), axis=1 + a)


def local_time_step3(U, setup, cfl: float):
    """3D: the faces of family a (``(3, ...)`` with one more along a)."""
    from . import physics3d as P3
    at, gamma = setup.atlas, setup.gamma
    q = np.stack(P3.primitive([U[k] for k in range(len(U))], gamma, 3))
    a = np.sqrt(gamma * q[4] / q[0])
    total, mean = None, []
    for fam, s in enumerate(at.faces):
        lo, hi = _lower_upper(s, fam)
        mags = [np.sqrt((f[0] * f[0] + f[1] * f[1]) + f[2] * f[2]) for f in (lo, hi)]
        lam = [np.abs((q[1] * f[0] + q[2] * f[1]) + q[3] * f[2]) + a * m for f, m in zip((lo, hi), mags)]
        total = lam[0] + lam[1] if total is None else total + lam[0] + lam[1]
        mean.append(0.5 * (mags[0] + mags[1]))
    if setup.viscous:
        squares = (mean[0] ** 2 + mean[1] ** 2) + mean[2] ** 2
        mu = P.viscosity(gamma * q[4] / q[0], setup.mu_inf, setup.sutherland)
        if setup.sa:
            mu = mu + P.eddy_viscosity(q[0], q[5] * setup.mu_inf, mu / q[0])
        total = total + 4.0 * (gamma / P.PR * mu / q[0]) * squares / at.volume
    return cfl * at.volume / np.maximum(total, 1e-300)


def rk_step(U, setup, alphas, cfl: float):
    """``(U^{n+1}, |R(U^n)|)``, 2D (``Setup``) or 3D (``Setup3D``)."""
    if getattr(setup.atlas, "dim", 2) == 3:
        from . import residual3d as R3
        factor = local_time_step3(U, setup, cfl) / setup.atlas.volume
        res, nrm_of = R3.residual, R3.norm
    else:
        factor = local_time_step(U, setup, cfl) / setup.atlas.area
        res, nrm_of = residual, norm
    R = res(U, setup)
    nrm = nrm_of(R, setup)
    new = U - alphas[0] * factor[None] * R
    for alpha in alphas[1:]:
        new = U - alpha * factor[None] * res(new, setup)
    return new, nrm
